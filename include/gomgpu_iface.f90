! gomgpu_iface.f90 - ISO_C_BINDING interface of the GPU step (include/gomgpu.h).
! Generated by `python -m gom_release_amd.fortran_binding`; do not edit by hand.
! Field order of the bind(C) types is the header's X-macro order.
module gomgpu_iface
  use, intrinsic :: iso_c_binding
  implicit none
  integer(c_int), parameter :: GOMGPU_STAGE_FRICTION = 0
  integer(c_int), parameter :: GOMGPU_STAGE_ELM = 1
  integer(c_int), parameter :: GOMGPU_STAGE_MOMENTUM = 2
  integer(c_int), parameter :: GOMGPU_STAGE_FREESURF = 3
  integer(c_int), parameter :: GOMGPU_STAGE_VELOCITY = 4
  integer(c_int), parameter :: GOMGPU_STAGE_LAYERS = 5
  integer(c_int), parameter :: GOMGPU_STAGE_NODEVEL = 6
  integer(c_int), parameter :: GOMGPU_STAGE_TRANSPORT = 7
  integer(c_int), parameter :: GOMGPU_STAGE_EOS = 8
  integer(c_int), parameter :: GOMGPU_STAGE_UPDATE = 9
  integer(c_int), parameter :: GOMGPU_STAGE_TURBULENCE = 10
  integer(c_int), parameter :: GOMGPU_STAGE_WIND = 11
  integer(c_int), parameter :: GOMGPU_STAGE_BOUNDARY = 12
  integer(c_int), parameter :: GOMGPU_STAGE_BARS = 13
  integer(c_int), parameter :: GOMGPU_FP64 = 0
  integer(c_int), parameter :: GOMGPU_E_ARG = 1
  integer(c_int), parameter :: GOMGPU_MIXED = 1
  integer(c_int), parameter :: GOMGPU_E_HIP = 2
  integer(c_int), parameter :: GOMGPU_E_NOMEM = 3
  integer(c_int), parameter :: GOMGPU_E_MODEL = 4
  integer(c_int), parameter :: GOMGPU_W_PCG_NOT_CONVERGED = 5

  type, bind(C) :: gomgpu_mesh
    type(c_ptr) :: x_node, y_node, h_node, x_cell, y_cell, h_cell, area, face_length, x_face, &
      & y_face, h_face, delta_j, cos_theta, sin_theta, cos_theta2, sin_theta2, coriolis_factor, &
      & manning, bf_height, z_level, delta_z, tri_or_quad, nodenum_at_cell, adj_cellnum_at_cell, &
      & facenum_at_cell, sign_in_outflow, adj_cellnum_at_face, nodenum_at_face, adj_cells_at_node, &
      & adj_cellnum_at_node, node_count_each_element, adj_nodes_at_node, adj_nodenum_at_node, &
      & boundary_type_of_face, isflowside3, ob_element_flag, Qb_element_flag, &
      & bottom_layer_at_node, bottom_layer_at_element, bottom_layer_at_face, initial_wetdry_node, &
      & ob_cell_id, Qb_cell_id = c_null_ptr
  end type gomgpu_mesh
  type, bind(C) :: gomgpu_state
    type(c_ptr) :: eta_cell, eta_cell_new, eta_node, Gamma_B, Cdb, un_face, vn_face, un_face_new, &
      & vn_face_new, un_ELM, vn_ELM, dz_face, dz_face_new, AinvG1, AinvG2, AinvDeltaZ1, &
      & AinvDeltaZ2, Kh, salt_face, temp_face, rho_face, dzhalf_face, dzhalf_face_new, Av, q2, xl, &
      & u_face_level, v_face_level, dz_cell, dz_cell_new, salt_cell, salt_cell_new, temp_cell, &
      & temp_cell_new, rho_cell, dzhalf_cell, dzhalf_cell_new, wn_cell, wn_cell_new, Kv, dz_node, &
      & dz_node_new, salt_node, temp_node, rho_node, dzhalf_node, dzhalf_node_new, u_node, v_node, &
      & w_node, ubar_node, vbar_node, sbar_node, tbar_node, rbar_node, top_layer_at_node, &
      & top_layer_at_element, top_layer_at_face, wetdry_node, num_sub_elm_iteration = c_null_ptr
  end type gomgpu_state
  type, bind(C) :: gomgpu_params
    integer(c_int32_t) :: maxnod, maxele, maxface, maxlayer, max_no_neighbor_node, num_ob_cell, &
      & num_Qb_cell, maxtran2, tran_id1, tran_id2, advection_flag, ELM_backtrace_flag, &
      & ELM_min_iter, ELM_max_iter, bf_flag, baroclinic_flag, ana_density, i_density_flag, &
      & transport_flag, trans_sub_iter, h_flux_limiter, v_flux_limiter, heat_option, &
      & max_iteration_pcg, precision, river_on_device, turbulence_flag, turbulence_model, &
      & stability_function, hurricane_flag, hurricane_read_interval, wind_formular, &
      & transport_solver = 0
    real(c_double) :: dt, theta, gravity, rho_o, MSL, dry_depth, adv_onoff_depth, &
      & smagorinsky_parameter, Ah_0, Kh_0, von_Karman, error_tolerance, &
      & denominator_min_for_matrix, delta_z_min, q2min, psimin, eps_min, bgdiff, diffmax, xlmin1, &
      & xlmin2, cmiu0, cpsi1, cpsi2, rpub, rmub, rnub, schk, schpsi, h1_pp, h2_pp, vdmax_pp1, &
      & vdmax_pp2, vdmin_pp1, vdmin_pp2, tdmin_pp, rho_a = 0.0d0
  end type gomgpu_params
  type, bind(C) :: gomgpu_forcing
    real(c_double) :: elapsed_time, spinup_function_tide, spinup_function_baroclinic, &
      & spinup_function_wind, hurricane_phase, julian_day, hurricane_centre_x, &
      & hurricane_centre_y = 0.0d0
    type(c_ptr) :: Gamma_T, wind_stress_normal, wind_stress_tangnt, air_p1, eta_at_ob_old, &
      & eta_at_ob_new, Qu_boundary, phi_sz, con_at_ob, con_at_Qb, Q_add, wind_u0, wind_v0, air_p0, &
      & wind_u1, wind_v1, wind_u2, wind_v2, air_p2, wind_u_next, wind_v_next, air_p_next, &
      & wind_u_at_face, wind_v_at_face = c_null_ptr
  end type gomgpu_forcing

  interface
    integer(c_int) function gomgpu_create(mesh, params, ctx) bind(C, name='gomgpu_create')
      import :: c_int, c_ptr, gomgpu_mesh, gomgpu_params
      type(gomgpu_mesh), intent(in) :: mesh
      type(gomgpu_params), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function gomgpu_create
    integer(c_int) function gomgpu_set_state(ctx, state) bind(C, name='gomgpu_set_state')
      import :: c_int, c_ptr, gomgpu_state
      type(c_ptr), value :: ctx
      type(gomgpu_state), intent(in) :: state
    end function gomgpu_set_state
    integer(c_int) function gomgpu_get_state(ctx, state) bind(C, name='gomgpu_get_state')
      import :: c_int, c_ptr, gomgpu_state
      type(c_ptr), value :: ctx
      type(gomgpu_state), intent(in) :: state
    end function gomgpu_get_state
    integer(c_int) function gomgpu_set_forcing(ctx, forcing) bind(C, name='gomgpu_set_forcing')
      import :: c_int, c_ptr, gomgpu_forcing
      type(c_ptr), value :: ctx
      type(gomgpu_forcing), intent(in) :: forcing
    end function gomgpu_set_forcing
    integer(c_int) function gomgpu_stage(ctx, stage_id) bind(C, name='gomgpu_stage')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: stage_id
    end function gomgpu_stage
    integer(c_int) function gomgpu_step(ctx) bind(C, name='gomgpu_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function gomgpu_step
    integer(c_int) function gomgpu_run(ctx, nsteps) bind(C, name='gomgpu_run')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function gomgpu_run
    integer(c_int) function gomgpu_budgets(ctx, volume, salt) bind(C, name='gomgpu_budgets')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: volume, salt
    end function gomgpu_budgets
    integer(c_int) function gomgpu_pcg_iterations(ctx) bind(C, name='gomgpu_pcg_iterations')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function gomgpu_pcg_iterations
    integer(c_int) function gomgpu_sync(ctx) bind(C, name='gomgpu_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function gomgpu_sync
    type(c_ptr) function gomgpu_last_error(ctx) bind(C, name='gomgpu_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function gomgpu_last_error
    integer(c_int) function gomgpu_timings(ctx, enable, names, ms, n) bind(C, name='gomgpu_timings')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      type(c_ptr), value :: names, ms
      integer(c_int), value :: n
    end function gomgpu_timings
    subroutine gomgpu_destroy(ctx) bind(C, name='gomgpu_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine gomgpu_destroy
  end interface
end module gomgpu_iface
