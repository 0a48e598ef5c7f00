! Reference-run harness for the CPU oracle (test infrastructure; never shipped).
!
! Built by `make -C oracle ref` against the GOM v1.1.0 objects compiled from the sources
! under /root/reference (this file is the only source of ours in that link). It drives the
! reference's own start-up and per-step subroutines in the order of the reference time loop
! (GOM_Rv1.1.0/source/main.f90:22-47 start-up, :53-304 loop) and writes the module state
! at full precision:
!   <dumpdir>/geom.bin             static geometry and parameters after prepare_gom
!   <dumpdir>/s<it>_<stage>.bin    state after each hot-path stage of step <it>
! Stage names: pre (after host forcing), elm, mom, fs, vel, lay, nodev, tvd, eos, post.
!
! Usage (run from a GOM project directory holding input/ and output/):
!   gom_ref_harness <dumpdir> <first_dump_step> <last_dump_step> [<extra_step> ...]
! With <dumpdir> = "-" nothing is written and per-stage wall times are printed instead,
! which is how the reference CPU baseline is timed.
!
! Record format (stream access): name(32 chars), kind (int32: 1 = real64, 2 = int32),
! rank (int32), lower(3) (int32), extent(3) (int32), then the data in Fortran order.
module harness_dump
   use mod_global_variables, only : dp
   implicit none
   integer :: hu = 77
contains
   subroutine hdr(name, kind, rank, lo, ex)
      character(len=*), intent(in) :: name
      integer, intent(in) :: kind, rank, lo(3), ex(3)
      character(len=32) :: nm
      nm = name
      write(hu) nm, kind, rank, lo, ex
   end subroutine hdr
   subroutine put_r0(name, v)
      character(len=*), intent(in) :: name
      real(dp), intent(in) :: v
      call hdr(name, 1, 0, [0,0,0], [1,1,1])
      write(hu) v
   end subroutine put_r0
   subroutine put_i0(name, v)
      character(len=*), intent(in) :: name
      integer, intent(in) :: v
      call hdr(name, 2, 0, [0,0,0], [1,1,1])
      write(hu) v
   end subroutine put_i0
   subroutine put_r1(name, v)
      character(len=*), intent(in) :: name
      real(dp), intent(in) :: v(:)
      call hdr(name, 1, 1, [1,0,0], [size(v),1,1])
      write(hu) v
   end subroutine put_r1
   subroutine put_i1(name, v)
      character(len=*), intent(in) :: name
      integer, intent(in) :: v(:)
      call hdr(name, 2, 1, [1,0,0], [size(v),1,1])
      write(hu) v
   end subroutine put_i1
   subroutine put_r2(name, v, lo1)
      character(len=*), intent(in) :: name
      real(dp), intent(in) :: v(:,:)
      integer, intent(in) :: lo1
      call hdr(name, 1, 2, [lo1,1,0], [size(v,1),size(v,2),1])
      write(hu) v
   end subroutine put_r2
   subroutine put_i2(name, v, lo1)
      character(len=*), intent(in) :: name
      integer, intent(in) :: v(:,:)
      integer, intent(in) :: lo1
      call hdr(name, 2, 2, [lo1,1,0], [size(v,1),size(v,2),1])
      write(hu) v
   end subroutine put_i2
end module harness_dump

program gom_ref_harness
   use omp_lib
   use mod_global_variables
   use mod_file_definition
   use harness_dump
   implicit none
   character(len=256) :: dumpdir, arg
   integer :: first_dump, last_dump, nextra, i, j, k
   integer, allocatable :: extra(:)
   logical :: dumping, dump_now, no_heat
   character(len=8) :: envv
   integer :: envlen, envst
   real(dp) :: t0, tstage(12)
   character(len=12), parameter :: stage_name(12) = [character(len=12) :: &
      'forcing', 'elm', 'momentum', 'freesurface', 'velocities', 'layers', &
      'nodevel', 'transport', 'eos', 'update', 'output', 'total']

   call get_command_argument(1, dumpdir)
   dumping = trim(dumpdir) /= '-'
   first_dump = 0
   last_dump = -1
   if (command_argument_count() >= 3) then
      call get_command_argument(2, arg); read(arg,*) first_dump
      call get_command_argument(3, arg); read(arg,*) last_dump
   end if
   nextra = max(0, command_argument_count() - 3)
   allocate(extra(max(1,nextra)))
   extra = -1
   do i = 1, nextra
      call get_command_argument(3+i, arg); read(arg,*) extra(i)
   end do
   tstage = 0.0_dp
   ! GOM_HARNESS_NO_HEAT=1 leaves out the surface heat model (heat_term_by_term reads the
   ! OpenMP-private copy of its vapour pressure ea before setting it, so its flux is
   ! undefined): phi_sz then keeps its allocation value 0 and the transport adds no heat.
   no_heat = .false.
   call get_environment_variable('GOM_HARNESS_NO_HEAT', envv, envlen, envst)
   if (envst == 0 .and. envlen > 0) no_heat = envv(1:1) == '1'

   ! start-up, as GOM_Rv1.1.0/source/main.f90:12-47
   call omp_set_nested(.true.)
   call welcome_message
   call scan_gom
   call set_geometry_1
   call allocate_variables
   call read_main_inp
   if (hurricane_flag == 1 .and. hurricane_interp_method == 2) then
      call find_neighbour_nodes
   end if
   call prepare_gom

   if (dumping) then
      call open_dump('geom')
      call dump_geometry
      close(hu)
      call open_dump('s0_post')
      call dump_state
      close(hu)
   end if

   ! time loop, as GOM_Rv1.1.0/source/main.f90:53-304
   do it = 1, ndt
      dump_now = dumping .and. ((it >= first_dump .and. it <= last_dump) .or. any(extra == it))
      t0 = omp_get_wtime()
      elapsed_time      = it*dt
      elapsed_time_hr   = elapsed_time/3600.0_dp
      julian_day        = julian_day + dt/86400.0_dp
      current_jday_1900 = current_jday_1900  + dt/86400.0_dp
      call julian_to_localtime(year, month, day, hour, minute, julian_day)
      call setup_spinup
      call calculate_bottom_friction
      if (wind_flag == 1 .or. airp_flag == 1) then
         if (num_windp_ser > 0) call get_windp_at_face
      end if
      if (hurricane_flag == 1) then
         if (julian_day >= hurricane_start_jday .and. julian_day <= hurricane_end_jday) then
            if (hurricane_interp_method == 1) then
               call read_hurricane_ser_1
            elseif (hurricane_interp_method == 2) then
               call read_hurricane_ser_2
            end if
         end if
      end if
      if (wind_flag == 0) then
         wind_stress_normal = 0.0_dp
         wind_stress_tangnt = 0.0_dp
      end if
      if (wind_flag == 1 .or. hurricane_flag == 1) call calculate_wind_stress
      if (wind_flag == 2) call ana_windp
      call compute_boundary_velocity
      call compute_boundary_eta
      call lap(1)
      if (dump_now) call stage_dump('pre')

      if (advection_flag == 0) then
         do j = 1, maxface
            do k = 1, maxlayer
               un_ELM(k,j) = un_face(k,j)
               vn_ELM(k,j) = vn_face(k,j)
            end do
         end do
      else
         call solve_nonlinear_advection
      end if
      call lap(2)
      if (dump_now) call stage_dump('elm')

      call solve_momentum_equation
      call lap(3)
      if (dump_now) call stage_dump('mom')

      call solve_free_surface_equation
      call lap(4)
      if (dump_now) call stage_dump('fs')

      call solve_velocities
      call lap(5)
      if (dump_now) call stage_dump('vel')

      call update_vertical_layers_1
      call lap(6)
      if (dump_now) call stage_dump('lay')

      call calculate_velocity_at_node
      call lap(7)
      if (dump_now) call stage_dump('nodev')

      if (transport_flag == 1 .and. is_temp == 1 .and. .not. no_heat) then
         if (heat_option == 1) then
            call heat_term_by_term
         else if (heat_option == 2) then
            call heat_equilibrium_temperature
         end if
      end if
      if (transport_flag == 1) then
         if (transport_solver == 1) then
            if (maxval(is_tran) > 0) call solve_transport_equation_v19
         else if (transport_solver == 2) then
            if (maxval(is_tran) > 0) call solve_transport_equation_ELM_v8
         end if
      end if
      call lap(8)
      if (dump_now) call stage_dump('tvd')

      if (baroclinic_flag == 1) then
         if (ana_density == 0) then
            call calculate_density_full
         else if (ana_density == 1) then
            call calculate_density_linear
         else if (ana_density == 2) then
            call calculate_analytical_density
         end if
      end if
      call lap(9)
      if (dump_now) call stage_dump('eos')

      call update_variables
      call lap(10)
      if (dump_now) call stage_dump('post')

      call write_output_files
      if (IS2D_switch == 1) then
         if (IS2D_flood_map > 0) call update_flood_map
      end if
      if (restart_out == -1) then
         if (it == ndt) call write_restart
      else if (restart_out == 1) then
         if (mod(it,restart_freq) == 0) call write_restart
      end if
      if (mod(it,terminate_check_freq) == 0) call model_termination_check
      call lap(11)
   end do

   if (.not. dumping) then
      tstage(12) = sum(tstage(1:11))
      write(*,'(A,I8,A,I4)') 'harness_timing steps=', ndt, ' threads=', omp_get_max_threads()
      do i = 1, 12
         write(*,'(A,A12,A,ES16.8)') 'harness_stage ', stage_name(i), ' seconds=', tstage(i)
      end do
   end if

contains

   subroutine lap(idx)
      integer, intent(in) :: idx
      real(dp) :: t1
      t1 = omp_get_wtime()
      tstage(idx) = tstage(idx) + (t1 - t0)
      t0 = t1
   end subroutine lap

   subroutine open_dump(tag)
      character(len=*), intent(in) :: tag
      open(hu, file=trim(dumpdir)//'/'//trim(tag)//'.bin', access='stream', &
         form='unformatted', status='replace')
   end subroutine open_dump

   subroutine stage_dump(stage)
      character(len=*), intent(in) :: stage
      character(len=32) :: tag
      write(tag,'(A,I0,A,A)') 's', it, '_', trim(stage)
      call open_dump(trim(tag))
      call dump_state
      close(hu)
      t0 = omp_get_wtime()
   end subroutine stage_dump

   subroutine dump_geometry
      call put_i0('maxnod', maxnod); call put_i0('maxele', maxele)
      call put_i0('maxface', maxface); call put_i0('maxlayer', maxlayer)
      call put_i0('maxtran2', maxtran2)
      call put_i0('max_no_neighbor_node', max_no_neighbor_node)
      call put_r0('ref_salt', ref_salt); call put_r0('ref_temp', ref_temp)
      call put_r0('MSL', MSL); call put_r0('dt', dt); call put_r0('theta', theta)
      call put_r0('gravity', gravity); call put_r0('rho_o', rho_o)
      call put_r0('dry_depth', dry_depth); call put_r0('delta_z_min', delta_z_min)
      call put_r0('adv_onoff_depth', adv_onoff_depth)
      call put_r0('smagorinsky_parameter', smagorinsky_parameter)
      call put_r0('Ah_0', Ah_0); call put_r0('Kh_0', Kh_0)
      call put_r0('Av_0', Av_0); call put_r0('Kv_0', Kv_0)
      call put_r0('error_tolerance', error_tolerance)
      call put_r0('denominator_min_for_matrix', denominator_min_for_matrix)
      call put_r0('von_Karman', von_Karman)
      call put_i0('max_iteration_pcg', max_iteration_pcg)
      call put_i0('advection_flag', advection_flag)
      call put_i0('ELM_backtrace_flag', ELM_backtrace_flag)
      call put_i0('ELM_min_iter', ELM_min_iter); call put_i0('ELM_max_iter', ELM_max_iter)
      call put_i0('bf_flag', bf_flag); call put_i0('baroclinic_flag', baroclinic_flag)
      call put_i0('ana_density', ana_density)
      call put_i0('transport_flag', transport_flag)
      call put_i0('trans_sub_iter', trans_sub_iter)
      call put_i0('h_flux_limiter', h_flux_limiter)
      call put_i0('v_flux_limiter', v_flux_limiter)
      call put_i0('heat_option', heat_option)
      call put_i0('no_Etide_species', no_Etide_species)
      call put_i0('num_ob_cell', num_ob_cell); call put_i0('num_Qb_cell', num_Qb_cell)
      call put_i0('num_WR_cell', num_WR_cell); call put_i0('num_SS_cell', num_SS_cell)
      call put_i0('i_density_flag', i_density_flag)
      call put_r0('rho_a', rho_a); call put_i0('wind_formular', wind_formular)
      call put_i0('transport_solver', transport_solver)
      call put_i0('hurricane_flag', hurricane_flag)
      if (hurricane_flag == 1) then
         call put_i0('hurricane_read_interval', hurricane_read_interval)
      else
         call put_i0('hurricane_read_interval', 0)
      end if
      call put_i1('tran_id', tran_id)
      call put_i2('start_end_node', reshape(start_end_node, [4, 12]), 1)
      call put_r1('z_level', z_level); call put_r1('delta_z', delta_z)
      call put_r1('x_node', x_node); call put_r1('y_node', y_node); call put_r1('h_node', h_node)
      call put_r1('x_cell', x_cell); call put_r1('y_cell', y_cell); call put_r1('h_cell', h_cell)
      call put_r1('area', area)
      call put_r1('face_length', face_length); call put_r1('x_face', x_face)
      call put_r1('y_face', y_face); call put_r1('h_face', h_face)
      call put_r1('delta_j', delta_j)
      call put_r1('cos_theta', cos_theta); call put_r1('sin_theta', sin_theta)
      call put_r1('cos_theta2', cos_theta2); call put_r1('sin_theta2', sin_theta2)
      call put_r1('coriolis_factor', coriolis_factor)
      call put_r1('manning', manning); call put_r1('bf_height', bf_height)
      call put_i1('tri_or_quad', tri_or_quad)
      call put_i2('nodenum_at_cell', nodenum_at_cell, 1)
      call put_i2('adj_cellnum_at_cell', adj_cellnum_at_cell, 1)
      call put_i2('facenum_at_cell', facenum_at_cell, 1)
      call put_i2('adj_cellnum_at_face', adj_cellnum_at_face, 1)
      call put_i2('nodenum_at_face', nodenum_at_face, 1)
      call put_i2('sign_in_outflow', sign_in_outflow, 1)
      call put_i1('adj_cells_at_node', adj_cells_at_node)
      call put_i2('adj_cellnum_at_node', adj_cellnum_at_node, 1)
      call put_i2('node_count_each_element', node_count_each_element, 1)
      call put_i1('adj_nodes_at_node', adj_nodes_at_node)
      call put_i2('adj_nodenum_at_node', adj_nodenum_at_node, 0)
      call put_i1('boundary_type_of_face', boundary_type_of_face)
      call put_i1('ob_element_flag', ob_element_flag)
      call put_i1('Qb_element_flag', Qb_element_flag)
      call put_i1('bottom_layer_at_node', bottom_layer_at_node)
      call put_i1('bottom_layer_at_element', bottom_layer_at_element(1:maxele))
      call put_i1('bottom_layer_at_face', bottom_layer_at_face)
      call put_i1('isflowside', isflowside); call put_i1('isflowside2', isflowside2)
      call put_i1('isflowside3', isflowside3)
      if (allocated(isflowside4)) call put_i1('isflowside4', isflowside4)
      call put_i1('initial_wetdry_node', initial_wetdry_node)
      if (num_ob_cell > 0) then
         call put_i1('ob_cell_id', ob_cell_id); call put_i1('ob_face_id', ob_face_id)
      end if
      if (allocated(Q_boundary)) call put_i2('Q_boundary', Q_boundary, 1)
      if (allocated(WR_element_flag)) call put_i1('WR_element_flag', WR_element_flag)
      if (allocated(SS_element_flag)) call put_i1('SS_element_flag', SS_element_flag)
   end subroutine dump_geometry

   subroutine dump_state
      call put_i0('it', it)
      call put_r0('elapsed_time', elapsed_time); call put_r0('julian_day', julian_day)
      call put_r0('spinup_function_tide', spinup_function_tide)
      call put_r0('spinup_function_baroclinic', spinup_function_baroclinic)
      call put_i1('top_layer_at_node', top_layer_at_node)
      call put_i1('top_layer_at_element', top_layer_at_element(1:maxele))
      call put_i1('top_layer_at_face', top_layer_at_face)
      call put_i1('wetdry_node', wetdry_node)
      call put_r1('eta_cell', eta_cell); call put_r1('eta_cell_new', eta_cell_new)
      call put_r1('eta_node', eta_node)
      call put_r1('Gamma_B', Gamma_B); call put_r1('Gamma_T', Gamma_T); call put_r1('Cdb', Cdb)
      call put_r1('wind_stress_normal', wind_stress_normal)
      call put_r1('wind_stress_tangnt', wind_stress_tangnt)
      call put_r1('air_p1', air_p1)
      call put_r0('spinup_function_wind', spinup_function_wind)
      if (hurricane_flag == 1) then
         call put_r1('wind_u0', wind_u0); call put_r1('wind_v0', wind_v0); call put_r1('air_p0', air_p0)
         call put_r1('wind_u1', wind_u1); call put_r1('wind_v1', wind_v1)
         call put_r1('wind_u2', wind_u2); call put_r1('wind_v2', wind_v2); call put_r1('air_p2', air_p2)
         call put_r1('wind_u_at_face', wind_u_at_face); call put_r1('wind_v_at_face', wind_v_at_face)
      end if
      if (num_ob_cell > 0) then
         call put_r1('eta_at_ob_old', eta_at_ob_old); call put_r1('eta_at_ob_new', eta_at_ob_new)
      end if
      if (allocated(Qu_boundary)) call put_r1('Qu_boundary', Qu_boundary)
      if (allocated(WRu_boundary)) call put_r1('WRu_boundary', WRu_boundary)
      call put_i2('num_sub_elm_iteration', num_sub_elm_iteration, 1)
      call put_r2('un_ELM', un_ELM, 1); call put_r2('vn_ELM', vn_ELM, 1)
      call put_r2('un_face', un_face, 1); call put_r2('vn_face', vn_face, 1)
      call put_r2('un_face_new', un_face_new, 1); call put_r2('vn_face_new', vn_face_new, 1)
      call put_r2('wn_cell', wn_cell, 0); call put_r2('wn_cell_new', wn_cell_new, 0)
      call put_r2('u_node', u_node, 0); call put_r2('v_node', v_node, 0)
      call put_r2('w_node', w_node, 0)
      call put_r2('u_face_level', u_face_level, 0); call put_r2('v_face_level', v_face_level, 0)
      call put_r2('dz_node', dz_node, 1); call put_r2('dzhalf_node', dzhalf_node, 0)
      call put_r2('dz_node_new', dz_node_new, 1); call put_r2('dzhalf_node_new', dzhalf_node_new, 0)
      call put_r2('dz_cell', dz_cell, 1); call put_r2('dzhalf_cell', dzhalf_cell, 0)
      call put_r2('dz_cell_new', dz_cell_new, 1); call put_r2('dzhalf_cell_new', dzhalf_cell_new, 0)
      call put_r2('dz_face', dz_face, 1); call put_r2('dzhalf_face', dzhalf_face, 0)
      call put_r2('dz_face_new', dz_face_new, 1); call put_r2('dzhalf_face_new', dzhalf_face_new, 0)
      call put_r2('Kh', Kh, 1); call put_r2('Av', Av, 0); call put_r2('Kv', Kv, 0)
      call put_r2('AinvDeltaZ1', AinvDeltaZ1, 1); call put_r2('AinvDeltaZ2', AinvDeltaZ2, 1)
      call put_r2('AinvG1', AinvG1, 1); call put_r2('AinvG2', AinvG2, 1)
      call put_r2('salt_cell', salt_cell, 1); call put_r2('salt_cell_new', salt_cell_new, 1)
      call put_r2('salt_node', salt_node, 1); call put_r2('salt_face', salt_face, 1)
      call put_r2('temp_cell', temp_cell, 1); call put_r2('temp_cell_new', temp_cell_new, 1)
      call put_r2('temp_node', temp_node, 1); call put_r2('temp_face', temp_face, 1)
      call put_r2('rho_face', rho_face, 1); call put_r2('rho_cell', rho_cell, 1)
      call put_r2('rho_node', rho_node, 1)
      call put_r2('phi_sz', phi_sz, 1)
      call put_r1('ubar_node', ubar_node); call put_r1('vbar_node', vbar_node)
      call put_r1('sbar_node', sbar_node); call put_r1('tbar_node', tbar_node)
      call put_r1('rbar_node', rbar_node)
   end subroutine dump_state

end program gom_ref_harness
