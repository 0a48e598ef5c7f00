/*
 * gomgpu.h - C ABI of the MI355X implementation of GOM's semi-implicit time step.
 *
 * The reference calls its step as argument-less Fortran subroutines that exchange data
 * through the module mod_global_variables (GOM_Rv1.1.0/source/main.f90:179-272). This ABI
 * replaces them one for one, with the module arrays passed explicitly:
 *
 *   gomgpu_stage(ctx, GOMGPU_STAGE_BOUNDARY)  compute_boundary_eta (ob_eta_type 1, 2), main.f90:84-161
 *                                             calculate_Q, and the transport's boundary
 *                                             concentrations (only after gomgpu_set_boundary;
 *                                             first in gomgpu_step)
 *   gomgpu_stage(ctx, GOMGPU_STAGE_WIND)      read_hurricane_ser_1 +      main.f90:113-137
 *                                             calculate_wind_stress (hurricane_flag = 1 only;
 *                                             first in gomgpu_step)
 *   gomgpu_stage(ctx, GOMGPU_STAGE_FRICTION)  calculate_bottom_friction   main.f90:95
 *   gomgpu_stage(ctx, GOMGPU_STAGE_ELM)       solve_nonlinear_advection   main.f90:167-180
 *   gomgpu_stage(ctx, GOMGPU_STAGE_MOMENTUM)  solve_momentum_equation     main.f90:193
 *   gomgpu_stage(ctx, GOMGPU_STAGE_FREESURF)  solve_free_surface_equation main.f90:201
 *   gomgpu_stage(ctx, GOMGPU_STAGE_VELOCITY)  solve_velocities            main.f90:207
 *   gomgpu_stage(ctx, GOMGPU_STAGE_LAYERS)    update_vertical_layers_1    main.f90:210
 *   gomgpu_stage(ctx, GOMGPU_STAGE_NODEVEL)   calculate_velocity_at_node  main.f90:213
 *   gomgpu_stage(ctx, GOMGPU_STAGE_TURBULENCE) subroutine turbulence      turbulence_closure_v1.f90:9-492
 *                                             (not called by main.f90; runs between NODEVEL and
 *                                             TRANSPORT when turbulence_flag != 0)
 *   gomgpu_stage(ctx, GOMGPU_STAGE_TRANSPORT) solve_transport_equation_v19 main.f90:231
 *                                             (transport_solver = 1; = 2: the ELM solver
 *                                             solve_transport_equation_ELM_v8)
 *   gomgpu_stage(ctx, GOMGPU_STAGE_EOS)       calculate_density_full/_linear main.f90:261,264
 *   gomgpu_stage(ctx, GOMGPU_STAGE_UPDATE)    update_variables            main.f90:272
 *   gomgpu_step(ctx)                          all of the above, in that order
 *   gomgpu_stage(ctx, GOMGPU_STAGE_BARS)      calculate_vertically_averaged_values, the node
 *                                             depth means (prepare_gom.f90:172, at start-up;
 *                                             in a step the UPDATE stage forms them)
 *
 * Layout conventions (identical to the reference module arrays, so a Fortran caller passes
 * its allocatables with c_loc): every 2-D array is Fortran column-major (layer, entity),
 * entity ids are 1-based int32 with 0 meaning "none". Arrays whose Fortran lower bound is
 * 0 (dzhalf_*, wn_cell*, u/v/w_node, u/v_face_level, Av, Kv) have leading dimension
 * maxlayer+1, the others maxlayer. Host arrays are only read (or written) during the call;
 * no pointer is kept. Device state is owned by the context and stays resident between
 * steps; host copies are refreshed only by gomgpu_get_state.
 *
 * Errors: every function returns 0 on success, a positive GOMGPU_E* code on failure. A
 * failing device check (the reference's `stop` conditions) is latched with the first
 * failing entity and reported by gomgpu_last_error. The PCG never fails silently: a solve
 * cut off by max_iteration_pcg is reported as GOMGPU_W_PCG_NOT_CONVERGED by the next
 * gomgpu_sync (once; the step carries on with the last iterate, as the reference does),
 * and the iteration count of the last solve is available from gomgpu_pcg_iterations.
 */
#ifndef GOMGPU_H
#define GOMGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------
 * Field lists. Each entry is X(name, extent) where extent names the array shape:
 *   P nodes, E cells, J faces, L maxlayer, L1 maxlayer+1, NB max_no_neighbor_node,
 *   NB1 max_no_neighbor_node+1, OB num_ob_cell, QB num_Qb_cell; HP and HJ are P and J
 *   when hurricane_flag = 1 and 1 otherwise.
 * ------------------------------------------------------------------------------------- */

/* Static geometry, from set_geometry_1.f90 / set_geometry_2.f90 / read_*_inp.f90. */
#define GOMGPU_MESH_REAL(X) \
  X(x_node, P) X(y_node, P) X(h_node, P) \
  X(x_cell, E) X(y_cell, E) X(h_cell, E) X(area, E) \
  X(face_length, J) X(x_face, J) X(y_face, J) X(h_face, J) X(delta_j, J) \
  X(cos_theta, J) X(sin_theta, J) X(cos_theta2, J) X(sin_theta2, J) \
  X(coriolis_factor, J) X(manning, J) X(bf_height, J) \
  X(z_level, L1) X(delta_z, L)

#define GOMGPU_MESH_INT(X) \
  X(tri_or_quad, E) X(nodenum_at_cell, 4E) X(adj_cellnum_at_cell, 4E) \
  X(facenum_at_cell, 4E) X(sign_in_outflow, 4E) \
  X(adj_cellnum_at_face, 2J) X(nodenum_at_face, 2J) \
  X(adj_cells_at_node, P) X(adj_cellnum_at_node, NBP) X(node_count_each_element, NBP) \
  X(adj_nodes_at_node, P) X(adj_nodenum_at_node, NB1P) \
  X(boundary_type_of_face, J) X(isflowside3, J) \
  X(ob_element_flag, E) X(Qb_element_flag, E) \
  X(bottom_layer_at_node, P) X(bottom_layer_at_element, E) X(bottom_layer_at_face, J) \
  X(initial_wetdry_node, P) X(ob_cell_id, OB) X(Qb_cell_id, QB)

/* Prognostic and diagnostic state (mod_global_variables.f90). */
#define GOMGPU_STATE_REAL(X) \
  X(eta_cell, E) X(eta_cell_new, E) X(eta_node, P) X(Gamma_B, J) X(Cdb, J) \
  X(un_face, LJ) X(vn_face, LJ) X(un_face_new, LJ) X(vn_face_new, LJ) \
  X(un_ELM, LJ) X(vn_ELM, LJ) X(dz_face, LJ) X(dz_face_new, LJ) \
  X(AinvG1, LJ) X(AinvG2, LJ) X(AinvDeltaZ1, LJ) X(AinvDeltaZ2, LJ) X(Kh, LJ) \
  X(salt_face, LJ) X(temp_face, LJ) X(rho_face, LJ) \
  X(dzhalf_face, L1J) X(dzhalf_face_new, L1J) X(Av, L1J) X(q2, L1J) X(xl, L1J) \
  X(u_face_level, L1J) X(v_face_level, L1J) \
  X(dz_cell, LE) X(dz_cell_new, LE) X(salt_cell, LE) X(salt_cell_new, LE) \
  X(temp_cell, LE) X(temp_cell_new, LE) X(rho_cell, LE) \
  X(dzhalf_cell, L1E) X(dzhalf_cell_new, L1E) X(wn_cell, L1E) X(wn_cell_new, L1E) X(Kv, L1E) \
  X(dz_node, LP) X(dz_node_new, LP) X(salt_node, LP) X(temp_node, LP) X(rho_node, LP) \
  X(dzhalf_node, L1P) X(dzhalf_node_new, L1P) \
  X(u_node, L1P) X(v_node, L1P) X(w_node, L1P) \
  X(ubar_node, P) X(vbar_node, P) X(sbar_node, P) X(tbar_node, P) X(rbar_node, P)

#define GOMGPU_STATE_INT(X) \
  X(top_layer_at_node, P) X(top_layer_at_element, E) X(top_layer_at_face, J) \
  X(wetdry_node, P) X(num_sub_elm_iteration, LP)

/* Per-step forcing computed by the host's 2-D boundary routines (main.f90:84-161,
 * heat_term_by_term.f90) and the boundary concentrations of the transport sub-cycles
 * (solve_transport_equation_v19.f90:229-534): con_at_ob is (Nt, maxtran2, OB) and
 * con_at_Qb (Nt, maxtran2, QB), fastest index first. With params.river_on_device = 1 the
 * river face velocity Qu_boundary is computed on the device from the discharge Q_add and
 * the current face depth (compute_boundary_velocity.f90:32-46) instead of being passed.
 * With transport_solver = 2 (the ELM solver, which carries salt and temperature both, so
 * maxtran2 = 2) sub-cycle 0 of con_at_ob and con_at_Qb holds the boundary values of
 * find_openboundary_salt_temp_v3.f90 (salt_at_obck / temp_at_obck, salt_at_Qbc /
 * temp_at_Qbc).
 * With params.hurricane_flag = 1 the wind stage keeps the hurricane fields on the device
 * (read_hurricane_ser_1.f90): snapshot pair wind_u0/v0/air_p0 and wind_u2/v2/air_p2, the
 * interpolated wind_u1/v1/air_p1 and the face winds; the host passes them once at the
 * start (u0 = first snapshot, u2 = second, the rest zero) and the next snapshot in
 * wind_u/v/air_p_next before a step whose hurricane_phase is 2. Gamma_T and the wind
 * stress are then outputs of that stage (calculate_wind_stress.f90:88-114). */
#define GOMGPU_FORCING_REAL(X) \
  X(Gamma_T, J) X(wind_stress_normal, J) X(wind_stress_tangnt, J) X(air_p1, P) \
  X(eta_at_ob_old, OB) X(eta_at_ob_new, OB) X(Qu_boundary, QB) X(phi_sz, LE) \
  X(con_at_ob, NTOB) X(con_at_Qb, NTQB) X(Q_add, QB) \
  X(wind_u0, HP) X(wind_v0, HP) X(air_p0, HP) X(wind_u1, HP) X(wind_v1, HP) \
  X(wind_u2, HP) X(wind_v2, HP) X(air_p2, HP) \
  X(wind_u_next, HP) X(wind_v_next, HP) X(air_p_next, HP) \
  X(wind_u_at_face, HJ) X(wind_v_at_face, HJ)

/* Scalars of main.inp and set-up (read_main_inp.f90). */
#define GOMGPU_PARAMS_INT(X) \
  X(maxnod) X(maxele) X(maxface) X(maxlayer) X(max_no_neighbor_node) \
  X(num_ob_cell) X(num_Qb_cell) X(maxtran2) X(tran_id1) X(tran_id2) \
  X(advection_flag) X(ELM_backtrace_flag) X(ELM_min_iter) X(ELM_max_iter) \
  X(bf_flag) X(baroclinic_flag) X(ana_density) X(i_density_flag) \
  X(transport_flag) X(trans_sub_iter) X(h_flux_limiter) X(v_flux_limiter) \
  X(heat_option) X(max_iteration_pcg) X(precision) X(river_on_device) \
  X(turbulence_flag) X(turbulence_model) X(stability_function) \
  X(hurricane_flag) X(hurricane_read_interval) X(wind_formular) X(transport_solver)

#define GOMGPU_PARAMS_REAL(X) \
  X(dt) X(theta) X(gravity) X(rho_o) X(MSL) X(dry_depth) X(adv_onoff_depth) \
  X(smagorinsky_parameter) X(Ah_0) X(Kh_0) X(von_Karman) X(error_tolerance) \
  X(denominator_min_for_matrix) X(delta_z_min) \
  X(q2min) X(psimin) X(eps_min) X(bgdiff) X(diffmax) X(xlmin1) X(xlmin2) \
  X(cmiu0) X(cpsi1) X(cpsi2) X(rpub) X(rmub) X(rnub) X(schk) X(schpsi) \
  X(h1_pp) X(h2_pp) X(vdmax_pp1) X(vdmax_pp2) X(vdmin_pp1) X(vdmin_pp2) X(tdmin_pp) \
  X(rho_a)

/* hurricane_phase of a step (main.f90:116-123): 0 outside the hurricane window (the face
 * winds are kept), 1 interpolate (read_hurricane_ser_1.f90:48-60), 2 take the next
 * snapshot (read_hurricane_ser_1.f90:18-46, every hurricane_read_interval steps). */
#define GOMGPU_FORCING_SCALARS(X) \
  X(elapsed_time) X(spinup_function_tide) X(spinup_function_baroclinic) \
  X(spinup_function_wind) X(hurricane_phase) X(julian_day) \
  X(hurricane_centre_x) X(hurricane_centre_y)

#define GOMGPU_DECL_RPTR(name, ext) double *name;
#define GOMGPU_DECL_IPTR(name, ext) int32_t *name;
#define GOMGPU_DECL_CRPTR(name, ext) const double *name;
#define GOMGPU_DECL_CIPTR(name, ext) const int32_t *name;
#define GOMGPU_DECL_INT(name) int32_t name;
#define GOMGPU_DECL_REAL(name) double name;

typedef struct gomgpu_mesh {
  GOMGPU_MESH_REAL(GOMGPU_DECL_CRPTR)
  GOMGPU_MESH_INT(GOMGPU_DECL_CIPTR)
} gomgpu_mesh;

typedef struct gomgpu_params {
  GOMGPU_PARAMS_INT(GOMGPU_DECL_INT)
  GOMGPU_PARAMS_REAL(GOMGPU_DECL_REAL)
} gomgpu_params;

/* Fields may be NULL in gomgpu_get_state (not copied) and in gomgpu_set_state (left as
 * they are on the device; a new context starts from zeros). */
typedef struct gomgpu_state {
  GOMGPU_STATE_REAL(GOMGPU_DECL_RPTR)
  GOMGPU_STATE_INT(GOMGPU_DECL_IPTR)
} gomgpu_state;

/* Fields may be NULL in gomgpu_set_forcing (left as they are on the device). */
typedef struct gomgpu_forcing {
  GOMGPU_FORCING_SCALARS(GOMGPU_DECL_REAL)
  GOMGPU_FORCING_REAL(GOMGPU_DECL_CRPTR)
} gomgpu_forcing;

/* Boundary forcing computed on the device by GOMGPU_STAGE_BOUNDARY, once
 * gomgpu_set_boundary has uploaded these tables: the open-boundary elevations eta_at_ob_old/new
 * of compute_boundary_eta.f90:72-124 (harmonic sets, ob_eta_type 1, and elevation series, 2),
 * the river discharge Q_add of calculate_Q.f90:21-45 (q_interp_method 1), and con_at_ob /
 * con_at_Qb of the transport: the sub-cycles of solve_transport_equation_v19.f90:229-397, or
 * find_openboundary_salt_temp_v3.f90:17-97 with transport_solver 2. The per-step inputs are
 * the forcing scalars (elapsed_time, julian_day after its increment, the tide and baroclinic
 * spin-up factors); the host then uploads no boundary array. Series are 1-based indices into
 * one table of all series (0 = none); a series is the reference's times [s] and scaled values.
 * The harmonic terms use a correctly rounded cosine; the C library's cos, which the reference
 * calls, rounds a few arguments in a thousand the other way (one ulp). */
typedef struct gomgpu_boundary {
  int32_t num_harmonic_sets;
  const int32_t *harmonic_first;     /* (num_harmonic_sets + 1) first constituent of each set, 0-based, then the end */
  const double *tidal_phase_shift;   /* (num_harmonic_sets) degrees */
  const double *constituents;        /* (5, constituents) period [h], amplitude [m], nodal factor,
                                        equilibrium argument [deg], phase [deg] */
  int32_t num_series;
  const int32_t *series_first;       /* (num_series + 1) first point of each series, 0-based, then the end */
  const double *series_time;         /* seconds (Series.seconds of the data year) */
  const double *series_value;        /* scaled values */
  int32_t num_ob;                    /* num_ob_cell */
  const int32_t *ob_eta_type;        /* (num_ob) 1 harmonic set, 2 elevation series */
  const int32_t *ob_eta_id;          /* (num_ob) harmonic set or series, 1-based */
  const int32_t *ob_tracer_series;   /* (2, num_ob) salinity and temperature series, 0 = none */
  int32_t num_qb;                    /* num_Qb_cell */
  const int32_t *qb_q_series;        /* (num_qb) discharge series, 0 = none */
  const double *qb_portion;          /* (num_qb) Q_portion */
  const int32_t *qb_tracer_series;   /* (2, num_qb) */
  int32_t is_salt, is_temp;          /* the ELM solver's boundaries: open ones without a series take */
  double ref_salt, ref_temp;         /*   ref_salt / ref_temp */
  int32_t reference_diff_days;       /* calculate_reference_diff_days.f90: series times shift by it */
} gomgpu_boundary;

/* A Holland (1980) parametric cyclone generated on the device (gomgpu_set_holland): at each
 * step that takes the next hurricane snapshot (hurricane_phase 2) the wind stage forms
 * wind_u_next, wind_v_next and air_p_next itself, for the centre the step's forcing scalars
 * hurricane_centre_x/_y give, instead of the host uploading 3 maxnod values. With
 * s = (r_max / max(r, 1 m))^b the pressure is p_min + (p_env - p_min) e^-s [hPa, stored x100
 * as read_hurricane_ser_0.f90:35 stores Pa], the gradient wind
 * V = sqrt(b dp s e^-s / rho_air + (r f / 2)^2) - r f / 2 with dp = 100 (p_env - p_min) and
 * f = coriolis, surface_factor V counter-clockwise about the centre (0 within 1 m of it) plus
 * the translation (vx, vy). No counterpart in the reference, whose Holland routine
 * (holland_storm_surge.f90) no call reaches; it reads the fields from hurricane_ser.inp. */
typedef struct gomgpu_holland {
  double p_min, p_env;          /* hPa */
  double r_max, b;              /* m, Holland B */
  double coriolis;              /* f [1/s] */
  double rho_air;               /* kg/m3 */
  double surface_factor;
  double vx, vy;                /* translation velocity [m/s] */
} gomgpu_holland;

/* Destination of gomgpu_get_forcing: the forcing arrays as they are on the device
 * (the outputs of the wind stage among them); NULL fields are not copied. */
typedef struct gomgpu_forcing_out {
  GOMGPU_FORCING_REAL(GOMGPU_DECL_RPTR)
} gomgpu_forcing_out;

/* Vertical turbulence closure (gomgpu_params.turbulence_flag; turbulence_closure_v1.f90):
 * 0 constant Av/Kv, 1 Richardson step function, 2 Pacanowski-Philander, 3 two-equation
 * GLS / Mellor-Yamada-Galperin. turbulence_model selects the second equation of scheme 3
 * (i_turbulence_model_name 'my','kl','ke','kw','ub'); stability_function the quasi-
 * equilibrium stability functions (Galperin et al. 1988 or Kantha & Clayson 1994). q2 is the
 * turbulent kinetic energy k and xl the length scale, at whole levels 0..maxlayer of faces. */
enum { GOMGPU_TURB_OFF = 0, GOMGPU_TURB_STEP = 1, GOMGPU_TURB_PP = 2, GOMGPU_TURB_GLS = 3 };
enum { GOMGPU_GLS_MY = 0, GOMGPU_GLS_KL = 1, GOMGPU_GLS_KE = 2, GOMGPU_GLS_KW = 3, GOMGPU_GLS_UB = 4 };
enum { GOMGPU_STAB_GA = 0, GOMGPU_STAB_KC = 1 };

/* Precision of the bulk arithmetic (gomgpu_params.precision). */
enum { GOMGPU_FP64 = 0, GOMGPU_MIXED = 1 };

enum gomgpu_stage_id {
  GOMGPU_STAGE_FRICTION = 0,
  GOMGPU_STAGE_ELM = 1,
  GOMGPU_STAGE_MOMENTUM = 2,
  GOMGPU_STAGE_FREESURF = 3,
  GOMGPU_STAGE_VELOCITY = 4,
  GOMGPU_STAGE_LAYERS = 5,
  GOMGPU_STAGE_NODEVEL = 6,
  GOMGPU_STAGE_TRANSPORT = 7,
  GOMGPU_STAGE_EOS = 8,
  GOMGPU_STAGE_UPDATE = 9,
  GOMGPU_STAGE_TURBULENCE = 10,
  GOMGPU_STAGE_WIND = 11,
  GOMGPU_STAGE_BOUNDARY = 12,
  GOMGPU_STAGE_BARS = 13,
  GOMGPU_NUM_STAGES = 14
};

enum gomgpu_status {
  GOMGPU_OK = 0,
  GOMGPU_E_ARG = 1,         /* bad argument or unsupported option */
  GOMGPU_E_HIP = 2,         /* HIP runtime error */
  GOMGPU_E_NOMEM = 3,       /* device allocation failed */
  GOMGPU_E_MODEL = 4,       /* a reference `stop` condition fired on the device */
  GOMGPU_W_PCG_NOT_CONVERGED = 5
};

/* Device error codes latched by the kernels (the reference's stop conditions). */
enum gomgpu_model_error {
  GOMGPU_ERR_NONE = 0,
  GOMGPU_ERR_TRIDIAG_PIVOT = 1,      /* tridiagonal_solver.f90:32-35,47-50 */
  GOMGPU_ERR_NOT_SPD = 2,            /* solve_free_surface_equation.f90:100-104 */
  GOMGPU_ERR_LAYER_CELL = 3,         /* update_vertical_layers_1.f90:66-81 */
  GOMGPU_ERR_LAYER_FACE = 4,         /* update_vertical_layers_1.f90:166-176 */
  GOMGPU_ERR_LAYER_NODE = 5,         /* update_vertical_layers_1.f90:224-234 */
  GOMGPU_ERR_WET_DRY_SIDES = 6,      /* update_vertical_layers_1.f90:273-276 */
  GOMGPU_ERR_DZHALF_MIN = 7,         /* update_vertical_layers_1.f90:281-330 */
  GOMGPU_ERR_ELM_DRY_START = 8,      /* search_straight_line.f90:37-46 */
  GOMGPU_ERR_ELM_NOT_IN_CELL = 9,    /* search_straight_line.f90:82-89 */
  GOMGPU_ERR_ELM_ZERO_PATH = 10,     /* search_straight_line.f90:129-139 */
  GOMGPU_ERR_ELM_NO_EDGE = 11,       /* search_straight_line.f90:162-169, 343-353 */
  GOMGPU_ERR_ELM_OVERSHOT = 12,      /* search_straight_line.f90:199-207 */
  GOMGPU_ERR_ELM_TARGET = 13,        /* search_straight_line.f90:218-222 */
  GOMGPU_ERR_ELM_WRONG_SIDE = 14,    /* search_straight_line.f90:237-242 */
  GOMGPU_ERR_ELM_DRY_END = 15,       /* search_straight_line.f90:363-372 */
  GOMGPU_ERR_ELM_VERTICAL = 16,      /* search_straight_line.f90:379-385, ELM_backtrace_v1.f90:107-113 */
  GOMGPU_ERR_ELM_OUTSIDE_QUAD = 17,  /* ELM_backtrace_v1.f90:164-170 */
  GOMGPU_ERR_BILINEAR = 18,          /* bilinear_interpolation.f90:41-112 */
  GOMGPU_ERR_EOS_RANGE = 19,         /* calculate_density_full.f90:43-53,103-107 */
  GOMGPU_ERR_EOS_DRY_NODES = 20,     /* calculate_density_full.f90:223-225 */
  GOMGPU_ERR_TURB_XL = 21,           /* turbulence_closure_v1.f90:163-170 */
  GOMGPU_ERR_TURB_DIST = 22,         /* turbulence_closure_v1.f90:209-214 */
  GOMGPU_ERR_TURB_NEGATIVE = 23,     /* turbulence_closure_v1.f90:384-388 */
  GOMGPU_ERR_ELM_ALL_DRY = 24,       /* ELM_backtrace_v1.f90:275-279 (tracer foot in a dry cell) */
  GOMGPU_ERR_ELM_NEG_SUBCELL = 25,   /* ELM_backtrace_v1.f90:420-424 */
  GOMGPU_ERR_ELM_NO_SUBCELL = 26     /* ELM_backtrace_v1.f90:463-468 */
};

typedef struct gomgpu_ctx gomgpu_ctx;

/* Create a context on the current HIP device; uploads the static mesh once. */
int gomgpu_create(const gomgpu_mesh *mesh, const gomgpu_params *params, gomgpu_ctx **out);
/* Upload state fields (non-NULL ones) after prepare_gom or a restart. */
int gomgpu_set_state(gomgpu_ctx *ctx, const gomgpu_state *state);
/* Upload the host forcing of the coming step (main.f90:84-161). */
int gomgpu_set_forcing(gomgpu_ctx *ctx, const gomgpu_forcing *forcing);
/* Upload the boundary tables: from now on each step computes its boundary forcing on the
 * device (GOMGPU_STAGE_BOUNDARY). No counterpart in the reference, whose host computes it
 * (main.f90:84-161); for callers that keep the forcing on the device. */
int gomgpu_set_boundary(gomgpu_ctx *ctx, const gomgpu_boundary *tables);
/* From now on the wind stage generates the next hurricane snapshot on the device
 * (gomgpu_holland). No counterpart in the reference; for cyclone scenarios that would
 * otherwise upload every snapshot. */
int gomgpu_set_holland(gomgpu_ctx *ctx, const gomgpu_holland *cyclone);
/* Run one stage (for per-stage parity) or the whole step. Both are asynchronous on the
 * context's stream except where a stage needs a device-to-host value (none in FP64). */
int gomgpu_stage(gomgpu_ctx *ctx, int stage_id);
int gomgpu_step(gomgpu_ctx *ctx);
/* Advance n steps with the forcing already uploaded (benchmark and ensemble use: the
 * forcing is held fixed). */
int gomgpu_run(gomgpu_ctx *ctx, int nsteps);
/* Synchronise and download the non-NULL fields. */
int gomgpu_get_state(gomgpu_ctx *ctx, gomgpu_state *state);
/* Snapshot of the non-NULL fields of `dst` as they stand after the work already submitted,
 * without waiting for it: the device copies them on the step's stream and downloads the copy
 * on a stream of its own, so the following steps can be submitted (and run) meanwhile.
 * gomgpu_snapshot_wait completes it: it waits for the download and writes the fields into
 * the arrays `dst` pointed to, which must stay valid until then. One snapshot is in flight
 * per context (a new one completes the previous first); the two calls may come from
 * different threads. No counterpart in the reference, which writes its output files from
 * the host arrays between steps (write_output_files.f90); for asynchronous output. */
int gomgpu_snapshot(gomgpu_ctx *ctx, const gomgpu_state *dst);
int gomgpu_snapshot_wait(gomgpu_ctx *ctx);
/* Synchronise and download the non-NULL forcing arrays (the wind stage writes Gamma_T,
 * the wind stress, air_p1 and the hurricane fields on the device; a Fortran host reads
 * them back for its output, as write_tser_out_files.f90:312 reads air_p1). */
int gomgpu_get_forcing(gomgpu_ctx *ctx, gomgpu_forcing_out *forcing);
/* Total volume sum(area*(h+eta)) over wet cells and salt sum(S*area*dz) in FP64. */
int gomgpu_budgets(gomgpu_ctx *ctx, double *volume, double *salt);
/* Iterations taken by the last PCG solve (0 if it did not converge). */
int gomgpu_pcg_iterations(gomgpu_ctx *ctx);
/* Departure points of the momentum ELM (solve_nonlinear_advection.f90:92-202 calling
 * ELM_backtrace_v1.f90:15-200; the reference keeps nnel and jlev in locals of that loop).
 * A departure point within 1e-12 of an edge of its cell, relative to the edge's length,
 * can fall in either cell sharing the edge on the last bit of its coordinates; such points
 * are counted. gomgpu_elm_record(ctx, 1) makes every later ELM stage keep, for face-layer
 * t = (k-1)*maxface + (j-1), the departure cell nnel (1-based, negated for a point within
 * 1e-12 of an edge, 0 where nothing was backtraced) and level jlev (0 likewise); record = 0
 * stops keeping them. gomgpu_elm_departures waits for the context's queued work and
 * downloads the last kept nnel and jlev (maxlayer*maxface each; either may be NULL).
 * gomgpu_elm_near_edge waits likewise and returns the points within 1e-12 of an edge since
 * creation or the last reset: counts[0] momentum ELM, counts[1] tracer ELM
 * (transport_solver 2); reset != 0 zeroes them after reading. */
int gomgpu_elm_record(gomgpu_ctx *ctx, int record);
int gomgpu_elm_departures(gomgpu_ctx *ctx, int32_t *nnel, int32_t *jlev);
int gomgpu_elm_near_edge(gomgpu_ctx *ctx, long long *counts, int reset);
/* Wait for all queued work; returns the latched model error as GOMGPU_E_MODEL. */
int gomgpu_sync(gomgpu_ctx *ctx);
/* Text of the last error (static storage of the context). With ctx NULL: why the last
 * gomgpu_create on the calling thread failed. */
const char *gomgpu_last_error(gomgpu_ctx *ctx);
/* Event timing. enable != 0 turns it on (turning it on resets the kernel averages);
 * fills up to n (name, milliseconds) pairs: the GOMGPU_NUM_STAGES stage times of the last
 * step, then the average launch time of each hot kernel (k_pcg, k_elm, k_momentum,
 * k_tvd_column) since timing was turned on. Timing synchronises after every timed kernel,
 * so it is for measurement passes only. */
#define GOMGPU_NUM_KTIMERS 4
int gomgpu_timings(gomgpu_ctx *ctx, int enable, const char **names, double *ms, int n);
void gomgpu_destroy(gomgpu_ctx *ctx);

/* ---------------------------------------------------------------------------------------
 * Batches (ensembles and the serving loop; no counterpart in the reference, which runs one
 * scenario per process). A batch advances several independent contexts of one device by
 * one step in one submission: each member's stages run on its own stream, the PCG solves of
 * all members run in fused launches, and the submission is captured as a hipGraph once per
 * membership and replayed (GOMGPU_BATCH_GRAPH=0 submits eagerly instead). A member's result
 * is bit for bit the same alone and in any batch. Forcing is uploaded per member with
 * gomgpu_set_forcing before gomgpu_batch_step; gomgpu_get_state on a member is ordered after
 * the batch steps submitted before it.
 * ------------------------------------------------------------------------------------- */
/* Launch shapes the autotuner may change: threads per block of the kernels whose threads
 * each compute their own outputs, so no choice changes any result (the PCG partition is a
 * fixed function of the system size and is not tunable). Powers of two from 64 to 1024
 * (256 for ELM, MOM_RHS and TURB_SOLVE). No counterpart in the reference. */
enum {
  GOMGPU_LAUNCH_ELM = 0,        /* k_elm, one thread per (layer, face) */
  GOMGPU_LAUNCH_MOM_RHS = 1,    /* k_mom_rhs, 64 faces per block, one wave per row stride */
  GOMGPU_LAUNCH_MOM_SOLVE = 2,  /* k_mom_solve, one thread per face column */
  GOMGPU_LAUNCH_TVD_RHS = 3,    /* k_tvd_rhs, one thread per (layer, cell) */
  GOMGPU_LAUNCH_TVD_SOLVE = 4,  /* k_tvd_solve, one thread per cell column */
  GOMGPU_LAUNCH_TVD_LIM = 5,    /* k_tvd_lim_cell, one thread per (level, cell) */
  GOMGPU_LAUNCH_NODE_VEL = 6,   /* k_node_vel */
  GOMGPU_LAUNCH_TURB_PREP = 7,  /* k_turb_ri, one thread per (layer, face) (closure schemes 1, 2) */
  GOMGPU_LAUNCH_TURB_SOLVE = 8, /* k_turb_column, one thread per face column (scheme 3) */
  GOMGPU_LAUNCH_STREAM = 9,     /* every other per-entity kernel */
  GOMGPU_NUM_LAUNCH = 10
};
int gomgpu_set_launch(gomgpu_ctx *ctx, int which, int block);
int gomgpu_get_launch(gomgpu_ctx *ctx, int which);
/* Rows per thread of the register-resident PCG partition: 0 (default) picks the fewest of
 * 1, 2, 4 that fit 256 blocks, spreading a lone system over the most compute units; 4 packs
 * more batch members into one launch. Unlike the launch shapes this changes how the dot
 * products are associated (results agree to rounding, not bit for bit, across settings); a
 * context gives the same results alone and in a batch for the same setting. No
 * counterpart in the reference. */
int gomgpu_set_pcg_rows(gomgpu_ctx *ctx, int rows);

/* Device bytes gomgpu_create would allocate for these parameters (admission by HBM
 * footprint), and the bytes a context holds (gomgpu_elm_record adds 8*maxlayer*maxface);
 * -1 on bad arguments. */
long long gomgpu_footprint(const gomgpu_params *params);
long long gomgpu_device_bytes(gomgpu_ctx *ctx);

/* ---------------------------------------------------------------------------------------
 * Standalone free-surface solve: replaces pre_conj_grad(na, nz, A, colidx, rowstr, b, x)
 * (pre_conj_grad.f90:9, called at solve_free_surface_equation.f90:321). Jacobi-
 * preconditioned CG on a 1-based CSR matrix (A, colidx: nz entries; rowstr: na+1 entries),
 * right-hand side b and initial guess x (overwritten with the solution), stopping when
 * sqrt(delta_new^2 / delta_0^2) < error_tolerance (pre_conj_grad.f90:199) or after
 * max_iteration iterations. All five arrays are DEVICE pointers (the matrix of a caller that
 * assembles on the GPU). Rows must be free-surface rows: one diagonal entry and at most 4
 * off-diagonal entries (solve_free_surface_equation.f90:233-258); any other matrix is refused
 * with GOMGPU_E_ARG. stream is a hipStream_t (NULL: the null stream). The call returns after
 * the solve; *iterations receives the reference's cgit2 (0 when it did not converge, with
 * GOMGPU_W_PCG_NOT_CONVERGED returned instead of the reference's silence). The workspace
 * (gomgpu_pcg_workspace_bytes) is allocated and freed inside the call. gomgpu_pcg_last_error
 * returns the text of the calling thread's last failure. */
int gomgpu_pcg_solve(int na, int nz, const double *A, const int32_t *colidx, const int32_t *rowstr,
                     const double *b, double *x, int max_iteration, double error_tolerance, int *iterations,
                     void *stream);
long long gomgpu_pcg_workspace_bytes(int na);
const char *gomgpu_pcg_last_error(void);

/* gomgpu_pcg_solve without any device allocation. No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same arguments, kernels and results (bit for bit) as
 * gomgpu_pcg_solve, for at most GOMGPU_PCG_NOALLOC_MAX_ROWS rows (GOMGPU_E_ARG above). The
 * workspace is a fixed device array of the library (about 30.6 MB, held on every device the
 * library runs on until the process ends), shared by all calls on a device: calls that
 * overlap on one device race, and each call clears what must start at zero. */
#define GOMGPU_PCG_NOALLOC_MAX_ROWS 262144
int gomgpu_pcg_solve_noalloc(int na, int nz, const double *A, const int32_t *colidx, const int32_t *rowstr,
                             const double *b, double *x, int max_iteration, double error_tolerance,
                             int *iterations, void *stream);

typedef struct gomgpu_batch gomgpu_batch;
/* Batched contexts: one call steps every member. Destroying a member context removes it from
 * its batches; gomgpu_batch_destroy leaves the members usable on their own. */
int gomgpu_batch_create(gomgpu_batch **out);
int gomgpu_batch_add(gomgpu_batch *batch, gomgpu_ctx *member);
int gomgpu_batch_remove(gomgpu_batch *batch, gomgpu_ctx *member);
int gomgpu_batch_size(gomgpu_batch *batch);
/* How the last step solved the members' free surfaces: *groups fused launches of
 * register-resident solves (members packed while their blocks fit the co-resident grid) and
 * *solo members solved alone (more blocks than fit, or GOMGPU_MIXED). 0 and 0 before the
 * first step. */
int gomgpu_batch_layout(gomgpu_batch *batch, int *groups, int *solo);
/* One step of every member (asynchronous). */
int gomgpu_batch_step(gomgpu_batch *batch);
/* Wait for the batch; returns the first member error (gomgpu_sync semantics). */
int gomgpu_batch_sync(gomgpu_batch *batch);
const char *gomgpu_batch_last_error(gomgpu_batch *batch);
void gomgpu_batch_destroy(gomgpu_batch *batch);

#ifdef __cplusplus
}
#endif

#endif /* GOMGPU_H */
