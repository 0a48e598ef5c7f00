/*
 * gom_oracle.h - CPU restatement of GOM v1.1.0's semi-implicit time step, FP64.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP implementation:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. It is
 * never part of the product path.
 *
 * It follows the reference subroutines statement by statement, on the reference's own
 * data layout (Fortran column-major (layer, entity), 1-based ids), serially, in the
 * reference's evaluation order, so that on x86-64 without FMA it reproduces the reference
 * run (OMP_NUM_THREADS=1) bit for bit. The structs are those of the product ABI
 * (include/gomgpu.h): they only describe the arrays.
 *
 * Where the reference reads memory it never wrote, the restatement defines the value:
 *  - TVD automatic arrays (Q/D/d/phi/PSI/r of solve_transport_equation_v19.f90:38-53) read
 *    outside their written ranges are zero, which is what the reference reads when its heap
 *    starts zero-filled (the harness runs it with MALLOC_PERTURB_=255);
 *  - the stale thread-private rtemp0 of solve_momentum_equation.f90:490-509 for an
 *    open-boundary face whose end nodes are not both wet is taken as 0.
 * Parity status of each function is listed in gom_oracle.c's header.
 */
#ifndef GOM_ORACLE_H
#define GOM_ORACLE_H

#include "../include/gomgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

int oracle_friction(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                    const gomgpu_forcing *f);
int oracle_elm(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
               const gomgpu_forcing *f);
int oracle_momentum(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                    const gomgpu_forcing *f);
int oracle_freesurf(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                    const gomgpu_forcing *f);
int oracle_velocity(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                    const gomgpu_forcing *f);
int oracle_layers(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                  const gomgpu_forcing *f);
int oracle_nodevel(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                   const gomgpu_forcing *f);
int oracle_transport(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                     const gomgpu_forcing *f);
int oracle_eos(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
               const gomgpu_forcing *f);
int oracle_update(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                  const gomgpu_forcing *f);
/* The node depth means alone (calculate_vertically_averaged_values, prepare_gom.f90:172). */
int oracle_bars(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s, const gomgpu_forcing *f);
/* Vertical turbulence closure (turbulence_closure_v1.f90:9-492); see gom_oracle.c. */
int oracle_turbulence(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                      const gomgpu_forcing *f);
/* Quasi-equilibrium stability functions Sm, Sh at Gh (GOMGPU_STAB_GA or _KC). */
void oracle_stability(int kind, double gh, double *sm, double *sh);
int oracle_wind(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                const gomgpu_forcing *f);
/* All stages in the order of main.f90:95-272 (friction first), with the turbulence
 * closure between the node velocities and the transport when turbulence_flag != 0. */
int oracle_step(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                const gomgpu_forcing *f);

/* Building blocks, exposed for the known-answer tests. */
int oracle_tridiag(int n, int nc, int ld, const double *a, const double *b, const double *c,
                   const double *rhs, double *sol);
/* Jacobi PCG on a 1-based CSR matrix (pre_conj_grad.f90:9-230); returns iterations to
 * convergence, 0 when max_iter was reached. */
int oracle_pcg(int n, const double *A, const int32_t *col, const int32_t *rowstr,
               const double *b, double *x, int max_iter, double tol);
int oracle_check_intersection(double x1, double x2, double x3, double x4, double y1,
                              double y2, double y3, double y4, double *xin, double *yin);
int oracle_bilinear(double x1, double x2, double x3, double x4, double y1, double y2,
                    double y3, double y4, double x, double y, double *xi, double *eta,
                    double *shapef);
double oracle_rho_unesco(double t, double s);
/* UNESCO-1980 secant-bulk-modulus correction, pres in bar (calculate_density_full.f90:75-101). */
double oracle_rho_pressure(double rho_surface, double t, double s, double pres);
/* Departures of the ELM: with dep non-NULL (2*L*J ints), oracle_elm writes per face-layer
 * [k-1][j-1] the departure cell nnel (1-based, negated for a point within 1e-12 of an edge
 * of its cell, 0 where nothing is backtraced), then the levels jlev; NULL stops it. The
 * points within 1e-12 of an edge are counted: which 0 momentum ELM, 1 tracer ELM. */
void oracle_elm_record(int32_t *dep);
long long oracle_elm_near_edge(int which);
void oracle_elm_near_reset(void);
/* Last PCG iteration count and the first failing entity of the last error. */
int oracle_last_pcg_iterations(void);
int oracle_last_error_entity(void);

#ifdef __cplusplus
}
#endif

#endif
