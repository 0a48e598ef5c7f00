/*
 * gom_oracle.c - CPU restatement of GOM v1.1.0's semi-implicit step (test oracle only).
 *
 * Every function cites the reference routine it restates (paths relative to
 * GOM_Rv1.1.0/source). Evaluation order, literal widths and branch structure follow the
 * reference; single-precision Fortran literals appear as float literals widened to double
 * (e.g. 1.e-4 is (double)1.e-4f), integer powers x**n are evaluated as the sequential
 * products amdflang emits, and `a - b` constant folds of two default-real literals are done
 * in float, as the Fortran compiler does.
 *
 * Parity pins (see tests/): every stage is compared bit for bit against the reference
 * harness dumps on the committed fixtures (channel and estuary cases, limiters on and off).
 * Building blocks are pinned independently: tridiag against numpy dense solves, PCG against
 * scipy spsolve and a literal numpy transcription of the update order on small matrices,
 * check_intersection/bilinear by closed forms, the EOS against UNESCO-1980 check values, the
 * linear EOS by its closed form and the stop it makes on unset face densities. The ELM
 * scalar transport (transport_solver 2) is pinned by its own fixture, wet and dry.
 * Not pinned by any reference output: the pressure correction of the full EOS (only below
 * 100 m, absent from every fixture) -- "parity unpinned", it is checked against the
 * published UNESCO-1980 high-pressure check value only.
 *
 * Threads: liboracle_omp.so is this file built with -fopenmp, liboracle.so without. The
 * parallel loops run over independent entities (faces, cells, nodes) with private scratch,
 * and every sum keeps its serial order, so both builds give the same bits
 * (tests/test_oracle_threads.py). A failing entity in a parallel loop is recorded and the
 * lowest one reported, which is the one the serial loop stops at.
 */
#include "gom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ index helpers ---- */
#define R1(a, i) ((a)[(i) - 1])
#define RL(a, k, e) ((a)[((k) - 1) + (size_t)((e) - 1) * L])
#define R0(a, k, e) ((a)[(k) + (size_t)((e) - 1) * (L + 1)])
#define I4(a, l, i) ((a)[((l) - 1) + (size_t)((i) - 1) * 4])
#define I2(a, l, j) ((a)[((l) - 1) + (size_t)((j) - 1) * 2])
#define INB(a, l, i) ((a)[((l) - 1) + (size_t)((i) - 1) * NB])
#define INB0(a, l, i) ((a)[(l) + (size_t)((i) - 1) * (NB + 1)])
#define ZL(k) (m->z_level[(k)])

static int g_last_iters = 0;
static _Thread_local int g_err_entity = 0; /* set by the calling thread */

int oracle_last_pcg_iterations(void) { return g_last_iters; }
int oracle_last_error_entity(void) { return g_err_entity; }

static int fail(int code, int entity) {
  g_err_entity = entity;
  return code;
}

/* The first failure of a parallel loop over entities: the lowest entity and its code. */
typedef struct {
  int code, entity;
} loop_err;

static void loop_fail(loop_err *e, int code, int entity) {
#pragma omp critical(oracle_loop_err)
  {
    if (e->code == 0 || entity < e->entity) {
      e->code = code;
      e->entity = entity;
    }
  }
}

/* start_end_node(N, l, w): set_geometry_1.f90:78-97 */
static inline int sen(int N, int l, int w) {
  int v = l + w;
  return v > N ? v - N : v;
}

/* calculate_area: mod_function_library.f90:49-60 */
static inline double carea(double x1, double x2, double x3, double y1, double y2, double y3) {
  return fabs((x1 - x3) * (y2 - y3) - (x2 - x3) * (y1 - y3)) / 2.0;
}

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
/* Fortran MAX/MIN: for x86 flang both return the second operand on ties and NaN
 * handling is that of maxsd/minsd with the operands in source order; all uses here see
 * ordered, finite operands. */

/* ---------------------------------------------------------- tridiagonal_solver ------ */
/* tridiagonal_solver.f90:15-90. a,b,c 1-based of length >= n; rhs/sol (ld, nc). */
int oracle_tridiag(int n, int nc, int ld, const double *a, const double *b, const double *c,
                   const double *rhs, double *sol) {
  double gam[64];
  if (n > 64) return GOMGPU_E_ARG;
  if (b[0] == 0.0) return GOMGPU_ERR_TRIDIAG_PIVOT;
  double den = b[0];
  for (int i = 0; i < nc; i++) sol[0 + i * ld] = rhs[0 + i * ld] / den;
  for (int j = 2; j <= n; j++) {
    gam[j - 1] = c[j - 2] / den;
    den = b[j - 1] - a[j - 1] * gam[j - 1];
    if (den == 0.0) return GOMGPU_ERR_TRIDIAG_PIVOT;
    for (int i = 0; i < nc; i++)
      sol[(j - 1) + i * ld] = (rhs[(j - 1) + i * ld] - a[j - 1] * sol[(j - 2) + i * ld]) / den;
  }
  for (int j = n - 1; j >= 1; j--)
    for (int i = 0; i < nc; i++)
      sol[(j - 1) + i * ld] = sol[(j - 1) + i * ld] - gam[j] * sol[j + i * ld];
  return 0;
}

/* ------------------------------------------------------------------ friction -------- */
/* calculate_bottom_friction.f90:9-73 */
int oracle_friction(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                    const gomgpu_forcing *f) {
  (void)f;
  const int J = p->maxface, L = p->maxlayer;
  for (int j = 1; j <= J; j++) {
    R1(s->Cdb, j) = 0.0;
    R1(s->Gamma_B, j) = 0.0;
  }
  if (p->bf_flag == 1) {
    const double expo = 1.0 / 6.0;
    for (int j = 1; j <= J; j++) {
      if (R1(m->h_face, j) > p->dry_depth) {
        double hr = dmax(1.0, R1(m->h_face, j));
        double chezy = pow(hr, expo) / R1(m->manning, j);
        double rt = p->gravity / (chezy * chezy);
        R1(s->Cdb, j) = rt;
        int k = R1(m->bottom_layer_at_face, j);
        double u = RL(s->un_face, k, j), v = RL(s->vn_face, k, j);
        R1(s->Gamma_B, j) = rt * sqrt(u * u + v * v);
      }
    }
  } else if (p->bf_flag == 2) {
    for (int j = 1; j <= J; j++) {
      if (R1(m->h_face, j) > p->dry_depth) {
        int k = R1(m->bottom_layer_at_face, j);
        double lg = log(dmax(0.5, R0(s->dzhalf_face, k - 1, j)) / R1(m->bf_height, j));
        double rt = p->von_Karman * p->von_Karman / (lg * lg);
        R1(s->Cdb, j) = rt;
        double u = RL(s->un_face, k, j), v = RL(s->vn_face, k, j);
        R1(s->Gamma_B, j) = rt * sqrt(u * u + v * v);
      }
    }
  }
  return 0;
}

/* ---------------------------------------------------------------- ELM pieces -------- */
/* check_intersection.f90:14-50 (small2 = 0, bounds inclusive) */
int oracle_check_intersection(double x1, double x2, double x3, double x4, double y1,
                              double y2, double y3, double y4, double *xin, double *yin) {
  double delta = (x2 - x1) * (y3 - y4) - (y2 - y1) * (x3 - x4);
  double delta1 = (x3 - x1) * (y3 - y4) - (y3 - y1) * (x3 - x4);
  double delta2 = (x2 - x1) * (y3 - y1) - (y2 - y1) * (x3 - x1);
  if (delta != 0.0) {
    double tt1 = delta1 / delta, tt2 = delta2 / delta;
    if (tt1 >= -0.0 && tt1 <= 1.0 && tt2 >= -0.0 && tt2 <= 1.0) {
      *xin = x1 + (x2 - x1) * tt1;
      *yin = y1 + (y2 - y1) * tt1;
      return 1;
    }
  }
  return 0;
}

/* bilinear_interpolation.f90:8-121; returns 0 or GOMGPU_ERR_BILINEAR */
int oracle_bilinear(double x1, double x2, double x3, double x4, double y1, double y2,
                    double y3, double y4, double x, double y, double *xi_out, double *eta_out,
                    double *shapef) {
  const double small3 = (double)1.e-5f;
  const double lim = (double)1.1f;
  double axi[2], aet[2], bxy[2], root_xi[2], root_et[2];
  double x0 = (x1 + x2 + x3 + x4) / 4.0;
  double y0 = (y1 + y2 + y3 + y4) / 4.0;
  double xi = 0.0, eta = 0.0, dd;
  axi[0] = x2 - x1 + x3 - x4;
  axi[1] = y2 - y1 + y3 - y4;
  aet[0] = x3 + x4 - x1 - x2;
  aet[1] = y3 + y4 - y1 - y2;
  bxy[0] = x1 - x2 + x3 - x4;
  bxy[1] = y1 - y2 + y3 - y4;
  double dxi = 2.0 * ((x3 - x4) * (y1 - y2) - (y3 - y4) * (x1 - x2));
  double deta = 2.0 * ((x4 - x1) * (y3 - y2) - (y4 - y1) * (x3 - x2));
  if ((fabs(bxy[0]) < small3 && fabs(bxy[1]) < small3) || (fabs(dxi) < small3 && fabs(deta) < small3)) {
    dd = axi[0] * aet[1] - axi[1] * aet[0];
    if (dd == 0.0) return GOMGPU_ERR_BILINEAR;
    xi = 4.0 * (aet[1] * (x - x0) - aet[0] * (y - y0)) / dd;
    eta = 4.0 * (axi[0] * (y - y0) - axi[1] * (x - x0)) / dd;
  } else if (fabs(dxi) < small3 && fabs(deta) >= small3) {
    eta = 4.0 * (bxy[1] * (x - x0) - bxy[0] * (y - y0)) / deta;
    double t1 = axi[0] + eta * bxy[0], t2 = axi[1] + eta * bxy[1];
    dd = t1 * t1 + t2 * t2;
    if (dd == 0.0) return GOMGPU_ERR_BILINEAR;
    xi = ((4.0 * (x - x0) - eta * aet[0]) * (axi[0] + eta * bxy[0]) +
          (4.0 * (y - y0) - eta * aet[1]) * (axi[1] + eta * bxy[1])) / dd;
  } else if (fabs(dxi) >= small3 && fabs(deta) < small3) {
    xi = 4.0 * (bxy[1] * (x - x0) - bxy[0] * (y - y0)) / dxi;
    double t1 = aet[0] + xi * bxy[0], t2 = aet[1] + xi * bxy[1];
    dd = t1 * t1 + t2 * t2;
    if (dd == 0.0) return GOMGPU_ERR_BILINEAR;
    eta = ((4.0 * (x - x0) - xi * axi[0]) * (aet[0] + xi * bxy[0]) +
           (4.0 * (y - y0) - xi * axi[1]) * (aet[1] + xi * bxy[1])) / dd;
  } else {
    double beta = aet[1] * axi[0] - aet[0] * axi[1] - 4 * (bxy[1] * (x - x0) - bxy[0] * (y - y0));
    double gamma = 4.0 * (aet[0] * (y - y0) - aet[1] * (x - x0));
    double delta = beta * beta - 4 * gamma * dxi;
    if (delta == 0.0) {
      xi = -beta / 2 / dxi;
      eta = (4.0 * (bxy[1] * (x - x0) - bxy[0] * (y - y0)) - xi * dxi) / deta;
    } else if (delta > 0) {
      root_xi[0] = (-beta + sqrt(delta)) / 2 / dxi;
      root_xi[1] = (-beta - sqrt(delta)) / 2 / dxi;
      int icount = 0;
      for (int i = 0; i < 2; i++) {
        root_et[i] = (4 * (bxy[1] * (x - x0) - bxy[0] * (y - y0)) - root_xi[i] * dxi) / deta;
        if (fabs(root_xi[i]) <= lim && fabs(root_et[i]) <= lim) {
          xi = root_xi[i];
          eta = root_et[i];
          icount++;
        }
      }
      if (icount == 2 && fabs(root_xi[0] - root_xi[1]) < (double)1.e-6f) {
        /* keep the second root, as the reference does */
      } else if (icount != 1) {
        return GOMGPU_ERR_BILINEAR;
      }
    } else {
      return GOMGPU_ERR_BILINEAR;
    }
  }
  if (fabs(xi) > lim || fabs(eta) > lim) return GOMGPU_ERR_BILINEAR;
  xi = dmin(1.0, dmax(xi, -1.0));
  eta = dmin(1.0, dmax(eta, -1.0));
  shapef[0] = (1 - xi) * (1 - eta) / 4;
  shapef[1] = (1 + xi) * (1 - eta) / 4;
  shapef[2] = (1 + xi) * (1 + eta) / 4;
  shapef[3] = (1 - xi) * (1 + eta) / 4;
  *xi_out = xi;
  *eta_out = eta;
  return 0;
}

typedef struct {
  const gomgpu_mesh *m;
  const gomgpu_params *p;
  gomgpu_state *s;
  int L;
} elmctx;

/* search_straight_line.f90:15-399 (mode 1). Returns 0 or an error code. */
static int search_line(const elmctx *c, int *nnel, int *jlev, double time, double x0,
                       double y0, double z0, double *xt, double *yt, double *zt, int *nfl,
                       int jface) {
  const gomgpu_mesh *m = c->m;
  const gomgpu_params *p = c->p;
  gomgpu_state *s = c->s;
  const int L = c->L;
  const double small06 = (double)1.e-6f;
  const double MSL = p->MSL;
  int nel = *nnel;
  if (R1(s->top_layer_at_element, nel) == 0 || MSL - R1(m->h_cell, nel) > MSL + R1(s->eta_cell, nel))
    return fail(GOMGPU_ERR_ELM_DRY_START, jface);
  *nfl = 0;
  double trm = time;
  double area1 = 0.0, area2 = 0.0;
  int N = R1(m->tri_or_quad, nel);
  for (int l = 1; l <= N; l++) {
    int n1 = I4(m->nodenum_at_cell, l, nel);
    int n2 = I4(m->nodenum_at_cell, sen(N, l, 1), nel);
    area1 = area1 + fabs(carea(R1(m->x_node, n1), R1(m->x_node, n2), x0, R1(m->y_node, n1), R1(m->y_node, n2), y0));
    area2 = area2 + fabs(carea(R1(m->x_node, n1), R1(m->x_node, n2), *xt, R1(m->y_node, n1), R1(m->y_node, n2), *yt));
  }
  double ae = fabs(area1 - R1(m->area, nel)) / R1(m->area, nel);
  if (ae > small06) return fail(GOMGPU_ERR_ELM_NOT_IN_CELL, jface);
  ae = fabs(area2 - R1(m->area, nel)) / R1(m->area, nel);
  if (ae < small06) {
    *nnel = nel;
  } else {
    double xcg = (1.0 - 1.0e-4) * x0 + 1.0e-4 * R1(m->x_cell, nel);
    double ycg = (1.0 - 1.0e-4) * y0 + 1.0e-4 * R1(m->y_cell, nel);
    double pathl = sqrt((*xt - xcg) * (*xt - xcg) + (*yt - ycg) * (*yt - ycg));
    if ((xcg == *xt && ycg == *yt) || pathl == 0.0) return fail(GOMGPU_ERR_ELM_ZERO_PATH, jface);
    int iflag = 0, nel_j = 0;
    double xin = 0.0, yin = 0.0;
    for (int l = 1; l <= N; l++) {
      int jd1 = I4(m->nodenum_at_cell, sen(N, l, 1), nel);
      int jd2 = I4(m->nodenum_at_cell, sen(N, l, 2), nel);
      iflag = oracle_check_intersection(xcg, *xt, R1(m->x_node, jd1), R1(m->x_node, jd2), ycg, *yt,
                                        R1(m->y_node, jd1), R1(m->y_node, jd2), &xin, &yin);
      if (iflag == 1) {
        nel_j = l;
        break;
      }
    }
    if (iflag == 0) return fail(GOMGPU_ERR_ELM_NO_EDGE, jface);
    double zin = z0;
    int iteration = 0;
    for (;;) {
      iteration++;
      if (iteration > 1000) {
        *nfl = 1;
        *xt = xin;
        *yt = yin;
        *zt = zin;
        *nnel = nel;
        break;
      }
      N = R1(m->tri_or_quad, nel);
      int md1 = I4(m->nodenum_at_cell, sen(N, nel_j, 1), nel);
      int md2 = I4(m->nodenum_at_cell, sen(N, nel_j, 2), nel);
      double dist = sqrt((xin - *xt) * (xin - *xt) + (yin - *yt) * (yin - *yt));
      if (dist / pathl > 1.0 + 1.0e-4) return fail(GOMGPU_ERR_ELM_OVERSHOT, jface);
      zin = *zt - dist / pathl * (*zt - zin);
      trm = trm * dist / pathl;
      pathl = sqrt((xin - *xt) * (xin - *xt) + (yin - *yt) * (yin - *yt));
      if (pathl == 0.0 || trm == 0.0) return fail(GOMGPU_ERR_ELM_TARGET, jface);
      int adj = I4(m->adj_cellnum_at_cell, nel_j, nel);
      if (adj == 0 || R1(s->top_layer_at_element, adj) == 0) {
        int isd = I4(m->facenum_at_cell, nel_j, nel);
        if (I2(m->nodenum_at_face, 1, isd) + I2(m->nodenum_at_face, 2, isd) != md1 + md2)
          return fail(GOMGPU_ERR_ELM_WRONG_SIDE, jface);
        xin = (1.0 - 1.0e-4) * xin + 1.0e-4 * R1(m->x_cell, nel);
        yin = (1.0 - 1.0e-4) * yin + 1.0e-4 * R1(m->y_cell, nel);
        /* The reference computes a sliding target from the face tangential velocity here
         * (search_straight_line.f90:251-257) and then discards it: the walk ends at the
         * nudged boundary point. */
        *nfl = 1;
        *xt = xin;
        *yt = yin;
        *zt = zin;
        *nnel = nel;
        break;
      }
      nel = adj;
      N = R1(m->tri_or_quad, nel);
      area1 = 0.0;
      for (int l = 1; l <= N; l++) {
        int k1 = I4(m->nodenum_at_cell, l, nel);
        int k2 = I4(m->nodenum_at_cell, sen(N, l, 1), nel);
        area1 = area1 + fabs(carea(R1(m->x_node, k1), R1(m->x_node, k2), *xt, R1(m->y_node, k1), R1(m->y_node, k2), *yt));
      }
      ae = fabs(area1 - R1(m->area, nel)) / R1(m->area, nel);
      if (ae < small06) {
        *nnel = nel;
        break;
      }
      int found = 0;
      for (int l = 1; l <= N; l++) {
        int jd1 = I4(m->nodenum_at_cell, sen(N, l, 1), nel);
        int jd2 = I4(m->nodenum_at_cell, sen(N, l, 2), nel);
        if ((jd1 == md1 && jd2 == md2) || (jd2 == md1 && jd1 == md2)) continue;
        iflag = oracle_check_intersection(xcg, *xt, R1(m->x_node, jd1), R1(m->x_node, jd2), ycg, *yt,
                                          R1(m->y_node, jd1), R1(m->y_node, jd2), &xin, &yin);
        if (iflag == 1) {
          nel_j = l;
          found = 1;
          break;
        }
      }
      if (!found) return fail(GOMGPU_ERR_ELM_NO_EDGE, jface);
    }
  }
  nel = *nnel;
  if (R1(s->top_layer_at_element, nel) == 0 || MSL - R1(m->h_cell, nel) > MSL + R1(s->eta_cell, nel))
    return fail(GOMGPU_ERR_ELM_DRY_END, jface);
  double z = dmin(dmax(*zt, MSL - R1(m->h_cell, nel)), MSL + R1(s->eta_cell, nel));
  *zt = z;
  if (z <= ZL(0) || z > ZL(L)) return fail(GOMGPU_ERR_ELM_VERTICAL, jface);
  for (int k = 0; k <= L - 1; k++) {
    if (z > ZL(k) && z <= ZL(k + 1)) {
      *jlev = k + 1;
      break;
    }
  }
  int jl = *jlev;
  int bl = R1(m->bottom_layer_at_element, nel), tl = R1(s->top_layer_at_element, nel);
  jl = jl < bl ? bl : jl;
  jl = jl > tl ? tl : jl;
  *jlev = jl;
  return 0;
}

/* Departure bookkeeping (no counterpart in the reference, where nnel and jlev are locals of
 * solve_nonlinear_advection.f90:92-202): points within 1e-12 of an edge of their cell,
 * counted per ELM ([0] momentum, [1] tracer), and the momentum ELM's departure cell
 * (negated for such a point) and level per face-layer, [k-1][j-1], when recording. */
static long long g_near[2];
static int32_t *g_dep;

void oracle_elm_record(int32_t *dep) { g_dep = dep; }
long long oracle_elm_near_edge(int which) { return which == 0 || which == 1 ? g_near[which] : -1; }
void oracle_elm_near_reset(void) { g_near[0] = g_near[1] = 0; }

/* Whether (px, py) is within 1e-12 of an edge of cell nel relative to the edge's length:
 * |(B - A) x (P - A)| <= 1e-12 |B - A|^2 for an edge A -> B, in node-slot order. */
static int near_edge(const gomgpu_mesh *m, int nel, double px, double py) {
  const int N = R1(m->tri_or_quad, nel);
  for (int l = 1; l <= N; l++) {
    int na = I4(m->nodenum_at_cell, l, nel), nb = I4(m->nodenum_at_cell, l < N ? l + 1 : 1, nel);
    double ax = R1(m->x_node, na), ay = R1(m->y_node, na);
    double ex = R1(m->x_node, nb) - ax, ey = R1(m->y_node, nb) - ay;
    double cr = ex * (py - ay) - ey * (px - ax);
    if (fabs(cr) <= 1.0e-12 * (ex * ex + ey * ey)) return 1;
  }
  return 0;
}

/* ELM_backtrace_v1.f90:15-200 (i_which_backtrack = 1) */
/* ELM_backtrace_v1.f90:15-214. With foot non-NULL, the foot point and its interpolation
 * weights are left there: xt, yt, zt, zup, staint[4]. *near is set when the end point of a
 * sub-step is within 1e-12 of an edge of its cell (near_edge). */
static int backtrace(const elmctx *c, int jface, int bt_step, double bt_dt, double *uu,
                     double *vv, double *ww, double x0, double y0, double z0, int *nnel,
                     int *jlev, double *foot, int *near) {
  const gomgpu_mesh *m = c->m;
  const gomgpu_params *p = c->p;
  gomgpu_state *s = c->s;
  const int L = c->L;
  double vxl[4][2], vyl[4][2], vzl[4][2], vxn[4], vyn[4], vzn[4], staint[4] = {0, 0, 0, 0};
  double xt = 0, yt = 0, zt = 0;
  for (int idt = 1; idt <= bt_step; idt++) {
    xt = x0 - bt_dt * *uu;
    yt = y0 - bt_dt * *vv;
    zt = z0 - bt_dt * *ww;
    int iflqs1 = 0;
    int err = search_line(c, nnel, jlev, bt_dt, x0, y0, z0, &xt, &yt, &zt, &iflqs1, jface);
    if (err) return err;
    if (near_edge(m, *nnel, xt, yt)) *near = 1;
    int nel = *nnel, N = R1(m->tri_or_quad, nel), jl = *jlev;
    for (int l = 1; l <= N; l++) {
      int nd = I4(m->nodenum_at_cell, l, nel);
      for (int k2 = 1; k2 <= 2; k2++) {
        int lev = jl + k2 - 2;
        vxl[l - 1][k2 - 1] = R0(s->u_node, lev, nd);
        vyl[l - 1][k2 - 1] = R0(s->v_node, lev, nd);
        vzl[l - 1][k2 - 1] = R0(s->w_node, lev, nd);
      }
    }
    double zup = (jl == R1(s->top_layer_at_element, nel)) ? p->MSL + R1(s->eta_cell, nel) : ZL(jl);
    if (zt > zup) return fail(GOMGPU_ERR_ELM_VERTICAL, jface);
    double zrat = (zup - zt) / RL(s->dz_cell, jl, nel);
    for (int l = 0; l < N; l++) {
      vxn[l] = vxl[l][1] * (1 - zrat) + vxl[l][0] * zrat;
      vyn[l] = vyl[l][1] * (1 - zrat) + vyl[l][0] * zrat;
      vzn[l] = vzl[l][1] * (1 - zrat) + vzl[l][0] * zrat;
    }
    int n1 = I4(m->nodenum_at_cell, 1, nel), n2 = I4(m->nodenum_at_cell, 2, nel), n3 = I4(m->nodenum_at_cell, 3, nel);
    double X1 = R1(m->x_node, n1), X2 = R1(m->x_node, n2), X3 = R1(m->x_node, n3);
    double Y1 = R1(m->y_node, n1), Y2 = R1(m->y_node, n2), Y3 = R1(m->y_node, n3);
    double ar = R1(m->area, nel);
    if (N == 3) {
      staint[0] = carea(xt, X2, X3, yt, Y2, Y3) / ar;
      staint[1] = carea(X1, xt, X3, Y1, yt, Y3) / ar;
      staint[2] = carea(X1, X2, xt, Y1, Y2, yt) / ar;
      staint[0] = dmax(0.0, dmin(1.0, staint[0]));
      staint[1] = dmax(0.0, dmin(1.0, staint[1]));
      if (staint[0] + staint[1] > 1) {
        staint[2] = 0;
        staint[1] = 1.0 - staint[0];
      } else {
        staint[2] = 1.0 - staint[0] - staint[1];
      }
    } else {
      int n4 = I4(m->nodenum_at_cell, 4, nel);
      double X4 = R1(m->x_node, n4), Y4 = R1(m->y_node, n4);
      double aa1 = carea(X1, X2, xt, Y1, Y2, yt);
      double aa2 = carea(X2, X3, xt, Y2, Y3, yt);
      double aa3 = carea(X3, X4, xt, Y3, Y4, yt);
      double aa4 = carea(X4, X1, xt, Y4, Y1, yt);
      double aa = fabs(aa1) + fabs(aa2) + fabs(aa3) + fabs(aa4);
      if (fabs(aa - ar) / ar > (double)1.e-6f) return fail(GOMGPU_ERR_ELM_OUTSIDE_QUAD, jface);
      double csi, etta;
      int err2 = oracle_bilinear(X1, X2, X3, X4, Y1, Y2, Y3, Y4, xt, yt, &csi, &etta, staint);
      if (err2) return fail(err2, jface);
    }
    *uu = 0.0;
    *vv = 0.0;
    *ww = 0.0;
    for (int l = 0; l < N; l++) {
      *uu = *uu + vxn[l] * staint[l];
      *vv = *vv + vyn[l] * staint[l];
      *ww = *ww + vzn[l] * staint[l];
    }
    if (foot) {
      foot[0] = xt;
      foot[1] = yt;
      foot[2] = zt;
      foot[3] = zup;
      for (int l = 0; l < 4; l++) foot[4 + l] = staint[l];
    }
    if (iflqs1 == 1) break;
    x0 = xt;
    y0 = yt;
    z0 = zt;
  }
  return 0;
}

/* solve_nonlinear_advection.f90:14-204 (and main.f90:167-176 when advection is off) */
int oracle_elm(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
               const gomgpu_forcing *f) {
  (void)f;
  const int P = p->maxnod, J = p->maxface, L = p->maxlayer;
  if (p->advection_flag == 0) {
    for (int j = 1; j <= J; j++)
      for (int k = 1; k <= L; k++) {
        RL(s->un_ELM, k, j) = RL(s->un_face, k, j);
        RL(s->vn_ELM, k, j) = RL(s->vn_face, k, j);
      }
    return 0;
  }
  for (size_t i = 0; i < (size_t)L * J; i++) {
    s->un_ELM[i] = 0.0;
    s->vn_ELM[i] = 0.0;
  }
  if (g_dep) memset(g_dep, 0, 2 * (size_t)L * J * sizeof(int32_t));
  const double dt = p->dt;
  if (p->ELM_backtrace_flag == 1) {
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= P; i++) {
      int nadj = R1(m->adj_cells_at_node, i);
      for (int k = 1; k <= L; k++) {
        double summ = 0.0;
        for (int j = 1; j <= nadj; j++) {
          int NB = p->max_no_neighbor_node;
          int ie = INB(m->adj_cellnum_at_node, j, i);
          int id = INB(m->node_count_each_element, j, i);
          int N = R1(m->tri_or_quad, ie);
          double devm = 0.0;
          for (int l = 1; l <= 2; l++) {
            int nd = (l == 1) ? I4(m->nodenum_at_cell, sen(N, id, 1), ie)
                              : I4(m->nodenum_at_cell, sen(N, id, N - 1), ie);
            double dx = R1(m->x_node, i) - R1(m->x_node, nd), dy = R1(m->y_node, i) - R1(m->y_node, nd);
            double rl = sqrt(dx * dx + dy * dy);
            double uvel1 = (R0(s->u_node, k, i) + R0(s->u_node, k - 1, i)) * 0.5;
            double uvel2 = (R0(s->u_node, k, nd) + R0(s->u_node, k - 1, nd)) * 0.5;
            double vvel1 = (R0(s->v_node, k, i) + R0(s->v_node, k - 1, i)) * 0.5;
            double vvel2 = (R0(s->v_node, k, nd) + R0(s->v_node, k - 1, nd)) * 0.5;
            double du = uvel1 - uvel2, dv = vvel1 - vvel2;
            double dev = sqrt(du * du + dv * dv) / rl * dt;
            if (dev > devm) devm = dev;
          }
          double vw = (R0(s->w_node, k, i) + R0(s->w_node, k - 1, i)) * 0.5;
          int iw = (int)(2 * vw * dt / p->delta_z_min);
          int i1 = (int)(devm / (double)1.e-1f);
          int idelta = i1 > iw ? i1 : iw;
          summ = summ + (double)idelta / nadj;
        }
        int ns = (int)summ;
        ns = ns < p->ELM_max_iter ? ns : p->ELM_max_iter;
        ns = ns > p->ELM_min_iter ? ns : p->ELM_min_iter;
        RL(s->num_sub_elm_iteration, k, i) = ns;
      }
    }
  }
  elmctx c = {m, p, s, L};
  const double vmin = (double)1.e-4f;
  loop_err e = {0, 0};
#pragma omp parallel for schedule(dynamic, 64)
  for (int j = 1; j <= J; j++) {
    int n1 = I2(m->nodenum_at_face, 1, j), n2 = I2(m->nodenum_at_face, 2, j);
    int kb = R1(m->bottom_layer_at_face, j), kt = R1(s->top_layer_at_face, j);
    double atod = dmax(R1(m->h_node, n1), R1(m->h_node, n2));
    if (fabs(atod) < p->adv_onoff_depth) {
      for (int k = kb; k <= kt; k++) {
        RL(s->un_ELM, k, j) = RL(s->un_face, k, j);
        RL(s->vn_ELM, k, j) = RL(s->vn_face, k, j);
      }
      continue;
    }
    int a1 = I2(m->adj_cellnum_at_face, 1, j), a2 = I2(m->adj_cellnum_at_face, 2, j);
    int ie0;
    if (R1(s->top_layer_at_element, a1) != 0) {
      ie0 = a1;
    } else if (a2 != 0 && R1(s->top_layer_at_element, a2) != 0) {
      ie0 = a2;
    } else {
      for (int k = kb; k <= kt; k++) {
        RL(s->un_ELM, k, j) = RL(s->un_face, k, j);
        RL(s->vn_ELM, k, j) = RL(s->vn_face, k, j);
      }
      continue;
    }
    double ct = R1(m->cos_theta, j), st = R1(m->sin_theta, j);
    for (int k = kb; k <= kt; k++) {
      int nnel = ie0, jlev = k;
      double x0 = R1(m->x_face, j), y0 = R1(m->y_face, j), z0;
      if (k == R1(s->top_layer_at_element, nnel))
        z0 = p->MSL + R1(s->eta_cell, nnel) - RL(s->dz_cell, k, nnel) * 0.5;
      else
        z0 = ZL(k) - RL(s->dz_cell, k, nnel) * 0.5;
      double un = RL(s->un_face, k, j), vn = RL(s->vn_face, k, j);
      double uu = un * ct - vn * st;
      double vv = un * st + vn * ct;
      double wdown = (R0(s->w_node, k - 1, n1) + R0(s->w_node, k - 1, n2)) * 0.5;
      double wup = (R0(s->w_node, k, n1) + R0(s->w_node, k, n2)) * 0.5;
      double ww = (wdown + wup) * 0.5;
      double vmag = sqrt(uu * uu + vv * vv + ww * ww);
      if (vmag <= vmin) {
        RL(s->un_ELM, k, j) = un;
        RL(s->vn_ELM, k, j) = vn;
      } else {
        int bt_step = (int)((RL(s->num_sub_elm_iteration, k, n1) + RL(s->num_sub_elm_iteration, k, n2)) * 0.5);
        double bt_dt = dt / bt_step;
        int near = 0;
        int err = backtrace(&c, j, bt_step, bt_dt, &uu, &vv, &ww, x0, y0, z0, &nnel, &jlev, NULL, &near);
        if (err) {
          loop_fail(&e, err, j);
          break;
        }
        if (near) {
#pragma omp atomic
          g_near[0]++;
        }
        if (g_dep) {
          g_dep[(size_t)(k - 1) * J + (j - 1)] = near ? -nnel : nnel;
          g_dep[(size_t)L * J + (size_t)(k - 1) * J + (j - 1)] = jlev;
        }
        RL(s->un_ELM, k, j) = uu * ct + vv * st;
        RL(s->vn_ELM, k, j) = -uu * st + vv * ct;
      }
    }
  }
  if (e.code) return fail(e.code, e.entity);
  return 0;
}

/* ------------------------------------------------------------------ momentum -------- */
/* solve_momentum_equation.f90:15-671 (no equilibrium tide, no WR points) */
int oracle_momentum(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                    const gomgpu_forcing *f) {
  const int J = p->maxface, L = p->maxlayer;
  const double dt = p->dt, g = p->gravity, th = p->theta;
  for (size_t i = 0; i < (size_t)L * J; i++) {
    s->AinvDeltaZ1[i] = 0.0;
    s->AinvDeltaZ2[i] = 0.0;
    s->AinvG1[i] = 0.0;
    s->AinvG2[i] = 0.0;
  }
  loop_err ferr = {0, 0};
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= J; j++) {
    int kt = R1(s->top_layer_at_face, j);
    if (kt == 0) continue;
    double a[64] = {0}, b[64] = {0}, cc[64] = {0}, rhs[3 * 64], sol[3 * 64], G1[64], G2[64], temp[13];
    int kb = R1(m->bottom_layer_at_face, j);
    int node1 = I2(m->nodenum_at_face, 1, j), node2 = I2(m->nodenum_at_face, 2, j);
    int n = kt - kb + 1;
    int a1 = I2(m->adj_cellnum_at_face, 1, j), a2 = I2(m->adj_cellnum_at_face, 2, j);
    int bt = R1(m->boundary_type_of_face, j);
    double dj = R1(m->delta_j, j), Lf = R1(m->face_length, j);
    double ct = R1(m->cos_theta, j), st = R1(m->sin_theta, j);
    double ct2 = R1(m->cos_theta2, j), st2 = R1(m->sin_theta2, j);
    for (int k = kb + 1; k <= kt - 1; k++) {
      int kk = kt - k + 1;
      a[kk] = -(dt * R0(s->Av, k, j) / R0(s->dzhalf_face, k, j));
      cc[kk] = -(dt * R0(s->Av, k - 1, j) / R0(s->dzhalf_face, k - 1, j));
      b[kk] = -a[kk] + RL(s->dz_face, k, j) - cc[kk];
    }
    if (kt == kb) {
      b[1] = RL(s->dz_face, kt, j) + R1(s->Gamma_B, j) * dt;
    } else {
      cc[1] = -(dt * R0(s->Av, kt - 1, j) / R0(s->dzhalf_face, kt - 1, j));
      b[1] = R1(f->Gamma_T, j) * dt + RL(s->dz_face, kt, j) - cc[1];
      a[n] = -(dt * R0(s->Av, kb, j) / R0(s->dzhalf_face, kb, j));
      b[n] = -a[n] + RL(s->dz_face, kb, j) + R1(s->Gamma_B, j) * dt;
    }
    double rtemp0 = 0.0, rtemp1 = 0.0, rtemp2 = 0.0, rtemp3 = 0.0, rtemp4 = 0.0;
    for (int k = 1; k <= n; k++) {
      int kk = kt - k + 1;
      double dz = RL(s->dz_face, kk, j);
      double unf = RL(s->un_face, kk, j), vnf = RL(s->vn_face, kk, j);
      double cor = R1(m->coriolis_factor, j);
      G1[k] = dz * (RL(s->un_ELM, kk, j) + cor * dt * vnf);
      G2[k] = dz * (RL(s->vn_ELM, kk, j) - cor * dt * unf);
      if (a2 != 0) {
        for (int l = 1; l <= 2; l++) {
          int nd = I2(m->nodenum_at_face, l, j);
          double ut = (R0(s->u_node, kk, nd) + R0(s->u_node, kk - 1, nd)) * 0.5;
          double vt = (R0(s->v_node, kk, nd) + R0(s->v_node, kk - 1, nd)) * 0.5;
          temp[2 * l - 1] = ut * ct + vt * st;
          temp[2 * l] = -ut * st + vt * ct;
        }
        for (int l = 1; l <= 2; l++) {
          int ie = I2(m->adj_cellnum_at_face, l, j);
          int N = R1(m->tri_or_quad, ie);
          double ut = 0.0, vt = 0.0;
          for (int ll = 1; ll <= N; ll++) {
            int nd = I4(m->nodenum_at_cell, ll, ie);
            ut = ut + (R0(s->u_node, kk, nd) + R0(s->u_node, kk - 1, nd)) * 0.5 / N;
            vt = vt + (R0(s->v_node, kk, nd) + R0(s->v_node, kk - 1, nd)) * 0.5 / N;
          }
          temp[2 * l + 3] = ut * ct + vt * st;
          temp[2 * l + 4] = -ut * st + vt * ct;
        }
        double dudx = (temp[7] - temp[5]) / dj;
        double dvdx = (temp[8] - temp[6]) / dj;
        double dudy = (temp[3] - temp[1]) / Lf;
        double dvdy = (temp[4] - temp[2]) / Lf;
        double sx = dvdx + dudy;
        double hdiff = p->smagorinsky_parameter * sqrt(dudx * dudx + dvdy * dvdy + 0.5 * (sx * sx));
        RL(s->Kh, k, j) = hdiff * R1(m->area, a1);
        hdiff = hdiff + p->Ah_0;
        RL(s->Kh, k, j) = RL(s->Kh, k, j) + p->Kh_0;
        double d2udy = 4.0 * (temp[3] + temp[1] - 2 * unf) / (Lf * Lf);
        double d2vdy = 4.0 * (temp[4] + temp[2] - 2 * vnf) / (Lf * Lf);
        for (int l = 1; l <= 2; l++) {
          int ie = I2(m->adj_cellnum_at_face, l, j);
          double ddx = R1(m->x_cell, ie) - R1(m->x_face, j), ddy = R1(m->y_cell, ie) - R1(m->y_face, j);
          double rl = sqrt(ddx * ddx + ddy * ddy);
          temp[2 * l + 7] = (temp[2 * l + 3] - unf) / rl;
          temp[2 * l + 8] = (temp[2 * l + 4] - vnf) / rl;
        }
        double d2udx = 2.0 * (temp[9] + temp[11]) / dj;
        double d2vdx = 2.0 * (temp[10] + temp[12]) / dj;
        G1[k] = G1[k] + dz * dt * hdiff * (d2udx + d2udy);
        G2[k] = G2[k] + dz * dt * hdiff * (d2vdx + d2vdy);
      }
      if (bt == 0) {
        rtemp0 = g * dt / dj * (1.0 - th) * (R1(s->eta_cell, a2) - R1(s->eta_cell, a1));
        rtemp1 = rtemp0 * ct2;
        rtemp2 = rtemp0 * st2;
        G1[k] = G1[k] - dz * rtemp1;
      } else if (bt > 0) {
        int ie = R1(m->ob_element_flag, a1);
        rtemp1 = g * dt / dj * (1.0 - th) * (R1(f->eta_at_ob_old, ie) - R1(s->eta_cell, a1));
        G1[k] = G1[k] - dz * rtemp1;
      } else if (bt == -1) {
        G1[k] = 0.0;
      }
      int wet_nodes = (R1(s->wetdry_node, node1) == 0 && R1(s->wetdry_node, node2) == 0);
      if (wet_nodes) {
        rtemp0 = g * dt / Lf * (1.0 - th) * (R1(s->eta_node, node2) - R1(s->eta_node, node1));
        if (bt == 0)
          G2[k] = G2[k] - dz * (rtemp0 + rtemp2);
        else
          G2[k] = G2[k] - dz * rtemp0;
      } else if (a2 == 0) {
        rtemp0 = 0.0; /* defined value for the reference's stale private variable */
      }
      if (p->baroclinic_flag == 1) {
        double sum1 = 0.0, sum2 = 0.0;
        for (int k2 = kk; k2 <= kt; k2++) {
          double fct = (k2 == kk) ? 0.5 : 1.0;
          if (a2 != 0 && k2 >= R1(m->bottom_layer_at_element, a1) && k2 <= R1(s->top_layer_at_element, a1) &&
              k2 >= R1(m->bottom_layer_at_element, a2) && k2 <= R1(s->top_layer_at_element, a2)) {
            sum1 = sum1 + fct * (RL(s->rho_cell, k2, a2) - RL(s->rho_cell, k2, a1)) * RL(s->dz_face, k2, j);
          }
          if (k2 >= R1(m->bottom_layer_at_node, node1) && k2 <= R1(s->top_layer_at_node, node1) &&
              k2 >= R1(m->bottom_layer_at_node, node2) && k2 <= R1(s->top_layer_at_node, node2)) {
            sum2 = sum2 + fct * (RL(s->rho_node, k2, node2) - RL(s->rho_node, k2, node1)) * RL(s->dz_face, k2, j);
          }
        }
        rtemp1 = RL(s->dz_face, k, j) * (-(g * dt / (p->rho_o * dj))) * sum1;
        rtemp2 = RL(s->dz_face, k, j) * (-(g * dt / (p->rho_o * Lf))) * sum2;
        rtemp3 = rtemp1 * ct2;
        rtemp4 = rtemp1 * st2;
        G1[k] = G1[k] + rtemp3 * f->spinup_function_baroclinic;
        G2[k] = G2[k] + (rtemp2 + rtemp4) * f->spinup_function_baroclinic;
      }
      if (k == 1) {
        G1[k] = G1[k] + dt * R1(f->wind_stress_normal, j);
        G2[k] = G2[k] + dt * R1(f->wind_stress_tangnt, j);
      }
      /* equilibrium tide term (detp_dx = detp_dy = 0 without tidal potential species) */
      G1[k] = G1[k] + 0.0;
      G2[k] = G2[k] + 0.0;
      if (a2 != 0) {
        double tp[3];
        for (int l = 1; l <= 2; l++) {
          int ie = I2(m->adj_cellnum_at_face, l, j);
          int N = R1(m->tri_or_quad, ie);
          tp[l] = 0.0;
          for (int ll = 1; ll <= N; ll++) tp[l] = tp[l] + R1(f->air_p1, I4(m->nodenum_at_cell, ll, ie)) / N;
        }
        rtemp0 = -(dt / p->rho_o * (tp[2] - tp[1]) / dj);
      }
      rtemp1 = rtemp0 * ct2;
      rtemp2 = rtemp0 * st2;
      G1[k] = G1[k] + dz * rtemp1;
      rtemp3 = -(dt / p->rho_o * ((R1(f->air_p1, node2) - R1(f->air_p1, node1)) / Lf));
      G2[k] = G2[k] + dz * (rtemp3 + rtemp2);
      if (bt == -1) G1[k] = 0.0;
      if (R1(m->isflowside3, j) > 0) G1[k] = G1[k] + dz * R1(f->Qu_boundary, R1(m->isflowside3, j));
      rhs[(k - 1) + 0 * 64] = G1[k];
      rhs[(k - 1) + 1 * 64] = G2[k];
      rhs[(k - 1) + 2 * 64] = dz;
    }
    int err = oracle_tridiag(n, 3, 64, a + 1, b + 1, cc + 1, rhs, sol);
    if (err) {
      loop_fail(&ferr, err, j);
      continue;
    }
    for (int k = 1; k <= n; k++) {
      RL(s->AinvG1, k, j) = sol[(k - 1)];
      RL(s->AinvG2, k, j) = sol[(k - 1) + 64];
    }
    if (bt == -1 && R1(m->isflowside3, j) == 0)
      for (int k = 1; k <= n; k++) RL(s->AinvG1, k, j) = 0.0;
    for (int k = 1; k <= n; k++) {
      RL(s->AinvDeltaZ1, k, j) = sol[(k - 1) + 128];
      RL(s->AinvDeltaZ2, k, j) = sol[(k - 1) + 128];
    }
    if (bt == -1)
      for (int k = 1; k <= n; k++) RL(s->AinvDeltaZ1, k, j) = 0.0;
  }
  if (ferr.code) return fail(ferr.code, ferr.entity);
  return 0;
}

/* ---------------------------------------------------------------------- PCG ------- */
/* pre_conj_grad.f90:9-230, serial order. */
int oracle_pcg(int n, const double *A, const int32_t *col, const int32_t *rowstr,
               const double *b, double *x, int max_iter, double tol) {
  double *q = calloc((size_t)n, sizeof(double)), *r = calloc((size_t)n, sizeof(double));
  double *d = calloc((size_t)n, sizeof(double)), *Minv = calloc((size_t)n, sizeof(double));
  double *sv = calloc((size_t)n, sizeof(double));
  for (int i = 1; i <= n; i++) {
    double sum1 = 0.0;
    for (int k = rowstr[i - 1]; k <= rowstr[i] - 1; k++) {
      if (col[k - 1] == i) Minv[i - 1] = 1.0 / A[k - 1];
      sum1 = sum1 + A[k - 1] * x[col[k - 1] - 1];
    }
    r[i - 1] = b[i - 1] - sum1;
  }
  for (int i = 0; i < n; i++) d[i] = 0.0 + Minv[i] * r[i];
  double delta_new = 0.0;
  for (int i = 0; i < n; i++) delta_new = delta_new + r[i] * d[i];
  double delta_0 = delta_new;
  int cgit2 = 0;
  for (int cgit = 1; cgit <= max_iter; cgit++) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= n; j++) {
      double sum1 = 0.0;
      for (int k = rowstr[j - 1]; k <= rowstr[j] - 1; k++) sum1 = sum1 + A[k - 1] * d[col[k - 1] - 1];
      q[j - 1] = sum1;
    }
    double dq = 0.0;
    for (int i = 0; i < n; i++) dq = dq + d[i] * q[i];
    double alpha = delta_new / dq;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      x[i] = x[i] + alpha * d[i];
      r[i] = r[i] - alpha * q[i];
      sv[i] = 0.0 + Minv[i] * r[i];
    }
    double delta_old = delta_new;
    delta_new = 0.0;
    for (int i = 0; i < n; i++) delta_new = delta_new + r[i] * sv[i];
    double beta = delta_new / delta_old;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) d[i] = sv[i] + beta * d[i];
    if (sqrt(delta_new * delta_new / (delta_0 * delta_0)) < tol) {
      cgit2 = cgit;
      break;
    }
  }
  free(q);
  free(r);
  free(d);
  free(Minv);
  free(sv);
  return cgit2;
}

/* -------------------------------------------------------------- free surface ------- */
/* solve_free_surface_equation.f90:17-436 (no sponge layer, no isflowside2) */
int oracle_freesurf(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                    const gomgpu_forcing *f) {
  const int E = p->maxele, P = p->maxnod, L = p->maxlayer, NB = p->max_no_neighbor_node;
  const double dt = p->dt, g = p->gravity, th = p->theta;
  double *sparsem = calloc((size_t)5 * E, sizeof(double));
  double *rhs1 = calloc((size_t)E, sizeof(double));
  double *a1 = calloc((size_t)5 * E, sizeof(double));
  int32_t *ja = calloc((size_t)5 * E, sizeof(int32_t));
  int32_t *ia = calloc((size_t)E + 1, sizeof(int32_t));
  double *x = calloc((size_t)E, sizeof(double));
  int ret = 0;
#define SPM(l, i) sparsem[(l) + (size_t)((i) - 1) * 5]
  loop_err aerr = {0, 0};
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= E; i++) {
    int N = R1(m->tri_or_quad, i);
    SPM(0, i) = R1(m->area, i);
    for (int l = 1; l <= N; l++) {
      int ii = I4(m->adj_cellnum_at_cell, l, i);
      int j = I4(m->facenum_at_cell, l, i);
      int kt = R1(s->top_layer_at_face, j);
      if (kt != 0) {
        int n = kt - R1(m->bottom_layer_at_face, j) + 1;
        double c = 0.0;
        for (int k = 1; k <= n; k++) c = c + RL(s->dz_face, kt + 1 - k, j) * RL(s->AinvDeltaZ1, k, j);
        c = g * (dt * dt) * (th * th) * R1(m->face_length, j) / R1(m->delta_j, j) * c;
        c = c * R1(m->cos_theta2, j);
        if (c < 0.0) {
          loop_fail(&aerr, GOMGPU_ERR_NOT_SPD, i);
          break;
        }
        SPM(0, i) = SPM(0, i) + c;
        if (ii != 0) SPM(l, i) = -c;
      }
    }
    rhs1[i - 1] = R1(m->area, i) * R1(s->eta_cell, i);
    for (int l = 1; l <= N; l++) {
      int j = I4(m->facenum_at_cell, l, i);
      int kt = R1(s->top_layer_at_face, j);
      if (kt != 0) {
        int n = kt - R1(m->bottom_layer_at_face, j) + 1;
        double c1 = 0.0, c2 = 0.0;
        for (int k = 1; k <= n; k++) {
          int kk = kt + 1 - k;
          c1 = c1 + RL(s->dz_face, kk, j) * RL(s->un_face, kk, j);
          c2 = c2 + RL(s->dz_face, kk, j) * RL(s->AinvG1, k, j);
        }
        double sg = (double)I4(m->sign_in_outflow, l, i), Lf = R1(m->face_length, j);
        rhs1[i - 1] = rhs1[i - 1] - (1.0 - th) * dt * sg * Lf * c1 - th * dt * sg * Lf * c2;
      }
    }
    int obf = R1(m->ob_element_flag, i);
    if (obf > 0) {
      for (int l = 1; l <= N; l++) {
        int j = I4(m->facenum_at_cell, l, i);
        int kt = R1(s->top_layer_at_face, j);
        if (R1(m->boundary_type_of_face, j) > 0 && kt != 0) {
          int n = kt - R1(m->bottom_layer_at_face, j) + 1;
          double c = 0.0;
          for (int k = 1; k <= n; k++) c = c + RL(s->dz_face, kt + 1 - k, j) * RL(s->AinvDeltaZ1, k, j);
          c = g * (dt * dt) * (th * th) * R1(m->face_length, j) / R1(m->delta_j, j) * c;
          rhs1[i - 1] = rhs1[i - 1] + c * R1(f->eta_at_ob_new, obf);
        }
      }
    }
  }
  if (aerr.code) ret = fail(aerr.code, aerr.entity);
  if (!ret) {
    int nz = 0;
    for (int i = 1; i <= E; i++) {
      nz++;
      ia[i - 1] = nz;
      ja[nz - 1] = i;
      a1[nz - 1] = SPM(0, i);
      x[i - 1] = R1(s->eta_cell_new, i);
      int N = R1(m->tri_or_quad, i);
      for (int l = 1; l <= N; l++) {
        int ii = I4(m->adj_cellnum_at_cell, l, i);
        if (ii != 0) {
          nz++;
          ja[nz - 1] = ii;
          a1[nz - 1] = SPM(l, i);
        }
      }
    }
    ia[E] = nz + 1;
    g_last_iters = oracle_pcg(E, a1, ja, ia, rhs1, x, p->max_iteration_pcg, p->error_tolerance);
    for (int i = 1; i <= E; i++) R1(s->eta_cell_new, i) = x[i - 1];
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= P; i++) {
      R1(s->wetdry_node, i) = 0;
      double sum1 = 0.0, sum0 = 0.0;
      for (int jj = 1; jj <= R1(m->adj_cells_at_node, i); jj++) {
        int ie = INB(m->adj_cellnum_at_node, jj, i);
        if (R1(s->top_layer_at_element, ie) == 0) {
          R1(s->wetdry_node, i) = 1;
        } else {
          sum1 = sum1 + R1(m->area, ie) * R1(s->eta_cell_new, ie);
          sum0 = sum0 + R1(m->area, ie);
        }
      }
      if (R1(s->wetdry_node, i) == 0) R1(s->eta_node, i) = sum1 / sum0;
    }
    /* reads the wet nodes, writes the dry ones */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= P; i++) {
      if (R1(s->wetdry_node, i) == 1) {
        double sum0 = 0.0;
        int icount = 0;
        for (int jj = 1; jj <= R1(m->adj_nodes_at_node, i); jj++) {
          int nd = INB0(m->adj_nodenum_at_node, jj, i);
          if (R1(s->wetdry_node, nd) == 0) {
            icount++;
            sum0 = sum0 + R1(s->eta_node, nd);
          }
        }
        if (icount != 0) R1(s->eta_node, i) = sum0 / icount;
      }
    }
  }
#undef SPM
  free(sparsem);
  free(rhs1);
  free(a1);
  free(ja);
  free(ia);
  free(x);
  return ret;
}

/* ------------------------------------------------------------------ velocities ----- */
/* solve_velocities.f90:8-281 (no isflowside2, no WR points) */
int oracle_velocity(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                    const gomgpu_forcing *f) {
  const int J = p->maxface, E = p->maxele, L = p->maxlayer;
  const double dt = p->dt, g = p->gravity, th = p->theta;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= J; j++) {
    int n1 = I2(m->nodenum_at_face, 1, j), n2 = I2(m->nodenum_at_face, 2, j);
    int kt = R1(s->top_layer_at_face, j);
    if (kt == 0) {
      for (int k = 1; k <= L; k++) {
        RL(s->un_face_new, k, j) = 0.0;
        RL(s->vn_face_new, k, j) = 0.0;
      }
      continue;
    }
    int kb = R1(m->bottom_layer_at_face, j);
    int bt = R1(m->boundary_type_of_face, j);
    int a1 = I2(m->adj_cellnum_at_face, 1, j), a2 = I2(m->adj_cellnum_at_face, 2, j);
    double rtemp0, rtemp1, rtemp2 = 0.0;
    for (int k = 1; k <= kt - kb + 1; k++) {
      int kk = kt + 1 - k;
      if (bt == -1) {
        RL(s->un_face_new, kk, j) = 0.0;
      } else if (bt == 0) {
        rtemp0 = th * g * dt / R1(m->delta_j, j) * (R1(s->eta_cell_new, a2) - R1(s->eta_cell_new, a1)) *
                 RL(s->AinvDeltaZ1, k, j);
        rtemp1 = rtemp0 * R1(m->cos_theta2, j);
        rtemp2 = rtemp0 * R1(m->sin_theta2, j);
        RL(s->un_face_new, kk, j) = RL(s->AinvG1, k, j) - rtemp1;
      } else {
        rtemp0 = th * g * dt / R1(m->delta_j, j) *
                 (R1(f->eta_at_ob_new, R1(m->ob_element_flag, a1)) - R1(s->eta_cell_new, a1)) *
                 RL(s->AinvDeltaZ1, k, j);
        RL(s->un_face_new, kk, j) = RL(s->AinvG1, k, j) - rtemp0;
      }
      if (R1(s->wetdry_node, n1) == 0 && R1(s->wetdry_node, n2) == 0) {
        rtemp1 = th * g * dt / R1(m->face_length, j) * (R1(s->eta_node, n2) - R1(s->eta_node, n1)) *
                 RL(s->AinvDeltaZ2, k, j);
        if (bt == 0)
          RL(s->vn_face_new, kk, j) = RL(s->AinvG2, k, j) - (rtemp1 + rtemp2);
        else
          RL(s->vn_face_new, kk, j) = RL(s->AinvG2, k, j) - rtemp1;
      }
      if (R1(m->initial_wetdry_node, n1) == 1 || R1(m->initial_wetdry_node, n2) == 1)
        RL(s->vn_face_new, kk, j) = dmax(-5.0, dmin(RL(s->vn_face_new, kk, j), 5.0));
    }
    if (R1(m->isflowside3, j) > 0) {
      double q = R1(f->Qu_boundary, R1(m->isflowside3, j));
      for (int k = kb; k <= kt; k++) RL(s->un_face_new, k, j) = q;
    }
    for (int k = 1; k <= kb - 1; k++) {
      RL(s->un_face_new, k, j) = 0.0;
      RL(s->vn_face_new, k, j) = 0.0;
    }
    for (int k = kt + 1; k <= L; k++) {
      RL(s->un_face_new, k, j) = RL(s->un_face_new, kt, j);
      RL(s->vn_face_new, k, j) = RL(s->vn_face_new, kt, j);
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= E; i++) {
    int t = R1(s->top_layer_at_element, i), b = R1(m->bottom_layer_at_element, i);
    if (t == 0) {
      for (int k = 0; k <= L; k++) R0(s->wn_cell_new, k, i) = 0.0;
      continue;
    }
    R0(s->wn_cell_new, b - 1, i) = 0.0;
    int N = R1(m->tri_or_quad, i);
    for (int k = b; k <= t; k++) {
      double sum1 = 0.0;
      for (int l = 1; l <= N; l++) {
        int jt = I4(m->facenum_at_cell, l, i);
        if (k >= R1(m->bottom_layer_at_face, jt) && k <= R1(s->top_layer_at_face, jt))
          sum1 = sum1 + (double)I4(m->sign_in_outflow, l, i) * R1(m->face_length, jt) * RL(s->dz_face, k, jt) *
                            RL(s->un_face_new, k, jt);
      }
      R0(s->wn_cell_new, k, i) = R0(s->wn_cell_new, k - 1, i) - sum1 / R1(m->area, i);
    }
    for (int k = 0; k <= b - 2; k++) R0(s->wn_cell_new, k, i) = 0.0;
    for (int k = t + 1; k <= L; k++) R0(s->wn_cell_new, k, i) = R0(s->wn_cell_new, t, i);
  }
  return 0;
}

/* -------------------------------------------------------------------- layers ------- */
/* update_vertical_layers_1.f90:18-333 */
int oracle_layers(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                  const gomgpu_forcing *f) {
  (void)f;
  const int E = p->maxele, J = p->maxface, P = p->maxnod, L = p->maxlayer;
  const double MSL = p->MSL, dry = p->dry_depth;
  for (int i = 1; i <= E; i++) {
    int tnew = 0;
    double tw = R1(m->h_cell, i) + R1(s->eta_cell_new, i);
    int bl = R1(m->bottom_layer_at_element, i);
    if (tw > dry) {
      double tmp = MSL + R1(s->eta_cell_new, i);
      for (int k = L - 1; k >= 0; k--) {
        if (tmp > ZL(k) && tmp <= ZL(k + 1)) {
          tnew = k + 1;
          RL(s->dz_cell_new, tnew, i) = tmp - ZL(k);
          break;
        }
      }
      if (tnew < bl) return fail(GOMGPU_ERR_LAYER_CELL, i);
      if (tnew == bl) {
        RL(s->dz_cell_new, tnew, i) = tw;
        R0(s->dzhalf_cell_new, tnew, i) = tw;
      } else {
        RL(s->dz_cell_new, bl, i) = ZL(bl) - (MSL - R1(m->h_cell, i));
        for (int k = bl + 1; k <= tnew - 1; k++) RL(s->dz_cell_new, k, i) = m->delta_z[k - 1];
        for (int k = bl; k <= tnew - 1; k++)
          R0(s->dzhalf_cell_new, k, i) = (RL(s->dz_cell_new, k + 1, i) + RL(s->dz_cell_new, k, i)) * 0.5;
        R0(s->dzhalf_cell_new, tnew, i) = RL(s->dz_cell_new, tnew, i);
        R0(s->dzhalf_cell_new, bl - 1, i) = RL(s->dz_cell_new, bl, i);
      }
    }
    if (R1(s->top_layer_at_element, i) == 0 && tnew != 0)
      for (int k = 0; k <= L; k++) R0(s->wn_cell_new, k, i) = 0.0;
    R1(s->top_layer_at_element, i) = tnew;
  }
  for (int j = 1; j <= J; j++) {
    int tnew = 0;
    int a1 = I2(m->adj_cellnum_at_face, 1, j), a2 = I2(m->adj_cellnum_at_face, 2, j);
    double eta_max = (a2 != 0) ? dmax(R1(s->eta_cell_new, a1), R1(s->eta_cell_new, a2)) : R1(s->eta_cell_new, a1);
    double tw = R1(m->h_face, j) + eta_max;
    int bl = R1(m->bottom_layer_at_face, j);
    if (tw > dry) {
      double tmp = MSL + eta_max;
      for (int k = L - 1; k >= 0; k--) {
        if (tmp > ZL(k) && tmp <= ZL(k + 1)) {
          tnew = k + 1;
          RL(s->dz_face_new, tnew, j) = tmp - ZL(k);
          break;
        }
      }
      if (tnew < bl) return fail(GOMGPU_ERR_LAYER_FACE, j);
      if (tnew == bl) {
        RL(s->dz_face_new, tnew, j) = tw;
        R0(s->dzhalf_face_new, tnew - 1, j) = tw;
        R0(s->dzhalf_face_new, tnew, j) = tw;
      } else {
        RL(s->dz_face_new, bl, j) = ZL(bl) - (MSL - R1(m->h_face, j));
        for (int k = bl + 1; k <= tnew - 1; k++) RL(s->dz_face_new, k, j) = m->delta_z[k - 1];
        for (int k = bl; k <= tnew - 1; k++)
          R0(s->dzhalf_face_new, k, j) = (RL(s->dz_face_new, k + 1, j) + RL(s->dz_face_new, k, j)) * 0.5;
        R0(s->dzhalf_face_new, tnew, j) = RL(s->dz_face_new, tnew, j);
        R0(s->dzhalf_face_new, bl - 1, j) = RL(s->dz_face_new, bl, j);
      }
    }
    R1(s->top_layer_at_face, j) = tnew;
  }
  for (int i = 1; i <= P; i++) {
    int tnew = 0;
    double tw = R1(m->h_node, i) + R1(s->eta_node, i);
    int bl = R1(m->bottom_layer_at_node, i);
    if (tw > dry) {
      double tmp = MSL + R1(s->eta_node, i);
      for (int k = L - 1; k >= 0; k--) {
        if (tmp > ZL(k) && tmp <= ZL(k + 1)) {
          tnew = k + 1;
          RL(s->dz_node_new, tnew, i) = tmp - ZL(k);
          break;
        }
      }
      if (tnew < bl) return fail(GOMGPU_ERR_LAYER_NODE, i);
      if (tnew == bl) {
        RL(s->dz_node_new, tnew, i) = tw;
        R0(s->dzhalf_node_new, tnew, i) = tw;
      } else {
        RL(s->dz_node_new, bl, i) = ZL(bl) - (MSL - R1(m->h_node, i));
        for (int k = bl + 1; k <= tnew - 1; k++) RL(s->dz_node_new, k, i) = m->delta_z[k - 1];
        for (int k = bl; k <= tnew - 1; k++)
          R0(s->dzhalf_node_new, k, i) = (RL(s->dz_node_new, k + 1, i) + RL(s->dz_node_new, k, i)) * 0.5;
      }
    }
    R1(s->top_layer_at_node, i) = tnew;
  }
  for (int i = 1; i <= E; i++) {
    int nsum1 = 0, nsum2 = 0, N = R1(m->tri_or_quad, i);
    for (int l = 1; l <= N; l++) {
      nsum1 += R1(s->top_layer_at_node, I4(m->nodenum_at_cell, l, i));
      nsum2 += R1(s->top_layer_at_face, I4(m->facenum_at_cell, l, i));
    }
    if (nsum1 == 0) R1(s->top_layer_at_element, i) = 0;
    if (nsum2 == 0 && R1(s->top_layer_at_element, i) != 0) return fail(GOMGPU_ERR_WET_DRY_SIDES, i);
  }
  const double dmin_m = p->denominator_min_for_matrix;
  for (int i = 1; i <= E; i++)
    for (int k = R1(m->bottom_layer_at_element, i); k <= R1(s->top_layer_at_element, i) - 1; k++)
      if (R0(s->dzhalf_cell_new, k, i) < dmin_m) return fail(GOMGPU_ERR_DZHALF_MIN, i);
  for (int j = 1; j <= J; j++)
    for (int k = R1(m->bottom_layer_at_face, j); k <= R1(s->top_layer_at_face, j) - 1; k++)
      if (R0(s->dzhalf_face_new, k, j) < dmin_m) return fail(GOMGPU_ERR_DZHALF_MIN, j);
  for (int i = 1; i <= P; i++)
    for (int k = R1(m->bottom_layer_at_node, i); k <= R1(s->top_layer_at_node, i) - 1; k++)
      if (R0(s->dzhalf_node_new, k, i) < dmin_m) return fail(GOMGPU_ERR_DZHALF_MIN, i);
  return 0;
}

/* ------------------------------------------------------------- node velocity ------ */
/* calculate_velocity_at_node.f90:10-152 */
int oracle_nodevel(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                   const gomgpu_forcing *f) {
  (void)f;
  const int J = p->maxface, P = p->maxnod, L = p->maxlayer, NB = p->max_no_neighbor_node;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= J; j++) {
    int kt = R1(s->top_layer_at_face, j);
    if (kt == 0) continue;
    int kb = R1(m->bottom_layer_at_face, j);
    double ct = R1(m->cos_theta, j), st = R1(m->sin_theta, j);
    double uf[66], vf[66];
    for (int k = kb; k <= kt; k++) {
      double un = RL(s->un_face_new, k, j), vn = RL(s->vn_face_new, k, j);
      uf[k] = un * ct - vn * st;
      vf[k] = un * st + vn * ct;
    }
    R0(s->u_face_level, kb - 1, j) = uf[kb];
    R0(s->v_face_level, kb - 1, j) = vf[kb];
    R0(s->u_face_level, kt, j) = uf[kt];
    R0(s->v_face_level, kt, j) = vf[kt];
    for (int k = kb; k <= kt - 1; k++) {
      double r = RL(s->dz_face, k, j) * 0.5 / R0(s->dzhalf_face, k, j);
      R0(s->u_face_level, k, j) = uf[k] + r * (uf[k + 1] - uf[k]);
      R0(s->v_face_level, k, j) = vf[k] + r * (vf[k + 1] - vf[k]);
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= P; i++) {
    int nadj = R1(m->adj_cells_at_node, i);
    for (int k = 0; k <= L; k++) {
      double u = 0.0, v = 0.0, weight = 0.0;
      int icount = 0;
      for (int jj = 1; jj <= nadj; jj++) {
        int isn = INB(m->adj_cellnum_at_node, jj, i);
        int nc = INB(m->node_count_each_element, jj, i);
        int N = R1(m->tri_or_quad, isn);
        for (int l = N - 2; l <= N - 1; l++) {
          int ifn = I4(m->facenum_at_cell, sen(N, nc, l), isn);
          double fw = (I2(m->adj_cellnum_at_face, 2, ifn) == 0) ? 2.0 : 1.0;
          int kt = R1(s->top_layer_at_face, ifn);
          if (kt != 0 && k >= R1(m->bottom_layer_at_face, ifn) - 1) {
            int kvt = k < kt ? k : kt;
            u = u + R0(s->u_face_level, kvt, ifn) / R1(m->face_length, ifn) * fw;
            v = v + R0(s->v_face_level, kvt, ifn) / R1(m->face_length, ifn) * fw;
            icount++;
          }
          weight = weight + fw / R1(m->face_length, ifn);
        }
      }
      if (icount != 0) {
        u = u / weight;
        v = v / weight;
      }
      R0(s->u_node, k, i) = u;
      R0(s->v_node, k, i) = v;
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= P; i++) {
    int nadj = R1(m->adj_cells_at_node, i);
    for (int k = 0; k <= L; k++) {
      double w = 0.0, tw = 0.0;
      for (int jj = 1; jj <= nadj; jj++) {
        int ite = INB(m->adj_cellnum_at_node, jj, i);
        int te = R1(s->top_layer_at_element, ite);
        if (te != 0 && k >= R1(m->bottom_layer_at_element, ite) - 1) {
          int kvt = k < te ? k : te;
          tw = tw + R1(m->area, ite);
          w = w + R0(s->wn_cell, kvt, ite) * R1(m->area, ite);
        }
      }
      if (tw != 0.0) w = w / tw;
      R0(s->w_node, k, i) = w;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ transport ------ */
/* Fortran MAX/MIN with a default-real first argument */
static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }

static double limiter(int which, double r, double phi) {
  double PHI;
  if (which == 1) {
    PHI = fmax2(phi, fmin2(1.0, r));
  } else if (which == 2) {
    PHI = fmax2(phi, (r + fabs(r)) / (1.0 + fabs(r)));
  } else {
    PHI = fmax2(fmax2(phi, fmin2(1.0, 2.0 * r)), fmin2(2.0, r));
  }
  return PHI - phi;
}

/* solve_transport_equation_v19.f90:7-1489 (TVD; no WR or SS points; heat_option 1) */
static int oracle_transport_elm(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                                const gomgpu_forcing *f);

int oracle_transport(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                     const gomgpu_forcing *f) {
  if (p->transport_solver == 2) return oracle_transport_elm(m, p, s, f);
  const int E = p->maxele, J = p->maxface, P = p->maxnod, L = p->maxlayer, NB = p->max_no_neighbor_node;
  const int NT = p->maxtran2, NOB = p->num_ob_cell, NQB = p->num_Qb_cell;
  const int tid[2] = {p->tran_id1, p->tran_id2};
  const double dt = p->dt, th = p->theta;
  const size_t LC = (size_t)(L + 2) * E; /* con_cell(0:L+1, E) per tracer */
  double *con = calloc(LC * NT, sizeof(double));
  double *con_new = calloc((size_t)L * E * NT, sizeof(double));
  double *node_new = calloc((size_t)L * P * NT, sizeof(double));
  double *face_new = calloc((size_t)L * J * NT, sizeof(double));
  double *Qj = calloc((size_t)L * J, sizeof(double)), *Dj = calloc((size_t)L * J, sizeof(double));
  double *dj = calloc((size_t)L * J, sizeof(double)), *phij = calloc((size_t)L * J, sizeof(double));
  double *Qi = calloc((size_t)(L + 1) * E, sizeof(double)), *Di = calloc((size_t)(L + 1) * E, sizeof(double));
  double *di = calloc((size_t)(L + 1) * E, sizeof(double)), *phii = calloc((size_t)(L + 1) * E, sizeof(double));
  double *rj = calloc((size_t)L * J * NT, sizeof(double)), *ri = calloc((size_t)(L + 1) * E * NT, sizeof(double));
  double *PSIj = calloc((size_t)L * J * NT, sizeof(double)), *PSIi = calloc((size_t)(L + 1) * E * NT, sizeof(double));
  double *obk = calloc((size_t)L * (NOB > 0 ? NOB : 1) * NT, sizeof(double));
  double *qbk = calloc((size_t)L * (NQB > 0 ? NQB : 1) * NT, sizeof(double));
  int ret = 0;
#define CON(k, i, t) con[(k) + (size_t)((i) - 1) * (L + 2) + (size_t)(t) * LC]
#define CNEW(k, i, t) con_new[((k) - 1) + (size_t)((i) - 1) * L + (size_t)(t) * L * E]
#define FJ(arr, k, j) arr[((k) - 1) + (size_t)((j) - 1) * L]
#define FI(arr, k, i) arr[(k) + (size_t)((i) - 1) * (L + 1)]
#define RJ(k, j, t) rj[((k) - 1) + (size_t)((j) - 1) * L + (size_t)(t) * L * J]
#define RI(k, i, t) ri[(k) + (size_t)((i) - 1) * (L + 1) + (size_t)(t) * (L + 1) * E]
#define PSJ(k, j, t) PSIj[((k) - 1) + (size_t)((j) - 1) * L + (size_t)(t) * L * J]
#define PSI_(k, i, t) PSIi[(k) + (size_t)((i) - 1) * (L + 1) + (size_t)(t) * (L + 1) * E]
#define OBK(k, i, t) obk[((k) - 1) + (size_t)((i) - 1) * L + (size_t)(t) * L * (NOB > 0 ? NOB : 1)]
#define QBK(k, i, t) qbk[((k) - 1) + (size_t)((i) - 1) * L + (size_t)(t) * L * (NQB > 0 ? NQB : 1)]
  for (int t = 0; t < NT; t++) {
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= E; i++) {
      const double *src = (tid[t] == 1) ? s->salt_cell : s->temp_cell;
      for (int k = 1; k <= L; k++) CON(k, i, t) = RL(src, k, i);
      CON(0, i, t) = 0.0;
      CON(L + 1, i, t) = 0.0;
      if (tid[t] == 2 && p->heat_option == 1 && f->phi_sz) {
        const double cp = 4186.0;
        int tl = R1(s->top_layer_at_element, i), bl = R1(m->bottom_layer_at_element, i);
        for (int k = bl; k <= tl; k++)
          CON(k, i, t) = CON(k, i, t) + dt * RL(f->phi_sz, k, i) / (p->rho_o * cp * RL(s->dz_cell, k, i));
      }
    }
  }
  int Nt = p->trans_sub_iter;
  if (Nt == 0) Nt = 1;
  const double dt2 = dt / Nt;
  for (int sub = 0; sub <= Nt - 1 && !ret; sub++) {
    memset(obk, 0, sizeof(double) * (size_t)L * (NOB > 0 ? NOB : 1) * NT);
    memset(qbk, 0, sizeof(double) * (size_t)L * (NQB > 0 ? NQB : 1) * NT);
    for (size_t q = 0; q < (size_t)L * J * NT; q++) rj[q] = 1.0;
    for (size_t q = 0; q < (size_t)(L + 1) * E * NT; q++) ri[q] = 1.0;
    /* boundary concentrations at the open and river cells (:232-397) */
    for (int i = 1; i <= NOB; i++) {
      int i2 = R1(m->ob_cell_id, i);
      int tl = R1(s->top_layer_at_element, i2), bl = R1(m->bottom_layer_at_element, i2);
      int N = R1(m->tri_or_quad, i2);
      for (int l = 1; l <= N; l++) {
        int j2 = I4(m->facenum_at_cell, l, i2);
        if (R1(m->boundary_type_of_face, j2) > 0) {
          for (int k = bl; k <= tl; k++) {
            double dzo = RL(s->dz_face, k, j2) + (RL(s->dz_face_new, k, j2) - RL(s->dz_face, k, j2)) * sub / Nt;
            double sg = (double)I4(m->sign_in_outflow, l, i2);
            double Q2 = R1(m->face_length, j2) * dzo *
                        ((1.0 - th) * RL(s->un_face, k, j2) * sg + th * RL(s->un_face_new, k, j2) * sg);
            for (int t = 0; t < NT; t++) {
              if (Q2 <= 0.0)
                OBK(k, i, t) = f->con_at_ob[sub + (size_t)Nt * (t + (size_t)NT * (i - 1))];
              else
                OBK(k, i, t) = CON(k, i2, t);
            }
          }
        }
      }
    }
    for (int i = 1; i <= NQB; i++) {
      int i2 = R1(m->Qb_cell_id, i);
      int tl = R1(s->top_layer_at_element, i2), bl = R1(m->bottom_layer_at_element, i2);
      for (int k = bl; k <= tl; k++)
        for (int t = 0; t < NT; t++) QBK(k, i, t) = f->con_at_Qb[sub + (size_t)Nt * (t + (size_t)NT * (i - 1))];
    }
    /* face and cell coefficients (:549-611) */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= J; j++) {
      int bl = R1(m->bottom_layer_at_face, j), tl = R1(s->top_layer_at_face, j);
      for (int k = bl; k <= tl; k++) {
        double dzo = RL(s->dz_face, k, j) + (RL(s->dz_face_new, k, j) - RL(s->dz_face, k, j)) * sub / Nt;
        double Q = R1(m->face_length, j) * dzo * ((1.0 - th) * RL(s->un_face, k, j) + th * RL(s->un_face_new, k, j));
        double D = R1(m->face_length, j) * dzo * RL(s->Kh, k, j) / R1(m->delta_j, j);
        FJ(Qj, k, j) = Q;
        FJ(Dj, k, j) = D;
        FJ(dj, k, j) = fmax2(0.0, D - 0.5 * fabs(Q));
        FJ(phij, k, j) = (Q == 0.0) ? 1.0 : fmin2(1.0, 2.0 * D / fabs(Q));
      }
    }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= E; i++) {
      int bl = R1(m->bottom_layer_at_element, i), tl = R1(s->top_layer_at_element, i);
      for (int k = bl; k <= tl - 1; k++) {
        double Q = R1(m->area, i) * ((1.0 - th) * R0(s->wn_cell, k, i) + th * R0(s->wn_cell_new, k, i));
        double dzo = R0(s->dzhalf_cell, k, i) + (R0(s->dzhalf_cell_new, k, i) - R0(s->dzhalf_cell, k, i)) * sub / Nt;
        double D = R1(m->area, i) * R0(s->Kv, k, i) / dzo;
        FI(Qi, k, i) = Q;
        FI(Di, k, i) = D;
        FI(di, k, i) = fmax2(0.0, D - 0.5 * fabs(Q));
        FI(phii, k, i) = (Q == 0.0) ? 1.0 : fmin2(1.0, 2.0 * D / fabs(Q));
      }
    }
    /* limiter ratios (:615-904) */
    if (p->h_flux_limiter > 0 || p->v_flux_limiter > 0) {
#pragma omp parallel for schedule(static)
      for (int i = 1; i <= E; i++) {
        int tl = R1(s->top_layer_at_element, i), bl = R1(m->bottom_layer_at_element, i);
        if (tl == 0) continue;
        int nv = tl - bl + 1, N = R1(m->tri_or_quad, i);
        double oqb[64][2];
        int iob = R1(m->ob_element_flag, i), iqb = R1(m->Qb_element_flag, i);
        for (int k = bl; k <= tl; k++) oqb[k][0] = oqb[k][1] = 0.0;
        if (iob > 0)
          for (int k = bl; k <= tl; k++)
            for (int t = 0; t < NT; t++) oqb[k][t] = OBK(k, iob, t);
        if (iqb > 0)
          for (int k = bl; k <= tl; k++)
            for (int t = 0; t < NT; t++) oqb[k][t] = QBK(k, iqb, t);
        double sum1[2] = {0, 0}, sum2[2] = {0, 0}, s1a[2], s2a[2], Q2[5];
        for (int k = bl; k <= tl; k++) {
          for (int t = 0; t < NT; t++) {
            sum1[t] = 0.0;
            sum2[t] = 0.0;
            s1a[t] = 0.0;
            s2a[t] = 0.0;
          }
          if (nv != 1) {
            for (int t = 0; t < NT; t++) {
              if (k == bl) {
                double r2 = fabs(0.5 * (FI(Qi, k, i) + fabs(FI(Qi, k, i))));
                s1a[t] = r2 * (CON(k, i, t) - CON(k + 1, i, t));
                s2a[t] = r2;
              } else if (k == tl) {
                double r1 = fabs(0.5 * (FI(Qi, k - 1, i) + fabs(FI(Qi, k - 1, i))));
                s1a[t] = r1 * (CON(k, i, t) - CON(k - 1, i, t));
                s2a[t] = r1;
              } else {
                double r1 = fabs(0.5 * (FI(Qi, k - 1, i) + fabs(FI(Qi, k - 1, i))));
                double r2 = fabs(0.5 * (FI(Qi, k, i) + fabs(FI(Qi, k, i))));
                s1a[t] = r1 * (CON(k, i, t) - CON(k - 1, i, t)) + r2 * (CON(k, i, t) - CON(k + 1, i, t));
                s2a[t] = r1 + r2;
              }
            }
          }
          for (int l = 1; l <= N; l++) {
            int j = I4(m->facenum_at_cell, l, i), ii = I4(m->adj_cellnum_at_cell, l, i);
            if (R1(m->boundary_type_of_face, j) != -1 || R1(m->isflowside3, j) > 0) {
              Q2[l] = FJ(Qj, k, j) * (double)I4(m->sign_in_outflow, l, i);
              if (Q2[l] < 0.0) {
                for (int t = 0; t < NT; t++) {
                  double ci = ii > 0 ? CON(k, ii, t) : oqb[k][t];
                  sum1[t] = sum1[t] + fabs(Q2[l]) * (CON(k, i, t) - ci);
                  sum2[t] = sum2[t] + fabs(Q2[l]);
                }
              }
            }
          }
          for (int t = 0; t < NT; t++) {
            s1a[t] = s1a[t] + sum1[t];
            s2a[t] = s2a[t] + sum2[t];
          }
          for (int l = 1; l <= N; l++) {
            int j = I4(m->facenum_at_cell, l, i), ii = I4(m->adj_cellnum_at_cell, l, i);
            if (R1(m->boundary_type_of_face, j) != -1 || R1(m->isflowside3, j) > 0) {
              /* without flow both cells of the face write it; the serial loop keeps the
                 higher-numbered cell's value */
              if (Q2[l] > 0.0 || (Q2[l] == 0.0 && ii < i)) {
                for (int t = 0; t < NT; t++) {
                  double ci = ii > 0 ? CON(k, ii, t) : oqb[k][t];
                  double rt = s2a[t] * (ci - CON(k, i, t));
                  RJ(k, j, t) = (rt == 0.0) ? 1.0 : s1a[t] / rt;
                }
              }
            }
          }
        }
        /* vertical ratios use the horizontal sums of the top layer (:806-840) */
        for (int t = 0; t < NT; t++) {
          if (FI(Qi, tl, i) >= 0.0) {
            for (int k = bl; k <= tl - 1; k++) {
              double rt1 = sum1[t] + fabs(FI(Qi, k - 1, i)) * (CON(k, i, t) - CON(k - 1, i, t));
              double rt2 = sum2[t] + fabs(FI(Qi, k - 1, i));
              double rt = (CON(k + 1, i, t) - CON(k, i, t)) * rt2;
              RI(k, i, t) = (rt == 0.0) ? 1.0 : rt1 / rt;
            }
          } else {
            for (int k = bl + 1; k <= tl; k++) {
              double rt1 = sum1[t] + fabs(FI(Qi, k, i)) * (CON(k, i, t) - CON(k + 1, i, t));
              double rt2 = sum2[t] + fabs(FI(Qi, k, i));
              double rt = (CON(k - 1, i, t) - CON(k, i, t)) * rt2;
              RI(k - 1, i, t) = (rt == 0.0) ? 1.0 : rt1 / rt;
            }
          }
          RI(tl, i, t) = 1.0;
          RI(bl - 1, i, t) = 1.0;
        }
      }
#pragma omp parallel for schedule(static)
      for (int i = 1; i <= E; i++) {
        int bl = R1(m->bottom_layer_at_element, i), tl = R1(s->top_layer_at_element, i);
        for (int k = bl; k <= tl - 1; k++)
          for (int t = 0; t < NT; t++)
            PSI_(k, i, t) = (p->h_flux_limiter == 0) ? 0.0 : limiter(p->h_flux_limiter, RI(k, i, t), FI(phii, k, i));
      }
#pragma omp parallel for schedule(static)
      for (int j = 1; j <= J; j++) {
        int bl = R1(m->bottom_layer_at_face, j), tl = R1(s->top_layer_at_face, j);
        for (int k = bl; k <= tl; k++)
          for (int t = 0; t < NT; t++)
            PSJ(k, j, t) = (p->h_flux_limiter == 0) ? 0.0 : limiter(p->h_flux_limiter, RJ(k, j, t), FJ(phij, k, j));
      }
    }
    /* per-column update (:909-1346) */
    loop_err col_err = {0, 0};
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= E; i++) {
      int bl = R1(m->bottom_layer_at_element, i), tl = R1(s->top_layer_at_element, i);
      if (tl == 0) continue;
      int nv = tl - bl + 1, N = R1(m->tri_or_quad, i);
      double ar = R1(m->area, i);
      double oqb[64][2], a[64], b[64], cc[64], rhs[2 * 64], sol[2 * 64];
      for (int k = bl + 1; k <= tl - 1; k++) {
        int kk = tl + 1 - k;
        a[kk] = -dt2 * FI(di, k, i);
        cc[kk] = -dt2 * FI(di, k - 1, i);
        double dzn = RL(s->dz_cell, k, i) + (RL(s->dz_cell_new, k, i) - RL(s->dz_cell, k, i)) * (sub + 1) / Nt;
        b[kk] = -a[kk] + ar * dzn - cc[kk];
      }
      if (tl == bl) {
        double dzn = RL(s->dz_cell, tl, i) + (RL(s->dz_cell_new, tl, i) - RL(s->dz_cell, tl, i)) * (sub + 1) / Nt;
        b[1] = ar * dzn;
      } else {
        cc[1] = -dt2 * FI(di, tl - 1, i);
        double dzn = RL(s->dz_cell, tl, i) + (RL(s->dz_cell_new, tl, i) - RL(s->dz_cell, tl, i)) * (sub + 1) / Nt;
        b[1] = ar * dzn - cc[1];
        a[nv] = -dt2 * FI(di, bl, i);
        dzn = RL(s->dz_cell, bl, i) + (RL(s->dz_cell_new, bl, i) - RL(s->dz_cell, bl, i)) * (sub + 1) / Nt;
        b[nv] = -a[nv] + ar * dzn;
      }
      int iob = R1(m->ob_element_flag, i), iqb = R1(m->Qb_element_flag, i);
      for (int k = bl; k <= tl; k++) oqb[k][0] = oqb[k][1] = 0.0;
      if (iob > 0)
        for (int k = bl; k <= tl; k++)
          for (int t = 0; t < NT; t++) oqb[k][t] = OBK(k, iob, t);
      if (iqb > 0)
        for (int k = bl; k <= tl; k++)
          for (int t = 0; t < NT; t++) oqb[k][t] = QBK(k, iqb, t);
      for (int k = bl; k <= tl; k++) {
        int kk = tl + 1 - k;
        double sqi[2] = {0, 0}, sqii[2] = {0, 0}, shd[2] = {0, 0}, shl[2] = {0, 0};
        for (int l = 1; l <= N; l++) {
          int j = I4(m->facenum_at_cell, l, i), ii = I4(m->adj_cellnum_at_cell, l, i);
          double ci[2];
          for (int t = 0; t < NT; t++) ci[t] = ii > 0 ? CON(k, ii, t) : oqb[k][t];
          if (R1(m->boundary_type_of_face, j) != -1 || R1(m->isflowside3, j) > 0) {
            double Q2 = FJ(Qj, k, j) * (double)I4(m->sign_in_outflow, l, i);
            if (Q2 > 0.0) {
              for (int t = 0; t < NT; t++) sqi[t] = sqi[t] + fabs(Q2) * CON(k, i, t);
            } else if (Q2 < 0.0) {
              for (int t = 0; t < NT; t++) sqii[t] = sqii[t] + fabs(Q2) * ci[t];
            }
            for (int t = 0; t < NT; t++) shd[t] = shd[t] + FJ(dj, k, j) * (ci[t] - CON(k, i, t));
            for (int t = 0; t < NT; t++) shl[t] = shl[t] + PSJ(k, j, t) * fabs(Q2) * (CON(k, i, t) - ci[t]);
          }
        }
        for (int t = 0; t < NT; t++) {
          double h_adv = dt2 * (sqii[t] - sqi[t]);
          double h_diff = dt2 * shd[t];
          double h_flux = 0.5 * dt2 * shl[t];
          double dzo = RL(s->dz_cell, k, i) + (RL(s->dz_cell_new, k, i) - RL(s->dz_cell, k, i)) * sub / Nt;
          rhs[(kk - 1) + t * 64] = ar * dzo * CON(k, i, t) + h_adv + h_diff;
          if (p->h_flux_limiter > 0) rhs[(kk - 1) + t * 64] = rhs[(kk - 1) + t * 64] + h_flux;
        }
      }
      if (nv != 1) {
        for (int k = bl; k <= tl; k++) {
          int kk = tl + 1 - k;
          for (int t = 0; t < NT; t++) {
            double v_adv, vft = 0.0;
            if (k == bl) {
              double Q = FI(Qi, k, i);
              double Ap = fabs(0.5 * (Q + fabs(Q))), Am = fabs(0.5 * (Q - fabs(Q)));
              v_adv = dt2 * (-(Ap * CON(k, i, t)) + Am * CON(k + 1, i, t));
              if (p->v_flux_limiter > 0) {
                double ps = PSI_(k, i, t);
                vft = 0.5 * dt2 * (-(ps * Ap * (CON(k + 1, i, t) - CON(k, i, t))) + ps * Am * (CON(k, i, t) - CON(k + 1, i, t)));
              }
            } else if (k == tl) {
              double Q = FI(Qi, k - 1, i);
              double Ap = fabs(0.5 * (Q + fabs(Q))), Am = fabs(0.5 * (Q - fabs(Q)));
              v_adv = dt2 * (Ap * CON(k - 1, i, t) - Am * CON(k, i, t));
              if (p->v_flux_limiter > 0) {
                double ps = PSI_(k - 1, i, t);
                vft = 0.5 * dt2 * (ps * Ap * (CON(k, i, t) - CON(k - 1, i, t)) - ps * Am * (CON(k - 1, i, t) - CON(k, i, t)));
              }
            } else {
              double Q1 = FI(Qi, k - 1, i), Q2 = FI(Qi, k, i);
              double A1p = fabs(0.5 * (Q1 + fabs(Q1))), A1m = fabs(0.5 * (Q1 - fabs(Q1)));
              double A2p = fabs(0.5 * (Q2 + fabs(Q2))), A2m = fabs(0.5 * (Q2 - fabs(Q2)));
              v_adv = dt2 * (A1p * CON(k - 1, i, t) - A2p * CON(k, i, t) + A2m * CON(k + 1, i, t) - A1m * CON(k, i, t));
              if (p->v_flux_limiter > 0) {
                double p1 = PSI_(k - 1, i, t), p2 = PSI_(k, i, t);
                vft = 0.5 * dt2 * (p1 * A1p * (CON(k, i, t) - CON(k - 1, i, t)) - p2 * A2p * (CON(k + 1, i, t) - CON(k, i, t)) +
                                   p2 * A2m * (CON(k, i, t) - CON(k + 1, i, t)) - p1 * A1m * (CON(k - 1, i, t) - CON(k, i, t)));
              }
            }
            rhs[(kk - 1) + t * 64] = rhs[(kk - 1) + t * 64] + v_adv;
            if (p->v_flux_limiter > 0) rhs[(kk - 1) + t * 64] = rhs[(kk - 1) + t * 64] + vft;
          }
        }
      }
      int err = oracle_tridiag(nv, NT, 64, a + 1, b + 1, cc + 1, rhs, sol);
      if (err) {
        loop_fail(&col_err, err, i);
        continue;
      }
      for (int k = bl; k <= tl; k++) {
        int kk = tl - k + 1;
        for (int t = 0; t < NT; t++) {
          double v = sol[(kk - 1) + t * 64];
          if (v >= 0.0 && v <= 50.0)
            CNEW(k, i, t) = v;
          else if (v < 0.0)
            CNEW(k, i, t) = 0.0;
          else if (v > 50.0)
            CNEW(k, i, t) = 50.0;
        }
      }
    }
    if (col_err.code) {
      ret = fail(col_err.code, col_err.entity);
      break;
    }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= E; i++) {
      int bl = R1(m->bottom_layer_at_element, i), tl = R1(s->top_layer_at_element, i);
      for (int t = 0; t < NT; t++) {
        for (int k = bl; k <= tl; k++) CON(k, i, t) = CNEW(k, i, t);
        for (int k = 0; k <= bl - 1; k++) CON(k, i, t) = 0.0;
        for (int k = tl + 1; k <= L + 1; k++) CON(k, i, t) = 0.0;
      }
    }
  }
  if (!ret) {
    /* node and face values (:1403-1468) */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= P; i++) {
      int tn = R1(s->top_layer_at_node, i), bn = R1(m->bottom_layer_at_node, i);
      int nadj = R1(m->adj_cells_at_node, i);
      for (int k = tn; k >= bn; k--) {
        for (int t = 0; t < NT; t++) {
          double s11 = 0.0, s22 = 0.0;
          int count1 = 0;
          for (int l = 1; l <= nadj; l++) {
            int ii = INB(m->adj_cellnum_at_node, l, i);
            if (k >= R1(m->bottom_layer_at_element, ii) && k <= R1(s->top_layer_at_element, ii)) {
              s11 = s11 + R1(m->area, ii) * RL(s->dz_cell_new, k, ii);
              s22 = s22 + R1(m->area, ii) * RL(s->dz_cell_new, k, ii) * CNEW(k, ii, t);
            } else {
              count1++;
            }
          }
          double v = (s11 == 0.0) ? 0.0 : s22 / s11;
          node_new[(k - 1) + (size_t)(i - 1) * L + (size_t)t * L * P] = v;
          if (count1 == nadj && k != tn)
            node_new[(k - 1) + (size_t)(i - 1) * L + (size_t)t * L * P] = node_new[k + (size_t)(i - 1) * L + (size_t)t * L * P];
        }
      }
    }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= J; j++) {
      int n1 = I2(m->nodenum_at_face, 1, j), n2 = I2(m->nodenum_at_face, 2, j);
      for (int k = R1(m->bottom_layer_at_face, j); k <= R1(s->top_layer_at_face, j); k++)
        for (int t = 0; t < NT; t++)
          face_new[(k - 1) + (size_t)(j - 1) * L + (size_t)t * L * J] =
              (node_new[(k - 1) + (size_t)(n1 - 1) * L + (size_t)t * L * P] +
               node_new[(k - 1) + (size_t)(n2 - 1) * L + (size_t)t * L * P]) * 0.5;
    }
    /* copy-out with the reference's fixed tracer slots: salt <- slot 1, temp <- slot 2 */
    for (int t = 0; t < NT; t++) {
      int slot = (tid[t] == 1) ? 0 : 1;
      if (slot >= NT) {
        ret = GOMGPU_E_ARG;
        break;
      }
      double *cn = (tid[t] == 1) ? s->salt_cell_new : s->temp_cell_new;
      double *nn = (tid[t] == 1) ? s->salt_node : s->temp_node;
      double *fn = (tid[t] == 1) ? s->salt_face : s->temp_face;
      memcpy(cn, con_new + (size_t)slot * L * E, sizeof(double) * L * E);
      memcpy(nn, node_new + (size_t)slot * L * P, sizeof(double) * L * P);
      memcpy(fn, face_new + (size_t)slot * L * J, sizeof(double) * L * J);
    }
  }
#undef CON
#undef CNEW
#undef FJ
#undef FI
#undef RJ
#undef RI
#undef PSJ
#undef PSI_
#undef OBK
#undef QBK
  free(con);
  free(con_new);
  free(node_new);
  free(face_new);
  free(Qj);
  free(Dj);
  free(dj);
  free(phij);
  free(Qi);
  free(Di);
  free(di);
  free(phii);
  free(rj);
  free(ri);
  free(PSIj);
  free(PSIi);
  free(obk);
  free(qbk);
  return ret;
}

/* ------------------------------------------------------------------------ EOS ------ */
/* UNESCO-1980 one-atmosphere density with the reference's single-precision coefficients
 * (calculate_density_full.f90:57-71). Powers are sequential products. */
double oracle_rho_unesco(double t, double s) {
  double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  double sq = sqrt(s), sq3 = sq * sq * sq;
  return (double)(1000.0f - 0.157406f) + (double)6.793952e-2f * t - (double)9.095290e-3f * t2 +
         (double)1.001685e-4f * t3 - (double)1.120083e-6f * t4 + (double)6.536332e-9f * t5 +
         s * ((double)0.824493f - (double)4.089900e-3f * t + (double)7.643800e-5f * t2 -
              (double)8.246700e-7f * t3 + (double)5.387500e-9f * t4) +
         sq3 * ((double)(-5.72466e-3f) + (double)1.022700e-4f * t - (double)1.654600e-6f * t2) +
         (double)4.831400e-4f * (s * s);
}

/* pressure correction, calculate_density_full.f90:75-101 (parity unpinned) */
double oracle_rho_pressure(double rho_surface, double t, double s, double pres) {
  double t2 = t * t, t3 = t2 * t, t4 = t3 * t;
  double sq = sqrt(s), sq3 = sq * sq * sq, p2 = pres * pres;
  double sbm = (double)19652.21f + (double)148.4206f * t - (double)2.327105f * t2 + (double)1.360477e-2f * t3 -
               (double)5.155288e-5f * t4 +
               pres * ((double)3.239908f + (double)1.43713e-3f * t + (double)1.16092e-4f * t2 - (double)5.77905e-7f * t3) +
               p2 * ((double)8.50935e-5f - (double)6.12293e-6f * t + (double)5.2787e-8f * t2) +
               s * ((double)54.6746f - (double)0.603459f * t + (double)1.09987e-2f * t2 - (double)6.1670e-5f * t3) +
               sq3 * ((double)7.944e-2f + (double)1.6483e-2f * t - (double)5.3009e-4f * t2) +
               pres * s * ((double)2.2838e-3f - (double)1.0981e-5f * t - (double)1.6078e-6f * t2) +
               (double)1.91075e-4f * pres * sq3 - (double)9.9348e-7f * p2 * s + (double)2.08160e-8f * t * p2 * s +
               (double)9.1697e-10f * t2 * p2 * s;
  double ratio = 1.0 / (1.0 - pres / sbm);
  return rho_surface * ratio;
}

static int eos_column(const gomgpu_mesh *m, const gomgpu_params *p, int L, int full, int bl, int tl,
                      const double *tcol, const double *scol, double *rcol, const double *dzcol,
                      int entity, int do_compute) {
  for (int k = bl; k <= tl; k++) {
    double rt = tcol[k - 1], rs = scol[k - 1];
    if (rt < -5.0 || rt > 100.0 || rs < -5.0 || rs > 100.0) return fail(GOMGPU_ERR_EOS_RANGE, entity);
    /* a face of the linear EOS keeps its density (i_density_flag, never set, is 0), and the
     * check below still reads it (calculate_density_linear.f90:101-110) */
    if (!do_compute) {
      if (rcol[k - 1] < 980.0) return fail(GOMGPU_ERR_EOS_RANGE, entity);
      continue;
    }
    if (full) {
      rcol[k - 1] = oracle_rho_unesco(rt, rs);
      if (p->MSL - ZL(k) > 100.0) {
        double pres = 0.0;
        for (int kk = k; kk <= tl; kk++) {
          double ptmp = (double)9.81f * rcol[kk - 1] * dzcol[kk - 1];
          if (kk == k)
            pres = pres + (double)1.0e-5f * 0.5 * ptmp;
          else
            pres = pres + (double)1.0e-5f * ptmp;
        }
        rcol[k - 1] = oracle_rho_pressure(rcol[k - 1], rt, rs, pres);
      }
    } else {
      rcol[k - 1] = 1000.0 + (double)0.7f * rs;
    }
    if (rcol[k - 1] < 980.0) return fail(GOMGPU_ERR_EOS_RANGE, entity);
  }
  return 0;
}

/* ---------------------------------------------------- ELM scalar transport ---------- */
/* The tracers at a backtraced foot point: the which = 2 branch of ELM_backtrace_v1.f90:
 * 216-473. Boundary cells take the boundary values; elsewhere the node values of the two
 * levels bracketing zt are blended, weighted by the foot's interpolation weights (or
 * averaged over the wet nodes), and a wet sub-element of the cell that holds the foot --
 * the four sub-triangles of a triangle through its face midpoints, the four sub-quads of a
 * quad through its face midpoints and centre -- replaces that by its own interpolation. */
static int elm_tracer_foot(const elmctx *c, const gomgpu_forcing *f, int nnel, int jlev,
                           const double *foot, double *ttint, double *ssint) {
  const gomgpu_mesh *m = c->m;
  const gomgpu_params *p = c->p;
  gomgpu_state *s = c->s;
  const int L = c->L, NT = p->maxtran2, Nt = p->trans_sub_iter > 0 ? p->trans_sub_iter : 1;
  const double xt = foot[0], yt = foot[1], zt = foot[2], zup = foot[3];
  const double small_06 = (double)1.e-6f;
  double staint[4] = {foot[4], foot[5], foot[6], foot[7]};
  const int N = R1(m->tri_or_quad, nnel);
  const int ibnd1 = R1(m->ob_element_flag, nnel), ibnd2 = R1(m->Qb_element_flag, nnel);
  *ttint = 0.0;
  *ssint = 0.0;
  if (ibnd1 > 0) {
    *ssint = f->con_at_ob[0 + (size_t)Nt * (0 + (size_t)NT * (ibnd1 - 1))];
    *ttint = f->con_at_ob[0 + (size_t)Nt * (1 + (size_t)NT * (ibnd1 - 1))];
    return 0;
  }
  if (ibnd2 > 0) {
    *ssint = f->con_at_Qb[0 + (size_t)Nt * (0 + (size_t)NT * (ibnd2 - 1))];
    *ttint = f->con_at_Qb[0 + (size_t)Nt * (1 + (size_t)NT * (ibnd2 - 1))];
    return 0;
  }
  const double zanchor = zup - RL(s->dz_cell, jlev, nnel) * 0.5;
  int k1, k2;
  double zrat2;
  if (jlev == R1(s->top_layer_at_element, nnel) && zt > zanchor) {
    k1 = jlev;
    k2 = jlev;
    zrat2 = 0.5;
  } else if (zt > zanchor) {
    k1 = jlev;
    k2 = jlev + 1;
    zrat2 = (zt - zanchor) / R0(s->dzhalf_cell, jlev, nnel);
  } else {
    k1 = jlev - 1 > R1(m->bottom_layer_at_element, nnel) ? jlev - 1 : R1(m->bottom_layer_at_element, nnel);
    k2 = jlev;
    zrat2 = 1.0 - (zanchor - zt) / R0(s->dzhalf_cell, k1, nnel);
  }
  zrat2 = dmax(0.0, dmin(1.0, zrat2));
#define ZB(arr, k, e) (RL(arr, k1, e) * (1 - zrat2) + RL(arr, k2, e) * zrat2)
  int icount = 0;
  double ta = 0.0, sa = 0.0, t_xi[4] = {0, 0, 0, 0}, s_xi[4] = {0, 0, 0, 0};
  for (int l = 1; l <= N; l++) {
    int nd = I4(m->nodenum_at_cell, l, nnel);
    if (R1(s->top_layer_at_node, nd) != 0) {
      icount++;
      t_xi[l - 1] = ZB(s->temp_node, k, nd);
      s_xi[l - 1] = ZB(s->salt_node, k, nd);
      ta = ta + t_xi[l - 1];
      sa = sa + s_xi[l - 1];
    }
  }
  if (icount == 0) return fail(GOMGPU_ERR_ELM_ALL_DRY, nnel);
  if (icount != N) {
    ta = ta / icount;
    sa = sa / icount;
    *ttint = ta;
    *ssint = sa;
  } else {
    ta = ta / icount;
    sa = sa / icount;
    *ttint = 0.0;
    *ssint = 0.0;
    for (int l = 0; l < N; l++) {
      *ttint = *ttint + t_xi[l] * staint[l];
      *ssint = *ssint + s_xi[l] * staint[l];
    }
  }
  int index_1 = 0;
  const double ar = R1(m->area, nnel);
  if (N == 3) {
    for (int l = 1; l <= 4; l++) {
      double aa1, aa2, aa3, X1, Y1, X2, Y2, X3, Y3;
      int n1 = 0, f2, f3;
      if (l <= 3) {
        n1 = I4(m->nodenum_at_cell, l, nnel);
        f2 = I4(m->facenum_at_cell, sen(N, l, 2), nnel);
        f3 = I4(m->facenum_at_cell, sen(N, l, 1), nnel);
        X1 = R1(m->x_node, n1);
        Y1 = R1(m->y_node, n1);
      } else {
        int f1 = I4(m->facenum_at_cell, 1, nnel);
        f2 = I4(m->facenum_at_cell, 2, nnel);
        f3 = I4(m->facenum_at_cell, 3, nnel);
        X1 = R1(m->x_face, f1);
        Y1 = R1(m->y_face, f1);
        n1 = f1;
      }
      X2 = R1(m->x_face, f2);
      Y2 = R1(m->y_face, f2);
      X3 = R1(m->x_face, f3);
      Y3 = R1(m->y_face, f3);
      aa1 = carea(xt, X2, X3, yt, Y2, Y3);
      aa2 = carea(X1, xt, X3, Y1, yt, Y3);
      aa3 = carea(X1, X2, xt, Y1, Y2, yt);
      double aa = fabs(aa1) + fabs(aa2) + fabs(aa3);
      double subrat = fabs(aa - ar / 4) * 4 / ar;
      if (subrat < 100 * small_06) {
        index_1 = 1;
        double sig[3];
        sig[0] = aa1 * 4 / ar;
        sig[1] = aa2 * 4 / ar;
        sig[0] = dmax(0.0, dmin(1.0, sig[0]));
        sig[1] = dmax(0.0, dmin(1.0, sig[1]));
        if (sig[0] + sig[1] > 1) {
          sig[2] = 0;
          sig[1] = 1 - sig[0];
        } else {
          sig[2] = 1 - sig[0] - sig[1];
        }
        int wet = l <= 3 ? (R1(s->top_layer_at_node, n1) != 0) : (R1(s->top_layer_at_face, n1) != 0);
        wet = wet && R1(s->top_layer_at_face, f2) != 0 && R1(s->top_layer_at_face, f3) != 0;
        if (wet) {
          if (l <= 3) {
            t_xi[0] = ZB(s->temp_node, k, n1);
            s_xi[0] = ZB(s->salt_node, k, n1);
          } else {
            t_xi[0] = ZB(s->temp_face, k, n1);
            s_xi[0] = ZB(s->salt_face, k, n1);
          }
          t_xi[1] = ZB(s->temp_face, k, f2);
          t_xi[2] = ZB(s->temp_face, k, f3);
          s_xi[1] = ZB(s->salt_face, k, f2);
          s_xi[2] = ZB(s->salt_face, k, f3);
          *ttint = t_xi[0] * sig[0] + t_xi[1] * sig[1] + t_xi[2] * sig[2];
          *ssint = s_xi[0] * sig[0] + s_xi[1] * sig[1] + s_xi[2] * sig[2];
          break;
        }
      }
    }
  } else {
    const double XC = R1(m->x_cell, nnel), YC = R1(m->y_cell, nnel);
    for (int l = 1; l <= 4; l++) {
      int n1 = I4(m->nodenum_at_cell, l, nnel);
      int f2 = I4(m->facenum_at_cell, sen(N, l, 3), nnel);
      int f4 = I4(m->facenum_at_cell, sen(N, l, 2), nnel);
      double X1 = R1(m->x_node, n1), Y1 = R1(m->y_node, n1);
      double X2 = R1(m->x_face, f2), Y2 = R1(m->y_face, f2);
      double X4 = R1(m->x_face, f4), Y4 = R1(m->y_face, f4);
      double aa1 = carea(X1, X2, xt, Y1, Y2, yt);
      double aa2 = carea(X2, XC, xt, Y2, YC, yt);
      double aa3 = carea(XC, X4, xt, YC, Y4, yt);
      double aa4 = carea(X4, X1, xt, Y4, Y1, yt);
      double bb1 = carea(X1, X2, X4, Y1, Y2, Y4);
      double bb2 = carea(X2, XC, X4, Y2, YC, Y4);
      if (bb1 <= 0 || bb2 <= 0) return fail(GOMGPU_ERR_ELM_NEG_SUBCELL, nnel);
      double bb = bb1 + bb2;
      double aa = fabs(aa1) + fabs(aa2) + fabs(aa3) + fabs(aa4);
      double subrat = fabs(aa - bb) / bb;
      if (subrat < 4.0 * small_06) {
        index_1 = 1;
        double csi, etta;
        int err2 = oracle_bilinear(X1, X2, XC, X4, Y1, Y2, YC, Y4, xt, yt, &csi, &etta, staint);
        if (err2) return fail(err2, nnel);
        if (R1(s->top_layer_at_node, n1) != 0 && R1(s->top_layer_at_face, f2) != 0 &&
            R1(s->top_layer_at_face, f4) != 0) {
          t_xi[0] = ZB(s->temp_node, k, n1);
          t_xi[1] = ZB(s->temp_face, k, f2);
          t_xi[2] = ta;
          t_xi[3] = ZB(s->temp_face, k, f4);
          s_xi[0] = ZB(s->salt_node, k, n1);
          s_xi[1] = ZB(s->salt_face, k, f2);
          s_xi[2] = sa;
          s_xi[3] = ZB(s->salt_face, k, f4);
          *ttint = 0.0;
          *ssint = 0.0;
          for (int j = 0; j < 4; j++) {
            *ttint = *ttint + t_xi[j] * staint[j];
            *ssint = *ssint + s_xi[j] * staint[j];
          }
        }
      }
    }
  }
#undef ZB
  if (index_1 == 0) return fail(GOMGPU_ERR_ELM_NO_SUBCELL, nnel);
  return 0;
}

/* solve_transport_equation_ELM_v8.f90:7-346: salt and temperature backtraced from each
 * wet cell-layer centre, explicit horizontal diffusion, implicit vertical diffusion (the
 * top row's upper coefficient is -dt + Kv/dzhalf as the reference writes it), the values
 * extended above and below the column, area*dz-weighted node means and face means. Dry
 * cells and nodes keep their new values. */
static int oracle_transport_elm(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                                const gomgpu_forcing *f) {
  const int E = p->maxele, P = p->maxnod, J = p->maxface, L = p->maxlayer, NB = p->max_no_neighbor_node;
  const double dt = p->dt;
  const double cp = 4186.0;
  double *tc = calloc((size_t)L * E, sizeof(double)), *sc = calloc((size_t)L * E, sizeof(double));
  double *trhs = calloc((size_t)L * E, sizeof(double)), *srhs = calloc((size_t)L * E, sizeof(double));
  double a[65], b[65], cc[65], rhs[2 * 65], sol[2 * 65];  /* rows 1..nv, right-hand sides 65 apart */
  int ret = 0;
  elmctx c = {m, p, s, L};
  for (int i = 1; i <= E && !ret; i++) {
    int tl = R1(s->top_layer_at_element, i), bl = R1(m->bottom_layer_at_element, i);
    if (tl == 0) continue;  /* tc, sc stay 0 */
    int N = R1(m->tri_or_quad, i);
    for (int k = bl; k <= tl; k++) {
      int nnel = i, jlev = k;
      double x0 = R1(m->x_cell, i), y0 = R1(m->y_cell, i), z0;
      if (k == R1(s->top_layer_at_element, nnel))
        z0 = p->MSL + R1(s->eta_cell, nnel) - RL(s->dz_cell, k, nnel) * 0.5;
      else
        z0 = ZL(k) - RL(s->dz_cell, k, nnel) * 0.5;
      double u_temp = 0.0, v_temp = 0.0;
      int bt_step = 0;
      for (int l = 1; l <= N; l++) {
        int nd = I4(m->nodenum_at_cell, l, i);
        u_temp = u_temp + (R0(s->u_node, k, nd) + R0(s->u_node, k - 1, nd)) * 0.5;
        v_temp = v_temp + (R0(s->v_node, k, nd) + R0(s->v_node, k - 1, nd)) * 0.5;
        bt_step = bt_step + RL(s->num_sub_elm_iteration, k, nd);
      }
      u_temp = u_temp / N;
      v_temp = v_temp / N;
      bt_step = bt_step / N;
      double uu = u_temp, vv = v_temp;
      double ww = (R0(s->wn_cell, k, i) + R0(s->wn_cell, k - 1, i)) * 0.5;
      double vmag = sqrt(uu * uu + vv * vv + ww * ww);
      double t_, s_;
      if (vmag <= (double)1.e-4f) {
        t_ = RL(s->temp_cell, k, i);
        s_ = RL(s->salt_cell, k, i);
      } else {
        double bt_dt = dt / bt_step, foot[8];
        int near = 0;
        int err = backtrace(&c, i, bt_step, bt_dt, &uu, &vv, &ww, x0, y0, z0, &nnel, &jlev, foot, &near);
        g_near[1] += near;
        if (!err) err = elm_tracer_foot(&c, f, nnel, jlev, foot, &t_, &s_);
        if (err) {
          ret = err;
          break;
        }
      }
      t_ = t_ + dt * RL(f->phi_sz, k, i) / (p->rho_o * cp * RL(s->dz_cell, k, i));
      RL(tc, k, i) = t_;
      RL(sc, k, i) = s_;
    }
  }
  if (!ret) {
    for (int i = 1; i <= E; i++) {
      int tl = R1(s->top_layer_at_element, i), bl = R1(m->bottom_layer_at_element, i), N = R1(m->tri_or_quad, i);
      for (int k = bl; k <= tl; k++) {
        double sum1 = 0.0, sum2 = 0.0;
        for (int l = 1; l <= N; l++) {
          int j = I4(m->facenum_at_cell, l, i);
          int ie = I2(m->adj_cellnum_at_face, 2, j);
          if (ie != 0) {
            double rt = R1(m->face_length, j) * RL(s->dz_face, k, j) * RL(s->Kh, k, j) / R1(m->delta_j, j);
            sum1 = sum1 + rt * (RL(s->salt_cell, k, ie) - RL(s->salt_cell, k, i));
            sum2 = sum2 + rt * (RL(s->temp_cell, k, ie) - RL(s->temp_cell, k, i));
          }
        }
        sum1 = sum1 * dt / (R1(m->area, i) * RL(s->dz_cell, k, i));
        sum2 = sum2 * dt / (R1(m->area, i) * RL(s->dz_cell, k, i));
        RL(trhs, k, i) = RL(tc, k, i) + sum1;
        RL(srhs, k, i) = RL(sc, k, i) + sum2;
      }
    }
    for (int i = 1; i <= E && !ret; i++) {
      int tl = R1(s->top_layer_at_element, i), bl = R1(m->bottom_layer_at_element, i);
      if (tl == 0) continue;
      int nv = tl - bl + 1;
      for (int k = bl + 1; k <= tl - 1; k++) {
        int kk = tl - k + 1;
        a[kk] = -dt * R0(s->Kv, k, i) / R0(s->dzhalf_cell, k, i);
        cc[kk] = -dt * R0(s->Kv, k - 1, i) / R0(s->dzhalf_cell, k - 1, i);
        b[kk] = -a[kk] + RL(s->dz_cell, k, i) - cc[kk];
      }
      if (tl == bl) {
        b[1] = RL(s->dz_cell, tl, i);
      } else {
        cc[1] = -dt + R0(s->Kv, tl - 1, i) / R0(s->dzhalf_cell, tl - 1, i);
        b[1] = RL(s->dz_cell, tl, i) - cc[1];
        a[nv] = -dt * R0(s->Kv, bl, i) / R0(s->dzhalf_cell, bl, i);
        b[nv] = -a[nv] + RL(s->dz_cell, bl, i);
      }
      for (int k = 1; k <= nv; k++) {
        int kk = tl - k + 1;
        rhs[k] = RL(s->dz_cell, kk, i) * RL(trhs, kk, i);
        rhs[65 + k] = RL(s->dz_cell, kk, i) * RL(srhs, kk, i);
      }
      int err = oracle_tridiag(nv, 2, 65, a + 1, b + 1, cc + 1, rhs + 1, sol + 1);
      if (err) {
        ret = fail(err, i);
        break;
      }
      for (int k = 1; k <= nv; k++) {
        int klev = tl + 1 - k;
        RL(s->temp_cell_new, klev, i) = sol[k];
        RL(s->salt_cell_new, klev, i) = sol[65 + k];
      }
      for (int k = 1; k <= bl - 1; k++) {
        RL(s->temp_cell_new, k, i) = RL(s->temp_cell_new, bl, i);
        RL(s->salt_cell_new, k, i) = RL(s->salt_cell_new, bl, i);
      }
      for (int k = tl + 1; k <= L; k++) {
        RL(s->temp_cell_new, k, i) = RL(s->temp_cell_new, tl, i);
        RL(s->salt_cell_new, k, i) = RL(s->salt_cell_new, tl, i);
      }
    }
  }
  if (!ret) {
    for (int i = 1; i <= P; i++) {
      int tn = R1(s->top_layer_at_node, i), bn = R1(m->bottom_layer_at_node, i);
      int nadj = R1(m->adj_cells_at_node, i);
      for (int k = tn; k >= bn; k--) {
        double sum0 = 0.0, sum1 = 0.0, sum2 = 0.0;
        int icount = 0;
        for (int l = 1; l <= nadj; l++) {
          int ie = INB(m->adj_cellnum_at_node, l, i);
          if (k >= R1(m->bottom_layer_at_element, ie) && k <= R1(s->top_layer_at_element, ie)) {
            sum0 = sum0 + R1(m->area, ie) * RL(s->dz_cell_new, k, ie);
            sum1 = sum1 + R1(m->area, ie) * RL(s->dz_cell_new, k, ie) * RL(s->temp_cell_new, k, ie);
            sum2 = sum2 + R1(m->area, ie) * RL(s->dz_cell_new, k, ie) * RL(s->salt_cell_new, k, ie);
          } else {
            icount++;
          }
        }
        if (sum0 == 0.0) {
          RL(s->temp_node, k, i) = 0.0;
          RL(s->salt_node, k, i) = 0.0;
        } else {
          RL(s->temp_node, k, i) = sum1 / sum0;
          RL(s->salt_node, k, i) = sum2 / sum0;
        }
        if (icount == nadj && k != tn) {
          RL(s->temp_node, k, i) = RL(s->temp_node, k + 1, i);
          RL(s->salt_node, k, i) = RL(s->salt_node, k + 1, i);
        }
      }
    }
    for (int i = 1; i <= P; i++) {
      int tn = R1(s->top_layer_at_node, i), bn = R1(m->bottom_layer_at_node, i);
      if (tn == 0) continue;
      for (int k = 1; k <= bn - 1; k++) {
        RL(s->temp_node, k, i) = RL(s->temp_node, bn, i);
        RL(s->salt_node, k, i) = RL(s->salt_node, bn, i);
      }
      for (int k = tn + 1; k <= L; k++) {
        RL(s->temp_node, k, i) = RL(s->temp_node, tn, i);
        RL(s->salt_node, k, i) = RL(s->salt_node, tn, i);
      }
    }
    for (int j = 1; j <= J; j++) {
      int n1 = I2(m->nodenum_at_face, 1, j), n2 = I2(m->nodenum_at_face, 2, j);
      for (int k = 1; k <= L; k++) {
        RL(s->salt_face, k, j) = (RL(s->salt_node, k, n1) + RL(s->salt_node, k, n2)) * 0.5;
        RL(s->temp_face, k, j) = (RL(s->temp_node, k, n1) + RL(s->temp_node, k, n2)) * 0.5;
      }
    }
  }
  (void)NB;
  free(tc);
  free(sc);
  free(trhs);
  free(srhs);
  return ret;
}

/* calculate_density_full.f90:13-247 / calculate_density_linear.f90:10-165 */
int oracle_eos(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
               const gomgpu_forcing *f) {
  (void)f;
  if (p->baroclinic_flag != 1) return 0;
  const int E = p->maxele, J = p->maxface, P = p->maxnod, L = p->maxlayer;
  const int full = (p->ana_density == 0);
  if (p->ana_density != 0 && p->ana_density != 1) return GOMGPU_E_ARG;
  loop_err e = {0, 0};
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= P; i++) {
    int tl = R1(s->top_layer_at_node, i);
    if (tl == 0) continue;
    int bl = R1(m->bottom_layer_at_node, i);
    double *r = &RL(s->rho_node, 1, i);
    int err = eos_column(m, p, L, full, bl, tl, &RL(s->temp_node, 1, i), &RL(s->salt_node, 1, i), r,
                         &RL(s->dz_node, 1, i), i, 1);
    if (err) {
      loop_fail(&e, err, i);
      continue;
    }
    for (int k = 1; k <= bl - 1; k++) r[k - 1] = r[bl - 1];
    for (int k = tl + 1; k <= L; k++) r[k - 1] = r[tl - 1];
  }
  if (e.code) return fail(e.code, e.entity);
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= J; j++) {
    int tl = R1(s->top_layer_at_face, j);
    if (tl == 0) continue;
    int bl = R1(m->bottom_layer_at_face, j);
    double *r = &RL(s->rho_face, 1, j);
    int err = eos_column(m, p, L, full, bl, tl, &RL(s->temp_face, 1, j), &RL(s->salt_face, 1, j), r,
                         &RL(s->dz_face, 1, j), j, full || p->i_density_flag == 1);
    if (err) {
      loop_fail(&e, err, j);
      continue;
    }
    for (int k = 1; k <= bl - 1; k++) r[k - 1] = r[bl - 1];
    for (int k = tl + 1; k <= L; k++) r[k - 1] = r[tl - 1];
  }
  if (e.code) return fail(e.code, e.entity);
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= E; i++) {
    int tl = R1(s->top_layer_at_element, i);
    if (tl == 0) continue;
    int bl = R1(m->bottom_layer_at_element, i), N = R1(m->tri_or_quad, i);
    for (int k = bl; k <= tl; k++) {
      double r = 0.0;
      int icount = 0;
      for (int l = 1; l <= N; l++) {
        int nd = I4(m->nodenum_at_cell, l, i);
        if (R1(s->top_layer_at_node, nd) != 0) {
          icount++;
          r = r + RL(s->rho_node, k, nd);
        }
      }
      if (icount == 0) {
        loop_fail(&e, GOMGPU_ERR_EOS_DRY_NODES, i);
        break;
      }
      RL(s->rho_cell, k, i) = r / icount;
    }
    for (int k = 1; k <= bl - 1; k++) RL(s->rho_cell, k, i) = RL(s->rho_cell, bl, i);
    for (int k = tl + 1; k <= L; k++) RL(s->rho_cell, k, i) = RL(s->rho_cell, tl, i);
  }
  if (e.code) return fail(e.code, e.entity);
  return 0;
}

/* ------------------------------------------------------------------- update -------- */
/* ------------------------------------------------------------------ turbulence ------ */
/* Vertical turbulence closure, `subroutine turbulence` of turbulence_closure_v1.f90:9-492.
 * The reference routine is not called by main.f90 and calls an undefined
 * algebraic_stress; here it is defined as the quasi-equilibrium stability functions of
 * Galperin et al. (1988) (stability_function GA) or Kantha & Clayson (1994) (KC), in the
 * GLS form where q2 is the turbulent kinetic energy k (q = sqrt(2 k) in MY notation):
 *   Gh = (g/rho_o) (drho/dz) xl^2 / (2 k), clipped to [-0.28, 0.0233];
 *   Km = Sm q xl, Kh = Sh q xl; k and psi diffusivities 0.2 q xl (MY, k-kl) or Km/schk,
 *   Km/schpsi (k-eps, k-omega, generic).
 * Quantities the reference reads from arrays it never sizes (diffmax(j), xlmin1(j),
 * xlmin2(j)) are scalar parameters; the diffusivities of level bottom-1 used by the
 * bottom layer are those of level bottom. Parity unpinned by any reference run (dead
 * code); pinned by closed forms in tests/test_turbulence.py. */
#define TK_A1 0.92
#define TK_A2 0.74
#define TK_B1 16.6
#define TK_B2 10.1
#define TK_C1 0.08
#define TK_C2 0.7
#define TK_C3 0.2

/* x**e for the GLS exponents: the ones the models use (0, +-1, 2, 3, +-0.5, 1.5) by exact
 * arithmetic, so host and device agree bit for bit; others through pow. */
static double gls_pow(double x, double e) {
  if (e == 0.0) return 1.0;
  if (e == 1.0) return x;
  if (e == -1.0) return 1.0 / x;
  if (e == 2.0) return x * x;
  if (e == 3.0) return x * x * x;
  if (e == 0.5) return sqrt(x);
  if (e == -0.5) return 1.0 / sqrt(x);
  if (e == 1.5) return x * sqrt(x);
  if (e == -1.5) return 1.0 / (x * sqrt(x));
  if (e == -3.0) return 1.0 / (x * x * x);
  return pow(x, e);
}

/* Sm, Sh of the quasi-equilibrium closure at Gh (clipped). */
void oracle_stability(int kind, double gh, double *sm, double *sh) {
  if (gh > 0.0233) gh = 0.0233;
  if (gh < -0.28) gh = -0.28;
  double b2 = kind == GOMGPU_STAB_KC ? TK_B2 * (1.0 - TK_C3) : TK_B2;
  double a2 = kind == GOMGPU_STAB_KC ? TK_A2 * (1.0 - TK_C2) : TK_A2;
  double h = TK_A2 * (1.0 - 6.0 * TK_A1 / TK_B1) / (1.0 - 3.0 * TK_A2 * gh * (6.0 * TK_A1 + b2));
  double m = (TK_A1 * (1.0 - 3.0 * TK_C1 - 6.0 * TK_A1 / TK_B1) + 9.0 * TK_A1 * (2.0 * TK_A1 + a2) * h * gh) /
             (1.0 - 9.0 * TK_A1 * TK_A2 * gh);
  *sm = m;
  *sh = h;
}

/* algebraic_stress: vd, td, qd, qd2 at one whole level, clipped to [bgdiff, diffmax]. */
static void alg_stress(const gomgpu_params *p, double q2, double xl, double rz, double *vd, double *td,
                       double *qd, double *qd2) {
  double gh = rz * (xl * xl) / (2.0 * q2);
  double sm, sh;
  oracle_stability(p->stability_function, gh, &sm, &sh);
  double q = sqrt(2.0 * q2);
  double v = sm * q * xl, t = sh * q * xl, a, b;
  if (p->turbulence_model == GOMGPU_GLS_MY || p->turbulence_model == GOMGPU_GLS_KL) {
    a = 0.2 * q * xl;
    b = a;
  } else {
    a = v / p->schk;
    b = v / p->schpsi;
  }
  *vd = dmin(p->diffmax, dmax(p->bgdiff, v));
  *td = dmin(p->diffmax, dmax(p->bgdiff, t));
  *qd = dmin(p->diffmax, dmax(p->bgdiff, a));
  *qd2 = dmin(p->diffmax, dmax(p->bgdiff, b));
}

/* (g/rho_o) drho/dz at whole level k of face j from the face densities (0 at the top). */
static double rz_level(const gomgpu_params *p, const gomgpu_state *s, int L, int j, int k, int kt) {
  if (k >= kt) return 0.0;
  double drhodz = (RL(s->rho_face, k + 1, j) - RL(s->rho_face, k, j)) / R0(s->dzhalf_face_new, k, j);
  return p->gravity / p->rho_o * drhodz;
}

/* Face diffusivity -> cell Kv: mean over the cell's wet faces of the face value at the
 * level clamped to the face's active range; a cell with no wet face gets 1.0e-2 (the
 * fallback of the node mean at turbulence_closure_v1.f90:456-460). */
static void kv_cells(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s, const double *kvf) {
  const int E = p->maxele, L = p->maxlayer;
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= E; i++) {
    if (R1(s->top_layer_at_element, i) == 0) continue;
    int N = R1(m->tri_or_quad, i);
    for (int k = 0; k <= L; k++) {
      double sum = 0.0;
      int cnt = 0;
      for (int l = 1; l <= N; l++) {
        int f = I4(m->facenum_at_cell, l, i);
        int kt = R1(s->top_layer_at_face, f);
        if (kt == 0) continue;
        int kb = R1(m->bottom_layer_at_face, f);
        int kk = k < kb ? kb : (k > kt ? kt : k);
        sum = sum + kvf[kk + (size_t)(f - 1) * (L + 1)];
        cnt++;
      }
      R0(s->Kv, k, i) = cnt == 0 ? (double)1.0e-2f : sum / cnt;
    }
  }
}

int oracle_turbulence(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                      const gomgpu_forcing *f) {
  const int J = p->maxface, L = p->maxlayer;
  const int flag = p->turbulence_flag;
  if (flag == 0) return 0;
  if (flag < 0 || flag > 3 || L + 1 > 65) return GOMGPU_E_ARG;
  double *kvf = calloc((size_t)(L + 1) * J, sizeof(double));
  if (!kvf) return GOMGPU_E_NOMEM;
  const double dt = p->dt, g = p->gravity;
  const double fs_coef = (double)powf(16.6f, 2.0f / 3.0f);  /* 16.6**(2.0/3), default real */
  int err = 0;
  loop_err e = {0, 0};
#pragma omp parallel for schedule(dynamic, 64)
  for (int j = 1; j <= J; j++) {
    int kt = R1(s->top_layer_at_face, j);
    if (kt == 0) continue;
    int ferr = 0;
    int kb = R1(m->bottom_layer_at_face, j);
    double *KV = &kvf[(size_t)(j - 1) * (L + 1)];
    if (flag == 1 || flag == 2) {
      /* Richardson number at whole levels (:75-89) */
      double vmax = 0.0, vmin = 0.0;
      if (flag == 2) { /* :120-134 */
        double h = R1(m->h_face, j);
        if (h <= p->h1_pp) {
          vmax = p->vdmax_pp1;
          vmin = p->vdmin_pp1;
        } else if (h < p->h2_pp) {
          vmax = p->vdmax_pp1 + (p->vdmax_pp2 - p->vdmax_pp1) * (h - p->h1_pp) / (p->h2_pp - p->h1_pp);
          vmin = p->vdmin_pp1 + (p->vdmin_pp2 - p->vdmin_pp1) * (h - p->h1_pp) / (p->h2_pp - p->h1_pp);
        } else {
          vmax = p->vdmax_pp2;
          vmin = p->vdmin_pp2;
        }
      }
      for (int k = kb; k <= kt; k++) {
        double ri;
        if (k == kt) {
          ri = 0.0;
        } else {
          double dzh = R0(s->dzhalf_face_new, k, j);
          double drhodz = (RL(s->rho_face, k + 1, j) - RL(s->rho_face, k, j)) / dzh;
          double bvf = -g * (drhodz / p->rho_o + g / 2250000.0);
          double dundz = (RL(s->un_face_new, k + 1, j) - RL(s->un_face_new, k, j)) / dzh;
          double dutdz = (RL(s->vn_face_new, k + 1, j) - RL(s->vn_face_new, k, j)) / dzh;
          double shear2 = dundz * dundz + dutdz * dutdz;
          shear2 = dmax(shear2, 1.0e-10);
          ri = dmax(bvf / shear2, 0.0);
        }
        double av, kv;
        if (flag == 1) { /* :94-107 */
          if (ri < 0.25) {
            av = 1.0e-3;
            kv = 1.0 * 1.0e-3;
          } else {
            av = 1.0e-5;
            kv = 1.0e-5;
          }
        } else { /* :136-141 */
          double t = 1.0 + 5.0 * ri;
          av = vmax / (t * t) + vmin;
          kv = av / (1.0 + 5.0 * ri) + p->tdmin_pp;
        }
        R0(s->Av, k, j) = av;
        KV[k] = kv;
      }
      continue;
    }
    /* scheme 3 (:143-429) */
    double avc[66] = {0}, kvc[66] = {0}, qdc[66], qd2c[66];
    double q2bt[66], xlbt[66], rzbt[66], shearbt[66], q2tmp[66], xltmp[66];
    double a[66], b[66], c[66], rhs[66], sol[66];
    for (int k = kb; k <= kt; k++) /* initial diffusivities (:56-66) */
      alg_stress(p, R0(s->q2, k, j), R0(s->xl, k, j), rz_level(p, s, L, j, k, kt), &avc[k], &kvc[k], &qdc[k], &qd2c[k]);
    avc[kb - 1] = avc[kb];
    kvc[kb - 1] = kvc[kb];
    int n1 = I2(m->nodenum_at_face, 1, j), n2 = I2(m->nodenum_at_face, 2, j);
    for (int k = kb; k <= kt; k++) { /* :158-187 */
      q2bt[k] = (R0(s->q2, k - 1, j) + R0(s->q2, k, j)) / 2.0;
      xlbt[k] = (R0(s->xl, k - 1, j) + R0(s->xl, k, j)) / 2.0;
      if (xlbt[k] < p->xlmin2) {
        ferr = GOMGPU_ERR_TURB_XL;
        break;
      }
      double udown = (R0(s->u_node, k - 1, n1) + R0(s->u_node, k - 1, n2)) / 2.0;
      double uup = (R0(s->u_node, k, n1) + R0(s->u_node, k, n2)) / 2.0;
      double vdown = (R0(s->v_node, k - 1, n1) + R0(s->v_node, k - 1, n2)) / 2.0;
      double vup = (R0(s->v_node, k, n1) + R0(s->v_node, k, n2)) / 2.0;
      double dzk = RL(s->dz_face_new, k, j);
      double dudz = (uup - udown) / dzk, dvdz = (vup - vdown) / dzk;
      shearbt[k] = dudz * dudz + dvdz * dvdz;
      double drhodz;
      if (k == kt || k == kb) {
        drhodz = 0.0;
      } else {
        double rho_up = RL(s->rho_face, k, j) + dzk / 2.0 / R0(s->dzhalf_face_new, k, j) *
                                                    (RL(s->rho_face, k + 1, j) - RL(s->rho_face, k, j));
        double rho_down = RL(s->rho_face, k - 1, j) + RL(s->dz_face_new, k - 1, j) / 2.0 / R0(s->dzhalf_face_new, k - 1, j) *
                                                          (RL(s->rho_face, k, j) - RL(s->rho_face, k - 1, j));
        drhodz = (rho_up - rho_down) / dzk;
      }
      rzbt[k] = g / p->rho_o * drhodz;
    }
    if (ferr) {
      loop_fail(&e, ferr, j);
      continue;
    }
    int a1 = I2(m->adj_cellnum_at_face, 1, j), a2 = I2(m->adj_cellnum_at_face, 2, j);
    double etam = a2 != 0 ? dmax(R1(s->eta_cell_new, a1), R1(s->eta_cell_new, a2)) : R1(s->eta_cell_new, a1);
    double hf = R1(m->h_face, j);
    for (int k = kb; k <= kt; k++) { /* :198-215 */
      double dzk = RL(s->dz_face_new, k, j);
      double zctr = (k == kt) ? p->MSL + etam - dzk / 2.0 : ZL(k) - dzk / 2.0;
      double dists = p->MSL + etam - zctr, distb = zctr - (p->MSL - hf);
      double xlmax = (double)0.4f * dists * distb / (hf + etam);
      if (xlmax <= 0.0) {
        ferr = GOMGPU_ERR_TURB_DIST;
        break;
      }
    }
    if (ferr) {
      loop_fail(&e, ferr, j);
      continue;
    }
    /* boundary values (:219-227) */
    double wn = f->wind_stress_normal[j - 1], wt = f->wind_stress_tangnt[j - 1];
    double q2fs = fs_coef * sqrt(wn * wn + wt * wt) / 2.0;
    q2fs = dmax(q2fs, p->q2min);
    double ubn = RL(s->un_face_new, kb, j), ubt = RL(s->vn_face_new, kb, j);
    double q2bot = fs_coef * R1(s->Cdb, j) * sqrt(ubn * ubn + ubt * ubt) / 2.0;
    q2bot = dmax(q2bot, p->q2min);
    double xlfs = dmax(p->xlmin1, RL(s->dz_face_new, kt, j) / 2.0 * (double)0.4f);
    double xlbot = dmax(p->xlmin2, RL(s->dz_face_new, kb, j) / 2.0 * (double)0.4f);
    int nq = kt - kb + 1;
    double cm3 = p->cmiu0 * p->cmiu0 * p->cmiu0;
    for (int k = 1; k <= nq; k++) { /* q2 matrix (:234-264) */
      int kl = kt - k + 1;
      double dzk = RL(s->dz_face_new, kl, j);
      double eleft = 0.0, eright = 0.0;
      if (k > 1) {
        eleft = -(qdc[kl] * dt / R0(s->dzhalf_face_new, kl, j));
        a[k - 1] = eleft;
      }
      if (k < nq) {
        eright = -(qdc[kl - 1] * dt / R0(s->dzhalf_face_new, kl - 1, j));
        c[k - 1] = eright;
      }
      double prod = (avc[kl] + avc[kl - 1]) / 2.0 * shearbt[kl];
      double buoy = (kvc[kl] + kvc[kl - 1]) / 2.0 * rzbt[kl];
      double diss = cm3 * sqrt(q2bt[kl]) / xlbt[kl];
      double pplus, pminus;
      if (prod + buoy > 0.0) {
        pplus = prod + buoy;
        pminus = 0.0;
      } else {
        pplus = 0.0;
        pminus = -prod - buoy;
      }
      b[k - 1] = dzk - eleft - eright + (diss + pminus / q2bt[kl]) * dt * dzk;
      rhs[k - 1] = dzk * (q2bt[kl] + dt * pplus);
    }
    if (nq >= 1) {
      a[0] = 0.0;
      c[nq - 1] = 0.0;
    }
    ferr = oracle_tridiag(nq, 1, 66, a, b, c, rhs, sol);
    if (ferr) {
      loop_fail(&e, ferr, j);
      continue;
    }
    for (int k = 1; k <= nq; k++) { /* :268-278 */
      int kl = kt - k + 1;
      q2tmp[kl] = kl == kt ? q2fs : (kl == kb ? q2bot : dmax(sol[k - 1], p->q2min));
    }
    for (int k = 1; k <= nq; k++) { /* psi matrix (:281-347) */
      int kl = kt - k + 1;
      double dzk = RL(s->dz_face_new, kl, j);
      double eleft = 0.0, eright = 0.0;
      if (k > 1) {
        eleft = -(qd2c[kl] * dt / R0(s->dzhalf_face_new, kl, j));
        a[k - 1] = eleft;
      }
      if (k < nq) {
        eright = -(qd2c[kl - 1] * dt / R0(s->dzhalf_face_new, kl - 1, j));
        c[k - 1] = eright;
      }
      double bb = dzk - eleft - eright;
      double cpsi2p;
      if (p->turbulence_model == GOMGPU_GLS_MY || p->turbulence_model == GOMGPU_GLS_KL) {
        double zctr = ZL(kl) - dzk / 2.0;
        double dists = dmax(p->MSL + etam - zctr, p->xlmin2);
        double distb = dmax(zctr - (p->MSL - hf), p->xlmin2);
        double r1 = xlbt[kl] / (double)0.4f / distb, r2 = xlbt[kl] / (double)0.4f / dists;
        double psi = 1.0 + (double)1.33f * (r1 * r1) + (double)0.25f * (r2 * r2);
        cpsi2p = psi * p->cpsi2;
      } else {
        cpsi2p = p->cpsi2;
      }
      double diss = cpsi2p * cm3 * sqrt(q2bt[kl]) / xlbt[kl];
      bb = bb + dt * dzk * diss;
      double prod = p->cpsi1 * (avc[kl] + avc[kl - 1]) / 2.0 * shearbt[kl];
      double cpsi3;
      if (p->turbulence_model == GOMGPU_GLS_MY) {
        cpsi3 = (double)0.9f;
      } else if (rzbt[kl] > 0.0) {
        cpsi3 = 1.0;
      } else {
        switch (p->turbulence_model) {
          case GOMGPU_GLS_KL: cpsi3 = (double)2.53f; break;
          case GOMGPU_GLS_KE: cpsi3 = (double)(-0.52f); break;
          case GOMGPU_GLS_KW: cpsi3 = (double)(-0.58f); break;
          default: cpsi3 = (double)0.1f; break;
        }
      }
      double buoy = cpsi3 * (kvc[kl] + kvc[kl - 1]) / 2.0 * rzbt[kl];
      double pplus, pminus;
      if (prod + buoy > 0.0) {
        pplus = prod + buoy;
        pminus = 0.0;
      } else {
        pplus = 0.0;
        pminus = -prod - buoy;
      }
      bb = bb + pminus / q2bt[kl] * dt * dzk;
      if (kl == kb || kl == kt) bb = bb + dt * (double)0.4f * p->rnub * qd2c[kl] / xlbt[kl];
      b[k - 1] = bb;
      double tmp = gls_pow(p->cmiu0, p->rpub) * gls_pow(q2bt[kl], p->rmub) * gls_pow(xlbt[kl], p->rnub);
      rhs[k - 1] = dzk * tmp * (1.0 + dt * pplus / q2bt[kl]);
    }
    if (nq >= 1) {
      a[0] = 0.0;
      c[nq - 1] = 0.0;
    }
    ferr = oracle_tridiag(nq, 1, 66, a, b, c, rhs, sol);
    if (ferr) {
      loop_fail(&e, ferr, j);
      continue;
    }
    for (int k = 1; k <= nq; k++) { /* :351-368 */
      int kl = kt - k + 1;
      double q2l = dmax(sol[k - 1], p->psimin);
      double x;
      if (kl == kt)
        x = xlfs;
      else if (kl == kb)
        x = xlbot;
      else
        x = gls_pow(q2l * gls_pow(p->cmiu0, -p->rpub) * gls_pow(q2tmp[kl], -p->rmub), 1.0 / p->rnub);
      if (rzbt[kl] < 0.0) {
        double upper = sqrt(-(double)0.56f * q2tmp[kl] / rzbt[kl]);
        x = dmin(x, upper);
      }
      double cq = p->cmiu0 * sqrt(q2tmp[kl]);
      double xl_max = cq * cq * cq / p->eps_min;
      xltmp[kl] = dmax(p->xlmin2, dmin(xl_max, x));
    }
    for (int k = kb - 1; k <= kt; k++) { /* whole levels (:372-390) */
      double q, x;
      if (k == kb - 1) {
        q = q2bot;
        x = xlbot;
      } else if (k == kt) {
        q = q2fs;
        x = xlfs;
      } else {
        double r = RL(s->dz_face_new, k, j) / 2.0 / R0(s->dzhalf_face_new, k, j);
        q = q2tmp[k] + r * (q2tmp[k + 1] - q2tmp[k]);
        x = xltmp[k] + r * (xltmp[k + 1] - xltmp[k]);
        if (q < 0.0 || x < p->xlmin2) {
          ferr = GOMGPU_ERR_TURB_NEGATIVE;
          break;
        }
      }
      R0(s->q2, k, j) = q;
      R0(s->xl, k, j) = x;
    }
    if (ferr) {
      loop_fail(&e, ferr, j);
      continue;
    }
    for (int k = kb; k <= kt; k++) { /* new diffusivities (:397-405) */
      double vd, td, qd, qd2;
      alg_stress(p, R0(s->q2, k, j), R0(s->xl, k, j), rz_level(p, s, L, j, k, kt), &vd, &td, &qd, &qd2);
      R0(s->Av, k, j) = vd;
      KV[k] = td;
    }
    for (int k = 0; k <= kb - 2; k++) { /* extension (:408-427) */
      R0(s->q2, k, j) = R0(s->q2, kb - 1, j);
      R0(s->xl, k, j) = R0(s->xl, kb - 1, j);
    }
    for (int k = 1; k <= kb - 1; k++) {
      R0(s->Av, k, j) = R0(s->Av, kb, j);
      KV[k] = KV[kb];
    }
    for (int k = kt + 1; k <= L; k++) {
      R0(s->q2, k, j) = R0(s->q2, kt, j);
      R0(s->xl, k, j) = R0(s->xl, kt, j);
      R0(s->Av, k, j) = R0(s->Av, kt, j);
      KV[k] = KV[kt];
    }
  }
  if (e.code) err = fail(e.code, e.entity);
  if (!err) kv_cells(m, p, s, kvf);
  free(kvf);
  return err;
}

/* update_variables.f90:4-77 */
int oracle_update(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                  const gomgpu_forcing *f) {
  (void)f;
  const int E = p->maxele, J = p->maxface, P = p->maxnod, L = p->maxlayer;
  memcpy(s->eta_cell, s->eta_cell_new, sizeof(double) * E);
  memcpy(s->wn_cell, s->wn_cell_new, sizeof(double) * (L + 1) * E);
  memcpy(s->dz_cell, s->dz_cell_new, sizeof(double) * L * E);
  memcpy(s->dzhalf_cell, s->dzhalf_cell_new, sizeof(double) * (L + 1) * E);
  memcpy(s->salt_cell, s->salt_cell_new, sizeof(double) * L * E);
  memcpy(s->temp_cell, s->temp_cell_new, sizeof(double) * L * E);
  memcpy(s->un_face, s->un_face_new, sizeof(double) * L * J);
  memcpy(s->vn_face, s->vn_face_new, sizeof(double) * L * J);
  memcpy(s->dz_face, s->dz_face_new, sizeof(double) * L * J);
  memcpy(s->dzhalf_face, s->dzhalf_face_new, sizeof(double) * (L + 1) * J);
  memcpy(s->dz_node, s->dz_node_new, sizeof(double) * L * P);
  memcpy(s->dzhalf_node, s->dzhalf_node_new, sizeof(double) * (L + 1) * P);
  return oracle_bars(m, p, s, f);
}

/* Depth means at nodes: update_variables.f90:37-74, the same as
 * calculate_vertically_averaged_values, which prepare_gom.f90:172 calls at start-up. */
int oracle_bars(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s, const gomgpu_forcing *f) {
  (void)f;
  const int P = p->maxnod, L = p->maxlayer;
  for (int i = 1; i <= P; i++) {
    int tl = R1(s->top_layer_at_node, i), bl = R1(m->bottom_layer_at_node, i);
    int nv = tl - bl + 1;
    if (nv == 0) {
      R1(s->ubar_node, i) = 0.0;
      R1(s->vbar_node, i) = 0.0;
      R1(s->sbar_node, i) = 0.0;
      R1(s->tbar_node, i) = 0.0;
      R1(s->rbar_node, i) = 0.0;
      continue;
    }
    double su = 0.0, sv = 0.0;
    for (int k = bl - 1; k <= tl; k++) {
      su = su + R0(s->u_node, k, i);
      sv = sv + R0(s->v_node, k, i);
    }
    R1(s->ubar_node, i) = su / (nv + 1);
    R1(s->vbar_node, i) = sv / (nv + 1);
    double ss = 0.0, st = 0.0, sr = 0.0;
    for (int k = bl; k <= tl; k++) {
      ss = ss + RL(s->salt_node, k, i);
      st = st + RL(s->temp_node, k, i);
      sr = sr + RL(s->rho_node, k, i);
    }
    R1(s->sbar_node, i) = ss / nv;
    R1(s->tbar_node, i) = st / nv;
    R1(s->rbar_node, i) = sr / nv;
  }
  return 0;
}

/* read_hurricane_ser_1.f90:13-63 (the snapshot read replaced by f->wind_*_next) and the
 * hurricane branch of calculate_wind_stress.f90:88-114; f->hurricane_phase selects what
 * main.f90:116-123 does this step (0 outside the window, 1 interpolate, 2 next snapshot).
 * The forcing arrays are this stage's state and outputs, written in place. */
int oracle_wind(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                const gomgpu_forcing *f) {
  (void)s;
  if (p->hurricane_flag != 1) return 0;
  const int P = p->maxnod, J = p->maxface;
  double *u0 = (double *)f->wind_u0, *v0 = (double *)f->wind_v0, *p0 = (double *)f->air_p0;
  double *u1 = (double *)f->wind_u1, *v1 = (double *)f->wind_v1, *p1 = (double *)f->air_p1;
  double *u2 = (double *)f->wind_u2, *v2 = (double *)f->wind_v2, *p2 = (double *)f->air_p2;
  double *uf = (double *)f->wind_u_at_face, *vf = (double *)f->wind_v_at_face;
  const double phase = f->hurricane_phase;
  if (phase == 2.0) {
    for (int i = 1; i <= P; i++) {
      R1(u0, i) = R1(u2, i);
      R1(v0, i) = R1(v2, i);
      R1(p0, i) = R1(p2, i);
      R1(u1, i) = R1(u2, i);
      R1(v1, i) = R1(v2, i);
      R1(p1, i) = R1(p2, i);
    }
    for (int i = 1; i <= P; i++) {
      R1(u2, i) = R1(f->wind_u_next, i);
      R1(v2, i) = R1(f->wind_v_next, i);
      R1(p2, i) = R1(f->air_p_next, i);
    }
  } else if (phase == 1.0) {
    /* real(hurricane_read_interval): a default real, exact for these counts */
    const double rn = (double)(float)p->hurricane_read_interval;
    for (int i = 1; i <= P; i++) {
      double inc = (R1(u2, i) - R1(u0, i)) / rn;
      R1(u1, i) = R1(u1, i) + inc;
      inc = (R1(v2, i) - R1(v0, i)) / rn;
      R1(v1, i) = R1(v1, i) + inc;
      inc = (R1(p2, i) - R1(p0, i)) / rn;
      R1(p1, i) = R1(p1, i) + inc;
    }
  }
  if (phase != 0.0) {
    for (int j = 1; j <= J; j++) {
      int n1 = I2(m->nodenum_at_face, 1, j), n2 = I2(m->nodenum_at_face, 2, j);
      R1(uf, j) = (R1(u1, n1) + R1(u1, n2)) * 0.5;
      R1(vf, j) = (R1(v1, n1) + R1(v1, n2)) * 0.5;
    }
  }
  const double rho_ratio = p->rho_a / p->rho_o;
  for (int j = 1; j <= J; j++) {
    double ws = sqrt(R1(uf, j) * R1(uf, j) + R1(vf, j) * R1(vf, j));
    double wn = R1(uf, j) * R1(m->cos_theta, j) + R1(vf, j) * R1(m->sin_theta, j);
    double wt = -R1(uf, j) * R1(m->sin_theta, j) + R1(vf, j) * R1(m->cos_theta, j);
    double cda;
    if (p->wind_formular == 1)
      cda = 0.001 * (0.75 + 0.067 * ws);
    else if (p->wind_formular == 3)
      cda = 0.001 * (0.80 + 0.065 * ws);
    else
      cda = 0.001 * (0.61 + 0.063 * ws);
    cda = dmin((double)0.003f, cda);
    double g = cda * rho_ratio * ws;
    ((double *)f->Gamma_T)[j - 1] = g;
    ((double *)f->wind_stress_normal)[j - 1] = g * wn * f->spinup_function_wind;
    ((double *)f->wind_stress_tangnt)[j - 1] = g * wt * f->spinup_function_wind;
  }
  return 0;
}

int oracle_step(const gomgpu_mesh *m, const gomgpu_params *p, gomgpu_state *s,
                const gomgpu_forcing *f) {
  if (p->hurricane_flag == 1) {
    int err = oracle_wind(m, p, s, f);
    if (err) return err;
  }
  int (*stages[])(const gomgpu_mesh *, const gomgpu_params *, gomgpu_state *, const gomgpu_forcing *) = {
      oracle_friction, oracle_elm, oracle_momentum, oracle_freesurf, oracle_velocity,
      oracle_layers, oracle_nodevel, oracle_turbulence, oracle_transport, oracle_eos, oracle_update};
  for (int i = 0; i < 11; i++) {
    if (i == 7 && p->turbulence_flag == 0) continue;
    if (i == 8 && p->transport_flag != 1) continue;
    int err = stages[i](m, p, s, f);
    if (err) return err;
  }
  return 0;
}
