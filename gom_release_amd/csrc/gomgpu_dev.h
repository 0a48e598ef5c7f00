// gomgpu_dev.h - device-side context and layout helpers of the HIP step.
//
// Device layout: every (layer, entity) field of the reference is stored layer-major with
// the entity index fastest, d[(k - lo) * N + e] with e 0-based, so that one thread per
// entity (column kernels) or per (layer, entity) reads consecutive addresses across a
// wavefront. Connectivity tables (4, E), (2, J), (NB, P) are stored the same way
// ([slot][entity]). Entity ids inside the tables stay 1-based with 0 = none, as in the
// reference, so the host uploads them unchanged.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gomgpu.h"

#define GOM_LMAX 64  // largest supported maxlayer (per-thread column arrays)
#define GOM_NSCAL 8  // per-step device scalars (DevForcing::scal)
#define GOM_MAX_FORCING 32  // forcing arrays the ABI can hold (GOMGPU_FORCING_REAL has fewer)
#define GOM_NV_LEVELS 3  // node levels per thread of k_node_vel
#define GOM_TVD_NODE_LEVELS 4  // node levels per thread of k_tvd_node
#define GOM_KV_LEVELS 3  // cell levels per thread of k_turb_kv
#define GOM_CUV_LAYERS 4  // cell layers per thread of k_cell_uv
#define PCG_MAX_SLOT_BLOCKS 256  // most blocks of one system in a slot-exchange PCG launch

struct DevMesh {
#define DM_R(name, ext) const double *name;
#define DM_I(name, ext) const int32_t *name;
  GOMGPU_MESH_REAL(DM_R)
  GOMGPU_MESH_INT(DM_I)
#undef DM_R
#undef DM_I
  const double *Cdb_static;  // Manning drag from the still-water depth (bf_flag 1)
  // Static node stencil of calculate_velocity_at_node.f90:67-138, built at create: for the
  // jj-th adjacent cell of node i, its two faces touching the node (slots 2jj, 2jj+1) with
  // their length, boundary factor and bottom layer, and the cell with its area and bottom
  // layer; [slot][node] like the other tables. ns_weight is the node's sum of fw/L.
  const int32_t *ns_face, *ns_fkb, *ns_cell, *ns_ckb;
  const double *ns_Lf, *ns_fw, *ns_area, *ns_weight;
  // Cell records of the ELM walk, built at create: the x and y of a cell's nodes, the node
  // ids and the neighbour cells in slot order, 4 per cell, contiguous per cell
  // ([cell][slot]), so a visited cell costs a few aligned loads instead of a chain of
  // indirect ones.
  const double *cg_x, *cg_y;
  const int32_t *cg_node, *cg_adj;  // node ids and neighbour cells (1-based, 0 = none)
  // Per cell slot (TB layout, like facenum_at_cell): sign_in_outflow of the slot's face when
  // the transport counts its flux (boundary_type_of_face != -1 or a river face), else 0 --
  // one coalesced load in place of the sign and two loads indexed by the face.
  const int32_t *slot_flow;
};

struct DevState {
#define DS_R(name, ext) double *name;
#define DS_I(name, ext) int32_t *name;
  GOMGPU_STATE_REAL(DS_R)
  GOMGPU_STATE_INT(DS_I)
#undef DS_R
#undef DS_I
};

struct DevForcing {
#define DF_R(name, ext) const double *name;
  GOMGPU_FORCING_REAL(DF_R)
#undef DF_R
  // per-step scalars in device memory (so a captured step stays valid): [0] baroclinic
  // spin-up factor, [1] tidal spin-up factor, [2] elapsed time, [4] wind spin-up factor,
  // [5] hurricane phase (0 keep, 1 interpolate, 2 next snapshot)
  const double *scal;
};

// Scalars every kernel needs, passed by value.
struct DevDims {
  int P, E, J, L, NB, NT, NOB, NQB;
  int tid[2];
};

// Work arrays of the step (device-owned, sized at create time).
// Register budget of the latency-bound column and gather kernels: the compiler keeps each
// within what the given number of waves per SIMD leaves (amdgpu_waves_per_eu), trading
// registers for occupancy.
#ifndef GOM_WAVES_ELM
#define GOM_WAVES_ELM 4
#endif
#ifndef GOM_WAVES_MOM_RHS
#define GOM_WAVES_MOM_RHS 1
#endif
#ifndef GOM_WAVES_MOM_SOLVE
#define GOM_WAVES_MOM_SOLVE 1
#endif
#ifndef GOM_WAVES_NODE_VEL
#define GOM_WAVES_NODE_VEL 1
#endif
#ifndef GOM_WAVES_TURB
#define GOM_WAVES_TURB 4
#endif
#define GOM_WAVES(n) __attribute__((amdgpu_waves_per_eu(n)))

struct DevWork {
  // free surface / PCG (E rows, ELL with the diagonal and 4 neighbour slots)
  double *fs_diag, *fs_off, *fs_rhs, *pcg_x, *pcg_r, *pcg_d0, *pcg_d1, *pcg_q, *pcg_s, *pcg_minv;
  double *pcg_part;       // per-block partial sums (4 buffers)
  double *pcg_pub;        // 6 E published row values (single-barrier PCG)
  unsigned long long *pcg_slots;  // 6 * PCG_MAX_SLOT_BLOCKS partial-sum slots
  int *pcg_info;          // [0] iterations of the last solve, [1] converged flag, [2] exchange abort
  unsigned *pcg_bar;      // grid barrier state
  double *face_sdz;       // per face sum_k dz * AinvDeltaZ1 (top down)
  double *face_sun;       // per face sum_k dz * un
  double *face_sg;        // per face sum_k dz * AinvG1
  // transport
  double *con, *con_new;  // con (0:L+1, E, NT) and con_new (L, E, NT), layer-major
  double *con2;           // the sub-cycles' second con buffer (con and con2 alternate)
  double *htop;           // (2, E, NT) horizontal upwind sums of each column's top layer
  double *fqd, *fps;      // (2, L, J) each: per face-layer (Q, d) and the PSI of both tracers
  double *phij;           // (L, J)
  double *Qi, *di, *phii; // (0:L, E)
  double *PSIi;
  double *s1a, *s2a;      // (L, E, NT) limiter sums per cell-layer
  double *obk, *qbk;      // (L, OB, NT), (L, QB, NT)
  double *node_new;       // (L, P, NT)
  double *face_new;       // (L, J, NT)
  // turbulence closure: face diffusivity (0:L, J) before the face -> cell map
  double *turb_kvf;
  // elimination factors of the column tridiagonal solves, (L, max(E, J))
  double *tri_gam;
  double *tvd_x;          // (L, E, NT) right-hand sides / solution of the TVD column solve
  double *cell_u, *cell_v;  // (L, E) cell means of the mid-layer node velocities (momentum)
  // two-equation closure (allocated when turbulence_flag == 3): level diffusivities of the
  // current state (0:L, J) and layer quantities / solutions (L, J)
  double *tk_shear, *tk_rz, *tk_q2tmp, *tk_x;
  // budgets
  double *budget;         // [0] volume, [1] salt
  int *err;               // [0] code, [1] entity
  // departure points within 1e-12 of a cell edge, counted since creation or the last reset:
  // [0] momentum ELM, [1] tracer ELM
  unsigned long long *elm_near;
  int32_t *elm_dep;       // departure cells then levels per face-layer (2 L J), NULL unless recorded
};

// ------------------------------------------------------------------ index helpers
__device__ __forceinline__ size_t ix(int krow, int e, int n) { return (size_t)krow * n + e; }
// 1-based layer k of an (L, N) field
#define AL(a, k, e, n) (a)[ix((k) - 1, (e), (n))]
// 0-based layer k of an (0:L, N) field
#define A0(a, k, e, n) (a)[ix((k), (e), (n))]
// 1-based slot l of a connectivity table [slot][entity]
#define TB(a, l, e, n) (a)[ix((l) - 1, (e), (n))]

__device__ __forceinline__ int sen(int N, int l, int w) {
  int v = l + w;
  return v > N ? v - N : v;
}

__device__ __forceinline__ double carea(double x1, double x2, double x3, double y1, double y2, double y3) {
  return fabs((x1 - x3) * (y2 - y3) - (x2 - x3) * (y1 - y3)) / 2.0;
}

// ---- correctly rounded natural logarithm (host and device)
// The reference evaluates log with the host C library, whose log is correctly rounded in
// all but vanishingly rare cases; the device library's log is not, and differs from it in
// the last bit for some arguments. gom_log computes log(x) to about 2^-100 relative in
// double-double arithmetic and rounds once, so it returns the correctly rounded value.
// Method: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log x = e ln2 + 2 atanh(s),
// s = (m - 1)/(m + 1), |s| <= 0.1716, atanh(s)/s = sum_k s^(2k)/(2k+1) to k = 22.
struct GomDD {
  double hi, lo;
};
__host__ __device__ inline GomDD dd_two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ inline GomDD dd_quick(double a, double b) {  // |a| >= |b|
  double s = a + b;
  return {s, b - (s - a)};
}
__host__ __device__ inline GomDD dd_add(GomDD a, GomDD b) {
  GomDD s = dd_two_sum(a.hi, b.hi), t = dd_two_sum(a.lo, b.lo);
  s.lo = s.lo + t.hi;
  s = dd_quick(s.hi, s.lo);
  s.lo = s.lo + t.lo;
  return dd_quick(s.hi, s.lo);
}
__host__ __device__ inline GomDD dd_mul(GomDD a, GomDD b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e = e + (a.hi * b.lo + a.lo * b.hi);
  return dd_quick(p, e);
}
__host__ __device__ inline GomDD dd_div(GomDD a, GomDD b) {
  double q1 = a.hi / b.hi;
  GomDD r = dd_add(a, dd_mul(b, GomDD{-q1, 0.0}));
  double q2 = r.hi / b.hi;
  r = dd_add(r, dd_mul(b, GomDD{-q2, 0.0}));
  double q3 = r.hi / b.hi;
  return dd_add(dd_quick(q1, q2), GomDD{q3, 0.0});
}
__host__ __device__ inline double gom_log(double x) {
  if (!(x > 0.0) || x == x + x || x < 2.2250738585072014e-308) return log(x);  // 0, <0, inf, NaN, subnormal
  int e;
  double m = frexp(x, &e);  // m in [0.5, 1)
  if (m < 0.70710678118654752) {
    m = m + m;
    e = e - 1;
  }
  const GomDD s = dd_div(GomDD{m - 1.0, 0.0}, dd_two_sum(m, 1.0));  // m - 1 is exact
  const GomDD s2 = dd_mul(s, s);
  GomDD sum = dd_div(GomDD{1.0, 0.0}, GomDD{45.0, 0.0});
  for (int k = 21; k >= 0; k--) sum = dd_add(dd_mul(sum, s2), dd_div(GomDD{1.0, 0.0}, GomDD{2.0 * k + 1.0, 0.0}));
  GomDD lm = dd_mul(s, sum);
  lm = GomDD{2.0 * lm.hi, 2.0 * lm.lo};
  const GomDD ln2{0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
  const GomDD r = dd_add(dd_mul(ln2, GomDD{(double)e, 0.0}), lm);
  return r.hi + r.lo;
}

__device__ __forceinline__ double dmaxf(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dminf(double a, double b) { return a < b ? a : b; }

__device__ __forceinline__ void latch(int *err, int code, int entity) {
  if (atomicCAS(err, 0, code) == 0) err[1] = entity;
}

// Kernel argument blocks shared by the launchers and the kernels.
struct MomParams {
  double dt, g, th, rho_o, smag, Ah0, Kh0;
  int baroclinic;
};

// Closure parameters (the module variables of turbulence_closure_v1.f90) and the
// single-precision constant 16.6**(2.0/3) evaluated on the host.
struct TurbParams {
  double dt, g, rho_o, MSL, fs_coef;
  double q2min, psimin, eps_min, bgdiff, diffmax, xlmin1, xlmin2;
  double cmiu0, cpsi1, cpsi2, rpub, rmub, rnub, schk, schpsi;
  double h1_pp, h2_pp, vdmax_pp1, vdmax_pp2, vdmin_pp1, vdmin_pp2, tdmin_pp;
  int flag, model, stab;
};

struct TvdArgs {
  double dt, th, rho_o, dt2;
  int Nt, sub, hlim, vlim, heat;
};

struct PcgArgs {
  const double *diag, *off, *b;
  const int32_t *col;     // adj_cellnum_at_cell, [slot][cell], 1-based, 0 = none
  const int32_t *nfaces;  // tri_or_quad
  double *x, *r, *d0, *d1, *q, *s, *minv, *partA, *partB, *partC, *partD;
  double *pub;        // 6n published row values of the single-barrier variant
  unsigned long long *slots;  // 3 x PCG_MAX_SLOT_BLOCKS x 2 partial-sum slots (slot exchange)
  int *info;
  unsigned *bar;      // [0] arrival counter, [1] generation (custom grid barrier)
  int n, maxit;
  int barrier_mode;   // 0 cooperative-groups grid sync, 1 custom barrier
  double tol;
  static constexpr bool kMixed = false;
};

// The FP32-storage inner solve of the mixed-precision PCG (k_pcg.hip): the matrix, the
// right-hand side and the published row values (r, w, s) in FP32; the correction it
// accumulates and the search direction in FP64, as is every operation and reduction. Its
// stop test is relative to the outer residual: gpart holds the FP64 partial sums of
// r.M^-1 r of the current outer residual, g0 that of the first.
struct PcgArgs32 {
  const float *diag, *off, *b;
  const int32_t *col, *nfaces;
  double *x, *d0;
  float *pub;
  double *partA, *partB;
  unsigned long long *slots;
  int *info;
  unsigned *bar;
  int n, maxit;
  double tol;
  const double *gpart;  // partial sums of the outer gamma (ngpart of them)
  int ngpart;
  int first;            // first refinement: the outer gamma is gamma_0, stored in *g0
  double *g0;
  static constexpr bool kMixed = true;
};

// Launch helpers
static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Blocks of the register-resident PCG for n rows at 512 threads. rows = 0 (automatic):
// the fewest rows per thread (1, 2 or 4) that cover the system with at most 256 blocks,
// which spreads a lone system's row work over the most CUs; rows = 1, 2 or 4 fixes it
// (4 packs more batch members into one launch). A function of n and the setting alone, so
// a system gets the same partition alone and in a batch. Above 4 x 256 x 512 rows the
// result exceeds 256 (the memory-resident solve).
static inline int pcg_blocks_for(long n, int rows) {
  long r = rows;
  if (r <= 0) {
    r = (n + 256L * 512 - 1) / (256L * 512);
    r = r <= 1 ? 1 : r <= 2 ? 2 : 4;
  }
  int nb = (int)((n + r * 512 - 1) / (r * 512));
  return nb < 1 ? 1 : nb;
}
