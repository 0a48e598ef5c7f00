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
#ifndef GOM_NV_LEVELS
#define GOM_NV_LEVELS 4  // node levels per thread of k_node_vel
#endif
#ifndef GOM_TVD_LIM_LEVELS
#define GOM_TVD_LIM_LEVELS 1  // levels per thread of k_tvd_lim_cell
#endif
#ifndef GOM_TVD_NODE_LEVELS
#define GOM_TVD_NODE_LEVELS 8  // node levels per thread of k_tvd_node
#endif
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
  // Static node stencil of calculate_velocity_at_node.f90:67-138, built at create, one
  // record per (jj-th adjacent cell, node i), [jj][node]: ns_face (face 0, face 1, bottom
  // layer 0, bottom layer 1) of the cell's two faces touching the node, ns_Lf (length 0,
  // length 1, boundary factor 0, boundary factor 1), ns_cell (the cell, its bottom layer)
  // and ns_area; each a 16- or 8-byte load. ns_weight is the node's sum of fw/L.
  const int32_t *ns_face, *ns_cell;
  const double *ns_Lf, *ns_area, *ns_weight;
  // Cell records of the ELM walk, built at create: the x and y of a cell's nodes, the node
  // ids and the neighbour cells in slot order, 4 per cell, contiguous per cell
  // ([cell][slot]), so a visited cell costs a few aligned loads instead of a chain of
  // indirect ones.
  const double *cg_x, *cg_y;
  const int32_t *cg_node, *cg_adj;  // node ids and neighbour cells (1-based, 0 = none)
  // Per cell, its 4 slots contiguous ([cell][slot], one 16-byte load): where the transport
  // counts the slot face's flux (boundary_type_of_face != -1 or a river face), the face id
  // (1-based) shifted left by one with bit 0 set when the face's sign_in_outflow is +1 (clear
  // for -1); 0 for a face the transport skips and past the cell's faces. With cg_adj it
  // gives a cell's transport stencil in two loads instead of twelve.
  const int32_t *slot_face;
  // Per face (int2): its two cells (1-based, 0 = none), the first shifted left by one with
  // bit 0 set where the transport counts the face's flux, as slot_face
  const int32_t *face_adj;
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
  // spin-up factor, [1] tidal spin-up factor, [2] elapsed time, [3] and [7] x and y of a
  // generated cyclone's centre, [4] wind spin-up factor, [5] hurricane phase (0 keep,
  // 1 interpolate, 2 next snapshot), [6] julian day
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
  double *htop;           // [cell][tracer] (sum1, sum2): horizontal upwind sums of each column's top layer
  double *fqd, *fps;      // (2, L, J) each: per face-layer (Q, d) and the PSI of both tracers
  double *phij;           // (L, J)
  double *Qi, *di, *phii; // (0:L, E)
  double *PSIi;           // (2, 0:L, E): the vertical limiter of both tracers side by side
  double *s12;            // (2, NT, L, E): the limiter sums (s1a, s2a) of each tracer per cell-layer
  double *obk, *qbk;      // (L, OB, NT), (L, QB, NT)
  double *node_new;       // (L, P, NT)
  double *face_new;       // (L, J, NT)
  // turbulence closure: face diffusivity (0:L, J) before the face -> cell map
  double *turb_kvf;
  // elimination factors of the column tridiagonal solves, (L, max(E, J))
  double *tri_gam;
  double *tvd_x;          // [layer][cell][tracer] right-hand sides / solution of the TVD column solve
                          // (the ELM transport keeps [tracer][layer][cell] in it)
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

// ---- correctly rounded cosine (host and device)
// The harmonic tide of compute_boundary_eta.f90:72-90 takes cos from the host C library,
// which rounds correctly in all but vanishingly rare cases; the device library's cos does
// not always. gom_cos reduces x = n pi/2 + r with pi/2 held in three doubles (the product
// n P1 exact by an FMA split, x - n P1 exact by Sterbenz), so r carries ~2^-150 absolute
// error for |n| < 2^30; then cos r or sin r (|r| <= pi/4) by their Taylor series to r^28
// in double-double, Horner in r^2, rounded once.
__host__ __device__ inline GomDD dd_two_prod(double a, double b) {
  double p = a * b;
  return {p, fma(a, b, -p)};
}
__host__ __device__ inline double gom_cos(double x) {
  if (!(fabs(x) < 1.0e9)) return cos(x);  // inf, NaN and arguments far beyond a tide's
  // 1/n! as double-double, n = 0..28 (the terms of both series)
  const double ch[29] = {0x1p+0, 0x1p+0, 0x1p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
                         0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
                         0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22,
                         0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33,
                         0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45,
                         0x1.952c77030ad4ap-49, 0x1.6827863b97d97p-53, 0x1.2f49b46814157p-57,
                         0x1.e542ba4020225p-62, 0x1.71b8ef6dcf572p-66, 0x1.0ce396db7f853p-70,
                         0x1.761b41316381ap-75, 0x1.f2cf01972f578p-80, 0x1.3f3ccdd165fa9p-84,
                         0x1.88e85fc6a4e5ap-89, 0x1.d1ab1c2dccea3p-94, 0x1.0a18a2635085dp-98};
  const double cl[29] = {0.0, 0.0, 0.0, 0x1.5555555555555p-57, 0x1.5555555555555p-59,
                         0x1.1111111111111p-63, -0x1.f49f49f49f49fp-65, 0x1.a01a01a01a01ap-73,
                         0x1.a01a01a01a01ap-76, -0x1.c154f8ddc6c00p-73, 0x1.cbbc05b4fa99ap-76,
                         -0x1.c062e06d1f209p-80, -0x1.2aec959e14c06p-83, 0x1.f28e0cc748ebep-87,
                         0x1.05d6f8a2efd1fp-92, 0x1.1d8656b0ee8cbp-97, 0x1.1d8656b0ee8cbp-101,
                         0x1.ac981465ddc6cp-103, 0x1.eec01221a8b0bp-107, 0x1.2650f61dbdcb4p-112,
                         0x1.ea72b4afe3c2fp-120, -0x1.d043ae40c4647p-120, -0x1.aebcdbd20331cp-124,
                         -0x1.3423c7d91404fp-130, -0x1.9ada5fcc1ab14p-135, -0x1.58ddadf344487p-139,
                         -0x1.71c37ebd16540p-143, 0x1.054d0c78aea14p-149, 0x1.b9e2e28e1aa54p-153};
  const double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c07p-54, P3 = -0x1.f1976b7ed8fbcp-110;
  const double n = rint(x * 0x1.45f306dc9c883p-1);  // x 2/pi
  GomDD p1 = dd_two_prod(n, P1), p2 = dd_two_prod(n, P2);
  GomDD r = dd_two_sum(x - p1.hi, -p1.lo);
  r = dd_add(r, GomDD{-p2.hi, -p2.lo});
  r = dd_add(r, GomDD{-(n * P3), 0.0});
  const GomDD z = dd_mul(r, r);
  long long q = (long long)n % 4;
  if (q < 0) q += 4;
  // cos r = sum_{j<=14} (-1)^j z^j/(2j)!, sin r = r sum_{j<=13} (-1)^j z^j/(2j+1)!: o = 1
  // (odd quadrants, sin r) takes the odd factorials; Horner from the top term j = 14 - o
  const int o = (q & 1) ? 1 : 0;
  GomDD s{(o ? -1.0 : 1.0) * ch[28 - o], (o ? -1.0 : 1.0) * cl[28 - o]};
  for (int j = 13 - o; j >= 0; j--) {
    const double sg = (j & 1) ? -1.0 : 1.0;
    s = dd_add(dd_mul(s, z), GomDD{sg * ch[2 * j + o], sg * cl[2 * j + o]});
  }
  if (o) s = dd_mul(s, r);
  const double v = s.hi + s.lo;
  return (q == 1 || q == 2) ? -v : v;
}

__device__ __forceinline__ double dmaxf(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dminf(double a, double b) { return a < b ? a : b; }

__device__ __forceinline__ void latch(int *err, int code, int entity) {
  if (atomicCAS(err, 0, code) == 0) err[1] = entity;
}

// The block's share of the grid, XCD-aware. The dispatcher hands block b to XCD b % 8, so
// consecutive blocks (neighbouring cells, faces and nodes of one layer plane) would land on
// eight different L2 caches and each would fetch the stencil values the others share. Here
// each group of 8 * GOM_XCD_RUN blocks is dealt out so that XCD x takes GOM_XCD_RUN
// consecutive blocks of the group; groups stay in dispatch order, so all XCDs still sweep
// the same part of the grid together (the work of a layer-major grid varies with the layer).
// A bijection of [0, gridDim.x) onto itself: the blocks of a last partial group keep their
// index. Only the placement changes.
#define GOM_XCDS 8
#ifndef GOM_XCD_RUN
#define GOM_XCD_RUN 32
#endif
__device__ __forceinline__ unsigned gom_block() {
  const unsigned b = blockIdx.x, g = GOM_XCDS * GOM_XCD_RUN;
  if (b >= gridDim.x - gridDim.x % g) return b;
  return b - b % g + (b % GOM_XCDS) * GOM_XCD_RUN + (b / GOM_XCDS) % GOM_XCD_RUN;
}

// Column tiles of a (level, entity) grid. Threads are numbered tile by tile, a tile being
// GOM_TILE consecutive entities over all nlev levels (level-major inside the tile), so the
// blocks that run together cover whole columns: a column's vertical neighbours and its
// per-entity tables are read by blocks of one XCD (gom_block) within a short time, from L2,
// instead of once per layer plane from memory. A wavefront is still 64 consecutive entities
// of one level, so every field access stays coalesced. The grid holds tile_threads(n, nlev)
// threads; gom_tile gives this thread's level k (0-based) and entity e, false past the end.
#define GOM_TILE 64
static inline long tile_threads(long n, int nlev) { return (n + GOM_TILE - 1) / GOM_TILE * GOM_TILE * nlev; }
__device__ __forceinline__ bool gom_tile(int n, int nlev, int &k, int &e) {
  const unsigned q = gom_block() * blockDim.x + threadIdx.x;
  const unsigned tsz = GOM_TILE * (unsigned)nlev;
  const unsigned t = q / tsz, i = q - t * tsz;
  e = (int)(t * GOM_TILE + i % GOM_TILE);
  k = (int)(i / GOM_TILE);
  return e < n;
}

// Device copy of the boundary tables (gomgpu_boundary) and the stage's work arrays.
struct DevBnd {
  int nh, ns, nob, nqb;
  const int *h_first, *s_first, *s_mono;   // s_mono[s]: the series' times never decrease
  const double *h_shift, *h_con, *s_time, *s_val;
  const int *ob_type, *ob_id, *ob_ser, *qb_qser, *qb_ser;
  const double *qb_portion;
  double *h_old, *h_new;  // each set's sum at elapsed - dt and elapsed
  double *job_val;        // each series at each time slot of the step
  int *job_found;
  double shift, ref_salt, ref_temp;  // shift: reference_diff_days * 86400 s
  int is_salt, is_temp, any_eta_series, tracers;
};

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

// A slot_face entry: false for a slot the transport skips, else the face (0-based) and its
// outflow sign seen from the cell
__device__ __forceinline__ bool slot_decode(int v, int &j, int &sign) {
  j = (int)((unsigned)v >> 1) - 1;
  sign = (v & 1) ? 1 : -1;
  return v != 0;
}
