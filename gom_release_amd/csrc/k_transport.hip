// k_transport.hip - TVD scalar transport (solve_transport_equation_v19.f90), equation of
// state (calculate_density_full.f90 / calculate_density_linear.f90), the new->old update
// (update_variables.f90) and FP64 budgets. Paths relative to GOM_Rv1.1.0/source.
//
// Work arrays are layer-major with the entity fastest. Reads the reference makes of its
// automatic arrays outside their written ranges are defined as zero here (coefficients
// are written for every layer, zero outside the active range).
#include "gomgpu_dev.h"


// The transport's work arrays, typed. W = double is the reference's FP64 storage. W = float
// (GOMGPU_MIXED) stores in FP32 the face limiter's phi, the per-cell upwind sums that form
// the face ratios and the node/face means of the output; every value is widened to FP64
// when loaded and the arithmetic is FP64. The concentrations, the fluxes and diffusion
// coefficients (horizontal and vertical), the limiters PSI and the column solve stay FP64:
// a thin partial layer's update divides by area * dz, which magnifies a rounding of its
// fluxes by the layer's Courant number. Per face-layer, Q and d are stored together (fqd) and
// so are the PSI of the two tracers (fps): the right-hand side gathers each with one 16-byte
// load.
// The allocations are the FP64 ones; an FP32 array uses their first half.
template <typename W>
struct TvdW {
  double *con, *con_new, *fqd, *fps, *Qi, *di, *phii, *htop, *PSIi, *obk, *qbk, *tri_gam, *tvd_x;
  W *phij, *s12, *node_new, *face_new;
};
template <typename W>
static TvdW<W> tvd_view(const DevWork &w) {
  TvdW<W> v;
  v.con = w.con;
  v.con_new = w.con_new;
  v.Qi = w.Qi;
  v.htop = w.htop;
  v.di = w.di;
  v.phii = w.phii;
  v.PSIi = w.PSIi;
  v.obk = w.obk;
  v.qbk = w.qbk;
  v.tri_gam = w.tri_gam;
  v.tvd_x = w.tvd_x;
  v.fqd = w.fqd;
  v.fps = w.fps;
  v.phij = (W *)w.phij;
  v.s12 = (W *)w.s12;
  v.node_new = (W *)w.node_new;
  v.face_new = (W *)w.face_new;
  return v;
}

__device__ __forceinline__ size_t cix(int t, int k, int e, int L2, int n) {  // con (0:L+1)
  return ((size_t)t * L2 + k) * n + e;
}
__device__ __forceinline__ size_t nix(int t, int k, int e, int L, int n) {  // (L) fields, k 1-based
  return ((size_t)t * L + (k - 1)) * n + e;
}
__device__ __forceinline__ size_t frx(int k, int j, int J) {  // (Q, d) or (PSI 1, PSI 2) of layer k
  return 2 * ((size_t)(k - 1) * J + j);
}
// The right-hand sides of the column solve, w.tvd_x: layer k's values of the NT tracers side
// by side, so that both tracers move in one 16-byte access
__device__ __forceinline__ void x_load(const double *x, int k, int e, int n, int NT, double v[2]) {
  const size_t i = (size_t)(k - 1) * n + e;
  if (NT == 2) {
    const double2 p = ((const double2 *)x)[i];
    v[0] = p.x;
    v[1] = p.y;
  } else {
    v[0] = x[i];
  }
}
__device__ __forceinline__ void x_store(double *x, int k, int e, int n, int NT, const double v[2]) {
  const size_t i = (size_t)(k - 1) * n + e;
  if (NT == 2)
    ((double2 *)x)[i] = make_double2(v[0], v[1]);
  else
    x[i] = v[0];
}
// The top layer's horizontal upwind sums (sum1, sum2) of tracer t of cell e, w.htop
__device__ __forceinline__ size_t hix(int t, int e, int NT) { return 2 * ((size_t)e * NT + t); }
// PSI_ik of level k (0:L) for the two tracers side by side: one 16-byte load for both
__device__ __forceinline__ size_t pix(int k, int e, int n) { return 2 * ((size_t)k * n + e); }
// The upwind sums (s1a, s2a) of layer k (1:L), per tracer t at 2 t, NT tracers per record
__device__ __forceinline__ size_t six(int k, int e, int n, int NT) { return 2 * (size_t)NT * ((size_t)(k - 1) * n + e); }
template <typename W> struct Pair;
template <> struct Pair<double> { typedef double2 T; };
template <> struct Pair<float> { typedef float2 T; };

// con_cell from salt/temp (+ surface heat source), solve_transport_equation_v19.f90:88-173
template <typename W>
__global__ void k_tvd_init(DevMesh m, DevState s, DevForcing f, TvdW<W> w, DevDims d, TvdArgs a) {
  int q = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, L = d.L, L2 = d.L + 2;
  if (q >= E * L2) return;
  int e = q % E, k = q / E;
  for (int t = 0; t < d.NT; t++) {
    double v = 0.0;
    if (k >= 1 && k <= L) {
      const double *src = d.tid[t] == 1 ? s.salt_cell : s.temp_cell;
      v = AL(src, k, e, E);
      if (d.tid[t] == 2 && a.heat == 1 && f.phi_sz) {
        int tl = s.top_layer_at_element[e], bl = m.bottom_layer_at_element[e];
        if (k >= bl && k <= tl) {
          const double cp = 4186.0;
          v = v + a.dt * AL(f.phi_sz, k, e, E) / (a.rho_o * cp * AL(s.dz_cell, k, e, E));
        }
      }
    }
    w.con[cix(t, k, e, L2, E)] = v;
  }
}

// Boundary concentrations of open-boundary and river cells for sub-cycle a.sub
// (solve_transport_equation_v19.f90:291-397); one thread per (layer, boundary cell).
template <typename W>
__global__ void k_tvd_bnd(DevMesh m, DevState s, DevForcing f, TvdW<W> w, DevDims d, TvdArgs a) {
  int q = gom_block() * blockDim.x + threadIdx.x;
  const int L = d.L, E = d.E, J = d.J, L2 = d.L + 2;
  int nob = d.NOB, nqb = d.NQB;
  if (q < L * nob) {
    int i = q % nob, k = q / nob + 1;
    int i2 = m.ob_cell_id[i] - 1;
    int tl = s.top_layer_at_element[i2], bl = m.bottom_layer_at_element[i2];
    double v[2] = {0.0, 0.0};
    if (k >= bl && k <= tl) {
      int N = m.tri_or_quad[i2];
      for (int l = 1; l <= N; l++) {
        int j2 = TB(m.facenum_at_cell, l, i2, E) - 1;
        if (m.boundary_type_of_face[j2] > 0) {
          double dzf = AL(s.dz_face, k, j2, J);
          double dzo = dzf + (AL(s.dz_face_new, k, j2, J) - dzf) * a.sub / a.Nt;
          double sg = (double)TB(m.sign_in_outflow, l, i2, E);
          double Q2 = m.face_length[j2] * dzo *
                      ((1.0 - a.th) * AL(s.un_face, k, j2, J) * sg + a.th * AL(s.un_face_new, k, j2, J) * sg);
          for (int t = 0; t < d.NT; t++)
            v[t] = (Q2 <= 0.0) ? f.con_at_ob[a.sub + (size_t)a.Nt * (t + (size_t)d.NT * i)]
                               : w.con[cix(t, k, i2, L2, E)];
        }
      }
    }
    for (int t = 0; t < d.NT; t++) w.obk[nix(t, k, i, L, nob)] = v[t];
  } else if (q < L * nob + L * nqb) {
    int r = q - L * nob;
    int i = r % nqb, k = r / nqb + 1;
    int i2 = m.Qb_cell_id[i] - 1;
    int tl = s.top_layer_at_element[i2], bl = m.bottom_layer_at_element[i2];
    for (int t = 0; t < d.NT; t++)
      w.qbk[nix(t, k, i, L, nqb)] =
          (k >= bl && k <= tl) ? f.con_at_Qb[a.sub + (size_t)a.Nt * (t + (size_t)d.NT * i)] : 0.0;
  }
}

// Face coefficients Q, d, phi (solve_transport_equation_v19.f90:552-581)
template <typename W>
__global__ void k_tvd_coef_face(DevMesh m, DevState s, TvdW<W> w, DevDims d, TvdArgs a) {
  int q = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J;
  if (q >= J * d.L) return;
  int j = q % J, k = q / J + 1;
  double Q = 0.0, dd = 0.0, phi = 0.0;
  const bool in = k >= m.bottom_layer_at_face[j] && k <= s.top_layer_at_face[j];
  if (!in && a.sub > 0) return;  // zero since the first sub-cycle (the layers are fixed)
  if (in) {
    double dzf = AL(s.dz_face, k, j, J);
    double dzo = dzf + (AL(s.dz_face_new, k, j, J) - dzf) * a.sub / a.Nt;
    Q = m.face_length[j] * dzo * ((1.0 - a.th) * AL(s.un_face, k, j, J) + a.th * AL(s.un_face_new, k, j, J));
    double D = m.face_length[j] * dzo * AL(s.Kh, k, j, J) / m.delta_j[j];
    dd = dmaxf(0.0, D - 0.5 * fabs(Q));
    phi = (Q == 0.0) ? 1.0 : dminf(1.0, 2.0 * D / fabs(Q));
  }
  *(double2 *)(w.fqd + frx(k, j, J)) = make_double2(Q, dd);
  AL(w.phij, k, j, J) = phi;
}

// Concentration across face slot l of cell e at layer k (boundary value for no neighbour).
template <typename W>
__device__ __forceinline__ double con_across(const DevMesh &m, const TvdW<W> &w, const DevDims &d, int e,
                                             int ii, int k, int t) {
  const int E = d.E, L = d.L, L2 = d.L + 2;
  if (ii > 0) return w.con[cix(t, k, ii - 1, L2, E)];
  int iqb = m.Qb_element_flag[e];
  if (iqb > 0) return w.qbk[nix(t, k, iqb - 1, L, d.NQB)];
  int iob = m.ob_element_flag[e];
  if (iob > 0) return w.obk[nix(t, k, iob - 1, L, d.NOB)];
  return 0.0;
}

__device__ __forceinline__ double limiter(int which, double r, double phi) {
  double PHI;
  if (which == 1)
    PHI = dmaxf(phi, dminf(1.0, r));
  else if (which == 2)
    PHI = dmaxf(phi, (r + fabs(r)) / (1.0 + fabs(r)));
  else
    PHI = dmaxf(dmaxf(phi, dminf(1.0, 2.0 * r)), dminf(2.0, r));
  return PHI - phi;
}

// Horizontal upwind sums of one cell-layer (solve_transport_equation_v19.f90:700-760): the
// inflow-weighted differences sum1 and the inflows sum2, over the cell's active faces in
// slot order; sf4 and ad4 are the cell's slot_face and cg_adj records, c0 the cell-layer's
// own concentrations.
template <typename W>
__device__ __forceinline__ void lim_hsums(const DevMesh &m, const TvdW<W> &w, const DevDims &d, int e, int k,
                                          const int4 &sf4, const int4 &ad4, const double *c0, double *sum1,
                                          double *sum2) {
  const int E = d.E, L2 = d.L + 2, J = d.J;
  sum1[0] = sum1[1] = 0.0;
  sum2[0] = sum2[1] = 0.0;
  const int sf[4] = {sf4.x, sf4.y, sf4.z, sf4.w}, ad[4] = {ad4.x, ad4.y, ad4.z, ad4.w};
#pragma unroll
  for (int l = 0; l < 4; l++) {  // slots past the cell's faces hold 0
    int j, sa;
    if (slot_decode(sf[l], j, sa)) {
      const int ii = ad[l];
      double Q2 = w.fqd[frx(k, j, J)] * (double)sa;
      if (Q2 < 0.0) {
        for (int t = 0; t < d.NT; t++) {
          double ci = con_across(m, w, d, e, ii, k, t);
          sum1[t] = sum1[t] + fabs(Q2) * (c0[t] - ci);
          sum2[t] = sum2[t] + fabs(Q2);
        }
      }
    }
  }
}

// Cell (vertical) coefficients (solve_transport_equation_v19.f90:586-609). With a flux
// limiter, the thread of a wet column's top layer also leaves that layer's horizontal upwind
// sums in w.htop, which every vertical limiter ratio of the column uses (k_tvd_lim_cell).
template <typename W>
__global__ void k_tvd_coef_cell(DevMesh m, DevState s, TvdW<W> w, DevDims d, TvdArgs a) {
  int q = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E;
  if (q >= E * (d.L + 1)) return;
  int e = q % E, k = q / E;
  double Q = 0.0, dd = 0.0, phi = 0.0;
  const int tl = s.top_layer_at_element[e];
  const bool in = k >= m.bottom_layer_at_element[e] && k <= tl - 1;
  // levels outside the range stay zero after the first sub-cycle; the top layer's thread
  // still runs for the limiter sums
  if (!in && a.sub > 0 && !((a.hlim > 0 || a.vlim > 0) && k >= 1 && k == tl)) return;
  if (in) {
    double ar = m.area[e];
    // Q does not change over the sub-cycles (the velocities and the layer range are those
    // of the step): formed in the first, read back after
    if (a.sub == 0)
      Q = ar * ((1.0 - a.th) * A0(s.wn_cell, k, e, E) + a.th * A0(s.wn_cell_new, k, e, E));
    else
      Q = A0(w.Qi, k, e, E);
    double dh = A0(s.dzhalf_cell, k, e, E);
    double dzo = dh + (A0(s.dzhalf_cell_new, k, e, E) - dh) * a.sub / a.Nt;
    double D = ar * A0(s.Kv, k, e, E) / dzo;
    dd = dmaxf(0.0, D - 0.5 * fabs(Q));
    phi = (Q == 0.0) ? 1.0 : dminf(1.0, 2.0 * D / fabs(Q));
  }
  if (in || a.sub == 0) {
    if (a.sub == 0) A0(w.Qi, k, e, E) = Q;
    A0(w.di, k, e, E) = dd;
    A0(w.phii, k, e, E) = phi;
  }
  if ((a.hlim > 0 || a.vlim > 0) && k >= 1 && k == tl) {
    double sum1[2], sum2[2], c0[2] = {0.0, 0.0};
    for (int t = 0; t < d.NT; t++) c0[t] = w.con[cix(t, k, e, d.L + 2, E)];
    const int4 sf4 = *(const int4 *)(m.slot_face + 4 * (size_t)e);
    const int4 ad4 = *(const int4 *)(m.cg_adj + 4 * (size_t)e);
    lim_hsums(m, w, d, e, k, sf4, ad4, c0, sum1, sum2);
    for (int t = 0; t < d.NT; t++) *(double2 *)(w.htop + hix(t, e, d.NT)) = make_double2(sum1[t], sum2[t]);
  }
}

// Limiter ratios (solve_transport_equation_v19.f90:624-842), the work of one level of a cell:
// the per-layer upwind sums s1a, s2a that the face ratios r_jk need (layer k), and the
// vertical ratio r_ik (level k) with its limiter PSI_ik (level k). As in the reference, the
// vertical ratios reuse the horizontal sums left over from the column's top layer, which
// k_tvd_coef_cell computed once per column (w.htop) in the same order.
template <typename W>
__device__ __forceinline__ void lim_cell_level(const DevMesh &m, const TvdW<W> &w, const DevDims &d, int hlim,
                                               int first, int e, int k, int tl, int bl, const int4 &sf4,
                                               const int4 &ad4) {
  const int E = d.E, L2 = d.L + 2;
  if (tl == 0 || k < bl - 1 || k > tl) {  // zero from the first sub-cycle on
    if (first) *(double2 *)(w.PSIi + pix(k, e, E)) = make_double2(0.0, 0.0);
    return;
  }
  if (k == bl - 1) {  // PSI_ik is 0 at bl - 1 (r_ik is 1.0 there)
    if (first) *(double2 *)(w.PSIi + pix(k, e, E)) = make_double2(0.0, 0.0);
    return;
  }
  // bl <= k <= tl from here. The concentrations of levels k-1, k and k+1 (con holds levels
  // 0..L+1) and the vertical flows at k-1 and k, read once for both parts.
  const int nv = tl - bl + 1;
  double cm[2] = {0.0, 0.0}, c0[2] = {0.0, 0.0}, cp[2] = {0.0, 0.0};
  for (int t = 0; t < d.NT; t++) {
    cm[t] = w.con[cix(t, k - 1, e, L2, E)];
    c0[t] = w.con[cix(t, k, e, L2, E)];
    cp[t] = w.con[cix(t, k + 1, e, L2, E)];
  }
  const double qm = A0(w.Qi, k - 1, e, E), qk = A0(w.Qi, k, e, E);
  {  // s1a, s2a of layer k
    double s1a[2] = {0, 0}, s2a[2] = {0, 0}, sum1[2], sum2[2];
    for (int t = 0; t < d.NT; t++) {
      if (nv != 1) {
        if (k == bl) {
          double r2 = fabs(0.5 * (qk + fabs(qk)));
          s1a[t] = r2 * (c0[t] - cp[t]);
          s2a[t] = r2;
        } else if (k == tl) {
          double r1 = fabs(0.5 * (qm + fabs(qm)));
          s1a[t] = r1 * (c0[t] - cm[t]);
          s2a[t] = r1;
        } else {
          double r1 = fabs(0.5 * (qm + fabs(qm))), r2 = fabs(0.5 * (qk + fabs(qk)));
          s1a[t] = r1 * (c0[t] - cm[t]) + r2 * (c0[t] - cp[t]);
          s2a[t] = r1 + r2;
        }
      }
    }
    lim_hsums(m, w, d, e, k, sf4, ad4, c0, sum1, sum2);
    typedef typename Pair<W>::T W2;
    for (int t = 0; t < d.NT; t++) {
      W2 v;
      v.x = (W)(s1a[t] + sum1[t]);
      v.y = (W)(s2a[t] + sum2[t]);
      *(W2 *)(w.s12 + six(k, e, E, d.NT) + 2 * t) = v;
    }
  }
  // r_ik of level k: computed for bl <= k <= tl - 1 (1.0 at tl, where PSI_ik is 0)
  if (k == tl) {
    if (first) *(double2 *)(w.PSIi + pix(k, e, E)) = make_double2(0.0, 0.0);
    return;
  }
  const bool up = A0(w.Qi, tl, e, E) >= 0.0;
  const double phi = A0(w.phii, k, e, E);
  double psi[2] = {0.0, 0.0};
  for (int t = 0; t < d.NT; t++) {
    const double2 h = *(const double2 *)(w.htop + hix(t, e, d.NT));  // (sum1, sum2)
    double r;
    if (up) {
      double rt1 = h.x + fabs(qm) * (c0[t] - cm[t]);
      double rt2 = h.y + fabs(qm);
      double rt = (cp[t] - c0[t]) * rt2;
      r = (rt == 0.0) ? 1.0 : rt1 / rt;
    } else {  // the reference's k runs bl+1..tl and writes level k-1: level k's ratio from layer k+1
      double qk1 = A0(w.Qi, k + 1, e, E);
      double rt1 = h.x + fabs(qk1) * (cp[t] - w.con[cix(t, k + 2, e, L2, E)]);
      double rt2 = h.y + fabs(qk1);
      double rt = (c0[t] - cp[t]) * rt2;
      r = (rt == 0.0) ? 1.0 : rt1 / rt;
    }
    psi[t] = hlim != 0 ? limiter(hlim, r, phi) : 0.0;
  }
  *(double2 *)(w.PSIi + pix(k, e, E)) = make_double2(psi[0], psi[1]);
}

// One thread per (chunk of TVD_LIM_LEVELS levels, cell): the cell's layer range and slot
// records are read once for the chunk's levels.
#define TVD_LIM_LEVELS GOM_TVD_LIM_LEVELS
template <typename W>
__global__ void k_tvd_lim_cell(DevMesh m, DevState s, TvdW<W> w, DevDims d, int hlim, int first) {
  long q = (long)gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, L1 = d.L + 1;
  const int nch = (L1 + TVD_LIM_LEVELS - 1) / TVD_LIM_LEVELS;
  if (q >= (long)nch * E) return;
  const int e = (int)(q % E), k0 = (int)(q / E) * TVD_LIM_LEVELS;
  const int tl = s.top_layer_at_element[e], bl = m.bottom_layer_at_element[e];
  int4 sf4 = make_int4(0, 0, 0, 0), ad4 = sf4;
  if (tl != 0 && k0 + TVD_LIM_LEVELS - 1 >= bl && k0 <= tl) {
    sf4 = *(const int4 *)(m.slot_face + 4 * (size_t)e);
    ad4 = *(const int4 *)(m.cg_adj + 4 * (size_t)e);
  }
  for (int z = 0; z < TVD_LIM_LEVELS && k0 + z < L1; z++)
    lim_cell_level(m, w, d, hlim, first, e, k0 + z, tl, bl, sf4, ad4);
}

// Face ratio r_jk and PSI_jk. In the reference each adjacent cell whose outflow the face
// carries (Q2 >= 0 from its side) writes r_jk in cell order, so the higher cell id wins
// a tie (Q = 0) -- the same writer is chosen here.
template <typename W>
__global__ void k_tvd_lim_face(DevMesh m, DevState s, TvdW<W> w, DevDims d, int hlim, int first) {
  int q = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J, E = d.E, L = d.L, L2 = d.L + 2;
  if (q >= J * L) return;
  int j = q % J, k = q / J + 1;
  const bool in = (k >= m.bottom_layer_at_face[j] && k <= s.top_layer_at_face[j]);
  if (!in || hlim == 0) {  // PSI_jk is zero there; written once, in the first sub-cycle
    if (first)
      *(double2 *)(w.fps + frx(k, j, J)) = make_double2(0.0, 0.0);
    return;
  }
  double r[2] = {1.0, 1.0};
  const int2 fa = *(const int2 *)(m.face_adj + 2 * (size_t)j);
  if (fa.x & 1) {  // a face whose flux the transport counts
    const int cell[2] = {fa.x >> 1, fa.y};
    double Q = w.fqd[frx(k, j, J)];
    for (int side = 1; side <= 2; side++) {
      int c = cell[side - 1];
      if (c == 0) continue;
      int e = c - 1;
      int tl = s.top_layer_at_element[e], bl = m.bottom_layer_at_element[e];
      if (tl == 0 || k < bl || k > tl) continue;
      double sg = side == 1 ? 1.0 : -1.0;  // sign_in_outflow of face j seen from this cell
      if (Q * sg >= 0.0) {
        int other = cell[2 - side];
        typedef typename Pair<W>::T W2;
        for (int t = 0; t < d.NT; t++) {
          double ci = con_across(m, w, d, e, other, k, t);
          const W2 sv = *(const W2 *)(w.s12 + six(k, e, E, d.NT) + 2 * t);  // (s1a, s2a)
          double rt = (double)sv.y * (ci - w.con[cix(t, k, e, L2, E)]);
          r[t] = (rt == 0.0) ? 1.0 : (double)sv.x / rt;
        }
      }
    }
  }
  double psi[2] = {0.0, 0.0};
  for (int t = 0; t < d.NT; t++) psi[t] = limiter(hlim, r[t], AL(w.phij, k, j, J));
  *(double2 *)(w.fps + frx(k, j, J)) = make_double2(psi[0], psi[1]);
}

// Per-column implicit update (solve_transport_equation_v19.f90:921-1303) in two passes.
// tvd_rhs_layer forms the explicit right-hand side of layer k: horizontal upwind advection,
// diffusion and limiter fluxes, and the vertical advection and limiter terms.
template <typename W>
__device__ __forceinline__ void tvd_rhs_layer(const DevMesh &m, const DevState &s, const TvdW<W> &w,
                                              const DevDims &d, const TvdArgs &a, int e, int k, int bl, int tl,
                                              double *x) {
  const int E = d.E, L = d.L, L1 = d.L + 1, L2 = d.L + 2, J = d.J;
  const int NT = d.NT, Nt = a.Nt, sub = a.sub;
  const double dt2 = a.dt2;
  int nv = tl - bl + 1;
  double ar = m.area[e];
  double sqi[2] = {0, 0}, sqii[2] = {0, 0}, shd[2] = {0, 0}, shl[2] = {0, 0};
  const int4 sf4 = *(const int4 *)(m.slot_face + 4 * (size_t)e);
  const int4 ad4 = *(const int4 *)(m.cg_adj + 4 * (size_t)e);
  const int sf[4] = {sf4.x, sf4.y, sf4.z, sf4.w}, ad[4] = {ad4.x, ad4.y, ad4.z, ad4.w};
  double cc[2] = {0.0, 0.0};
  for (int t = 0; t < NT; t++) cc[t] = w.con[cix(t, k, e, L2, E)];
#pragma unroll
  for (int l = 0; l < 4; l++) {  // slots past the cell's faces hold 0
    int j, sa;
    if (!slot_decode(sf[l], j, sa)) continue;
    const int ii = ad[l];
    const double2 qd = *(const double2 *)(w.fqd + frx(k, j, J));   // Q, d
    const double2 ps = *(const double2 *)(w.fps + frx(k, j, J));   // PSI of the two tracers
    double Q2 = qd.x * (double)sa;
    double dfj = qd.y;
    for (int t = 0; t < NT; t++) {
      double c0 = cc[t];
      double ci = con_across(m, w, d, e, ii, k, t);
      if (Q2 > 0.0)
        sqi[t] = sqi[t] + fabs(Q2) * c0;
      else if (Q2 < 0.0)
        sqii[t] = sqii[t] + fabs(Q2) * ci;
      shd[t] = shd[t] + dfj * (ci - c0);
      shl[t] = shl[t] + (t == 0 ? ps.x : ps.y) * fabs(Q2) * (c0 - ci);
    }
  }
  double dz0 = AL(s.dz_cell, k, e, E);
  double dzo = dz0 + (AL(s.dz_cell_new, k, e, E) - dz0) * sub / Nt;
  for (int t = 0; t < NT; t++) {
    double h_adv = dt2 * (sqii[t] - sqi[t]);
    double h_diff = dt2 * shd[t];
    double h_flux = 0.5 * dt2 * shl[t];
    double rhs = ar * dzo * cc[t] + h_adv + h_diff;
    if (a.hlim > 0) rhs = rhs + h_flux;
    x[t] = rhs;
  }
  if (nv != 1) {
    // the vertical limiter of both tracers at levels k-1 and k, one 16-byte load each
    double2 pm = make_double2(0.0, 0.0), pk = make_double2(0.0, 0.0);
    if (a.vlim > 0) {
      if (k > bl) pm = *(const double2 *)(w.PSIi + pix(k - 1, e, E));
      if (k < tl) pk = *(const double2 *)(w.PSIi + pix(k, e, E));
    }
    for (int t = 0; t < NT; t++) {
      double ck = cc[t];
      double v_adv, vft = 0.0;
      if (k == bl) {
        double Q = A0(w.Qi, k, e, E), ckp = w.con[cix(t, k + 1, e, L2, E)];
        double Ap = fabs(0.5 * (Q + fabs(Q))), Am = fabs(0.5 * (Q - fabs(Q)));
        v_adv = dt2 * (-(Ap * ck) + Am * ckp);
        if (a.vlim > 0) {
          double ps = t == 0 ? pk.x : pk.y;
          vft = 0.5 * dt2 * (-(ps * Ap * (ckp - ck)) + ps * Am * (ck - ckp));
        }
      } else if (k == tl) {
        double Q = A0(w.Qi, k - 1, e, E), ckm = w.con[cix(t, k - 1, e, L2, E)];
        double Ap = fabs(0.5 * (Q + fabs(Q))), Am = fabs(0.5 * (Q - fabs(Q)));
        v_adv = dt2 * (Ap * ckm - Am * ck);
        if (a.vlim > 0) {
          double ps = t == 0 ? pm.x : pm.y;
          vft = 0.5 * dt2 * (ps * Ap * (ck - ckm) - ps * Am * (ckm - ck));
        }
      } else {
        double Q1 = A0(w.Qi, k - 1, e, E), Q2 = A0(w.Qi, k, e, E);
        double ckm = w.con[cix(t, k - 1, e, L2, E)], ckp = w.con[cix(t, k + 1, e, L2, E)];
        double A1p = fabs(0.5 * (Q1 + fabs(Q1))), A1m = fabs(0.5 * (Q1 - fabs(Q1)));
        double A2p = fabs(0.5 * (Q2 + fabs(Q2))), A2m = fabs(0.5 * (Q2 - fabs(Q2)));
        v_adv = dt2 * (A1p * ckm - A2p * ck + A2m * ckp - A1m * ck);
        if (a.vlim > 0) {
          double p1 = t == 0 ? pm.x : pm.y, p2 = t == 0 ? pk.x : pk.y;
          vft = 0.5 * dt2 * (p1 * A1p * (ck - ckm) - p2 * A2p * (ckp - ck) + p2 * A2m * (ck - ckp) - p1 * A1m * (ckm - ck));
        }
      }
      x[t] = x[t] + v_adv;
      if (a.vlim > 0) x[t] = x[t] + vft;
    }
  }
}

// Layer k's thickness at the sub-cycle's time: dz_cell moved towards dz_cell_new by
// (sub+1)/Nt (solve_transport_equation_v19.f90:957-1025)
__device__ __forceinline__ double tvd_dzn(const DevState &s, int k, int e, int E, int sub, int Nt) {
  const double dz0 = AL(s.dz_cell, k, e, E);
  return dz0 + (AL(s.dz_cell_new, k, e, E) - dz0) * (sub + 1) / Nt;
}

// The clamp of solve_transport_equation_v19.f90:1310-1322.
__device__ __forceinline__ double tvd_clamp(double v) {
  if (v >= 0.0 && v <= 50.0) return v;
  return v < 0.0 ? 0.0 : 50.0;
}

// The explicit right-hand sides, one thread per (layer, cell), into w.tvd_x (row of the layer).
template <typename W>
__global__ void k_tvd_rhs(DevMesh m, DevState s, TvdW<W> w, DevDims d, TvdArgs a) {
  long q = (long)gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, L = d.L;
  if (q >= (long)L * E) return;
  int e = (int)(q % E), k = (int)(q / E) + 1;
  int bl = m.bottom_layer_at_element[e], tl = s.top_layer_at_element[e];
  if (tl == 0 || k < bl || k > tl) return;
  double x[2];
  tvd_rhs_layer(m, s, w, d, a, e, k, bl, tl, x);
  x_store(w.tvd_x, k, e, E, d.NT, x);
}

// One thread per cell column: vertical diffusion implicit, the Thomas recurrence of
// tridiagonal_solver.f90 with NT right-hand sides in place in w.tvd_x (factors in
// w.tri_gam), then the clamp to [0, 50]. The result goes to the next sub-cycle's
// concentration buffer con_next (every level in the first two sub-cycles, so both buffers
// hold the reference's zeros outside the wet range, solve_transport_equation_v19.f90:
// 1352-1371; the wet range after that) or, in the last sub-cycle, to con_new.
template <typename W>
__global__ void k_tvd_solve(DevMesh m, DevState s, TvdW<W> w, DevDims d, TvdArgs a, double *con_next, int full,
                            int last, int *err) {
  int e = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, L = d.L, L2 = d.L + 2;
  if (e >= E) return;
  const int bl = m.bottom_layer_at_element[e], tl = s.top_layer_at_element[e];
  const int NT = d.NT;
  if (!last && full) {  // the levels outside the wet range (all of a dry column)
    for (int k = 0; k <= L + 1; k++)
      if (tl == 0 || k < bl || k > tl)
        for (int t = 0; t < NT; t++) con_next[cix(t, k, e, L2, E)] = 0.0;
  }
  if (tl == 0) return;
  const int Nt = a.Nt, sub = a.sub;
  const double dt2 = a.dt2;
  int nv = tl - bl + 1;
  double ar = m.area[e];
  // Rows top down (row r is layer tl - r), solve_transport_equation_v19.f90:957-1025: row r
  // has a = -dt2 di(k) below the surface row, c = -dt2 di(k-1) above the bottom row and
  // b = area dz - a - c. Row r+1's a is row r's c, carried.
  double cv = 0.0, bv, x[2];
  if (nv == 1) {
    bv = ar * tvd_dzn(s, tl, e, E, sub, Nt);
  } else {
    cv = -dt2 * A0(w.di, tl - 1, e, E);
    bv = ar * tvd_dzn(s, tl, e, E, sub, Nt) - cv;
  }
  if (bv == 0.0) {
    latch(err, GOMGPU_ERR_TRIDIAG_PIVOT, e + 1);
    return;
  }
  double den = bv;
  x_load(w.tvd_x, tl, e, E, NT, x);
  for (int t = 0; t < NT; t++) x[t] = x[t] / den;
  x_store(w.tvd_x, tl, e, E, NT, x);
  for (int r = 1; r < nv; r++) {
    const int k = tl - r;
    const double av = cv, cprev = cv;
    if (r == nv - 1) {
      cv = 0.0;
      bv = -av + ar * tvd_dzn(s, k, e, E, sub, Nt);
    } else {
      cv = -dt2 * A0(w.di, k - 1, e, E);
      bv = -av + ar * tvd_dzn(s, k, e, E, sub, Nt) - cv;
    }
    double gam = cprev / den;
    den = bv - av * gam;
    if (den == 0.0) {
      latch(err, GOMGPU_ERR_TRIDIAG_PIVOT, e + 1);
      return;
    }
    double xr[2];
    x_load(w.tvd_x, k, e, E, NT, xr);
    for (int t = 0; t < NT; t++) x[t] = (xr[t] - av * x[t]) / den;
    x_store(w.tvd_x, k, e, E, NT, x);
    AL(w.tri_gam, r + 1, e, E) = gam;
  }
  // back substitution, then the clamp (rows top down = layers tl .. bl)
  for (int t = 0; t < NT; t++) {
    double v = tvd_clamp(x[t]);
    if (last)
      w.con_new[nix(t, bl, e, L, E)] = v;
    else
      con_next[cix(t, bl, e, L2, E)] = v;
  }
  for (int r = nv - 2; r >= 0; r--) {
    double gam = AL(w.tri_gam, r + 2, e, E);
    double xr[2];
    x_load(w.tvd_x, tl - r, e, E, NT, xr);
    for (int t = 0; t < NT; t++) {
      x[t] = xr[t] - gam * x[t];
      double v = tvd_clamp(x[t]);
      if (last)
        w.con_new[nix(t, tl - r, e, L, E)] = v;
      else
        con_next[cix(t, tl - r, e, L2, E)] = v;
    }
  }
}

// Node values: area*dz weighted cell means, top down (solve_transport_equation_v19.f90:1403-1453).
// The mean of one level; `none` is set when no adjacent cell is wet at that level.
template <typename W>
__device__ __forceinline__ double tvd_node_mean(const DevMesh &m, const DevState &s, const TvdW<W> &w,
                                                const DevDims &d, int i, int k, int t, bool &none) {
  const int P = d.P, E = d.E, L = d.L;
  const int nadj = m.adj_cells_at_node[i];
  double s11 = 0.0, s22 = 0.0;
  int count1 = 0;
  for (int l = 1; l <= nadj; l++) {
    int ii = TB(m.adj_cellnum_at_node, l, i, P) - 1;
    if (k >= m.bottom_layer_at_element[ii] && k <= s.top_layer_at_element[ii]) {
      double wgt = m.area[ii] * AL(s.dz_cell_new, k, ii, E);
      s11 = s11 + wgt;
      s22 = s22 + wgt * w.con_new[nix(t, k, ii, L, E)];
    } else {
      count1++;
    }
  }
  none = count1 == nadj;
  return (s11 == 0.0) ? 0.0 : s22 / s11;
}

// One thread per (node, chunk of TVD_NODE_LEVELS levels), the chunk's levels top down. A
// level whose adjacent cells are all dry takes the value of the level above it (unless it
// is the node's top level); at the top of a chunk that value is found by walking up.
#define TVD_NODE_LEVELS GOM_TVD_NODE_LEVELS
template <typename W>
__global__ void k_tvd_node(DevMesh m, DevState s, TvdW<W> w, DevDims d) {
  const int P = d.P, L = d.L;
  const int nch = (L + TVD_NODE_LEVELS - 1) / TVD_NODE_LEVELS;
  long q = (long)gom_block() * blockDim.x + threadIdx.x;
  if (q >= (long)P * nch) return;
  const int i = (int)(q % P), c = (int)(q / P);
  const int khi = L - c * TVD_NODE_LEVELS;                  // the chunk's top level
  const int klo = khi - TVD_NODE_LEVELS + 1 > 1 ? khi - TVD_NODE_LEVELS + 1 : 1;
  const int tn = s.top_layer_at_node[i], bn = m.bottom_layer_at_node[i];
  for (int t = 0; t < d.NT; t++) {
    double above = 0.0;
    if (khi < tn && khi >= bn) {  // the value of level khi + 1, as the loop above would leave it
      int ku = khi + 1;
      bool none;
      double v = tvd_node_mean(m, s, w, d, i, ku, t, none);
      while (none && ku != tn) {
        ku++;
        v = tvd_node_mean(m, s, w, d, i, ku, t, none);
      }
      above = v;
    }
    for (int k = khi; k >= klo; k--) {
      double v = 0.0;
      if (k <= tn && k >= bn) {
        bool none;
        v = tvd_node_mean(m, s, w, d, i, k, t, none);
        if (none && k != tn) v = above;
      }
      w.node_new[nix(t, k, i, L, P)] = v;
      above = v;
    }
  }
}

// Face values and the copy to the model arrays (solve_transport_equation_v19.f90:1457-1482)
template <typename W>
__global__ void k_tvd_face_out(DevMesh m, DevState s, TvdW<W> w, DevDims d) {
  int q = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J, P = d.P, L = d.L;
  if (q >= J * L) return;
  int j = q % J, k = q / J + 1;
  int n1 = TB(m.nodenum_at_face, 1, j, J) - 1, n2 = TB(m.nodenum_at_face, 2, j, J) - 1;
  bool in = k >= m.bottom_layer_at_face[j] && k <= s.top_layer_at_face[j];
  for (int t = 0; t < d.NT; t++) {
    double v = in ? ((double)w.node_new[nix(t, k, n1, L, P)] + w.node_new[nix(t, k, n2, L, P)]) * 0.5 : 0.0;
    double *dst = d.tid[t] == 1 ? s.salt_face : s.temp_face;
    AL(dst, k, j, J) = v;
  }
}

template <typename W>
__global__ void k_tvd_cellnode_out(DevState s, TvdW<W> w, DevDims d) {
  int q = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, P = d.P, L = d.L;
  long nc = (long)E * L, nn = (long)P * L;
  for (int t = 0; t < d.NT; t++) {
    if (q < nc) {
      double *dst = d.tid[t] == 1 ? s.salt_cell_new : s.temp_cell_new;
      dst[q] = w.con_new[(size_t)t * nc + q];
    }
    if (q < nn) {
      double *dst = d.tid[t] == 1 ? s.salt_node : s.temp_node;
      dst[q] = w.node_new[(size_t)t * nn + q];
    }
  }
}

// ================================================================ equation of state
__device__ __forceinline__ double rho_unesco(double t, double s) {
  double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  double sq = sqrt(s), sq3 = sq * sq * sq;
  return (double)(1000.0f - 0.157406f) + (double)6.793952e-2f * t - (double)9.095290e-3f * t2 +
         (double)1.001685e-4f * t3 - (double)1.120083e-6f * t4 + (double)6.536332e-9f * t5 +
         s * ((double)0.824493f - (double)4.089900e-3f * t + (double)7.643800e-5f * t2 - (double)8.246700e-7f * t3 +
              (double)5.387500e-9f * t4) +
         sq3 * ((double)(-5.72466e-3f) + (double)1.022700e-4f * t - (double)1.654600e-6f * t2) +
         (double)4.831400e-4f * (s * s);
}

__device__ double rho_pressure(double rho_surface, double t, double s, double pres) {
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

// One column of calculate_density_full.f90:36-117 / calculate_density_linear.f90:33-71.
__device__ void eos_column(const DevMesh &m, double MSL, int full, int compute, int bl, int tl, const double *T,
                           const double *S, double *R, const double *dz, int e, int n, int L, int *err, int ent) {
  for (int k = bl; k <= tl; k++) {
    double rt = AL(T, k, e, n), rs = AL(S, k, e, n);
    if (rt < -5.0 || rt > 100.0 || rs < -5.0 || rs > 100.0) {
      latch(err, GOMGPU_ERR_EOS_RANGE, ent);
      return;
    }
    if (compute) {
      double r;
      if (full) {
        r = rho_unesco(rt, rs);
        if (MSL - m.z_level[k] > 100.0) {
          double pres = 0.0;
          for (int kk = k; kk <= tl; kk++) {
            double rr = (kk == k) ? r : AL(R, kk, e, n);
            double ptmp = (double)9.81f * rr * AL(dz, kk, e, n);
            pres = (kk == k) ? pres + (double)1.0e-5f * 0.5 * ptmp : pres + (double)1.0e-5f * ptmp;
          }
          r = rho_pressure(r, rt, rs, pres);
        }
      } else {
        r = 1000.0 + (double)0.7f * rs;
      }
      AL(R, k, e, n) = r;
    }
    if (AL(R, k, e, n) < 980.0) {
      latch(err, GOMGPU_ERR_EOS_RANGE, ent);
      return;
    }
  }
  double rb = AL(R, bl, e, n), rtp = AL(R, tl, e, n);
  for (int k = 1; k <= bl - 1; k++) AL(R, k, e, n) = rb;
  for (int k = tl + 1; k <= L; k++) AL(R, k, e, n) = rtp;
}

__global__ void k_eos_nodeface(DevMesh m, DevState s, DevDims d, double MSL, int full, int idens, int *err) {
  int q = gom_block() * blockDim.x + threadIdx.x;
  const int P = d.P, J = d.J, L = d.L;
  if (q < P) {
    int tl = s.top_layer_at_node[q];
    if (tl == 0) return;
    eos_column(m, MSL, full, 1, m.bottom_layer_at_node[q], tl, s.temp_node, s.salt_node, s.rho_node, s.dz_node, q, P,
               L, err, q + 1);
  } else if (q < P + J) {
    int j = q - P;
    int tl = s.top_layer_at_face[j];
    if (tl == 0) return;
    eos_column(m, MSL, full, full || idens == 1, m.bottom_layer_at_face[j], tl, s.temp_face, s.salt_face, s.rho_face,
               s.dz_face, j, J, L, err, j + 1);
  }
}

__global__ void k_eos_cell(DevMesh m, DevState s, DevDims d, int *err) {
  int e = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, P = d.P, L = d.L;
  if (e >= E) return;
  int tl = s.top_layer_at_element[e];
  if (tl == 0) return;
  int bl = m.bottom_layer_at_element[e], N = m.tri_or_quad[e];
  int nd[4], wet[4];
  for (int l = 0; l < N; l++) {
    nd[l] = TB(m.nodenum_at_cell, l + 1, e, E) - 1;
    wet[l] = s.top_layer_at_node[nd[l]] != 0;
  }
  for (int k = bl; k <= tl; k++) {
    double r = 0.0;
    int icount = 0;
    for (int l = 0; l < N; l++)
      if (wet[l]) {
        icount++;
        r = r + AL(s.rho_node, k, nd[l], P);
      }
    if (icount == 0) {
      latch(err, GOMGPU_ERR_EOS_DRY_NODES, e + 1);
      return;
    }
    AL(s.rho_cell, k, e, E) = r / icount;
  }
  double rb = AL(s.rho_cell, bl, e, E), rtp = AL(s.rho_cell, tl, e, E);
  for (int k = 1; k <= bl - 1; k++) AL(s.rho_cell, k, e, E) = rb;
  for (int k = tl + 1; k <= L; k++) AL(s.rho_cell, k, e, E) = rtp;
}

// ================================================================ update
// Depth averages at nodes, update_variables.f90:37-74
__global__ void k_bars(DevMesh m, DevState s, DevDims d) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int P = d.P;
  if (i >= P) return;
  int tl = s.top_layer_at_node[i], bl = m.bottom_layer_at_node[i];
  int nv = tl - bl + 1;
  if (nv == 0) {
    s.ubar_node[i] = s.vbar_node[i] = s.sbar_node[i] = s.tbar_node[i] = s.rbar_node[i] = 0.0;
    return;
  }
  double su = 0.0, sv = 0.0;
  for (int k = bl - 1; k <= tl; k++) {
    su = su + A0(s.u_node, k, i, P);
    sv = sv + A0(s.v_node, k, i, P);
  }
  s.ubar_node[i] = su / (nv + 1);
  s.vbar_node[i] = sv / (nv + 1);
  double ss = 0.0, st = 0.0, sr = 0.0;
  for (int k = bl; k <= tl; k++) {
    ss = ss + AL(s.salt_node, k, i, P);
    st = st + AL(s.temp_node, k, i, P);
    sr = sr + AL(s.rho_node, k, i, P);
  }
  s.sbar_node[i] = ss / nv;
  s.tbar_node[i] = st / nv;
  s.rbar_node[i] = sr / nv;
}

// FP64 volume and salt over wet cell layers; one block partial per block, summed on the
// host in block order (deterministic).
__global__ void k_budget(DevMesh m, DevState s, DevDims d, double *part) {
  __shared__ double lv[256], ls[256];
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int E = d.E;
  double vol = 0.0, salt = 0.0;
  if (e < E) {
    int tl = s.top_layer_at_element[e], bl = m.bottom_layer_at_element[e];
    for (int k = bl; k <= tl; k++) {
      double v = m.area[e] * AL(s.dz_cell, k, e, E);
      vol = vol + v;
      salt = salt + v * AL(s.salt_cell, k, e, E);
    }
  }
  lv[threadIdx.x] = vol;
  ls[threadIdx.x] = salt;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      lv[threadIdx.x] += lv[threadIdx.x + st];
      ls[threadIdx.x] += ls[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lv[0];
    part[2 * blockIdx.x + 1] = ls[0];
  }
}

// ---- launchers: the FP64 or the FP32-storage instantiation of each transport kernel
#define TVD_GRID(n, bs) dim3(nblk((n), (bs))), dim3(bs), 0, st
void tvd_launch_init(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevForcing f, DevWork w,
                     DevDims d, TvdArgs a) {
  if (f32) k_tvd_init<float><<<TVD_GRID(n, bs)>>>(m, s, f, tvd_view<float>(w), d, a);
  else k_tvd_init<double><<<TVD_GRID(n, bs)>>>(m, s, f, tvd_view<double>(w), d, a);
}
void tvd_launch_bnd(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevForcing f, DevWork w,
                    DevDims d, TvdArgs a) {
  if (f32) k_tvd_bnd<float><<<TVD_GRID(n, bs)>>>(m, s, f, tvd_view<float>(w), d, a);
  else k_tvd_bnd<double><<<TVD_GRID(n, bs)>>>(m, s, f, tvd_view<double>(w), d, a);
}
void tvd_launch_coef_face(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevWork w, DevDims d,
                          TvdArgs a) {
  if (f32) k_tvd_coef_face<float><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<float>(w), d, a);
  else k_tvd_coef_face<double><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<double>(w), d, a);
}
void tvd_launch_coef_cell(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevWork w, DevDims d,
                          TvdArgs a) {
  if (f32) k_tvd_coef_cell<float><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<float>(w), d, a);
  else k_tvd_coef_cell<double><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<double>(w), d, a);
}
void tvd_launch_lim_cell(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevWork w, DevDims d,
                         int hlim, int first) {
  if (f32) k_tvd_lim_cell<float><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<float>(w), d, hlim, first);
  else k_tvd_lim_cell<double><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<double>(w), d, hlim, first);
}
void tvd_launch_lim_face(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevWork w, DevDims d,
                         int hlim, int first) {
  if (f32) k_tvd_lim_face<float><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<float>(w), d, hlim, first);
  else k_tvd_lim_face<double><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<double>(w), d, hlim, first);
}
void tvd_launch_rhs(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevWork w, DevDims d,
                    TvdArgs a) {
  if (f32) k_tvd_rhs<float><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<float>(w), d, a);
  else k_tvd_rhs<double><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<double>(w), d, a);
}
void tvd_launch_solve(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevWork w, DevDims d,
                      TvdArgs a, double *con_next, int full, int last, int *err) {
  if (f32) k_tvd_solve<float><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<float>(w), d, a, con_next, full, last, err);
  else k_tvd_solve<double><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<double>(w), d, a, con_next, full, last, err);
}
void tvd_launch_node(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevWork w, DevDims d) {
  if (f32) k_tvd_node<float><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<float>(w), d);
  else k_tvd_node<double><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<double>(w), d);
}
void tvd_launch_face_out(bool f32, long n, int bs, hipStream_t st, DevMesh m, DevState s, DevWork w, DevDims d) {
  if (f32) k_tvd_face_out<float><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<float>(w), d);
  else k_tvd_face_out<double><<<TVD_GRID(n, bs)>>>(m, s, tvd_view<double>(w), d);
}
void tvd_launch_cellnode_out(bool f32, long n, int bs, hipStream_t st, DevState s, DevWork w, DevDims d) {
  if (f32) k_tvd_cellnode_out<float><<<TVD_GRID(n, bs)>>>(s, tvd_view<float>(w), d);
  else k_tvd_cellnode_out<double><<<TVD_GRID(n, bs)>>>(s, tvd_view<double>(w), d);
}
#undef TVD_GRID
