// k_boundary.hip - boundary forcing of one step on the device (GOMGPU_STAGE_BOUNDARY):
//   the open-boundary elevations of compute_boundary_eta.f90:72-124 (harmonic sets,
//     ob_eta_type 1, and elevation series, type 2),
//   the river discharge Q_add of calculate_Q.f90:21-45 (q_interp_method 1),
//   the boundary concentrations of the transport sub-cycles, con_at_ob and con_at_Qb, of
//     solve_transport_equation_v19.f90:229-397 (TVD) or find_openboundary_salt_temp_v3.f90:
//     17-97 (ELM transport).
// Paths relative to GOM_Rv1.1.0/source. The step's time and spin-up factors are the device
// scalars the host uploads with the forcing; the tables (harmonic constituents, series,
// which boundary reads what) are uploaded once by gomgpu_set_boundary.
//
// One block. Its threads first evaluate every harmonic set at both times and every series
// at every time the step needs (the first interval that brackets the time, as the reference
// scans for it), then walk the boundaries in the reference's loop order. Where the reference
// carries the interpolated value of one boundary to the next (a series that does not cover
// the time leaves its local v2 as the previous boundary set it), the walk is a block-wide
// scan for the last position that set the value.
#include "gomgpu_dev.h"

// Time slots of a step's series lookups: julian_day*86400 - dt, julian_day*86400, then
// (julian_day*86400 - dt) + (sub + 1) dt/Nt for the TVD sub-cycles.
#define BND_SLOTS(Nt) (2 + (Nt))

__device__ __forceinline__ double bnd_time(int slot, double jd, double dt, int Nt) {
  if (slot == 0) return jd * 86400.0 - dt;
  if (slot == 1) return jd * 86400.0;
  return (jd * 86400.0 - dt) + (slot - 1) * (dt / Nt);
}

// First interval [t(i-1), t(i)] of series s whose ends (shifted by reference_diff_days)
// bracket u; the interpolated value, or false when none does. A series whose times never
// decrease is searched by bisection: its first bracketing interval is the one ending at
// the first time >= u (or the first interval when u equals the first time).
__device__ bool bnd_series(const DevBnd &b, int s, double u, double *v) {
  const int i0 = b.s_first[s], n = b.s_first[s + 1] - i0;
  const double *t = b.s_time + i0, *y = b.s_val + i0;
  const double sh = b.shift;
  int i = -1;
  if (b.s_mono[s]) {
    int lo = 0, hi = n;  // first index with t - sh >= u
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (t[mid] - sh >= u) hi = mid;
      else lo = mid + 1;
    }
    if (lo == 0) {
      if (n > 1 && t[0] - sh == u) i = 1;
    } else if (lo < n) {
      i = lo;
    }
  } else {
    for (int k = 1; k < n; k++) {
      if (u >= t[k - 1] - sh && u <= t[k] - sh) {
        i = k;
        break;
      }
    }
  }
  if (i < 0) return false;
  const double u1 = t[i - 1] - sh, u3 = t[i] - sh;
  *v = (u - u1) * (y[i] - y[i - 1]) / (u3 - u1) + y[i - 1];
  return true;
}

// The walk of one carried sequence, positions 0..np-1 in the reference's loop order. For
// position p, Seq::at(p, &val) tells whether p sets the carried value (and to what);
// Seq::out(p, v2) receives the value in force at p (0 before the first set, as the host's
// local starts).
template <typename Seq>
__device__ void bnd_scan(const Seq &q, int np) {
  __shared__ int sidx[2][256];
  __shared__ int carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = -1;
  __syncthreads();
  for (int base = 0; base < np; base += 256) {
    const int p = base + tid;
    double val;
    int cur = 0;
    sidx[0][tid] = (p < np && q.at(p, &val)) ? p : -1;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {  // inclusive max-scan (Hillis-Steele)
      int v = sidx[cur][tid];
      if (tid >= off) v = max(v, sidx[cur][tid - off]);
      sidx[cur ^ 1][tid] = v;
      cur ^= 1;
      __syncthreads();
    }
    const int last = max(sidx[cur][tid], carry);
    if (p < np) {
      double v2 = 0.0;
      if (last >= 0) q.at(last, &v2);
      q.out(p, v2);
    }
    __syncthreads();
    if (tid == 255) carry = max(carry, sidx[cur][255]);
    __syncthreads();
  }
}

struct BndCtx {
  DevBnd b;
  double *eta_old, *eta_new, *ob, *qb;
  int NOB, NQB, NT, Nt, nslot, elm;
  int tid[2];
  double sp_tide, sp_bc;
  __device__ bool job(int s1, int slot, double *v) const {  // series s1 (1-based) at a slot
    const int j = (s1 - 1) * nslot + slot;
    if (!b.job_found[j]) return false;
    *v = b.job_val[j];
    return true;
  }
};

// Elevation-series boundaries at one time slot (compute_boundary_eta.f90:92-124).
struct SeqEta {
  BndCtx c;
  int slot;
  __device__ bool at(int i, double *v) const {
    return c.b.ob_type[i] == 2 && c.job(c.b.ob_id[i], slot, v);
  }
  __device__ void out(int i, double v2) const {
    if (c.b.ob_type[i] != 2) return;
    (slot == 0 ? c.eta_old : c.eta_new)[i] = c.sp_tide * v2;
  }
};

// TVD boundary concentrations: sub-cycle, then open boundary, then tracer; then river
// cell, then tracer (solve_transport_equation_v19.f90:229-397).
struct SeqTvd {
  BndCtx c;
  __device__ void decode(int p, int &sub, int &riv, int &i, int &t) const {
    const int per = (c.NOB + c.NQB) * c.NT;
    sub = p / per;
    int r = p - sub * per;
    riv = r >= c.NOB * c.NT;
    if (riv) r -= c.NOB * c.NT;
    i = r / c.NT;
    t = r - i * c.NT;
  }
  __device__ int sid(int riv, int i, int t) const {
    const int *ser = riv ? c.b.qb_ser : c.b.ob_ser;
    const int id = c.tid[t];
    return id == 1 ? ser[2 * i] : (id == 2 ? ser[2 * i + 1] : 0);
  }
  __device__ bool at(int p, double *v) const {
    int sub, riv, i, t;
    decode(p, sub, riv, i, t);
    const int s = sid(riv, i, t);
    return s > 0 && c.job(s, 2 + sub, v);
  }
  __device__ void out(int p, double v2) const {
    int sub, riv, i, t;
    decode(p, sub, riv, i, t);
    double *dst = riv ? c.qb : c.ob;
    dst[sub + (size_t)c.Nt * (t + (size_t)c.NT * i)] = sid(riv, i, t) > 0 ? c.sp_bc * v2 : 0.0;
  }
};

// ELM boundary values: open boundary, then salt and temperature; then river cell, then
// salt and temperature (find_openboundary_salt_temp_v3.f90:17-97). An open boundary with
// no series takes ref_salt / ref_temp (which the carried value then holds); a river cell
// with no series keeps its previous value.
struct SeqElm {
  BndCtx c;
  __device__ bool at(int p, double *v) const {
    const int riv = p >= 2 * c.NOB;
    const int r = riv ? p - 2 * c.NOB : p, i = r >> 1, t = r & 1;
    const int on = t == 0 ? c.b.is_salt : c.b.is_temp;
    const int s = (riv ? c.b.qb_ser : c.b.ob_ser)[2 * i + t];
    if (on == 1 && s > 0) return c.job(s, 1, v);
    if (riv) return false;
    *v = t == 0 ? c.b.ref_salt : c.b.ref_temp;
    return true;
  }
  __device__ void out(int p, double v2) const {
    const int riv = p >= 2 * c.NOB;
    const int r = riv ? p - 2 * c.NOB : p, i = r >> 1, t = r & 1;
    const int on = t == 0 ? c.b.is_salt : c.b.is_temp;
    const int s = (riv ? c.b.qb_ser : c.b.ob_ser)[2 * i + t];
    if (riv && !(on == 1 && s > 0)) return;
    (riv ? c.qb : c.ob)[t + 2 * (size_t)i] = c.sp_bc * v2;
  }
};

__global__ void __launch_bounds__(256) k_boundary(DevBnd b, DevForcing f, DevDims d, double dt, int Nt, int elm) {
  const int tid = threadIdx.x;
  const double sp_bc = f.scal[0], sp_tide = f.scal[1], elapsed = f.scal[2], jd = f.scal[6];
  const int nslot = BND_SLOTS(Nt);
  // harmonic sets at elapsed - dt and elapsed (compute_boundary_eta.f90:74-89)
  const double deg2rad = M_PI / 180.0;
  for (int h = tid; h < b.nh; h += blockDim.x) {
    double o = 0.0, n = 0.0;
    const double sh = b.h_shift[h];
    for (int k = b.h_first[h]; k < b.h_first[h + 1]; k++) {
      const double *cn = b.h_con + 5 * (size_t)k;  // period [h], amplitude, nodal factor, V0, phase
      const double per = cn[0], amp = cn[1], fac = cn[2], v0 = cn[3], ph = cn[4];
      const double w = 2.0 * M_PI / (per * 3600.0);
      o = o + sp_tide * amp * fac * gom_cos(w * (elapsed - dt) + v0 * deg2rad - ph * deg2rad - sh * deg2rad);
      n = n + sp_tide * amp * fac * gom_cos(w * elapsed + v0 * deg2rad - ph * deg2rad - sh * deg2rad);
    }
    b.h_old[h] = o;
    b.h_new[h] = n;
  }
  // every series at every time slot of the step
  for (int j = tid; j < b.ns * nslot; j += blockDim.x) {
    const int s = j / nslot, slot = j - s * nslot;
    double v = 0.0;
    const bool ok = bnd_series(b, s, bnd_time(slot, jd, dt, Nt), &v);
    b.job_found[j] = ok ? 1 : 0;
    b.job_val[j] = v;
  }
  __syncthreads();
  BndCtx c;
  c.b = b;
  c.eta_old = (double *)f.eta_at_ob_old;
  c.eta_new = (double *)f.eta_at_ob_new;
  c.ob = (double *)f.con_at_ob;
  c.qb = (double *)f.con_at_Qb;
  c.NOB = d.NOB;
  c.NQB = d.NQB;
  c.NT = d.NT;
  c.Nt = Nt;
  c.nslot = nslot;
  c.elm = elm;
  c.tid[0] = d.tid[0];
  c.tid[1] = d.tid[1];
  c.sp_tide = sp_tide;
  c.sp_bc = sp_bc;
  // harmonic boundaries take their set's sums
  for (int i = tid; i < b.nob; i += blockDim.x)
    if (b.ob_type[i] == 1) {
      c.eta_old[i] = b.h_old[b.ob_id[i] - 1];
      c.eta_new[i] = b.h_new[b.ob_id[i] - 1];
    }
  // river discharge: a cell whose series does not cover the time keeps its previous Q_add
  for (int i = tid; i < b.nqb; i += blockDim.x) {
    double v;
    if (b.qb_qser[i] > 0 && c.job(b.qb_qser[i], 1, &v)) ((double *)f.Q_add)[i] = v * b.qb_portion[i];
  }
  if (b.any_eta_series) {
    bnd_scan(SeqEta{c, 0}, b.nob);
    bnd_scan(SeqEta{c, 1}, b.nob);
  }
  if (b.tracers) {
    if (elm) bnd_scan(SeqElm{c}, 2 * (b.nob + b.nqb));
    else bnd_scan(SeqTvd{c}, Nt * (b.nob + b.nqb) * d.NT);
  }
}

extern "C" hipError_t gom_launch_boundary(const DevBnd *b, const DevForcing *f, const DevDims *d, double dt, int Nt,
                                          int elm, hipStream_t st) {
  k_boundary<<<1, 256, 0, st>>>(*b, *f, *d, dt, Nt, elm);
  return hipGetLastError();
}
