// k_pcg.hip - Jacobi-preconditioned conjugate gradient of the free-surface system
// (pre_conj_grad.f90:9-230) as one persistent cooperative kernel.
//
// The iteration loop runs on the device: no host round trip per iteration. Two grid-wide
// barriers per iteration: the direction update d <- s + beta*d is folded into the next
// SpMV (q = A (s + beta d), with d double-buffered), so only the two dot products need a
// barrier. Each dot product is reduced deterministically: every thread sums its rows in a
// fixed order, blocks reduce in a fixed tree, and every block reduces the block partials in
// the same fixed tree, so all blocks agree bit for bit on alpha, beta and the stop test.
// Row sums run diagonal first, then the neighbours in the cell's face order, which is the
// order of the reference's CSR rows (solve_free_surface_equation.f90:233-258), so every
// SpMV row is bit-identical to the reference; only the dot-product association differs.
#include <hip/hip_cooperative_groups.h>

#include "gomgpu_dev.h"

namespace cg = cooperative_groups;

#define PCG_BS 256

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum of v over the block, valid in thread 0 (fixed order).
__device__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += lds[w];
  __syncthreads();
  return r;
}

// Every block reduces the nb block partials in the same order; result broadcast.
__device__ double grid_total(const double *part, int nb, double *lds) {
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[b];
  double r = block_sum(v, lds);
  if (threadIdx.x == 0) lds[8] = r;
  __syncthreads();
  r = lds[8];
  __syncthreads();
  return r;
}


__device__ __forceinline__ double row_apply(const PcgArgs &a, int i, const double *v) {
  const int n = a.n;
  double sum = a.diag[i] * v[i];
  int N = a.nfaces[i];
  for (int l = 0; l < N; l++) {
    int c = a.col[ix(l, i, n)];
    if (c != 0) sum = sum + a.off[ix(l, i, n)] * v[c - 1];
  }
  return sum;
}

// q = A (s + beta d_old) with the fused direction update written to d_new.
__device__ __forceinline__ double row_apply_dir(const PcgArgs &a, int i, const double *s, const double *dold,
                                                double beta, double &dnew_i) {
  const int n = a.n;
  double di = s[i] + beta * dold[i];
  dnew_i = di;
  double sum = a.diag[i] * di;
  int N = a.nfaces[i];
  for (int l = 0; l < N; l++) {
    int c = a.col[ix(l, i, n)];
    if (c != 0) sum = sum + a.off[ix(l, i, n)] * (s[c - 1] + beta * dold[c - 1]);
  }
  return sum;
}

// Grid-wide barrier for a co-resident grid: one device-scope atomic per block on arrival,
// the last block to arrive resets the counter and bumps the generation; the others spin
// on the generation with a short sleep. Releases and acquires are agent scope.
__device__ __forceinline__ void bar_sync(unsigned *bar, unsigned nb, unsigned &gen, int mode) {
  if (mode == 0) {
    cg::this_grid().sync();
    return;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned g = gen;
    __threadfence();
    unsigned arrived = atomicAdd(&bar[0], 1u);
    if (arrived == nb - 1) {
      atomicExch(&bar[0], 0u);
      __threadfence();
      atomicAdd(&bar[1], 1u);
    } else {
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g)
        __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
  }
  gen = gen + 1;
  __syncthreads();
}

__global__ void __launch_bounds__(PCG_BS) k_pcg(PcgArgs a) {
  __shared__ double lds[16];
  __shared__ unsigned gen0;
  if (threadIdx.x == 0) gen0 = __hip_atomic_load(&a.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  unsigned gen = gen0;
  struct {
    unsigned *bar;
    unsigned nb;
    unsigned *gen;
    int mode;
    __device__ void sync() { bar_sync(bar, nb, *gen, mode); }
  } grid{a.bar, gridDim.x, &gen, a.barrier_mode};
  const int n = a.n, nb = gridDim.x;
  const int stride = gridDim.x * blockDim.x;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  // r = b - A x, M^-1 = 1/diag, d = M^-1 r, delta = r.d   (pre_conj_grad.f90:46-92)
  double part = 0.0;
  for (int i = t0; i < n; i += stride) {
    double ax = row_apply(a, i, a.x);
    double mi = 1.0 / a.diag[i];
    double ri = a.b[i] - ax;
    double di = 0.0 + mi * ri;
    a.minv[i] = mi;
    a.r[i] = ri;
    a.d0[i] = di;
    part = part + ri * di;
  }
  double bs = block_sum(part, lds);
  if (threadIdx.x == 0) a.partB[blockIdx.x] = bs;
  grid.sync();
  double delta_new = grid_total(a.partB, nb, lds);
  double delta_0 = delta_new;
  double beta = 0.0;
  const double *dcur = a.d0;
  double *dnext = a.d1;
  bool first = true;
  int cgit2 = 0;
  for (int cgit = 1; cgit <= a.maxit; cgit++) {
    // q = A d ; dq = d.q   (:117-135)
    part = 0.0;
    for (int i = t0; i < n; i += stride) {
      double di, qi;
      if (first) {
        di = dcur[i];
        qi = row_apply(a, i, dcur);
      } else {
        qi = row_apply_dir(a, i, a.s, dcur, beta, di);
        dnext[i] = di;
      }
      a.q[i] = qi;
      part = part + di * qi;
    }
    if (!first) {
      const double *tmp = dnext;
      dnext = (double *)dcur;
      dcur = tmp;
    }
    bs = block_sum(part, lds);
    if (threadIdx.x == 0) a.partA[blockIdx.x] = bs;
    grid.sync();
    double dq = grid_total(a.partA, nb, lds);
    double alpha = delta_new / dq;
    // x, r update, s = M^-1 r, delta = r.s   (:150-183)
    part = 0.0;
    for (int i = t0; i < n; i += stride) {
      double xi = a.x[i] + alpha * dcur[i];
      double ri = a.r[i] - alpha * a.q[i];
      double si = 0.0 + a.minv[i] * ri;
      a.x[i] = xi;
      a.r[i] = ri;
      a.s[i] = si;
      part = part + ri * si;
    }
    bs = block_sum(part, lds);
    if (threadIdx.x == 0) a.partB[blockIdx.x] = bs;
    grid.sync();
    double delta_old = delta_new;
    delta_new = grid_total(a.partB, nb, lds);
    beta = delta_new / delta_old;
    first = false;
    // stop test on the ratio of r.M^-1 r (:199)
    if (sqrt(delta_new * delta_new / (delta_0 * delta_0)) < a.tol) {
      cgit2 = cgit;
      break;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.info[0] = cgit2;
    a.info[1] = cgit2 > 0 ? 1 : 0;
  }
}

__global__ void k_copy(double *dst, const double *src, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_fill(double *dst, double v, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// Host launcher: sizes the grid to the co-resident capacity (cooperative launch).
// Host launcher. The grid is at most max_blocks (the co-resident capacity, or a smaller
// tuned cap). cooperative != 0 launches through hipLaunchCooperativeKernel, which checks
// co-residency; otherwise a plain launch is used with a grid of at most one block per
// compute unit, which an otherwise idle device always holds at once.
extern "C" hipError_t gom_launch_pcg(const PcgArgs *args, int max_blocks, int cooperative, hipStream_t st) {
  PcgArgs a = *args;
  int nb = (a.n + PCG_BS - 1) / PCG_BS;
  if (nb > max_blocks) nb = max_blocks;
  if (nb < 1) nb = 1;
  if (cooperative) {
    void *kargs[] = {&a};
    return hipLaunchCooperativeKernel((const void *)k_pcg, dim3(nb), dim3(PCG_BS), kargs, 0, st);
  }
  if (a.barrier_mode == 0) return hipErrorInvalidValue;  // grid sync needs a cooperative launch
  hipLaunchKernelGGL(k_pcg, dim3(nb), dim3(PCG_BS), 0, st, a);
  return hipGetLastError();
}

extern "C" int gom_pcg_max_blocks(int device) {
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pcg, PCG_BS, 0) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  return per_cu * ncu;
}

// ---------------------------------------------------------------------------------------
// Register-resident PCG. Each thread owns up to R rows for the whole solve and keeps their
// matrix entries and x, r, M^-1, d (current direction) in registers; only the quantities
// the neighbouring rows read -- s and the direction -- go through memory (d is
// double-buffered because a row's new direction is read by its neighbours in the next
// SpMV). Per iteration and row this moves ~32 bytes instead of ~170. Rows are assigned in
// contiguous chunks, grouped so that the blocks the dispatcher places on one XCD (block id
// modulo 8) own one contiguous eighth of the rows, which keeps most neighbour reads in that
// XCD's cache. The per-row arithmetic is the same as k_pcg, statement for statement.
//
// Barrier: two levels -- the blocks of one XCD group arrive on their own counter (own
// 128-byte line), the last of them arrives on the top counter, the last group bumps the
// generation; waiters spin on the generation.
#define PCG_XCD 8

__device__ __forceinline__ void bar2_sync(unsigned *bar, unsigned nb, unsigned &gen, unsigned blk) {
  // every wave's global stores acknowledged by L2 before thread 0's agent-scope release
  // writes L2 back (CDNA counts stores in vmcnt; the workgroup barrier does not wait)
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned g = gen;
    unsigned grp = blk % PCG_XCD;
    unsigned ngrp = nb < PCG_XCD ? nb : PCG_XCD;
    unsigned gsize = nb / PCG_XCD + (grp < nb % PCG_XCD ? 1u : 0u);
    unsigned *cnt = bar + 32 * (grp + 1);  // 128-byte apart
    __threadfence();
    unsigned a = atomicAdd(cnt, 1u);
    bool last = false;
    if (a == gsize - 1) {
      atomicExch(cnt, 0u);
      __threadfence();
      unsigned b = atomicAdd(&bar[0], 1u);
      if (b == ngrp - 1) {
        atomicExch(&bar[0], 0u);
        __threadfence();
        atomicAdd(&bar[1], 1u);
        last = true;
      }
    }
    if (!last)
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) __builtin_amdgcn_s_sleep(1);
    __threadfence();
  }
  gen = gen + 1;
  __syncthreads();
}

__device__ double block_sum_bs(double v, double *lds) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += lds[w];
  __syncthreads();
  return r;
}

__device__ double grid_total_bs(const double *part, int nb, double *lds) {
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[b];
  double r = block_sum_bs(v, lds);
  if (threadIdx.x == 0) lds[16] = r;
  __syncthreads();
  r = lds[16];
  __syncthreads();
  return r;
}

template <int R, int BS>
__global__ void __launch_bounds__(BS) k_pcg_reg(PcgArgs a) {
  __shared__ double lds[24];
  __shared__ unsigned gen0;
  if (threadIdx.x == 0) gen0 = __hip_atomic_load(&a.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  unsigned gen = gen0;
  const int n = a.n, nb = gridDim.x, bs = blockDim.x;
  // chunk of this block: XCD group g = b % 8 owns a contiguous run of chunks
  const int b = blockIdx.x;
  const int ng = nb < PCG_XCD ? nb : PCG_XCD;
  const int g = b % ng, slot = b / ng;
  int chunk = 0;
  for (int q = 0; q < g; q++) chunk += nb / ng + (q < nb % ng ? 1 : 0);
  chunk += slot;
  const int rows_per_block = (n + nb - 1) / nb;
  const int row0 = chunk * rows_per_block;
  int rid[R];
  double dg[R], of[R][4], xr[R], rr[R], mi[R], dc[R], qv[R];
  int cl[R][4], nf[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    int i = row0 + threadIdx.x + k * bs;
    bool in = (threadIdx.x + k * bs < rows_per_block) && i < n;
    rid[k] = in ? i : -1;
    nf[k] = in ? a.nfaces[i] : 0;
    dg[k] = in ? a.diag[i] : 1.0;
#pragma unroll
    for (int l = 0; l < 4; l++) {
      cl[k][l] = (in && l < nf[k]) ? a.col[ix(l, i, n)] : 0;
      of[k][l] = (in && l < nf[k]) ? a.off[ix(l, i, n)] : 0.0;
    }
    xr[k] = in ? a.x[i] : 0.0;
  }
  // r = b - A x ; M^-1 ; d = M^-1 r ; delta = r.d
  double part = 0.0;
#pragma unroll
  for (int k = 0; k < R; k++) {
    if (rid[k] < 0) continue;
    int i = rid[k];
    double sum = dg[k] * xr[k];
#pragma unroll
    for (int l = 0; l < 4; l++)
      if (cl[k][l] != 0) sum = sum + of[k][l] * a.x[cl[k][l] - 1];
    mi[k] = 1.0 / dg[k];
    rr[k] = a.b[i] - sum;
    dc[k] = 0.0 + mi[k] * rr[k];
    a.d0[i] = dc[k];
    part = part + rr[k] * dc[k];
  }
  double bsum = block_sum_bs(part, lds);
  if (threadIdx.x == 0) a.partB[b] = bsum;
  bar2_sync(a.bar, nb, gen, b);
  double delta_new = grid_total_bs(a.partB, nb, lds);
  double delta_0 = delta_new;
  double beta = 0.0;
  double *dcur = a.d0, *dnext = a.d1;
  bool first = true;
  int cgit2 = 0;
  for (int cgit = 1; cgit <= a.maxit; cgit++) {
    part = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      if (rid[k] < 0) continue;
      double di = first ? dc[k] : a.s[rid[k]] + beta * dc[k];
      double sum = dg[k] * di;
#pragma unroll
      for (int l = 0; l < 4; l++) {
        int c = cl[k][l];
        if (c != 0) {
          double dn = first ? dcur[c - 1] : a.s[c - 1] + beta * dcur[c - 1];
          sum = sum + of[k][l] * dn;
        }
      }
      if (!first) {
        dnext[rid[k]] = di;
        dc[k] = di;
      }
      qv[k] = sum;
      part = part + di * sum;
    }
    if (!first) {
      double *t = dcur;
      dcur = dnext;
      dnext = t;
    }
    bsum = block_sum_bs(part, lds);
    if (threadIdx.x == 0) a.partA[b] = bsum;
    bar2_sync(a.bar, nb, gen, b);
    double dq = grid_total_bs(a.partA, nb, lds);
    double alpha = delta_new / dq;
    part = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      if (rid[k] < 0) continue;
      xr[k] = xr[k] + alpha * dc[k];
      rr[k] = rr[k] - alpha * qv[k];
      double si = 0.0 + mi[k] * rr[k];
      a.s[rid[k]] = si;
      part = part + rr[k] * si;
    }
    bsum = block_sum_bs(part, lds);
    if (threadIdx.x == 0) a.partB[b] = bsum;
    bar2_sync(a.bar, nb, gen, b);
    double delta_old = delta_new;
    delta_new = grid_total_bs(a.partB, nb, lds);
    beta = delta_new / delta_old;
    first = false;
    if (sqrt(delta_new * delta_new / (delta_0 * delta_0)) < a.tol) {
      cgit2 = cgit;
      break;
    }
  }
#pragma unroll
  for (int k = 0; k < R; k++)
    if (rid[k] >= 0) a.x[rid[k]] = xr[k];
  if (b == 0 && threadIdx.x == 0) {
    a.info[0] = cgit2;
    a.info[1] = cgit2 > 0 ? 1 : 0;
  }
}

// Launch of the register-resident PCG with bs threads per block (256, 512 or 1024) and the
// smallest R (rows per thread) that lets nb blocks cover the system; R is limited so the
// per-thread rows stay in registers (R <= 2 at 1024 threads, 4 at 512, 8 at 256). Returns
// hipErrorInvalidValue when the shape cannot cover the system (the caller then uses k_pcg).
static const void *pcg_reg_kernel(int bs, int R) {
  if (bs == 1024) {
    if (R <= 1) return (const void *)k_pcg_reg<1, 1024>;
    if (R <= 2) return (const void *)k_pcg_reg<2, 1024>;
  } else if (bs == 512) {
    if (R <= 1) return (const void *)k_pcg_reg<1, 512>;
    if (R <= 2) return (const void *)k_pcg_reg<2, 512>;
    if (R <= 4) return (const void *)k_pcg_reg<4, 512>;
  } else if (bs == 256) {
    if (R <= 1) return (const void *)k_pcg_reg<1, 256>;
    if (R <= 2) return (const void *)k_pcg_reg<2, 256>;
    if (R <= 4) return (const void *)k_pcg_reg<4, 256>;
    if (R <= 8) return (const void *)k_pcg_reg<8, 256>;
  }
  return nullptr;
}

extern "C" hipError_t gom_launch_pcg_reg(const PcgArgs *args, int nb, int bs, int cooperative, hipStream_t st) {
  PcgArgs a = *args;
  long per_block = ((long)a.n + nb - 1) / nb;
  int R = (int)((per_block + bs - 1) / bs);
  const void *kfun = pcg_reg_kernel(bs, R);
  if (!kfun) return hipErrorInvalidValue;
  void *kargs[] = {&a};
  if (cooperative) return hipLaunchCooperativeKernel(kfun, dim3(nb), dim3(bs), kargs, 0, st);
  return hipLaunchKernel(kfun, dim3(nb), dim3(bs), kargs, 0, st);
}

// Co-resident capacity (blocks) of the variant used for n rows at bs threads and nb blocks.
extern "C" int gom_pcg_reg_capacity(int device, int n, int nb, int bs) {
  long per_block = ((long)n + nb - 1) / nb;
  int R = (int)((per_block + bs - 1) / bs);
  const void *k = pcg_reg_kernel(bs, R);
  if (!k) return 0;
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, bs, 0) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  return per_cu * ncu;
}

// ---------------------------------------------------------------------------------------
// Single-barrier variant (Chronopoulos-Gear form of the same Jacobi-preconditioned CG).
// In exact arithmetic it produces the iterates of pre_conj_grad.f90 (same x_i, r_i, the
// same gamma_i = r_i . M^-1 r_i that the stop test uses); the two dot products of an
// iteration are fused into one reduction, and a row forms the new preconditioned residual
// u_j = M^-1_j (r_j - alpha (w_j + beta s_j)) of each neighbour itself from the neighbour's
// published (r, w, s), so the iteration needs one grid barrier instead of two.
//   p = u + beta p ; s = w + beta s ; x += alpha p ; r -= alpha s ; u = M^-1 r ; w = A u
//   gamma = r.u ; delta = w.u ; beta' = gamma'/gamma ; alpha' = gamma'/(delta' - beta' gamma'/alpha)
// Slot exchange of the per-block partial sums, used instead of a grid barrier inside the
// single-barrier PCG loop. slots holds 3 x nb pairs of 64-bit words; a pair is either the
// sentinel or the block's (gamma, delta) partials of one iteration. In iteration it block b
// (1) resets its pair of iteration it+1 to the sentinel, (2) waits until its row values and
// the reset are performed, (3) stores its pair of iteration it, then polls the nb pairs of
// iteration it until none is the sentinel. A pair is reset only after every block has
// published the iteration after the one it held, i.e. after every block has read it.
#define PCG_SLOT_SENTINEL 0x7ff4c0ffee0f1a6bull
#define PCG_CANONICAL_NAN 0x7ff8000000000000ull

__device__ __forceinline__ unsigned long long slot_bits(double v) {
  unsigned long long u = __double_as_longlong(v);
  return u == PCG_SLOT_SENTINEL ? PCG_CANONICAL_NAN : u;
}

// A poll gives up after PCG_SLOT_SPIN_LIMIT spins (seconds) and raises the abort word;
// every poll then stops waiting, the solve runs out its iterations on NaN totals and the
// host reports the abort. With all blocks co-resident this never triggers.
#define PCG_SLOT_SPIN_LIMIT (1u << 22)

// Row values published for the neighbouring rows (r, w, s) go through relaxed agent-scope
// atomics. On gfx950 these are the agent-coherent (sc1) forms of the plain load and store:
// a store is performed at agent scope once the issuing wave's vmcnt reaches zero, and a load
// reads the agent-coherent value. Ordering them by that wait (writer) and by the control
// dependence on the polled slot (reader) is what an agent-scope release/acquire pair does for
// atomic data; the L2 write-back and invalidate that such fences add for non-atomic data are
// not needed, and they cost more than the rest of the exchange.
__device__ __forceinline__ void pub_st(double *p, double v) {
  __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pub_ld(const double *p) {
  return __longlong_as_double(
      (long long)__hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void pub_st(float *p, double v) {
  __hip_atomic_store((unsigned *)p, (unsigned)__float_as_uint((float)v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pub_ld(const float *p) {
  return (double)__uint_as_float(__hip_atomic_load((const unsigned *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Pairs polled per lane of wave 0: at most 64 * PCG_SLOT_POLL blocks per system.
#define PCG_SLOT_POLL (PCG_MAX_SLOT_BLOCKS / 64)

// Resets this block's pair of slot (it+1)%3 at the top of iteration it (thread 0). Every
// block has read that pair's previous value (iteration it-2): this block has seen every
// block's iteration it-1 pair, which a block publishes only after reading all pairs of it-2.
__device__ __forceinline__ void cg1_reset_slot(unsigned long long *slots, int b, int nb, int it) {
  if (threadIdx.x == 0) {
    unsigned long long *rst = slots + 2 * ((size_t)((it + 1) % 3) * nb + b);
    __hip_atomic_store(rst, PCG_SLOT_SENTINEL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(rst + 1, PCG_SLOT_SENTINEL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// One iteration's reduction and exchange. (1) Every wave waits for its stores (the slot
// reset and the row publications of this iteration are then performed at agent scope) and
// leaves its partial sums in LDS; block barrier S1. (2) Thread 0 sums the wave partials and
// publishes the block's (gamma, delta) pair in slot it%3. (3) Wave 0 polls the nb pairs of
// that slot, lane l owning pairs l, l+64, ..., until none is the sentinel, adds its pairs in
// that order and reduces over the wave; block barrier S2 broadcasts the totals. Every block
// adds the same values in the same order, so all agree bit for bit. LDS: [0,2NW) wave
// partials, [2NW,2NW+2) totals; each is rewritten only after a barrier that follows its last
// read.
template <int BS>
__device__ __forceinline__ void cg1_exchange(unsigned long long *slots, int b, int nb, int it, double pg, double pd,
                                             double *lds, int *abort, double &gamma, double &delta) {
  constexpr int NW = BS / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  pg = wave_sum(pg);
  pd = wave_sum(pd);
  if (lane == 0) {
    lds[wid] = pg;
    lds[NW + wid] = pd;
  }
  __builtin_amdgcn_s_waitcnt(0);  // this wave's slot reset and row publications performed
  __syncthreads();                // S1
  const int cur = it % 3;
  if (wid == 0) {
    if (lane == 0) {
      double bg = 0.0, bd = 0.0;
      for (int w = 0; w < NW; w++) bg += lds[w];
      for (int w = 0; w < NW; w++) bd += lds[NW + w];
      unsigned long long *out = slots + 2 * ((size_t)cur * nb + b);
      __hip_atomic_store(out, slot_bits(bg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(out + 1, slot_bits(bd), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const unsigned long long *src = slots + 2 * (size_t)cur * nb;
    unsigned long long xs[PCG_SLOT_POLL], ys[PCG_SLOT_POLL];
    unsigned pending = 0;
#pragma unroll
    for (int j = 0; j < PCG_SLOT_POLL; j++) {
      xs[j] = ys[j] = 0;
      if (lane + 64 * j < nb) pending |= 1u << j;
    }
    unsigned spins = 0;
    while (pending) {
#pragma unroll
      for (int j = 0; j < PCG_SLOT_POLL; j++)
        if (pending >> j & 1u) {
          xs[j] = __hip_atomic_load(src + 2 * (lane + 64 * j), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ys[j] = __hip_atomic_load(src + 2 * (lane + 64 * j) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
      for (int j = 0; j < PCG_SLOT_POLL; j++)
        if ((pending >> j & 1u) && xs[j] != PCG_SLOT_SENTINEL && ys[j] != PCG_SLOT_SENTINEL) pending &= ~(1u << j);
      if (!pending) break;
      if ((++spins & 255u) == 0) {
        if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (spins >= PCG_SLOT_SPIN_LIMIT) {
          __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
      __builtin_amdgcn_s_sleep(1);
    }
    double vg = 0.0, vd = 0.0;
#pragma unroll
    for (int j = 0; j < PCG_SLOT_POLL; j++)
      if (lane + 64 * j < nb) {
        vg = vg + __longlong_as_double(xs[j]);
        vd = vd + __longlong_as_double(ys[j]);
      }
    vg = wave_sum(vg);
    vd = wave_sum(vd);
    if (lane == 0) {
      lds[2 * NW] = vg;
      lds[2 * NW + 1] = vd;
    }
  }
  __syncthreads();  // S2
  gamma = lds[2 * NW];
  delta = lds[2 * NW + 1];
}

#define CG1_LDS(BS) (4 * ((BS) / 64) + 2 > 48 ? 4 * ((BS) / 64) + 2 : 48)

// Block b of nb of one system: the body shared by the single-system kernel and the batched
// kernel (where b and nb are member-local). The rows a thread owns are row0 + t + k*BS for
// k < R; slices k*BS >= rows_per_block are empty for the whole block, so a launch with a
// larger R than the partition needs runs the same rows with the same arithmetic (the batched
// kernel relies on this). The neighbour values of an iteration are loaded before any of the
// row arithmetic (all rows at once for R <= 2, one row at a time for R = 4 to stay within
// the register budget), so a thread waits for a few memory latencies per iteration instead
// of one per neighbour.
template <int R, int BS, bool SLOTS>
__device__ __forceinline__ void cg1_body(const PcgArgs &a, const int b, const int nb, double *lds,
                                         unsigned *gen0) {
  if (threadIdx.x == 0) *gen0 = __hip_atomic_load(&a.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  unsigned gen = *gen0;
  const int n = a.n, bs = BS;
  const int ng = nb < PCG_XCD ? nb : PCG_XCD;
  const int g = b % ng, slot = b / ng;
  int chunk = 0;
  for (int q = 0; q < g; q++) chunk += nb / ng + (q < nb % ng ? 1 : 0);
  chunk += slot;
  const int rows_per_block = (n + nb - 1) / nb;
  const int row0 = chunk * rows_per_block;
  if (SLOTS && b == 0 && threadIdx.x == 0) a.info[2] = 0;  // abort word, ordered by the first barrier
  if (SLOTS && threadIdx.x < 6)  // this block's pair in each of the 3 slots
    __hip_atomic_store(a.slots + 2 * ((size_t)(threadIdx.x >> 1) * nb + b) + (threadIdx.x & 1),
                       PCG_SLOT_SENTINEL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // published per-row values, double-buffered by iteration parity: [r | w | s] x 2
  double *pub = a.pub;
  int rid[R], cl[R][4];
  double dg[R], of[R][4], mn[R][4], mi[R], xr[R], rr[R], uu[R], ww[R], pp[R], ss[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    int i = row0 + threadIdx.x + k * bs;
    bool in = (threadIdx.x + k * bs < rows_per_block) && i < n;
    rid[k] = in ? i : -1;
    int nf = in ? a.nfaces[i] : 0;
    dg[k] = in ? a.diag[i] : 1.0;
#pragma unroll
    for (int l = 0; l < 4; l++) {
      cl[k][l] = (in && l < nf) ? a.col[ix(l, i, n)] : 0;
      of[k][l] = (in && l < nf) ? a.off[ix(l, i, n)] : 0.0;
      mn[k][l] = cl[k][l] != 0 ? 1.0 / a.diag[cl[k][l] - 1] : 0.0;  // the neighbour's M^-1
    }
    xr[k] = in ? a.x[i] : 0.0;
    pp[k] = 0.0;
    ss[k] = 0.0;
  }
  // r0 = b - A x0 (pre_conj_grad.f90:48-60); publish r0
#pragma unroll
  for (int k = 0; k < R; k++) {
    if (rid[k] < 0) continue;
    int i = rid[k];
    double sum = dg[k] * xr[k];
#pragma unroll
    for (int l = 0; l < 4; l++)
      if (cl[k][l] != 0) sum = sum + of[k][l] * a.x[cl[k][l] - 1];
    mi[k] = 1.0 / dg[k];
    rr[k] = a.b[i] - sum;
    pub[i] = rr[k];
  }
  bar2_sync(a.bar, nb, gen, b);
  // u0 = M^-1 r0 ; w0 = A u0 ; gamma0 = r0.u0 ; delta0 = w0.u0
  double pg = 0.0, pd = 0.0;
#pragma unroll
  for (int k = 0; k < R; k++) {
    if (rid[k] < 0) continue;
    uu[k] = 0.0 + mi[k] * rr[k];
    double sum = dg[k] * uu[k];
#pragma unroll
    for (int l = 0; l < 4; l++) {
      int c = cl[k][l];
      if (c != 0) sum = sum + of[k][l] * (0.0 + mn[k][l] * pub[c - 1]);
    }
    ww[k] = sum;
    pg = pg + rr[k] * uu[k];
    pd = pd + ww[k] * uu[k];
  }
  int par = 1;  // buffer the next publication goes to
#pragma unroll
  for (int k = 0; k < R; k++) {
    if (rid[k] < 0) continue;
    size_t o = (size_t)par * 3 * n;
    pub[o + rid[k]] = rr[k];
    pub[o + n + rid[k]] = ww[k];
    pub[o + 2 * (size_t)n + rid[k]] = ss[k];
  }
  double bg = block_sum_bs(pg, lds), bd = block_sum_bs(pd, lds + 16);
  if (threadIdx.x == 0) {
    a.partA[b] = bg;
    a.partB[b] = bd;
  }
  bar2_sync(a.bar, nb, gen, b);
  double gamma = grid_total_bs(a.partA, nb, lds);
  double delta = grid_total_bs(a.partB, nb, lds + 24);
  const double gamma_0 = gamma;
  double alpha = gamma / delta, beta = 0.0;
  int cgit2 = 0;
  constexpr int H = R >= 4 ? R : 1, RH = R / H;  // gather phases and rows per phase
  for (int cgit = 1; cgit <= a.maxit; cgit++) {
    if (SLOTS) cg1_reset_slot(a.slots, b, nb, cgit);
    const size_t oin = (size_t)par * 3 * n;  // neighbours' (r, w, s) of this iteration
    par ^= 1;
    const size_t oout = (size_t)par * 3 * n;
    pg = 0.0;
    pd = 0.0;
#pragma unroll
    for (int h = 0; h < H; h++) {
      double nr[RH][4], nw[RH][4], ns[RH][4];
#pragma unroll
      for (int q = 0; q < RH; q++) {
        const int k = h * RH + q;
        if (k * bs < rows_per_block) {  // uniform over the block
#pragma unroll
          for (int l = 0; l < 4; l++) {
            const size_t j = cl[k][l] != 0 ? (size_t)(cl[k][l] - 1) : 0;  // row 0 stands in for none
            nr[q][l] = pub_ld(pub + oin + j);
            nw[q][l] = pub_ld(pub + oin + n + j);
            ns[q][l] = pub_ld(pub + oin + 2 * (size_t)n + j);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < RH; q++) {
        const int k = h * RH + q;
        if (rid[k] < 0) continue;
        pp[k] = uu[k] + beta * pp[k];
        ss[k] = ww[k] + beta * ss[k];
        xr[k] = xr[k] + alpha * pp[k];
        rr[k] = rr[k] - alpha * ss[k];
        uu[k] = 0.0 + mi[k] * rr[k];
        double sum = dg[k] * uu[k];
#pragma unroll
        for (int l = 0; l < 4; l++) {
          double sn = nw[q][l] + beta * ns[q][l];
          double rn = nr[q][l] - alpha * sn;
          double un = 0.0 + mn[k][l] * rn;
          double t = sum + of[k][l] * un;
          sum = cl[k][l] != 0 ? t : sum;
        }
        ww[k] = sum;
        pg = pg + rr[k] * uu[k];
        pd = pd + ww[k] * uu[k];
        pub_st(pub + oout + rid[k], rr[k]);
        pub_st(pub + oout + n + rid[k], ww[k]);
        pub_st(pub + oout + 2 * (size_t)n + rid[k], ss[k]);
      }
    }
    double gamma_new, delta_new;
    if (SLOTS) {
      cg1_exchange<BS>(a.slots, b, nb, cgit, pg, pd, lds, a.info + 2, gamma_new, delta_new);
    } else {
      bg = block_sum_bs(pg, lds);
      bd = block_sum_bs(pd, lds + 16);
      // alternate partial buffers so a slow block still reading the previous totals is safe
      double *PA = (cgit & 1) ? a.partC : a.partA, *PB = (cgit & 1) ? a.partD : a.partB;
      if (threadIdx.x == 0) {
        PA[b] = bg;
        PB[b] = bd;
      }
      bar2_sync(a.bar, nb, gen, b);
      gamma_new = grid_total_bs(PA, nb, lds);
      delta_new = grid_total_bs(PB, nb, lds + 24);
    }
    double beta_new = gamma_new / gamma;
    double alpha_new = gamma_new / (delta_new - beta_new * gamma_new / alpha);
    gamma = gamma_new;
    beta = beta_new;
    alpha = alpha_new;
    if (sqrt(gamma * gamma / (gamma_0 * gamma_0)) < a.tol) {
      cgit2 = cgit;
      break;
    }
  }
#pragma unroll
  for (int k = 0; k < R; k++)
    if (rid[k] >= 0) a.x[rid[k]] = xr[k];
  if (b == 0 && threadIdx.x == 0) {
    a.info[0] = cgit2;
    a.info[1] = cgit2 > 0 ? 1 : 0;
  }
}

template <int R, int BS, bool SLOTS>
__global__ void __launch_bounds__(BS) k_pcg_cg1(PcgArgs a) {
  __shared__ double lds[CG1_LDS(BS)];
  __shared__ unsigned gen0;
  cg1_body<R, BS, SLOTS>(a, blockIdx.x, gridDim.x, lds, &gen0);
}

// Memory-resident form of the same single-reduction solve, for systems whose rows do not
// fit in registers (more than 256 blocks x 2048 rows). A thread walks its rows row0 + t +
// k*BS (the assignment of the register-resident kernel) each iteration; x and the search
// direction p stay in a.x and a.d0, the published (r, w, s) in a.pub, the matrix is
// streamed. Per row the arithmetic is that of cg1_body statement for statement.
template <int BS, typename A>
__device__ __forceinline__ void cg1m_body(const A &a, const int b, const int nb, double *lds, unsigned *gen0) {
  if (threadIdx.x == 0) *gen0 = __hip_atomic_load(&a.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  unsigned gen = *gen0;
  const int n = a.n;
  const int ng = nb < PCG_XCD ? nb : PCG_XCD;
  const int g = b % ng, slot = b / ng;
  int chunk = 0;
  for (int q = 0; q < g; q++) chunk += nb / ng + (q < nb % ng ? 1 : 0);
  chunk += slot;
  const int rows_per_block = (n + nb - 1) / nb;
  const int row0 = chunk * rows_per_block;
  if (b == 0 && threadIdx.x == 0) a.info[2] = 0;
  if (threadIdx.x < 6)
    __hip_atomic_store(a.slots + 2 * ((size_t)(threadIdx.x >> 1) * nb + b) + (threadIdx.x & 1),
                       PCG_SLOT_SENTINEL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  auto *pub = a.pub;
  const size_t n1 = (size_t)n, n2 = 2 * (size_t)n;
#define CG1M_ROWS(i)                                                                     \
  for (int loc = threadIdx.x, i = row0 + loc; loc < rows_per_block && i < n; loc += BS, i += BS)
  double tol = a.tol;
  if constexpr (A::kMixed) {
    // the outer gamma of this refinement (every block sums the partials in one order) and the
    // inner tolerance that brings the outer ratio below tol; nothing to do once it is there
    double gcur = 0.0;
    for (int q = 0; q < a.ngpart; q++) gcur = gcur + a.gpart[q];
    const double g0 = a.first ? gcur : *a.g0;
    if (a.first && b == 0 && threadIdx.x == 0) *a.g0 = gcur;
    if (!(gcur > 0.0) || sqrt(gcur * gcur / (g0 * g0)) < a.tol) {
      if (b == 0 && threadIdx.x == 0) a.info[1] = 1;
      return;  // the correction stays 0 (zeroed with the residual)
    }
    tol = a.tol * (g0 / gcur);
  }
  // r0 = b - A x0 (pre_conj_grad.f90:48-60); publish r0; p = 0
  CG1M_ROWS(i) {
    double dg = a.diag[i];
    int nf = a.nfaces[i];
    double sum = dg * a.x[i];
    for (int l = 0; l < nf; l++) {
      int c = a.col[ix(l, i, n)];
      if (c != 0) sum = sum + a.off[ix(l, i, n)] * a.x[c - 1];
    }
    pub[i] = a.b[i] - sum;
    a.d0[i] = 0.0;
  }
  bar2_sync(a.bar, nb, gen, b);
  // u0, w0 = A u0, gamma0 = r0.u0, delta0 = w0.u0; publish (r0, w0, 0) to buffer 1
  double pg = 0.0, pd = 0.0;
  CG1M_ROWS(i) {
    double dg = a.diag[i], mi = 1.0 / dg, rr = pub[i];
    double uu = 0.0 + mi * rr;
    double sum = dg * uu;
    int nf = a.nfaces[i];
    for (int l = 0; l < nf; l++) {
      int c = a.col[ix(l, i, n)];
      if (c != 0) sum = sum + a.off[ix(l, i, n)] * (0.0 + (1.0 / a.diag[c - 1]) * pub[c - 1]);
    }
    double ww = sum;
    pg = pg + rr * uu;
    pd = pd + ww * uu;
    size_t o = 3 * n1;
    pub[o + i] = rr;
    pub[o + n1 + i] = ww;
    pub[o + n2 + i] = 0.0;
  }
  double bg = block_sum_bs(pg, lds), bd = block_sum_bs(pd, lds + 16);
  if (threadIdx.x == 0) {
    a.partA[b] = bg;
    a.partB[b] = bd;
  }
  bar2_sync(a.bar, nb, gen, b);
  double gamma = grid_total_bs(a.partA, nb, lds);
  double delta = grid_total_bs(a.partB, nb, lds + 24);
  const double gamma_0 = gamma;
  double alpha = gamma / delta, beta = 0.0;
  int cgit2 = 0, par = 1;
  for (int cgit = 1; cgit <= a.maxit; cgit++) {
    cg1_reset_slot(a.slots, b, nb, cgit);
    const size_t oin = (size_t)par * 3 * n1;
    par ^= 1;
    const size_t oout = (size_t)par * 3 * n1;
    pg = 0.0;
    pd = 0.0;
    CG1M_ROWS(i) {
      double dg = a.diag[i], mi = 1.0 / dg;
      double rr = pub_ld(pub + oin + i), ww = pub_ld(pub + oin + n1 + i), ss = pub_ld(pub + oin + n2 + i);
      double uu = 0.0 + mi * rr;
      double pp = uu + beta * a.d0[i];
      ss = ww + beta * ss;
      double xr = a.x[i] + alpha * pp;
      rr = rr - alpha * ss;
      uu = 0.0 + mi * rr;
      double sum = dg * uu;
      int nf = a.nfaces[i];
      for (int l = 0; l < nf; l++) {
        int c = a.col[ix(l, i, n)];
        if (c != 0) {
          double rj = pub_ld(pub + oin + c - 1), wj = pub_ld(pub + oin + n1 + c - 1),
                 sj = pub_ld(pub + oin + n2 + c - 1);
          double sn = wj + beta * sj;
          double rn = rj - alpha * sn;
          double un = 0.0 + (1.0 / a.diag[c - 1]) * rn;
          sum = sum + a.off[ix(l, i, n)] * un;
        }
      }
      ww = sum;
      pg = pg + rr * uu;
      pd = pd + ww * uu;
      a.d0[i] = pp;
      a.x[i] = xr;
      pub_st(pub + oout + i, rr);
      pub_st(pub + oout + n1 + i, ww);
      pub_st(pub + oout + n2 + i, ss);
    }
    double gamma_new, delta_new;
    cg1_exchange<BS>(a.slots, b, nb, cgit, pg, pd, lds, a.info + 2, gamma_new, delta_new);
    double beta_new = gamma_new / gamma;
    double alpha_new = gamma_new / (delta_new - beta_new * gamma_new / alpha);
    gamma = gamma_new;
    beta = beta_new;
    alpha = alpha_new;
    if constexpr (A::kMixed) cgit2 = cgit;  // iterations run, converged or not
    if (sqrt(gamma * gamma / (gamma_0 * gamma_0)) < tol) {
      cgit2 = cgit;
      break;
    }
  }
#undef CG1M_ROWS
  if (b == 0 && threadIdx.x == 0) {
    if constexpr (A::kMixed) {
      a.info[0] = a.info[0] + cgit2;  // the outer check (k_pcg_mixed_final) sets info[1]
    } else {
      a.info[0] = cgit2;
      a.info[1] = cgit2 > 0 ? 1 : 0;
    }
  }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_pcg_cg1m(PcgArgs a) {
  __shared__ double lds[CG1_LDS(BS)];
  __shared__ unsigned gen0;
  cg1m_body<BS>(a, blockIdx.x, gridDim.x, lds, &gen0);
}

template <int BS>
__global__ void __launch_bounds__(BS) k_pcg_cg1m32(PcgArgs32 a) {
  __shared__ double lds[CG1_LDS(BS)];
  __shared__ unsigned gen0;
  cg1m_body<BS>(a, blockIdx.x, gridDim.x, lds, &gen0);
}

#define CG1M_BS 1024
// Blocks of the memory-resident solve for n rows: a function of n alone (at most 256).
extern "C" int gom_pcg_cg1m_blocks(long n) {
  long nb = (n + 2047) / 2048;
  return (int)(nb < 1 ? 1 : (nb > 256 ? 256 : nb));
}

extern "C" hipError_t gom_launch_pcg_cg1m(const PcgArgs *args, int nb, int cooperative, hipStream_t st) {
  PcgArgs a = *args;
  if (nb > CG1M_BS || nb > PCG_MAX_SLOT_BLOCKS) return hipErrorInvalidValue;
  void *kargs[] = {&a};
  const void *k = (const void *)k_pcg_cg1m<CG1M_BS>;
  if (cooperative) return hipLaunchCooperativeKernel(k, dim3(nb), dim3(CG1M_BS), kargs, 0, st);
  return hipLaunchKernel(k, dim3(nb), dim3(CG1M_BS), kargs, 0, st);
}

extern "C" int gom_pcg_cg1m_capacity(int device) {
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pcg_cg1m<CG1M_BS>, CG1M_BS, 0) !=
      hipSuccess)
    return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  return per_cu * ncu;
}

// Several independent systems in one launch: member m owns blocks base[m] .. base[m] +
// nbm[m] - 1 (base[m] a multiple of 8, so the member-local block index keeps the XCD
// grouping); blocks in the padding between members return at once. Each member runs the
// slot-exchange solve among its own blocks with its own work arrays, exactly as it would
// alone with the same block count.
template <int R, int BS>
__global__ void __launch_bounds__(BS) k_pcg_cg1_batch(const PcgArgs *members, const int *base, const int *nbm,
                                                      int nmem) {
  __shared__ double lds[CG1_LDS(BS)];
  __shared__ unsigned gen0;
  __shared__ int mm;
  if (threadIdx.x == 0) {
    int m = 0;
    while (m + 1 < nmem && base[m + 1] <= (int)blockIdx.x) m++;
    mm = m;
  }
  __syncthreads();
  const int m = mm;
  const int b = (int)blockIdx.x - base[m];
  const int nb = nbm[m];
  if (b >= nb) return;
  const PcgArgs a = members[m];
  cg1_body<R, BS, true>(a, b, nb, lds, &gen0);
}

extern "C" hipError_t gom_launch_pcg_cg1_batch(const PcgArgs *d_members, const int *d_base, const int *d_nbm,
                                               int nmem, int total_blocks, hipStream_t st) {
  void *kargs[] = {&d_members, &d_base, &d_nbm, &nmem};
  return hipLaunchKernel((const void *)k_pcg_cg1_batch<4, 512>, dim3(total_blocks), dim3(512), kargs, 0, st);
}

extern "C" int gom_pcg_cg1_batch_capacity(int device) {
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pcg_cg1_batch<4, 512>, 512, 0) !=
      hipSuccess)
    return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  return per_cu * ncu;
}

template <bool SLOTS>
static const void *pcg_cg1_kernel_t(int bs, int R) {
  if (bs == 1024) {
    if (R <= 1) return (const void *)k_pcg_cg1<1, 1024, SLOTS>;
    if (R <= 2) return (const void *)k_pcg_cg1<2, 1024, SLOTS>;
  } else if (bs == 512) {
    if (R <= 1) return (const void *)k_pcg_cg1<1, 512, SLOTS>;
    if (R <= 2) return (const void *)k_pcg_cg1<2, 512, SLOTS>;
    if (R <= 4) return (const void *)k_pcg_cg1<4, 512, SLOTS>;
  } else if (bs == 256) {
    if (R <= 1) return (const void *)k_pcg_cg1<1, 256, SLOTS>;
    if (R <= 2) return (const void *)k_pcg_cg1<2, 256, SLOTS>;
    if (R <= 4) return (const void *)k_pcg_cg1<4, 256, SLOTS>;
    if (R <= 8) return (const void *)k_pcg_cg1<8, 256, SLOTS>;
  }
  return nullptr;
}

static const void *pcg_cg1_kernel(int bs, int R, int slots) {
  return slots ? pcg_cg1_kernel_t<true>(bs, R) : pcg_cg1_kernel_t<false>(bs, R);
}

extern "C" hipError_t gom_launch_pcg_cg1(const PcgArgs *args, int nb, int bs, int slots, int cooperative,
                                         hipStream_t st) {
  PcgArgs a = *args;
  long per_block = ((long)a.n + nb - 1) / nb;
  int R = (int)((per_block + bs - 1) / bs);
  if (slots && nb > PCG_MAX_SLOT_BLOCKS) return hipErrorInvalidValue;
  const void *kfun = pcg_cg1_kernel(bs, R, slots);
  if (!kfun) return hipErrorInvalidValue;
  void *kargs[] = {&a};
  if (cooperative) return hipLaunchCooperativeKernel(kfun, dim3(nb), dim3(bs), kargs, 0, st);
  return hipLaunchKernel(kfun, dim3(nb), dim3(bs), kargs, 0, st);
}

extern "C" int gom_pcg_cg1_capacity(int device, int n, int nb, int bs, int slots) {
  long per_block = ((long)n + nb - 1) / nb;
  int R = (int)((per_block + bs - 1) / bs);
  const void *k = pcg_cg1_kernel(bs, R, slots);
  if (!k) return 0;
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, bs, 0) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  return per_cu * ncu;
}

// ---------------------------------------------------------------------------------------
// Mixed-precision PCG (GOMGPU_MIXED): iterative refinement of the same Jacobi-preconditioned
// system. Each refinement: the FP64 residual r = b - A x and its gamma = r.M^-1 r partials
// (k_pcg_mixed_res, which also zeroes the correction and, the first time, writes the FP32
// matrix); the FP32-storage single-reduction solve of A e = r to the inner tolerance that
// takes the outer ratio below tol (k_pcg_cg1m32); x = x + e in FP64 (k_pcg_mixed_update).
// A final residual decides convergence by the reference's test sqrt(g^2/g0^2) < tol
// (pre_conj_grad.f90:199) on the FP64 residual. The sequence is fixed (no host round trip):
// once converged, further refinements return at once.
#define PCG_MIXED_REFINE 3
#define PCG_RES_BS 256
#define PCG_RES_BLOCKS 1024

__global__ void __launch_bounds__(PCG_RES_BS) k_pcg_mixed_res(PcgArgs a, PcgArgs32 m, int first, int write_r) {
  __shared__ double lds[PCG_RES_BS / 64];
  const int n = a.n;
  double g = 0.0;
  for (int i = blockIdx.x * PCG_RES_BS + threadIdx.x; i < n; i += gridDim.x * PCG_RES_BS) {
    const double dg = a.diag[i];
    double sum = dg * a.x[i];
    const int nf = a.nfaces[i];
    for (int l = 0; l < nf; l++) {
      const int c = a.col[ix(l, i, n)];
      if (c != 0) sum = sum + a.off[ix(l, i, n)] * a.x[c - 1];
    }
    const double r = a.b[i] - sum;
    g = g + r * (r / dg);
    if (write_r) {
      ((float *)m.b)[i] = (float)r;
      m.x[i] = 0.0;
    }
    if (first) {
      ((float *)m.diag)[i] = (float)dg;
      for (int l = 0; l < 4; l++) ((float *)m.off)[ix(l, i, n)] = (float)a.off[ix(l, i, n)];
    }
  }
  g = wave_sum(g);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < PCG_RES_BS / 64; w++) t += lds[w];
    ((double *)m.gpart)[blockIdx.x] = t;
  }
}

__global__ void k_pcg_mixed_update(PcgArgs a, PcgArgs32 m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n) a.x[i] = a.x[i] + m.x[i];
}

__global__ void k_pcg_mixed_final(PcgArgs32 m, double tol) {
  if (threadIdx.x != 0) return;
  double g = 0.0;
  for (int q = 0; q < m.ngpart; q++) g = g + m.gpart[q];
  const double g0 = *m.g0;
  m.info[1] = (g0 > 0.0 && sqrt(g * g / (g0 * g0)) < tol) || g0 == 0.0 ? 1 : 0;
}

__global__ void k_pcg_mixed_init(int *info) {
  if (threadIdx.x == 0) {
    info[0] = 0;
    info[1] = 0;
    info[2] = 0;
  }
}

// The sequence on the buffers of a PCG context: m holds the FP32 arrays and the partials.
extern "C" hipError_t gom_launch_pcg_mixed(const PcgArgs *a64, const PcgArgs32 *m32, int cooperative,
                                           hipStream_t st) {
  PcgArgs a = *a64;
  PcgArgs32 m = *m32;
  const int nbr = nblk(a.n, PCG_RES_BS) < PCG_RES_BLOCKS ? nblk(a.n, PCG_RES_BS) : PCG_RES_BLOCKS;
  m.ngpart = nbr;
  const int nb = gom_pcg_cg1m_blocks(a.n);
  if (nb > CG1M_BS || nb > PCG_MAX_SLOT_BLOCKS) return hipErrorInvalidValue;
  k_pcg_mixed_init<<<1, 64, 0, st>>>(m.info);
  for (int r = 0; r < PCG_MIXED_REFINE; r++) {
    m.first = r == 0;
    k_pcg_mixed_res<<<nbr, PCG_RES_BS, 0, st>>>(a, m, r == 0, 1);
    void *kargs[] = {&m};
    const void *k = (const void *)k_pcg_cg1m32<CG1M_BS>;
    hipError_t e = cooperative ? hipLaunchCooperativeKernel(k, dim3(nb), dim3(CG1M_BS), kargs, 0, st)
                               : hipLaunchKernel(k, dim3(nb), dim3(CG1M_BS), kargs, 0, st);
    if (e != hipSuccess) return e;
    k_pcg_mixed_update<<<nblk(a.n, 256), 256, 0, st>>>(a, m);
  }
  k_pcg_mixed_res<<<nbr, PCG_RES_BS, 0, st>>>(a, m, 0, 0);
  k_pcg_mixed_final<<<1, 64, 0, st>>>(m, a.tol);
  return hipGetLastError();
}

extern "C" int gom_pcg_cg1m32_capacity(int device) {
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pcg_cg1m32<CG1M_BS>, CG1M_BS, 0) !=
      hipSuccess)
    return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  return per_cu * ncu;
}
