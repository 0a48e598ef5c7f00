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
