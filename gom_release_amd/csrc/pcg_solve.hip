// pcg_solve.hip - the standalone Jacobi-preconditioned CG of pre_conj_grad
// (GOM_Rv1.1.0/source/pre_conj_grad.f90:9-230) behind the C ABI: a 1-based CSR system that
// is already in device memory, solved by the single-reduction PCG kernels of k_pcg.hip.
//
// The reference builds its matrix with the diagonal first and then the cell's neighbours in
// face order, leaving out absent neighbours (solve_free_surface_equation.f90:233-258), so a
// row holds at most 5 entries. The solve converts the CSR once into the row layout the PCG
// kernels read (diagonal, then up to 4 off-diagonal slots in the CSR's order, 0 = none),
// then runs the register-resident solve (k_pcg_cg1, up to 524,288 rows) or the
// memory-resident one (k_pcg_cg1m) above that. Row sums keep the CSR's order, so with the
// diagonal first every matrix-vector row is the reference's sum; only the dot products are
// associated differently (a fixed tree instead of a serial sum).
#include <stdio.h>
#include <string.h>

#include "pcg_solve.h"

extern "C" hipError_t gom_launch_pcg_cg1(const PcgArgs *args, int nb, int bs, int slots, int cooperative,
                                         hipStream_t st);
extern "C" int gom_pcg_cg1_capacity(int device, int n, int nb, int bs, int slots);
extern "C" int gom_pcg_cg1m_blocks(long n);
extern "C" hipError_t gom_launch_pcg_cg1m(const PcgArgs *args, int nb, int cooperative, hipStream_t st);
extern "C" int gom_pcg_cg1m_capacity(int device);

size_t gom_pcg_ws_carve(long n, char *base, PcgWs *w) {
  size_t off = 0;
  auto seg = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t N = (size_t)n;
  w->diag = (double *)seg(N * sizeof(double));
  w->off = (double *)seg(4 * N * sizeof(double));
  w->d0 = (double *)seg(N * sizeof(double));
  w->pub = (double *)seg(6 * N * sizeof(double));
  w->part = (double *)seg(4 * PCG_PART_STRIDE * sizeof(double));
  w->slots = (unsigned long long *)seg(6 * PCG_MAX_SLOT_BLOCKS * sizeof(unsigned long long));
  w->col = (int32_t *)seg(4 * N * sizeof(int32_t));
  w->nf = (int32_t *)seg(N * sizeof(int32_t));
  w->info = (int *)seg(4 * sizeof(int));
  w->bad = (int *)seg(2 * sizeof(int));
  w->bar = (unsigned *)seg(512 * sizeof(unsigned));
  return off;
}

// Words that must start at zero in every solve: the barrier counters and generation, the
// result words and the malformed-row flag (a kernel, not a memset, so the solve can sit in
// a captured graph).
__global__ void k_pcg_ws_clear(unsigned *bar, int *info, int *bad) {
  const int t = threadIdx.x;
  for (int i = t; i < 512; i += blockDim.x) bar[i] = 0u;
  if (t < 4) info[t] = 0;
  if (t < 2) bad[t] = 0;
}

// One thread per row: the CSR row (rowstr 1-based, colidx 1-based) into the diagonal and up
// to 4 off-diagonal slots. A row with no diagonal, two diagonals, more than 4 off-diagonal
// entries, a column outside 1..n or a pointer outside 1..nz+1 is marked bad (and its slots
// hold what was read before the fault, never used: the solve is refused).
__global__ void k_csr_rows(int n, int nz, const double *A, const int32_t *colidx, const int32_t *rowstr,
                           double *diag, double *off, int32_t *col, int32_t *nf, int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k0 = rowstr[i] - 1, k1 = rowstr[i + 1] - 1;
  bool ok = k0 >= 0 && k1 >= k0 && k1 <= nz && k1 - k0 <= 5;
  double dv = 1.0, o[4] = {0.0, 0.0, 0.0, 0.0};
  int c[4] = {0, 0, 0, 0}, cnt = 0;
  bool hasd = false;
  if (ok) {
    for (int k = k0; k < k1; k++) {
      const int cc = colidx[k];
      if (cc < 1 || cc > n) {
        ok = false;
        break;
      }
      if (cc == i + 1) {
        if (hasd) {
          ok = false;
          break;
        }
        hasd = true;
        dv = A[k];
      } else {
        if (cnt == 4) {
          ok = false;
          break;
        }
        o[cnt] = A[k];
        c[cnt] = cc;
        cnt++;
      }
    }
  }
  if (!hasd) ok = false;
  diag[i] = dv;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    off[ix(l, i, n)] = o[l];
    col[ix(l, i, n)] = c[l];
  }
  nf[i] = cnt;
  if (!ok) {
    atomicOr(&bad[0], 1);
    atomicMax(&bad[1], i + 1);
  }
}

static int pcg_fail(char *msg, size_t cap, int rc, const char *text) {
  if (msg && cap) snprintf(msg, cap, "%s", text);
  return rc;
}

// The solve on a workspace of at least gom_pcg_ws_carve(n) bytes. Synchronous: returns after the
// solve, with the iteration count of the reference's cgit2 (0 when it did not converge).
int gom_pcg_solve_on(void *ws, int na, int nz, const double *A, const int32_t *colidx, const int32_t *rowstr,
                     const double *b, double *x, int max_iteration, double error_tolerance, int *iterations,
                     hipStream_t st, char *msg, size_t msgcap) {
  PcgWs w;
  gom_pcg_ws_carve(na, (char *)ws, &w);
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return pcg_fail(msg, msgcap, GOMGPU_E_HIP, "hipGetDevice failed");
  k_pcg_ws_clear<<<1, 256, 0, st>>>(w.bar, w.info, w.bad);
  k_csr_rows<<<nblk(na, 256), 256, 0, st>>>(na, nz, A, colidx, rowstr, w.diag, w.off, w.col, w.nf, w.bad);
  if (hipGetLastError() != hipSuccess) return pcg_fail(msg, msgcap, GOMGPU_E_HIP, "CSR conversion launch failed");
  int hbad[2] = {0, 0};
  if (hipMemcpyAsync(hbad, w.bad, sizeof hbad, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return pcg_fail(msg, msgcap, GOMGPU_E_HIP, "CSR conversion failed");
  if (hbad[0]) {
    if (msg && msgcap)
      snprintf(msg, msgcap,
               "pcg_solve: CSR row %d is not a free-surface row (needs exactly one diagonal entry, at most 4 "
               "off-diagonal entries, columns in 1..na and row pointers in 1..nz+1)",
               hbad[1]);
    return GOMGPU_E_ARG;
  }
  PcgArgs a;
  memset(&a, 0, sizeof a);
  a.diag = w.diag;
  a.off = w.off;
  a.b = b;
  a.col = w.col;
  a.nfaces = w.nf;
  a.x = x;
  a.d0 = w.d0;
  a.partA = w.part;
  a.partB = w.part + PCG_PART_STRIDE;
  a.partC = w.part + 2 * PCG_PART_STRIDE;
  a.partD = w.part + 3 * PCG_PART_STRIDE;
  a.pub = w.pub;
  a.slots = w.slots;
  a.info = w.info;
  a.bar = w.bar;
  a.barrier_mode = 1;
  a.n = na;
  a.maxit = max_iteration;
  a.tol = error_tolerance;
  hipError_t le = hipErrorInvalidValue;
  const int nb = pcg_blocks_for(na, 0);
  if (nb <= PCG_MAX_SLOT_BLOCKS && nb <= gom_pcg_cg1_capacity(device, na, nb, 512, 1))
    le = gom_launch_pcg_cg1(&a, nb, 512, 1, 1, st);
  if (le == hipErrorInvalidValue) {
    const int nbm = gom_pcg_cg1m_blocks(na);
    if (nbm <= gom_pcg_cg1m_capacity(device)) le = gom_launch_pcg_cg1m(&a, nbm, 1, st);
  }
  if (le != hipSuccess) return pcg_fail(msg, msgcap, GOMGPU_E_HIP, "PCG launch failed");
  int info[4] = {0, 0, 0, 0};
  if (hipMemcpyAsync(info, w.info, sizeof info, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return pcg_fail(msg, msgcap, GOMGPU_E_HIP, "PCG solve failed");
  if (info[2]) return pcg_fail(msg, msgcap, GOMGPU_E_HIP, "PCG block exchange timed out");
  if (iterations) *iterations = info[1] ? info[0] : 0;
  if (!info[1]) return pcg_fail(msg, msgcap, GOMGPU_W_PCG_NOT_CONVERGED, "PCG did not converge");
  return GOMGPU_OK;
}

thread_local char g_pcg_msg[256] = "";

int gom_pcg_check(int na, int nz, const double *A, const int32_t *colidx, const int32_t *rowstr,
                     const double *b, double *x, int max_iteration, int *iterations) {
  if (iterations) *iterations = 0;
  if (na < 1 || (long)na > PCG_SOLVE_MAX_ROWS || nz < na || !A || !colidx || !rowstr || !b || !x ||
      max_iteration < 0) {
    snprintf(g_pcg_msg, sizeof g_pcg_msg, "pcg_solve: bad arguments (na %d, nz %d, max_iteration %d)", na, nz,
             max_iteration);
    return GOMGPU_E_ARG;
  }
  return GOMGPU_OK;
}

extern "C" long long gomgpu_pcg_workspace_bytes(int na) {
  if (na < 1 || (long)na > PCG_SOLVE_MAX_ROWS) return -1;
  PcgWs w;
  return (long long)gom_pcg_ws_carve(na, nullptr, &w);
}

extern "C" int gomgpu_pcg_solve(int na, int nz, const double *A, const int32_t *colidx, const int32_t *rowstr,
                                const double *b, double *x, int max_iteration, double error_tolerance,
                                int *iterations, void *stream) {
  if (int rc = gom_pcg_check(na, nz, A, colidx, rowstr, b, x, max_iteration, iterations)) return rc;
  PcgWs w;
  const size_t bytes = gom_pcg_ws_carve(na, nullptr, &w);
  void *ws = nullptr;
  if (hipMalloc(&ws, bytes) != hipSuccess) {
    snprintf(g_pcg_msg, sizeof g_pcg_msg, "pcg_solve: hipMalloc of %zu bytes failed", bytes);
    return GOMGPU_E_NOMEM;
  }
  int rc = gom_pcg_solve_on(ws, na, nz, A, colidx, rowstr, b, x, max_iteration, error_tolerance, iterations,
                            (hipStream_t)stream, g_pcg_msg, sizeof g_pcg_msg);
  if (hipFree(ws) != hipSuccess && rc == GOMGPU_OK) {
    snprintf(g_pcg_msg, sizeof g_pcg_msg, "pcg_solve: hipFree failed");
    rc = GOMGPU_E_HIP;
  }
  if (rc == GOMGPU_OK) g_pcg_msg[0] = 0;
  return rc;
}

extern "C" const char *gomgpu_pcg_last_error(void) { return g_pcg_msg; }
