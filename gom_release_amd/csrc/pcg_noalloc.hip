// pcg_noalloc.hip - gomgpu_pcg_solve_noalloc: the standalone free-surface solve of
// pcg_solve.hip without any device allocation.
#include <stdio.h>

#include <mutex>

#include "pcg_solve.h"

// The same solve without any allocation, for callers that cannot allocate (code under
// graph capture, tight serving loops). The workspace is the file-scope device array
// `scratch`, sized for GOMGPU_PCG_NOALLOC_MAX_ROWS rows; each device holds its own copy
// from the library's load to the end of the process. Host code never passes `scratch`
// itself (its host-side name is not a device address): the device address comes from
// hipGetSymbolAddress, once per device, with that device current. Every call on one device
// uses the one array, so calls that overlap on a device race; the words that must start at
// zero are cleared by a kernel at the start of every call.
static constexpr size_t ws_round(size_t b) { return (b + 255) & ~(size_t)255; }
static constexpr size_t ws_bytes_for(size_t n) {
  return ws_round(n * 8) + ws_round(4 * n * 8) + ws_round(n * 8) + ws_round(6 * n * 8) +
         ws_round(4 * PCG_PART_STRIDE * 8) + ws_round(6 * PCG_MAX_SLOT_BLOCKS * 8) + ws_round(4 * n * 4) +
         ws_round(n * 4) + ws_round(4 * 4) + ws_round(2 * 4) + ws_round(512 * 4);
}
#define PCG_NOALLOC_BYTES ws_bytes_for(GOMGPU_PCG_NOALLOC_MAX_ROWS)

static __device__ __align__(256) unsigned char scratch[PCG_NOALLOC_BYTES];

#define PCG_MAX_DEVICES 64
static void *g_scratch_ptr[PCG_MAX_DEVICES];
static std::mutex g_scratch_mu;

// The device address of this device's `scratch` (resolved on the first call on a device).
static int scratch_for(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= PCG_MAX_DEVICES) return GOMGPU_E_ARG;
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return GOMGPU_E_HIP;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return GOMGPU_OK;
}

extern "C" int gomgpu_pcg_solve_noalloc(int na, int nz, const double *A, const int32_t *colidx,
                                        const int32_t *rowstr, const double *b, double *x, int max_iteration,
                                        double error_tolerance, int *iterations, void *stream) {
  if (int rc = gom_pcg_check(na, nz, A, colidx, rowstr, b, x, max_iteration, iterations)) return rc;
  if (na > GOMGPU_PCG_NOALLOC_MAX_ROWS) {
    snprintf(g_pcg_msg, sizeof g_pcg_msg,
             "pcg_solve_noalloc: %d rows exceed the fixed workspace of %d rows (use gomgpu_pcg_solve)", na,
             GOMGPU_PCG_NOALLOC_MAX_ROWS);
    return GOMGPU_E_ARG;
  }
  PcgWs w;
  if (gom_pcg_ws_carve(na, nullptr, &w) > PCG_NOALLOC_BYTES) {
    snprintf(g_pcg_msg, sizeof g_pcg_msg, "pcg_solve_noalloc: workspace layout exceeds the fixed array");
    return GOMGPU_E_ARG;
  }
  int device = 0;
  void *ws = nullptr;
  if (hipGetDevice(&device) != hipSuccess || scratch_for(device, &ws) != GOMGPU_OK) {
    snprintf(g_pcg_msg, sizeof g_pcg_msg, "pcg_solve_noalloc: cannot resolve the device workspace");
    return GOMGPU_E_HIP;
  }
  int rc = gom_pcg_solve_on(ws, na, nz, A, colidx, rowstr, b, x, max_iteration, error_tolerance, iterations,
                            (hipStream_t)stream, g_pcg_msg, sizeof g_pcg_msg);
  if (rc == GOMGPU_OK) g_pcg_msg[0] = 0;
  return rc;
}
