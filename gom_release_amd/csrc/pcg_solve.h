// pcg_solve.h - the workspace of the standalone PCG solve (pcg_solve.hip), shared by its
// entry points.
#pragma once

#include <stddef.h>

#include "gomgpu_dev.h"

// Most rows of one standalone solve (the row pointers and the 6n published values stay
// well inside 32-bit and size_t ranges).
#define PCG_SOLVE_MAX_ROWS (1L << 27)
#define PCG_PART_STRIDE 4096

// Workspace of a solve of n rows, carved in 256-byte aligned segments.
struct PcgWs {
  double *diag, *off, *d0, *pub, *part;
  unsigned long long *slots;
  int32_t *col, *nf;
  int *info;  // [0] iterations, [1] converged, [2] abort word
  int *bad;   // [0] a malformed row was seen, [1] the last such row + 1
  unsigned *bar;
};

// Carves the workspace of an n-row solve from base (null: only the size) and returns its
// size in bytes.
size_t gom_pcg_ws_carve(long n, char *base, PcgWs *w);
// Argument checks shared by the entry points (the message goes to g_pcg_msg).
int gom_pcg_check(int na, int nz, const double *A, const int32_t *colidx, const int32_t *rowstr,
                  const double *b, double *x, int max_iteration, int *iterations);
// The solve on a carved workspace; synchronous on stream st.
int gom_pcg_solve_on(void *ws, int na, int nz, const double *A, const int32_t *colidx, const int32_t *rowstr,
                     const double *b, double *x, int max_iteration, double error_tolerance, int *iterations,
                     hipStream_t st, char *msg, size_t msgcap);
// Text of the calling thread's last failure of a standalone solve.
extern thread_local char g_pcg_msg[256];
