/*
 * fastexpm testing ABI -- component entry points for parity tests.
 *
 * These have no counterpart in the reference (whose engine body is absent,
 * CMakeLists.txt:23-25); they expose single kernels and internal stages of
 * the engine so that tests can pin each stage against the CPU oracle:
 *   - fexpm_test_gemm: one launch of the MFMA GEMM family on host data;
 *   - fexpm_test_lu_solve: the FP64 blocked LU + solve on host data;
 *   - fexpm_debug_plane: copy an internal working plane after a run;
 *   - option FEXPM_OPT_MAX_SQUARINGS: perform at most that many squarings
 *     (-1 = all), so that the Pade value X can be inspected.
 * All arrays are row-major, FP64 on the host side.
 */
#ifndef FASTEXPM_TESTING_H
#define FASTEXPM_TESTING_H

#include <stdint.h>

#include "fastexpm.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { FEXPM_OPT_MAX_SQUARINGS = 100 };

/* Batch option: run grouped programs with autotuner candidate `value` for
 * every class of GEMM launches (tile configuration * 3 + dispatch, 0..5) and
 * no tuning; -1 restores the autotuner.  Lets tests check every candidate. */
enum { FEXPM_OPT_GROUP_CANDIDATE = 101 };

/* Internal planes readable by fexpm_debug_plane (working precision, n x n). */
enum fexpm_plane {
  FEXPM_PLANE_A = 0,   /* (A - mu I) / 2^s; structured: its imaginary part G */
  FEXPM_PLANE_A2 = 1,  /* A^2 (structured: G^2) */
  FEXPM_PLANE_A4 = 2,
  FEXPM_PLANE_A6 = 3,
  FEXPM_PLANE_T1 = 4,
  FEXPM_PLANE_T2 = 5,
  FEXPM_PLANE_U = 6,   /* Pade odd part U (structured: U = i * plane) */
  FEXPM_PLANE_V = 7,   /* Pade even part V */
  FEXPM_PLANE_X0 = 8,
  FEXPM_PLANE_X1 = 9,
  FEXPM_PLANE_RESULT = 10 /* X after the squarings, before the phase e^mu */
};

/* C = A B^T (mode 0 real, 1 complex 4M, 2 complex 3M), sym != 0: symmetric
 * half-tile output (requires m == n).  wtype 0 = FP32 MFMA, 1 = FP64 MFMA,
 * 2 = fp16x2 and 3 = bf16x3 split-precision squaring kernel (mode 2, sym,
 * B = A, complex 4M on 16-bit MFMA; |a| must stay below 65504 / 2^14 for
 * fp16x2).  A is m x k, B is n x k, C is m x n; imaginary pointers unused
 * for mode 0. */
int fexpm_test_gemm(int wtype, int mode, int sym, int m, int n, int k, const double* ar, const double* ai,
                    const double* br, const double* bi, double* cr, double* ci);

/* X = M^{-1} B for complex n x n M and B (planar re/im); *info != 0 if an
 * exactly zero pivot was met. */
int fexpm_test_lu_solve(int n, const double* mr, const double* mi, const double* br, const double* bi,
                        double* xr, double* xi, int* info);

int fexpm_debug_plane(fexpm_engine* e, int which, double* re, double* im);

/* Average device time (ms) of one GEMM launch over reps launches on
 * device-resident operands (m x k, n x k -> m x n); cfg selects an FP64
 * tiling (0 4-wave 32x32, 1 8-wave 32x16 = default, 2 128x64, 3 64x128;
 * 2 and 3 are not square and are invalid with sym).  wtype 2 / 3 time the
 * split-precision squaring kernel alone (mode 2, sym). */
int fexpm_test_gemm_time(int wtype, int mode, int sym, int cfg, int m, int n, int k, int reps, double* ms);

/* reps + 1 launches of a plain device copy of `bytes` (multiple of 16);
 * average ms of the last reps.  Reference traffic for the profiler's
 * FETCH_SIZE / WRITE_SIZE counters (each launch reads and writes bytes). */
int fexpm_test_copy(int64_t bytes, int reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif
