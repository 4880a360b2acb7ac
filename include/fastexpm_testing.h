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

/* Engine option: stages whose symmetric FP64 products use the int8 slices
 * of FEXPM_OPT_FP64_SLICES: 1 the Pade products, 2 the Cholesky solve
 * (S, X = Z Z^T), 3 both, 0 automatic (default for complex64 engines: the
 * Pade products, and the solve's too for plans that deflate the eigenvector
 * 1 at n >= 4096).  For precision studies. */
enum { FEXPM_OPT_FP64_SLICES_SCOPE = 102 };

/* Engine option: 1 (default) runs the last, partly filled wave of 128 x 128
 * tiles of an FP32 symmetric squaring as 64 x 64 sub-tiles; 0 keeps one
 * launch of 128 x 128 tiles.  Both give the same values bit for bit. */
enum { FEXPM_OPT_TAIL_SPLIT = 103 };

/* Engine option: 1 (default) runs the Z updates of the structured Cholesky
 * solve on a second stream, overlapped with the factorisation of S; 0 keeps
 * one stream.  Both give the same values bit for bit. */
enum { FEXPM_OPT_Z_OVERLAP = 104 };

/* Engine option: width of the outer blocks of the structured Cholesky solve
 * (a multiple of 64 in [64, 4096], default 512): 64-wide panels update only
 * their outer block, which then updates the trailing matrix and the
 * remaining columns of Z with K equal to this width. */
enum { FEXPM_OPT_CHOL_OUTER = 105 };

/* Engine option: tile edge of the FP32 structured squarings: 0 (default,
 * automatic: 64 up to n = 6144, else 128 with the tail split), 64 or 128.
 * Every choice gives the same values bit for bit. */
enum { FEXPM_OPT_SQUARE_TILE = 106 };

/* Engine option: 1 applies the Z updates of the structured Cholesky solve
 * left-looking (one product with K = all finished columns at the start of
 * each outer block), 0 (default) right-looking (one product with K = the
 * outer width after each block).  Same flops, different summation order. */
enum { FEXPM_OPT_Z_LEFT = 107 };

/* Engine option (fexpm_set_option): host copies of 64 MiB and more through
 * the pinned staging buffers (1, default) or the runtime's pageable path (0,
 * also taken when pinned host memory cannot be allocated). */
enum { FEXPM_OPT_HOST_STAGING = 108 };

/* Batch option: 1 makes the next fexpm_batch_upload that must grow the
 * matrix buffers fail with FEXPM_ENOMEM after the old buffers are released
 * (host-side injection; no device call fails), to test the recovery path. */
enum { FEXPM_OPT_FAIL_ALLOC = 109 };

/* Engine option: block map of the batched int8 residue launches.  1
 * (default): the 8 XCDs work on one problem at a time, each on an eighth of
 * its tiles (at N = 10^4 a squaring launch takes 12.21 ms against 12.61);
 * 0: the launch's blocks spread over the XCDs in eighths of all tiles of all
 * problems.  The same values either way. */
enum { FEXPM_OPT_CRT_MAP = 110 };

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

/* C = A B^T (mode 0 real, 1 complex 4M, 2 complex 3M; mode 3 with sym: the
 * two-product squaring form, Re = lower((Ar + Ai)(Br - Bi)^T), Im =
 * lower(2 Ar Bi^T), mirrored), sym != 0: symmetric half-tile output
 * (requires m == n).  wtype 0 = FP32 MFMA, 1 = FP64 MFMA,
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

/* Copy of the walk Hamiltonian H (n x n, FP64) kept by fexpm_set_graph,
 * fexpm_set_graph_edges or fexpm_set_graph_er. */
int fexpm_debug_hamiltonian(fexpm_engine* e, double* h);

/* C = A B^T (FP64, host row-major A m x k, B n x k, C m x n) computed from
 * exact int8 slices of the rows of A and B (`slices` in 3..6) with int32
 * accumulation on the int8 MFMA and FP64 combination; with sym, B must be A
 * and only lower tiles are computed (mirrored).  slices == 13: the residue
 * form of the engine (rows rounded to beta bits, exact products modulo 13
 * moduli, Chinese-remainder reconstruction), sym and m == n only; the lower
 * triangle of A B^T is computed and mirrored (B may differ from A: the
 * engine uses it for products that are symmetric in exact arithmetic). */
int fexpm_test_ozaki_gemm(int slices, int sym, int m, int n, int k, const double* a, const double* b, double* c);

/* MFMA flops one FP32 structured (complex symmetric, 3M) squaring of order n
 * issues under the automatic tile choice, including the tail sub-tiles: the
 * figure an issued-rate report divides by the launch time. */
int fexpm_test_square_issued(int64_t n, double* flops);

/* Polynomial of exp(ix) on [-theta, theta] that spectral plans use in place
 * of the Pade approximant (FEXPM_OPT_CHEBYSHEV), evaluated on the host,
 * 0 <= theta <= 16: the degrees dc (even, cos) and ds (odd, sin) and the
 * coefficients in w = x^2 / theta^2, cos x = sum_j cw[j] w^j (j <= dc/2),
 * sin x = (x / theta) sum_j sw[j] w^j (j <= (ds-1)/2); cw and sw hold 40
 * values each (zeros beyond the degree). */
int fexpm_test_cheb_series(double theta, int* dc, int* ds, double* cw, double* sw);

/* Tile order of the symmetric (lower-triangular) GEMM launches, evaluated on
 * the host: out[b] = (ti << 16) | tj for every b < T(T+1)/2, 1 <= T <= 4096.
 * Groups of 8 x 8 tiles, group rows in order, within one the groups left to
 * right, within a group row-major with tj <= ti. */
int fexpm_test_grouped_tile_order(int T, int* out);

/* Tile order of the rectangular GEMM launches (nbm x nbn tiles, each at
 * most 4096), evaluated on the host: out[lin] = (ti << 16) | tj.  Groups of
 * 8 tile rows; a group sweeps its columns, visiting its rows per column. */
int fexpm_test_grouped_rect_order(int nbm, int nbn, int* out);

/* Grouped batch program for `count` CTQW walks with the given orders and
 * plans (Pade degree m, squarings s, trace shift, and the 1-norm `norm` of
 * the shifted matrix: with norm non-null the Chebyshev polynomial of each
 * walk's X = (A - mu I) / 2^s at theta = norm / 2^s replaces the Pade
 * products and the solve, as in batch runs by default), built by the same code as
 * a grouped batch run but on a virtual address layout (no device memory is
 * touched): address = region << 40 | byte offset, regions 1 FP64 planes,
 * 2 squaring planes, 3 per-walk scalars and diagonals, 4 split parts,
 * 5 input, 6 output, 7 psi0, 8 psi, 9 walk scalars, 10 int8 slices.
 * fp64_slices: -1 the batch default (complex64: Pade products from 6 int8
 * slices), 0 FP64 MFMA, 6 int8 slices (complex64).  Writes a text listing of
 * every stage and problem (NUL-terminated) into buf and its length into
 * *len; with cap too small returns FEXPM_EINVAL and only sets *len.  Lets a
 * CPU interpreter check the program against the oracle. */
int fexpm_test_group_program(int dtype, int precision, int fp64_slices, int64_t count, const int64_t* n,
                             const int32_t* m, const int32_t* s, const int32_t* shifted, const double* mu_re,
                             const double* mu_im, const double* norm, int with_psi, char* buf, int64_t cap,
                             int64_t* len);

#ifdef __cplusplus
}
#endif

#endif
