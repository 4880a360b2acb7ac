/*
 * fastexpm -- C ABI of the MI355X (gfx950) matrix-exponential engine.
 *
 * This is the drop-in boundary behind FastCTQW's `fastexpm` Python classes.
 * In the reference, the pybind11 module `_fastexpm_core` (src/main.cu:8-64)
 * binds `MatrixExpCalculator<T>` (include/MatrixExpCalculator.h:17-99),
 * whose bodies live in a prebuilt library that is absent from the reference
 * checkout (CMakeLists.txt:23-25).  Here the engine is a plain C ABI
 * (`extern "C"`, plain pointers and sizes, no torch or pybind types); the
 * module `_fastexpm_core` is a thin pybind11 layer over it.
 *
 * Status codes: every entry point returns FEXPM_OK (0) or a negative code;
 * fexpm_last_error(e) gives the message of the last failure on engine e,
 * fexpm_last_error(NULL) that of the calling thread's last failed call
 * without an engine (fexpm_create, testing entry points).
 * An engine serialises its own calls (one mutex per engine); separate
 * engines are independent.
 */
#ifndef FASTEXPM_H
#define FASTEXPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fexpm_engine fexpm_engine;

/* Element types, matching the reference's four bound classes
 * ExpMatFloat32 / ExpMatFloat64 / ExpMatComplex64 / ExpMatComplex128
 * (src/main.cu:12, 25, 38, 51). */
enum fexpm_dtype {
  FEXPM_FLOAT32 = 0,
  FEXPM_FLOAT64 = 1,
  FEXPM_COMPLEX64 = 2,
  FEXPM_COMPLEX128 = 3
};

enum fexpm_status {
  FEXPM_OK = 0,
  FEXPM_EINVAL = -1,     /* bad argument (size, dtype, null pointer, option) */
  FEXPM_ENOMEM = -2,     /* device allocation failed */
  FEXPM_EHIP = -3,       /* HIP runtime error */
  FEXPM_ESINGULAR = -4,  /* exactly singular Pade denominator */
  FEXPM_EFREED = -5,     /* engine memory was released by fexpm_free */
  FEXPM_ENODEVICE = -6   /* no usable gfx950 device */
};

/* Precision of the Pade products and squarings (option FEXPM_OPT_PRECISION). */
enum fexpm_precision {
  FEXPM_PREC_NATIVE = 0, /* FP32 MFMA for complex64/float32, FP64 for 128/64 */
  FEXPM_PREC_FP16X2 = 1, /* complex64 engines, structured squarings: fp16 hi/lo parts
                            (scale 2^14), 3 products on 16-bit MFMA, FP32 accumulation */
  FEXPM_PREC_BF16X3 = 2, /* complex64 engines, structured squarings: three bf16 parts,
                            6 products on 16-bit MFMA, FP32 accumulation */
  FEXPM_PREC_INT8_CRT = 3, /* complex64 engines, structured squarings of order n <= 16384
                              (larger orders use FP32 MFMA): each row rounded to beta bits of its
                              largest element, exact int8 products modulo K coprime moduli
                              (FEXPM_OPT_CRT_MODULI), Chinese-remainder reconstruction (no
                              accumulation rounding) */
  FEXPM_PREC_AUTO = 4      /* default of complex64 engines: FEXPM_PREC_INT8_CRT for orders
                              768 <= n <= 16384 (where it is the faster form, and more accurate
                              than FP32 MFMA), FEXPM_PREC_NATIVE otherwise; fexpm_stats.precision
                              reports the form a run used */
};

enum fexpm_option {
  FEXPM_OPT_PRECISION = 1,       /* enum fexpm_precision */
  FEXPM_OPT_STRUCTURE = 2,       /* 1 = detect purely-imaginary symmetric A (default), 0 = general path */
  FEXPM_OPT_KBLOCK = 3,          /* FP32 squarings: MFMA accumulation block in units of 16 k
                                    (default 8; 0 = one block over all k) */
  FEXPM_OPT_FP64_SQUARINGS = 4,  /* FP32 engines: number of leading squarings run in FP64 (default 0) */
  FEXPM_OPT_PADE_FP64 = 5,       /* FP32 engines: 1 = Pade products and solve in FP64 (default),
                                    0 = Pade products in FP32 */
  FEXPM_OPT_PANEL_COOP = 6,      /* LU panels: 1 = one cooperative launch per panel (default),
                                    0 = one launch per column (process-wide setting) */
  FEXPM_OPT_SPD_SOLVE = 7,       /* structured A = iG: 1 = real Cholesky solve of the Pade system
                                    (default), 0 = pivoted complex LU as for general A */
  FEXPM_OPT_BATCH_GROUPED = 8,   /* batches: 1 = grouped launches over all walks (default),
                                    0 = concurrent lanes */
  FEXPM_OPT_AUTOTUNE = 9,        /* grouped batches: 1 = time the tile-shape candidates on the first
                                    run of a new layout and keep the fastest (default), 0 = default tiles */
  FEXPM_OPT_GRAPH = 10,          /* grouped batches: 1 = capture the program in a hipGraph and replay it
                                    while the plans stay the same (default), 0 = eager launches */
  FEXPM_OPT_FP64_SLICES = 11,    /* complex64 / float32 engines, structured input: the symmetric FP64
                                    Pade products (and the solve's S and X = Z Z^T where the scope
                                    allows) on the int8 MFMA: 13 (default) = residue products modulo 13
                                    coprime moduli (each row rounded to 43 bits of its largest value,
                                    exact int32 products, Chinese-remainder reconstruction; products
                                    with k > 20480 take 6 slices); S = 4..6 = S exact int8 slices per
                                    value (about 7S - 1 bits) with int32 accumulation; 0 = FP64 MFMA.
                                    Batches (fexpm_batch_set_option): 5 (default; grouped products
                                    from 5 slices) or 0 */
  FEXPM_OPT_SPECTRAL = 12,       /* structured input (A = iG, G real symmetric, hence normal), degree 13:
                                    1 = s from the spectral-radius bound min_k ||B^k||_1^(1/k), k = 1, 2,
                                    4, 6, of the shifted B (default; single runs), 0 = s from ||B||_1
                                    (Higham 2005).  Batches always use the 1-norm rule. */
  FEXPM_OPT_TWO_PRODUCT = 14,    /* structured input: the first k squarings (at most s - 1) as two real
                                    products, Re = (Xr + Xi)(Xr - Xi)^T, Im = 2 Xr Xi^T, which needs
                                    Xr Xi = Xi Xr and issues 2/3 of the 3M flops; its non-commuting
                                    rounding errors grow in later squarings.  Default 0 (3M throughout):
                                    at N = 10^4 a two-product squaring takes 23.6 vs 25.4 ms (its
                                    two-MFMA loop is bound by the operand loads) */
  FEXPM_OPT_DEFLATE = 13,        /* with FEXPM_OPT_SPECTRAL: 1 = when every row sum of G is the same c
                                    (double-double sums), deflate the exact eigenvector 1 first:
                                    B = G - mu I + (mu - c)/n J and exp(A) = e^{i mu} (exp(iB) +
                                    (e^{i(c - mu)} - 1) J / n) (default); 0 = no deflation */
  FEXPM_OPT_CRT_MODULI = 16,     /* FEXPM_PREC_INT8_CRT squarings: number K of moduli, 7, 8 or 9
                                    (3 K int8 products per squaring); each row keeps beta bits of its
                                    largest element, the largest beta <= 27 with
                                    n 2^(2 beta + 1) < M / 2 (M the product of the moduli):
                                    K = 9: beta = 27 (n <= 16384); K = 8: beta = 24 for
                                    2729 < n <= 10915; K = 7: beta = 20 for 3051 < n <= 12203.
                                    Default 8: at N = 10^4 (ER(N, 0.5) Laplacians, t = 1 .. 20) 2-4x
                                    more accurate than FP32 MFMA squarings; 9 is 5-13x more accurate
                                    again at 6-8% more time per expm, 7 is 7-8% faster at 50-100x the
                                    infidelity of 8.  fexpm_stats reports K and beta */
  FEXPM_OPT_CRT64_MODULI = 17,   /* FP64-accurate residue products (FEXPM_OPT_FP64_SLICES = 13): number
                                    of moduli, 8 .. 13 (each row keeps beta = min(45,
                                    floor((log2 M - 1 - log2 k) / 2)) bits), or 0 = automatic
                                    (default): 11 for spectral plans (FEXPM_OPT_SPECTRAL; 36 bits at
                                    k = 10^4, no measurable change of the infidelity against 13 at
                                    N = 10^4, 5% less time per expm), 13 for 1-norm plans */
  FEXPM_OPT_CHEBYSHEV = 15       /* structured input with an FP64 stage (complex128, and complex64 with
                                    FEXPM_OPT_PADE_FP64): 1 = exp(iX), X = B / 2^s, from the Chebyshev
                                    (Jacobi-Anger) expansion of e^{ix} on [-theta, theta], theta =
                                    R / 2^s with R the spectral bound of B (spectral plans) or
                                    ||A - mu I||_1 (1-norm plans, batches), cos x = J0(theta) +
                                    2 sum (-1)^k J2k(theta) T2k(x/theta), sin x = 2 sum (-1)^k
                                    J2k+1(theta) T2k+1(x/theta), each truncated where the bound
                                    2 sum (theta/2)^m/m! of its tail reaches 2^-54, evaluated in X^2
                                    by symmetric products only (Paterson-Stockmeyer), in place of the
                                    degree-13 Pade approximant and its solve (default); 0 = Pade and
                                    solve.  fexpm_stats.chebyshev reports the degree. */
};

/* Per-run report (filled by fexpm_get_stats after a run). */
typedef struct fexpm_stats {
  int64_t n;
  int32_t dtype;
  int32_t m;            /* Pade degree (3, 5, 7, 9 or 13) */
  int32_t s;            /* number of squarings */
  int32_t structured;   /* 1: A = i*G with G real symmetric (CTQW form) */
  int32_t shifted;      /* 1: trace shift mu applied */
  int32_t nonfinite;    /* 1: input had NaN/Inf, output is NaN */
  double mu_re, mu_im;  /* trace shift tr(A)/n */
  double norm;          /* 1-norm used to choose m and s */
  double norm_plain;    /* ||A||_1 */
  double ms_total;      /* device time of the whole run (ms) */
  double ms_analyze;    /* norm / shift / structure analysis */
  double ms_pade;       /* scaling + Pade products + U/V assembly */
  double ms_solve;      /* FP64 LU + solve */
  double ms_square;     /* s squarings */
  double ms_output;     /* phase + write of U */
  double flops_issued;  /* real flops issued by the MFMA kernels this run */
  int32_t s_fp64;       /* squarings performed in FP64 by an FP32 engine */
  int32_t kblock;       /* accumulation block of the FP32 squarings (16-k tiles) */
  int32_t pade_fp64;    /* 1: Pade products and solve ran in FP64 */
  double ms_propagate;  /* device time of the last fexpm_propagate (ms) */
  int32_t precision;    /* enum fexpm_precision used by the squarings of this run */
  int32_t fp64_slices;  /* int8 slices per value of the FP64 products (0 = FP64 MFMA) */
  int32_t fp64_slices_scope;  /* stages using them: 1 Pade products, 2 solve, 3 both */
  int32_t spectral;     /* 1: s came from the spectral-radius bound (FEXPM_OPT_SPECTRAL) */
  int32_t deflated;     /* 1: the eigenvector 1 was deflated (FEXPM_OPT_DEFLATE) */
  int32_t s_onenorm;    /* s of the 1-norm rule (equal to s when spectral = 0) */
  int32_t chebyshev;    /* degree of the Chebyshev polynomial that replaced the Pade approximant
                           and its solve (FEXPM_OPT_CHEBYSHEV); 0 = Pade (m is then its degree) */
  double rho_bound;     /* spectral-radius bound of the shifted (deflated) B, when spectral = 1 */
  int32_t s_two_product;  /* leading squarings done in the two-product form (FEXPM_OPT_TWO_PRODUCT) */
  int32_t crt_moduli;   /* FEXPM_PREC_INT8_CRT squarings: number of moduli (FEXPM_OPT_CRT_MODULI), else 0 */
  int64_t device_bytes;  /* device memory the engine holds now (work planes, slices, I/O buffers, the
                            walk Hamiltonian); the FP64 planes of a complex64 engine are allocated
                            at its first run */
  double ms_square_kernel;  /* device time of the squarings' product launches alone (HIP events
                               around them on the squaring stream): the batched int8 residue
                               products of FEXPM_PREC_INT8_CRT squarings, else the squaring GEMMs */
  int32_t square_kernel_launches;  /* number of those timed launches (one per squaring) */
  int32_t crt_beta;     /* FEXPM_PREC_INT8_CRT squarings: bits kept of each row's largest element, else 0 */
  int32_t fp64_moduli;  /* FP64-accurate residue products (fp64_slices = 13): moduli used
                           (FEXPM_OPT_CRT64_MODULI), else 0 */
  int32_t reserved3;
} fexpm_stats;

/* Create an engine for n x n matrices of the given dtype and allocate its
 * device scratch (replaces MatrixExpCalculator<T>(size_t n),
 * include/MatrixExpCalculator.h:37,65,93; bound at src/main.cu:14). */
int fexpm_create(int64_t n, int dtype, fexpm_engine** out);

/* Release device memory; idempotent.  Later runs fail with FEXPM_EFREED
 * (replaces MatrixExpCalculator<T>::free, include/MatrixExpCalculator.h:42,70,98;
 * bound at src/main.cu:21). */
int fexpm_free(fexpm_engine* e);

/* Free everything and destroy the handle (replaces ~MatrixExpCalculator). */
int fexpm_destroy(fexpm_engine* e);

/* Matrix order (replaces MatrixExpCalculator<T>::getN,
 * include/MatrixExpCalculator.h:41,69,97). */
int64_t fexpm_get_n(const fexpm_engine* e);

/* U = expm(A) from host memory to host memory: `a` and `u` are row-major
 * n x n arrays of the engine's dtype (complex = interleaved re, im).  `a` is
 * not modified (replaces MatrixExpCalculator<T>::run,
 * include/MatrixExpCalculator.h:40,68,96; bound at src/main.cu:17). */
int fexpm_run(fexpm_engine* e, const void* a, void* u);

/* Same computation on device memory: `d_a`, `d_u` are device pointers to
 * row-major n x n arrays of the engine's dtype; `stream` is a hipStream_t
 * (NULL = the engine's own stream).  No host transfer of matrix data. */
int fexpm_run_device(fexpm_engine* e, const void* d_a, void* d_u, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The fexpm_run_device computation without an engine:
 * U = expm(A) for device arrays d_a, d_u of order n <= FEXPM_NOALLOC_NMAX
 * (row-major, dtype as in fexpm_create), on `stream` (NULL = the default
 * stream).  Every work buffer lives in a static device array of
 * FEXPM_NOALLOC_BYTES that each GPU holds until the process ends; the call
 * allocates no device memory (it creates no engine, stream or event).  A
 * larger n fails with FEXPM_EINVAL.  Calls in one process are serialised
 * and each returns after its work is complete.  Errors: fexpm_last_error(NULL). */
#define FEXPM_NOALLOC_NMAX 256
#define FEXPM_NOALLOC_BYTES ((size_t)48 << 20)
int fexpm_run_device_noalloc(int64_t n, int dtype, const void* d_a, void* d_u, void* stream);

/* Engine-resident input/output for callers without device memory of their
 * own: upload A once, run any number of times, download U. */
int fexpm_upload(fexpm_engine* e, const void* a);
int fexpm_run_resident(fexpm_engine* e);
int fexpm_download(fexpm_engine* e, void* u);

/* Keep a walk Hamiltonian on the device (complex engines): H = -gamma * A,
 * or -gamma * (D - A) with D = diag(row sums of A) when laplacian != 0.
 * `adj` is a row-major n x n host array of dtype FEXPM_FLOAT32 or
 * FEXPM_FLOAT64; H is formed in FP64 (row sums included).  Replaces the host
 * construction in Operator._buildHamiltonian (FastCTQW/Operator.py:49-53). */
int fexpm_set_graph(fexpm_engine* e, const void* adj, int adj_dtype, double gamma, int laplacian);

/* The walk Hamiltonian from an undirected edge list: vertices src[k], dst[k]
 * in [0, n), weight w[k] (w NULL: unit weights); A_ij = A_ji = w, a self-loop
 * on the diagonal once, repeated pairs add up (nx.to_numpy_array of the
 * graph); then H as in fexpm_set_graph, all in FP64 on the device.  Only
 * the 24 bytes per edge cross PCIe.  Replaces the stubs
 * Operator.from_adjacency_list / reset_adjList (FastCTQW/Operator.py:114-116,
 * 147-148) and Ctqw.from_adjacency_list (FastCTQW/Ctqw.py:60-71), whose
 * bodies are empty in the reference. */
int fexpm_set_graph_edges(fexpm_engine* e, int64_t m, const int64_t* src, const int64_t* dst, const double* w,
                          double gamma, int laplacian);

/* The walk Hamiltonian of the seeded Erdos-Renyi graph G(n, p) generated on
 * the device: edge {i, j} (i < j) iff u_1(i n + j) < p, weight 1 or (with
 * weighted) 1 - u_2(i n + j), where u_s(c) = 53 bits of splitmix64(c ^ key_s),
 * key_s = splitmix64(splitmix64(seed) ^ s), scaled to [0, 1): the same graph
 * as fastctqw_amd.graphs.erdos_renyi(n, p, seed, weighted) on the host, bit
 * for bit; no matrix crosses PCIe.  No counterpart in the reference (whose
 * graphs come from networkx on the host). */
int fexpm_set_graph_er(fexpm_engine* e, double p, uint64_t seed, int weighted, double gamma, int laplacian);

/* U = expm(-i t H) for the Hamiltonian kept by fexpm_set_graph, resident
 * (download / apply / propagate as after fexpm_run_resident).  Replaces
 * Operator._buildUnitory + buildExpmOperator (FastCTQW/Operator.py:55-56,
 * 70-78): A_exp is formed on the device, no matrix crosses PCIe. */
int fexpm_run_graph(fexpm_engine* e, double t);

/* psi[k] = U psi0[k] for k < nstates with the U of the last run, FP64
 * accumulation on the device; psi0 and psi are host complex128 arrays of
 * nstates x n (replaces the host product at FastCTQW/Ctqw.py:94). */
int fexpm_apply(fexpm_engine* e, const void* psi0, int64_t nstates, void* psi);

/* Propagation of many states by repeated application of the U of the last
 * run: psi[k][s] = U^k psi0[s], k = 1..nsteps, s < nstates (host complex128,
 * psi0: nstates x n, psi: nsteps x nstates x n).  Each step is one MFMA
 * GEMM over all states followed by an FP64 refinement (norm restored for
 * CTQW propagators, phase e^{k mu}).  With U = expm(-i dt H) this gives the
 * walk at t_k = k dt.  No counterpart in the reference (which re-runs the
 * whole expm per time, FastCTQW/Ctqw.py:90-94). */
int fexpm_propagate(fexpm_engine* e, const void* psi0, int64_t nstates, int64_t nsteps, void* psi);

int fexpm_set_option(fexpm_engine* e, int key, int64_t value);
int fexpm_get_stats(const fexpm_engine* e, fexpm_stats* out);
const char* fexpm_last_error(const fexpm_engine* e);
const char* fexpm_status_string(int status);

/* ---- Batches of independent expms of uneven order (many walks) ----
 * No counterpart in the reference (which runs one expm per Operator).  A
 * batch owns `lanes` pipelines of capacity nmax, each with its own stream;
 * walks are assigned to lanes by decreasing n^3 and each lane runs on its
 * own host thread, so small walks overlap on the device. */
typedef struct fexpm_batch fexpm_batch;
int fexpm_batch_create(int dtype, int64_t nmax, int lanes, fexpm_batch** out);
int fexpm_batch_destroy(fexpm_batch* b);
const char* fexpm_batch_last_error(const fexpm_batch* b);
/* FEXPM_OPT_PRECISION, FEXPM_OPT_KBLOCK, FEXPM_OPT_SPD_SOLVE (every lane) and
 * FEXPM_OPT_BATCH_GROUPED, FEXPM_OPT_AUTOTUNE, FEXPM_OPT_GRAPH (grouped runs). */
int fexpm_batch_set_option(fexpm_batch* b, int key, int64_t value);
/* Grouped batches (complex dtypes, option FEXPM_OPT_BATCH_GROUPED): every stage
 * of the structured pipeline is one launch over all walks that need it (the
 * analysis, each Pade product, each 64-wide Cholesky panel, each squaring
 * round r over the walks with s_i > r); the stages after the analysis are
 * captured once in a hipGraph and replayed while the walks' plans (m_i, s_i,
 * trace shift) stay the same; tile shapes come from an autotuner that times
 * each candidate on the batch itself.  Walks that are not of the CTQW form
 * (A = iG, G real symmetric) or have non-finite entries run on a lane
 * pipeline instead.  Split precision (complex64) runs as grouped split squaring
 * rounds; the LU solve option (FEXPM_OPT_SPD_SOLVE = 0) uses lanes. */
typedef struct fexpm_batch_report {
  int32_t grouped;         /* 1: the last run used grouped launches */
  int32_t grouped_walks;   /* walks run by the grouped program */
  int32_t fallback_walks;  /* walks run on a lane pipeline */
  int32_t launches;        /* kernel launches of the grouped program (after the analysis) */
  int32_t graph;           /* 1: the program ran as a hipGraph */
  int32_t graph_built;     /* 1: this run captured a new graph */
  int32_t tuned;           /* 1: this run ran the autotuner */
  int32_t candidate[3];    /* autotuner choice per class (squarings, symmetric FP64 products, solve
                              updates): tile configuration * 3 + dispatch */
  double class_ms[3];      /* per class: GEMM device time of the chosen candidate in the tuning runs */
  double tune_ms;          /* host time of the tuning (0 if it did not run) */
  double device_ms;        /* device time of the grouped program in the last run */
  double plan_ms;          /* host time from the start of the run to the plans (analysis + read-back) */
} fexpm_batch_report;
int fexpm_batch_get_report(const fexpm_batch* b, fexpm_batch_report* out);

/* `count` row-major matrices of orders n[i] (engine dtype), packed back to
 * back in host memory, copied to the device.  An order outside [1, nmax]
 * is rejected (FEXPM_EINVAL) before anything changes; a failure after that
 * (device allocation, copy) leaves the batch empty. */
int fexpm_batch_upload(fexpm_batch* b, int64_t count, const int64_t* n, const void* a);
/* Number of matrices of the current batch (-1 for a null handle). */
int64_t fexpm_batch_count(const fexpm_batch* b);
/* U_i = expm(A_i) for every uploaded matrix (U stays on the device); with
 * psi0 != NULL also psi_i = U_i psi0_i (packed complex128, length n_i each). */
int fexpm_batch_run_resident(fexpm_batch* b, const void* psi0, void* psi);
int fexpm_batch_download(fexpm_batch* b, int64_t i, void* u);
int fexpm_batch_get_stats(const fexpm_batch* b, int64_t i, fexpm_stats* out);
double fexpm_batch_last_ms(const fexpm_batch* b);  /* wall time of the last batch run */

/* Free device memory in GiB (replaces get_current_gpu_memory_gb,
 * include/MatrixExpCalculator.h:101). Returns a negative value on error. */
double fexpm_gpu_memory_gb(void);

#ifdef __cplusplus
}
#endif

#endif /* FASTEXPM_H */
