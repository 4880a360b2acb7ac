// Scaling-and-squaring Pade matrix exponential on one MI355X (gfx950).
//
// Pipeline of one run (stage names follow the reference's declared buffers
// d_A, d_A2, d_A4, d_A6, d_u*, d_v*, d_mu, d_nrmA,
// include/MatrixExpCalculator.h:52-61):
//   1. analysis (FP64): trace -> mu, exact 1-norms of A and A - mu I,
//      structure flags (purely imaginary + exactly symmetric), non-finite;
//   2. degree m and scaling s from the 1-norm (Higham 2005, Alg. 2.3);
//   3. A <- (A - mu I) / 2^s in planar working precision W;
//   4. Pade products A2, A4, A6 and U, V on MFMA;
//   5. FP64 LU with partial pivoting of [V - U | V + U], back substitution;
//   6. s squarings on MFMA;
//   7. U = e^mu X written in the caller's layout.
//
// CTQW inputs A = -i t H (H real symmetric) are detected in step 1 and take
// a structured path: A = i G with G real symmetric, every Pade product is a
// real symmetric product (half the tiles), and every squaring of the
// complex-symmetric X = Xr + i Xi is X X^T computed by the 3M rule with
// three half products.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/fastexpm.h"
#include "../../include/fastexpm_testing.h"
#include "fexpm_gemm.hpp"
#include "fexpm_kernels.hpp"
#include "fexpm_split.hpp"
#include "fexpm_ozaki.hpp"
#include "fexpm_crt.hpp"

namespace fexpm {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FX_HIP(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess)                                                                          \
      throw ::fexpm::Error(e_ == hipErrorOutOfMemory ? FEXPM_ENOMEM : FEXPM_EHIP,                  \
                           std::string(#x) + ": " + hipGetErrorString(e_));                        \
  } while (0)

#define FX_LAUNCH() FX_HIP(hipGetLastError())

// theta_m (double precision) and Pade coefficients b_j, Higham 2005.
static const int kDegrees[5] = {3, 5, 7, 9, 13};
static const double kTheta[5] = {1.495585217958292e-2, 2.539398330063230e-1, 9.504178996162932e-1,
                                 2.097847961257068e0, 5.371920351148152e0};

static double pade_b(int m, int j) {
  // b_j = (2m-j)! / (j! (m-j)!), exact in double for m <= 13
  long double num = 1.0L;
  for (int k = m - j + 1; k <= 2 * m - j; ++k) num *= (long double)k;  // (2m-j)!/(m-j)!
  long double den = 1.0L;
  for (int k = 2; k <= j; ++k) den *= (long double)k;
  return (double)(num / den);
}

static void choose_degree(double norm, int& m, int& s) {
  for (int q = 0; q < 4; ++q) {
    if (norm <= kTheta[q]) {
      m = kDegrees[q];
      s = 0;
      return;
    }
  }
  m = 13;
  s = std::max(0, (int)std::ceil(std::log2(norm / kTheta[4])));
  while (norm / std::ldexp(1.0, s) > kTheta[4]) ++s;
}

// Smallest of the spectral-radius bounds of a real symmetric B:
//   ||B^k||_1^(1/k), k = 1, 2, 4, 6 (h[0..3], the 1-norms), and
//   trace(B^2k)^(1/2k), 2k = 4, 6, ..., 12, and 14, 16 when B^8 was formed
//   (h[4..10], negative when absent; the sum of the 2k-th powers of the
//   eigenvalues is at least rho^2k),
// widened by 1e-9 to cover the rounding of the norms, sums and roots.  The
// trace bounds are the tight ones when the spectrum has no isolated
// extreme eigenvalue (ER(5000, 0.5) Laplacian, deflated: 1.12 rho at 2k = 12,
// 1.05 rho at 2k = 16).
constexpr int kSpecNorms = 11;
static double spectral_bound(const double h[kSpecNorms]) {
  static const double e[kSpecNorms] = {1.0, 2.0, 4.0, 6.0, 4.0, 6.0, 8.0, 10.0, 12.0, 14.0, 16.0};
  double rho = INFINITY;
  for (int k = 0; k < kSpecNorms; ++k)
    if (std::isfinite(h[k]) && h[k] >= 0.0) rho = std::min(rho, std::pow(h[k], 1.0 / e[k]));
  return rho * (1.0 + 1e-9);
}

// Squarings of degree 13 for a spectral-radius bound rho (the rule of
// choose_degree with rho in place of the 1-norm).
static int scaling_for(double rho) {
  if (!(rho > kTheta[4])) return 0;
  int s = std::max(0, (int)std::ceil(std::log2(rho / kTheta[4])));
  while (rho / std::ldexp(1.0, s) > kTheta[4]) ++s;
  return s;
}

// Polynomial of exp(i x) on [-theta, theta] from the Jacobi-Anger expansion
//     cos x = J0(theta) + 2 sum_{k>=1} (-1)^k J2k(theta) T2k(x / theta),
//     sin x = 2 sum_{k>=0} (-1)^k J2k+1(theta) T2k+1(x / theta),
// truncated at the least even degree dc and odd degree ds whose tails,
// bounded by |Jm(theta)| <= (theta/2)^m / m!, are each at most 2^-54: for a
// real symmetric X with spectrum in [-theta, theta] the 2-norm error of the
// polynomial at X is then at most 2^-53.  Returned in w = x^2 / theta^2 in
// [0, 1]: cos x = sum_j cw[j] w^j (j <= dc/2) and
// sin x = (x / theta) sum_j sw[j] w^j (j <= (ds-1)/2), from
// T2k(x/theta) = Tk(2w - 1) and T2k+1(x/theta) = (x/theta) Wk(w) with W0 = 1,
// W1 = 4w - 3, Wk+1 = 2 (2w - 1) Wk - Wk-1 (T(n+2) = 2 T2 Tn - T(n-2)).  The
// Bessel values come from their power series and the basis change runs in
// long double (80-bit on the host).
constexpr int kChebMax = 40;  // coefficients per part (dc/2 + 1, (ds-1)/2 + 1)
struct ChebSeries {
  int dc = 0, ds = 1;
  double cw[kChebMax], sw[kChebMax];
};

static long double bessel_j(int m, long double x) {
  // sum_j (-1)^j (x/2)^(2j+m) / (j! (j+m)!); 60 terms reach far below the
  // last bit for x <= 16
  const long double h = x / 2;
  long double term = 1.0L;
  for (int i = 1; i <= m; ++i) term *= h / (long double)i;
  long double sum = term;
  for (int j = 1; j < 60; ++j) {
    term *= -h * h / ((long double)j * (long double)(j + m));
    sum += term;
  }
  return sum;
}

// 2 sum_{m >= m0, m - m0 even} (theta/2)^m / m!
static double bessel_tail_bound(double theta, int m0) {
  const long double h = theta / 2;
  long double t = 1.0L;
  for (int i = 1; i <= m0; ++i) t *= h / (long double)i;
  long double s = 0.0L;
  for (int m = m0; m < m0 + 80; m += 2) {
    s += t;
    t *= h * h / ((long double)(m + 1) * (long double)(m + 2));
  }
  return (double)(2.0L * s);
}

// Tail bound of the truncated series for a result squared s times: 2^-54
// for FP64 results; for FP32 results (complex64 / float32 engines and
// batches) 2^-(24 + s), so that the truncation error of each eigenvalue's
// factor, multiplied by at most 2^s in the squarings, stays at the FP32
// unit roundoff of the output (as Higham's single-precision theta_m do for
// the Pade approximant).  At N = 10^4 (theta = 4.08, s = 8): degree 24 -> 16,
// two fewer products.
static double cheb_tail(int s, bool fp32_result) {
  return fp32_result ? std::max(0x1p-54, std::ldexp(1.0, -(24 + std::max(s, 0)))) : 0x1p-54;
}

static ChebSeries cheb_series(double theta, double tail = 0x1p-54) {
  ChebSeries c;
  const int dmax = 2 * (kChebMax - 1);
  while (c.dc < dmax && bessel_tail_bound(theta, c.dc + 2) > tail) c.dc += 2;
  while (c.ds < dmax - 1 && bessel_tail_bound(theta, c.ds + 2) > tail) c.ds += 2;
  const long double th = theta;
  const int nc = c.dc / 2 + 1, ns = (c.ds - 1) / 2 + 1;
  long double acc[kChebMax] = {}, prev[kChebMax + 1] = {}, cur[kChebMax + 1] = {}, next[kChebMax + 1] = {};
  // cos: sum_k b_k Tk(2w - 1), b_0 = J0, b_k = 2 (-1)^k J2k
  prev[0] = 1.0L;                 // T0
  cur[0] = -1.0L, cur[1] = 2.0L;  // T1(2w - 1)
  for (int k = 0; k < nc; ++k) {
    const long double b = k == 0 ? bessel_j(0, th) : 2.0L * (k & 1 ? -1.0L : 1.0L) * bessel_j(2 * k, th);
    const long double* t = k == 0 ? prev : cur;
    for (int j = 0; j <= k; ++j) acc[j] += b * t[j];
    if (k >= 1) {  // T(k+1) = 2 (2w - 1) Tk - T(k-1)
      for (int j = 0; j <= k + 1; ++j)
        next[j] = 2.0L * ((j > 0 ? 2.0L * cur[j - 1] : 0.0L) - cur[j]) - prev[j];
      for (int j = 0; j <= k + 1; ++j) {
        prev[j] = cur[j];
        cur[j] = next[j];
      }
    }
  }
  for (int j = 0; j < nc; ++j) c.cw[j] = (double)acc[j];
  // sin / (x / theta): sum_k b_k Wk(w), b_k = 2 (-1)^k J2k+1
  for (int j = 0; j <= kChebMax; ++j) prev[j] = cur[j] = 0.0L;
  for (int j = 0; j < kChebMax; ++j) acc[j] = 0.0L;
  prev[0] = 1.0L;                 // W0
  cur[0] = -3.0L, cur[1] = 4.0L;  // W1
  for (int k = 0; k < ns; ++k) {
    const long double b = 2.0L * (k & 1 ? -1.0L : 1.0L) * bessel_j(2 * k + 1, th);
    const long double* t = k == 0 ? prev : cur;
    for (int j = 0; j <= k; ++j) acc[j] += b * t[j];
    if (k >= 1) {
      for (int j = 0; j <= k + 1; ++j)
        next[j] = 2.0L * ((j > 0 ? 2.0L * cur[j - 1] : 0.0L) - cur[j]) - prev[j];
      for (int j = 0; j <= k + 1; ++j) {
        prev[j] = cur[j];
        cur[j] = next[j];
      }
    }
  }
  for (int j = 0; j < ns; ++j) c.sw[j] = (double)acc[j];
  return c;
}

static inline int64_t round_up(int64_t x, int64_t r) { return (x + r - 1) / r * r; }

// Tile edge of the FP32 structured squarings of order n: 64 x 64 tiles of 4
// waves up to n = 6144, 128 x 128 tiles of 16 waves with the tail split
// beyond (ms per squaring, 64 vs 128: 0.89 vs 0.93 at 3072, 6.30 vs 6.31 at
// 6144, 9.55 vs 9.31 at 7168, 14.9 vs 13.95 at 8192, 26.15 vs 25.24 at 10^4,
// 57.05 vs 54.39 at 13000; the same values); forced = 64 or 128 overrides.
static int square_tile(int64_t n, int forced) {
  if (forced == 64 || forced == 128) return forced;
  return n <= 6144 ? 64 : 128;
}

// Launch plan of one FP32 structured squaring of order m in tile x tile
// lower tiles: nb tiles; with 128 tiles and the tail split, when the last
// wave of resident workgroups would be at most half full, `full` tiles run
// in one launch and the last `rest` as 64 x 64 sub-tiles, `sub` of them
// (the upper sub-tile of each diagonal tile is skipped); rest = 0 otherwise.
struct SquarePlan {
  int nb = 0, full = 0, rest = 0, sub = 0;
};
static SquarePlan square_plan(int m, int tile, bool tail_split) {
  SquarePlan p;
  const int T = (m + tile - 1) / tile;
  p.nb = T * (T + 1) / 2;
  p.full = p.nb;
  if (tile != 128 || !tail_split) return p;
  // resident workgroups of the 128-tile kernel on the device (thread-safe one-time initialisation)
  static const int cap = [] {
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void*)gemm_nt_kernel<float, MODE_C3M, true, true, FCfg16w>, FCfg16w::NT, 0) ==
            hipSuccess)
      return std::max(1, per_cu) * prop.multiProcessorCount;
    return 0;
  }();
  static_assert(FCfg16w::BM == 2 * FCfgSmall::BM && FCfg16w::BN == 2 * FCfgSmall::BN, "tail sub-tiles halve the tile");
  const int rest = cap > 0 ? p.nb % cap : 0;
  if (!(cap > 0 && p.nb - rest > 0 && rest > 0 && 2 * rest <= cap)) return p;
  p.rest = rest;
  p.full = p.nb - rest;
  p.sub = 4 * rest;
  for (int t = p.full; t < p.nb; ++t) {
    int tm = (int)((std::sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((tm + 1) * (tm + 2) / 2 <= t) ++tm;
    while (tm * (tm + 1) / 2 > t) --tm;
    if (t - tm * (tm + 1) / 2 == tm) --p.sub;
  }
  return p;
}

// ---------------------------------------------------------------------------
// Device memory of the pipelines comes from hipMalloc, or, while an arena
// is active on the calling thread, from that fixed region (the
// no-allocation entry point, fexpm_run_device_noalloc): a bump allocator
// that never frees and fails with FEXPM_ENOMEM when the region is full.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
};
static thread_local Arena* g_arena = nullptr;

// Accounting of the device memory an engine holds: the sizes of the live
// hipMalloc blocks of dmalloc_bytes, and, while an engine call runs on a
// thread, the counter of that engine (g_acct), which every allocation and
// release through dmalloc_bytes / dfree adjusts.
static std::mutex g_dev_mtx;
static std::unordered_map<void*, size_t> g_dev_sizes;
static thread_local int64_t* g_acct = nullptr;

static void* dmalloc_bytes(size_t bytes) {
  if (g_arena) {
    const size_t off = (g_arena->used + 255) / 256 * 256;
    if (off > g_arena->cap || bytes > g_arena->cap - off)
      throw Error(FEXPM_ENOMEM, "no-allocation scratch array exhausted");
    g_arena->used = off + bytes;
    return g_arena->base + off;
  }
  void* p = nullptr;
  FX_HIP(hipMalloc(&p, bytes));
  std::lock_guard<std::mutex> lk(g_dev_mtx);
  g_dev_sizes[p] = bytes;
  if (g_acct) *g_acct += (int64_t)bytes;
  return p;
}
// Zeroed device array (arena memory is cleared once when the arena is set up).
template <typename T> static T* dalloc(size_t count) {
  if (count == 0) count = 1;
  void* p = dmalloc_bytes(count * sizeof(T));
  if (!g_arena) FX_HIP(hipMemset(p, 0, count * sizeof(T)));
  return static_cast<T*>(p);
}
static void dfree(void* p) {
  if (!p || g_arena) return;
  {
    std::lock_guard<std::mutex> lk(g_dev_mtx);
    auto it = g_dev_sizes.find(p);
    if (it != g_dev_sizes.end()) {
      if (g_acct) *g_acct -= (int64_t)it->second;
      g_dev_sizes.erase(it);
    }
  }
  (void)hipFree(p);
}

// Moduli and reconstruction constants of the residue squarings: the first K
// (7, 8 or 9) of the nine largest pairwise coprime odd moduli below 256,
// u_m = (M / m)^-1 mod m with M the product of those K.
static CrtTable crt_make_table(int K) {
  static const uint32_t mods[CRT_K] = {255, 253, 251, 247, 241, 239, 233, 229, 227};
  CrtTable t;
  std::memset(&t, 0, sizeof(t));
  t.k = K;
  double M = 1.0;
  for (int i = 0; i < K; ++i) M *= (double)mods[i];
  t.M = M;
  for (int i = 0; i < K; ++i) {
    const uint32_t m = mods[i];
    t.m[i] = m;
    t.magic[i] = (uint32_t)(((1ull << 39) + m - 1) / m);
    t.rcp[i] = 1.0 / (double)m;
    uint64_t mi = 1;  // (M / m) mod m
    for (int j = 0; j < K; ++j)
      if (j != i) mi = (mi * mods[j]) % m;
    uint32_t u = 0;  // its inverse mod m (m < 256: by search)
    for (uint32_t c = 1; c < m; ++c)
      if ((mi * c) % m == 1) {
        u = c;
        break;
      }
    t.inv[i] = u;
  }
  return t;
}
static const CrtTable& crt_table(int K) {
  static const CrtTable tabs[CRT_K - CRT_KMIN + 1] = {crt_make_table(7), crt_make_table(8), crt_make_table(9)};
  static_assert(CRT_KMIN == 7 && CRT_K == 9, "one table per moduli count");
  return tabs[K - CRT_KMIN];
}
// Bits kept of each row's largest element for products of length k with
// the moduli of tab: the largest beta <= CRT_BETA with k 2^(2 beta + 1) <
// M / 2, so that |Re|, |Im| <= k 2^(2 beta + 1) are recovered exactly.  The
// margin 2^-20 covers the rounding of M (relative 2^-50 at most).
static int crt_beta(const CrtTable& tab, int64_t k) {
  int beta = CRT_BETA;
  while (beta > 0 && std::ldexp((double)k, 2 * beta + 1) >= std::ldexp(tab.M, -1) * (1.0 - 0x1p-20)) --beta;
  return beta;
}

// Constants of the FP64 residue products with the first K (8 .. 13) of the
// nine moduli above and 223, 217, 211, 199 (pairwise coprime), M exact in
// 128-bit arithmetic.
static CrtTable64 crt_make_table64(int K) {
  static const uint32_t mods[CRT64_K] = {255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199};
  CrtTable64 t;
  std::memset(&t, 0, sizeof(t));
  t.k = K;
  unsigned __int128 M = 1;
  for (int i = 0; i < K; ++i) M *= mods[i];
  t.M_hi = (double)M;
  const __int128 rest = (__int128)M - (__int128)(unsigned __int128)t.M_hi;
  t.M_lo = (double)rest;
  t.log2M = std::log2(t.M_hi);
  for (int i = 0; i < K; ++i) {
    const uint32_t m = mods[i];
    t.m[i] = m;
    t.magic[i] = (uint32_t)(((1ull << 39) + m - 1) / m);
    t.c24[i] = (uint32_t)((1u << 24) % m);
    t.rcp_hi[i] = 1.0 / (double)m;
    t.rcp_lo[i] = std::fma(-t.rcp_hi[i], (double)m, 1.0) / (double)m;
    uint64_t mi = 1;
    for (int j = 0; j < K; ++j)
      if (j != i) mi = (mi * mods[j]) % m;
    uint32_t u = 0;
    for (uint32_t c = 1; c < m; ++c)
      if ((mi * c) % m == 1) {
        u = c;
        break;
      }
    t.inv[i] = u;
  }
  return t;
}
static const CrtTable64& crt_table64(int K = CRT64_K) {
  static const CrtTable64 tabs[CRT64_K - CRT64_KMIN + 1] = {crt_make_table64(8),  crt_make_table64(9),
                                                             crt_make_table64(10), crt_make_table64(11),
                                                             crt_make_table64(12), crt_make_table64(13)};
  static_assert(CRT64_KMIN == 8 && CRT64_K == 13, "one table per moduli count");
  return tabs[K - CRT64_KMIN];
}

// A device buffer that grows on demand (freeing its previous contents; the
// device synchronises in hipFree, so no launch still uses them) and that
// several pipelines may share when they use it one after another on one
// stream: the residue products of an FP32 pipeline's squarings and of its
// FP64 companion's products.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { dfree(p); }
  void* get(size_t need) {
    if (need > bytes) {
      dfree(p);
      p = nullptr;
      bytes = 0;
      p = dmalloc_bytes(need);
      bytes = need;
    }
    return p;
  }
};

template <typename W> struct Plane2 {
  W* r = nullptr;
  W* i = nullptr;
};

struct PipelineBase {
  virtual ~PipelineBase() {}
  virtual void run(const void* d_a, void* d_u, hipStream_t st, fexpm_stats& stats, int opt_structure) = 0;
  virtual void apply(const double2* d_psi0, int nstates, double2* d_out, hipStream_t st) = 0;
  virtual void propagate(const double2* d_psi0, int nstates, int nsteps, double2* d_out, hipStream_t st) = 0;
  virtual bool has_result() const = 0;
  virtual void debug_plane(int which, double* re, double* im) = 0;
  virtual void set_size(int n_new, hipStream_t st) = 0;
  virtual void analyze(const void* d_a, Scalars* slot, hipStream_t st) = 0;
  virtual void run_planned(const void* d_a, void* d_u, const Scalars& hs, hipStream_t st, fexpm_stats& S,
                           int opt_structure, int* info_slot) = 0;
  int max_squarings = -1;
  double last_prop_ms = 0.0;  // device time of the last propagate()
  int kblock = 8;          // FP32 squarings: k-tiles (of 16) per accumulation block, 0 = off
  int fp64_squarings = 0;  // FP32 engines: leading squarings done in FP64
  int pade_fp64 = 1;       // FP32 engines: Pade + solve in FP64 (companion FP64 pipeline)
  int panel_coop = 1;      // LU panels by one cooperative launch (0: one launch per column)
  int spd_solve = 1;       // structured inputs: Cholesky solve (0: pivoted LU as for general A)
  int precision = 0;       // enum fexpm_precision for the structured squarings of FP32 engines
  int oz_slices = 0;       // FP64 companion: symmetric FP64 products from int8 slices (0 = FP64 MFMA)
  int oz_scope = 3;        // stages using them: bit 0 the Pade products, bit 1 the solve (S, X = Z Z^T);
                           // 0 = automatic (slice_scope)
  int tail_split = 1;      // FP32 squarings: last partial wave of 128-tiles as 64-tiles
  int z_overlap = 1;       // Cholesky solve: Z updates on a second stream
  int ch_outer = 512;      // Cholesky solve: outer block width (a multiple of 64)
  int sq_tile = 0;         // FP32 structured squarings: 0 auto, 64 or 128 (tile edge)
  int z_left = 0;          // Cholesky solve: left-looking Z updates (one long-K product per outer block)
  int sq2_lead = 0;        // structured squarings: leading squarings in the two-product form
  int spectral = 1;        // structured inputs: s from a spectral-radius bound (0: the 1-norm rule)
  int deflate = 1;         // spectral scaling: deflate an exact eigenvector 1 (equal row sums)
  int cheb = 1;            // spectral plans: exp(iX) from its Chebyshev polynomial (0: Pade + solve)
  int crt_map = 1;         // block map of the residue launches (crt_launch_blocks; 1: 185.7 vs 191.3 ms at C2)
  // moduli of the residue squarings (7, 8 or 9; FEXPM_OPT_CRT_MODULI).  8
  // (rows to 24 bits at n = 10^4): 2-4x more accurate than the FP32 MFMA
  // squarings in every case of the N = 10^4 precision study (ER(N, 0.5)
  // Laplacians, unweighted and weighted, t = 1, 4.5, 20), 6-8% faster per
  // expm than 9 moduli
  int crt_moduli = 8;
  // moduli of the FP64-accurate residue products: crt64_opt (8 .. 13;
  // FEXPM_OPT_CRT64_MODULI) for every plan, or with crt64_opt = 0 the
  // automatic count: 11 (rows to 36 bits at k = 10^4) for spectral plans,
  // 13 for 1-norm plans.  At N = 10^4 (spectral plans: C2, the six
  // precision-study cases, the 32-time propagation of a weighted ER(8192,
  // 0.3)) 13, 12, 11 and 10 moduli gave the same infidelity to within 2.3x
  // (10 moduli, unweighted t = 20) and 9 up to 30x more, with 8-moduli
  // squarings.  crt64_moduli is the count of the current run's FP64 pipeline
  int crt64_opt = 0;
  int crt64_moduli = CRT64_K;
  int crt64_for(bool spectral) const { return crt64_opt ? crt64_opt : (spectral ? 11 : CRT64_K); }
  int last_crt_beta = 0;   // bits per row of the last residue squarings
};

// ---------------------------------------------------------------------------
// Two-level blocked right-looking LU with partial pivoting of the augmented
// FP64 system S = [M | B] (np x 2np, leading dimension ldl; np a multiple of
// 128), followed by block back substitution.  On return the right block
// holds X = M^{-1} B.
//
// Outer blocks of LU_OUTER columns are factored with 64-wide inner panels
// whose updates stay inside the outer panel; the rows they interchange are
// also interchanged in the outer panel's earlier L columns, so that the
// outer L21 is consistent when it is used.  Then the outer block row
// U12 = L11^{-1} A12 is formed slab by slab and the trailing matrix,
// including the right-hand side, is updated by one complex GEMM with
// K = LU_OUTER (4M on FP64 MFMA; measured faster than 3M here because the
// 3M variant needs more registers per wave).  The forward substitution thus happens
// inside the factorisation.  Back substitution uses the same two levels.
// temp (tr/ti) holds 2np x LU_OUTER values per plane.  Returns the real
// flops issued by the MFMA GEMMs.
constexpr int LU_OUTER = 256;
constexpr int CH_OUTER = 256;  // outer block of the grouped structured (Cholesky) solve
constexpr int SQ_GROUP = 8;    // tile group of the split squaring order

// Panel factorisation: one cooperative launch per 64-wide panel (default)
// or one launch per column.
static int g_panel_coop = 1;

// Largest cooperative grid of lu_panel_coop that the device can hold.
static int panel_coop_capacity() {
  static const int cap = [] {  // thread-safe one-time initialisation
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)lu_panel_coop, 256, 0) != hipSuccess)
      return 0;
    return per_cu * prop.multiProcessorCount;
  }();
  return cap;
}

// Complex FP64 trailing products of the LU factorisation and solve: 3M
// (three real products, Re = ArBr - AiBi, Im = (Ar+Ai)(Br+Bi) - ArBr -
// AiBi), a quarter fewer flops than the four-product form at the same
// norm-wise error bound.
static double lu_gemm(const GemmArgs<double>& g, hipStream_t st) {
  using C = GemmCfg<double>;
  constexpr int NT = (C::BM / C::WM) * (C::BN / C::WN) * 64;
  if (g.m <= 0 || g.n <= 0) return 0.0;
  dim3 grid((g.n + C::BN - 1) / C::BN, (g.m + C::BM - 1) / C::BM);
  gemm_nt_kernel<double, MODE_C3M, false><<<grid, dim3(NT), 0, st>>>(g);
  FX_LAUNCH();
  return 3.0 * 2.0 * (double)grid.x * grid.y * C::BM * C::BN * (double)g.k;
}

// FP64 product with the default tile shape; returns the real flops issued.
template <int MODE, bool SYM>
static double gemm64(const GemmArgs<double>& g, hipStream_t st) {
  using C = GemmCfg<double>;
  constexpr int NT = (C::BM / C::WM) * (C::BN / C::WN) * 64;
  if (g.m <= 0 || g.n <= 0) return 0.0;
  double tiles;
  if (SYM) {
    const int T = (g.m + C::BM - 1) / C::BM;
    tiles = (double)(T * (T + 1) / 2);
    gemm_nt_kernel<double, MODE, true><<<dim3(T * (T + 1) / 2), dim3(NT), 0, st>>>(g);
  } else {
    dim3 grid((g.n + C::BN - 1) / C::BN, (g.m + C::BM - 1) / C::BM);
    tiles = (double)grid.x * grid.y;
    gemm_nt_kernel<double, MODE, false><<<grid, dim3(NT), 0, st>>>(g);
  }
  FX_LAUNCH();
  const double mult = MODE == MODE_REAL ? 1.0 : (MODE == MODE_C4M ? 4.0 : 3.0);
  return mult * 2.0 * tiles * C::BM * C::BN * (double)g.k;
}

// Real FP64 NT product C = alpha A B^T + beta C on plain pointers.
static double dgemm_real(const double* a, int64_t lda, const double* b, int64_t ldb, double* c, int64_t ldc, int m,
                         int n, int k, double alpha, double beta, bool sym, hipStream_t st) {
  GemmArgs<double> g;
  std::memset(&g, 0, sizeof(g));
  g.ar = a;
  g.lda = lda;
  g.br = b;
  g.ldb = ldb;
  g.m = m;
  g.n = n;
  g.k = k;
  g.epi.cr = c;
  g.epi.ldc = ldc;
  g.epi.ld_add = ldc;
  g.epi.alpha = alpha;
  g.epi.beta = beta;
  return sym ? gemm64<MODE_REAL, true>(g, st) : gemm64<MODE_REAL, false>(g, st);
}

// C[rows, cols] -= A[rows, k] * B, with B given transposed (bt: ncols x k, ldb).
static double lu_update(double* lr, double* li, int64_t ldl, int r0, int nrows, int kcol0, int k, int c0,
                        int ncols, const double* btr, const double* bti, int ldb, hipStream_t st) {
  GemmArgs<double> g;
  std::memset(&g, 0, sizeof(g));
  g.ar = lr + (int64_t)r0 * ldl + kcol0;
  g.ai = li + (int64_t)r0 * ldl + kcol0;
  g.lda = ldl;
  g.br = btr;
  g.bi = bti;
  g.ldb = ldb;
  g.m = nrows;
  g.n = ncols;
  g.k = k;
  g.epi.cr = lr + (int64_t)r0 * ldl + c0;
  g.epi.ci = li + (int64_t)r0 * ldl + c0;
  g.epi.ldc = ldl;
  g.epi.alpha = -1.0;
  g.epi.beta = 1.0;
  return lu_gemm(g, st);
}

static void lu_swap_cols(double* lr, double* li, int64_t ldl, int k0, const int* ipiv, int c0, int c1,
                         hipStream_t st) {
  if (c1 <= c0) return;
  lu_laswp<<<(c1 - c0 + 255) / 256, 256, 0, st>>>(lr, li, ldl, k0, LU_NB, ipiv, c0, c1);
  FX_LAUNCH();
}

// S[k0:k0+64, c0:c1] <- T^{-1} S[k0:k0+64, c0:c1] with T the unit lower
// (UPPER = false) or upper (UPPER = true) diagonal block at k0; transposed
// copy to tr/ti.  inv holds 2 x 64 x 64 doubles of scratch.
template <bool UPPER>
static void lu_block_row_solve(double* lr, double* li, int64_t ldl, int k0, int c0, int c1, double* tr, double* ti,
                               int ldt, int toff, double* inv, hipStream_t st) {
  if (c1 <= c0) return;
  lu_invert_diag_block<UPPER><<<1, 256, 0, st>>>(lr, li, ldl, k0, inv, inv + LU_NB * LU_NB);
  FX_LAUNCH();
  lu_apply_inverse<<<(c1 - c0 + 63) / 64, 256, 0, st>>>(lr, li, ldl, k0, c0, c1, inv, inv + LU_NB * LU_NB, tr, ti,
                                                        ldt, toff);
  FX_LAUNCH();
}

static double lu_factor_solve(double* lr, double* li, int64_t ldl, int inp, double* tr, double* ti, int* ipiv,
                              PanelCand* cand, int gmax, double* top, Scalars* scal, double* inv,
                              unsigned* ctrs, hipStream_t st, bool coop = true) {
  double flops = 0.0;
  // one zeroed arrival counter per panel for the cooperative panel kernel
  FX_HIP(hipMemsetAsync(ctrs, 0, (size_t)(inp / LU_NB) * sizeof(unsigned), st));
  const int nb = LU_NB;
  const int ncol = 2 * inp;
  for (int K0 = 0; K0 < inp; K0 += LU_OUTER) {
    const int nbo = std::min(LU_OUTER, inp - K0);
    const int K1 = K0 + nbo;
    // --- outer panel S[K0:, K0:K1] with 64-wide inner panels ---
    for (int k0 = K0; k0 < K1; k0 += nb) {
      const int G = (inp - k0 + PANEL_R - 1) / PANEL_R;
      if (inp - k0 <= PANEL_R1) {  // the panel's rows fit one workgroup
        lu_panel_single<<<1, 256, 0, st>>>(lr, li, ldl, inp, k0, ipiv, scal);
        FX_LAUNCH();
      } else if (coop && g_panel_coop && G <= panel_coop_capacity()) {
        unsigned* ctr = ctrs + k0 / LU_NB;
        void* kargs[] = {&lr, &li, (void*)&ldl, &inp, (void*)&k0, (void*)&G, &cand, &top, &ipiv, &scal, &ctr};
        FX_HIP(hipLaunchCooperativeKernel((const void*)lu_panel_coop, dim3(G), dim3(256), kargs, 0, st));
      } else {
        lu_panel_init<<<G, 256, 0, st>>>(lr, li, ldl, inp, k0, cand, top);
        FX_LAUNCH();
        for (int j = 0; j < nb; ++j) {
          const int par = j & 1;
          lu_panel_step<<<G, 256, 0, st>>>(lr, li, ldl, inp, k0, j, G, cand + par * gmax,
                                           cand + (par ^ 1) * gmax, top + par * 2 * LU_NB,
                                           top + (par ^ 1) * 2 * LU_NB, ipiv, scal);
          FX_LAUNCH();
        }
      }
      lu_swap_cols(lr, li, ldl, k0, ipiv, K0, k0, st);          // earlier L columns of this outer panel
      lu_swap_cols(lr, li, ldl, k0, ipiv, k0 + nb, K1, st);     // rest of the outer panel
      if (k0 + nb < K1) {
        lu_block_row_solve<false>(lr, li, ldl, k0, k0 + nb, K1, tr, ti, nb, 0, inv, st);
        flops += lu_update(lr, li, ldl, k0 + nb, inp - k0 - nb, k0, nb, k0 + nb, K1 - k0 - nb, tr, ti, nb, st);
      }
    }
    if (K1 >= ncol) continue;
    // --- interchanges of the whole outer panel on the columns to its right ---
    for (int k0 = K0; k0 < K1; k0 += nb) lu_swap_cols(lr, li, ldl, k0, ipiv, K1, ncol, st);
    // --- block row U12 = L11^{-1} A12, slab by slab; transposed into tr (ld LU_OUTER) ---
    for (int r = K0; r < K1; r += nb) {
      lu_block_row_solve<false>(lr, li, ldl, r, K1, ncol, tr, ti, LU_OUTER, r - K0, inv, st);
      if (r + nb < K1)
        flops += lu_update(lr, li, ldl, r + nb, K1 - r - nb, r, nb, K1, ncol - K1, tr + (r - K0), ti + (r - K0),
                           LU_OUTER, st);
    }
    // --- trailing update with K = nbo ---
    if (K1 < inp) flops += lu_update(lr, li, ldl, K1, inp - K1, K0, nbo, K1, ncol - K1, tr, ti, LU_OUTER, st);
  }
  // --- back substitution U X = Y on the right block ---
  const int nouter = (inp + LU_OUTER - 1) / LU_OUTER;
  for (int ob = nouter - 1; ob >= 0; --ob) {
    const int I0 = ob * LU_OUTER;
    const int nbo = std::min(LU_OUTER, inp - I0);
    for (int r = I0 + nbo - nb; r >= I0; r -= nb) {
      lu_block_row_solve<true>(lr, li, ldl, r, inp, ncol, tr, ti, LU_OUTER, r - I0, inv, st);
      if (r > I0)
        flops += lu_update(lr, li, ldl, I0, r - I0, r, nb, inp, inp, tr + (r - I0), ti + (r - I0), LU_OUTER, st);
    }
    if (I0 > 0) flops += lu_update(lr, li, ldl, 0, I0, I0, nbo, inp, inp, tr, ti, LU_OUTER, st);
  }
  return flops;
}

template <typename W> struct Pipeline : PipelineBase {
  // Buffers are sized for the capacity ncap; a run works on the current
  // order n <= ncap (set_size), np = round_up(n, 128).  Leading dimensions
  // are those of the capacity.
  int n, dtype;
  int ncap;
  int64_t np, ld;        // current padded order / leading dimension of W planes
  int64_t npcap;
  int64_t ldl;           // leading dimension of the FP64 system [M | B]
  Plane2<W> a, p2, p4, p6, t1, t2, u, v, x0, x1;
  double* lr = nullptr;  // FP64 system, np x 2np
  double* li = nullptr;
  double* tr = nullptr;  // transposed solved block rows (2np x LU_OUTER)
  double* ti = nullptr;
  int* ipiv = nullptr;
  PanelCand* cand = nullptr;  // 2 x gmax
  double* top = nullptr;      // 2 x 2*LU_NB
  double* inv = nullptr;      // 2 x 64 x 64 triangular-block inverse
  unsigned* ctrs = nullptr;   // grid-barrier counters, one per panel
  int gmax = 0;
  double* colpart = nullptr;
  double* off = nullptr;
  double2* diag = nullptr;
  double2* colsig = nullptr;  // double-double row-block partial row sums of Im(A)
  double2* rsum = nullptr;    // double-double row sums of Im(A)
  double* npart = nullptr;    // 1-norm partial column sums (spectral bounds)
  double* d_norms = nullptr;  // ||B^k||_1, k = 1, 2, 4, 6; trace B^4, B^6, ..., B^12
  double* ncmax = nullptr;    // per-workgroup column-sum maxima of the 1-norms
  double2* dsq[2] = {nullptr, nullptr};  // split-form diagonals (squarings)
  int* flags = nullptr;
  Scalars* scal = nullptr;
  int* info_buf = nullptr;  // exception bits of the last run (device)
  void* sbuf = nullptr;  // k-blocked 16-bit split parts (split-precision squarings)
  size_t sbuf_bytes = 0;
  // residue squarings (FEXPM_PREC_INT8_CRT): row exponents, operand residues
  // (3 CRT_K planes of rows x kpad), product residues (3 CRT_K planes of
  // rows x rows), the launch's problem list for the current order
  int* crt_e = nullptr;
  // FP64 residue products (FP64 pipelines; oz_slices == CRT64_K): row
  // exponents of A and B
  int* c64_e[2] = {nullptr, nullptr};
  // operand and product residues of both: one buffer, shared with the FP64
  // companion (companion())
  std::shared_ptr<DevBuf> crt_buf = std::make_shared<DevBuf>();
  std::map<int, int*> sq_orders;  // grouped lower-triangular tile orders, per tile count
  hipEvent_t ev[8];
  bool events = false;
  // events around each squaring's product launch (kernel-level timing of
  // single runs): pairs, created on first use
  std::vector<hipEvent_t> kev;
  bool time_kernels = false;
  int kev_used = 0;
  void kernel_mark(int q, bool end, hipStream_t st) {
    if (!time_kernels) return;
    const size_t need = 2 * (size_t)q + 2;
    while (kev.size() < need) {
      hipEvent_t e;
      FX_HIP(hipEventCreate(&e));
      kev.push_back(e);
    }
    FX_HIP(hipEventRecord(kev[2 * q + (end ? 1 : 0)], st));
    if (end) kev_used = std::max(kev_used, q + 1);
  }
  // result of the last run (for apply)
  Plane2<W> res;
  double ph_re = 1.0, ph_im = 0.0;
  bool have_result = false;
  double flops = 0.0;

  bool in_arena = false;  // buffers carved from an arena: never freed

  Pipeline(int n_, int dtype_, bool with_events = true) : n(n_), dtype(dtype_), ncap(n_) {
    in_arena = g_arena != nullptr;
    // FP32 engines: the Pade products of their FP64 stages from 6 int8
    // slices (same accuracy as FP64 MFMA, measured); the solve's products
    // too when the plan deflates (automatic scope, see slice_scope)
    if (std::is_same<W, float>::value) {
      oz_slices = CRT64_K;
      oz_scope = 0;
    }
    npcap = round_up(std::max(ncap, 1), 128);
    np = npcap;
    ld = npcap;
    ldl = 2 * npcap;
    const size_t cnt = (size_t)npcap * (size_t)ld;
    for (Plane2<W>* p : {&t1, &x0, &x1}) {
      p->r = dalloc<W>(cnt);
      p->i = dalloc<W>(cnt);
    }
    // The Pade and solve buffers come at the first run that needs them
    // (ensure_stage_buffers): an FP32 pipeline runs those stages on its
    // FP64 companion by default, and structured inputs use only the real
    // planes.  A pipeline in an arena holds all of them from the start.
    if (in_arena) {
      ensure_stage_buffers(true);
      ensure_system();
    }
    ipiv = dalloc<int>(npcap);
    gmax = (int)((npcap + PANEL_R - 1) / PANEL_R);
    cand = dalloc<PanelCand>((size_t)2 * gmax);
    top = dalloc<double>(4 * LU_NB);
    inv = dalloc<double>(2 * LU_NB * LU_NB);
    ctrs = dalloc<unsigned>(npcap / LU_NB);
    const int nt = (n + 31) / 32;
    colpart = dalloc<double>((size_t)std::max(nt, 1) * std::max(n, 1));
    colsig = dalloc<double2>((size_t)std::max(nt, 1) * std::max(n, 1));
    rsum = dalloc<double2>(std::max(n, 1));
    off = dalloc<double>(std::max(n, 1));
    diag = dalloc<double2>(std::max(n, 1));
    dsq[0] = dalloc<double2>(std::max(n, 1));
    dsq[1] = dalloc<double2>(std::max(n, 1));
    flags = dalloc<int>(4);
    scal = dalloc<Scalars>(1);
    info_buf = dalloc<int>(1);
    if (with_events) {
      for (int k = 0; k < 8; ++k) FX_HIP(hipEventCreate(&ev[k]));
      events = true;
    }
  }

  // The Pade (or Chebyshev) and solve stages of this pipeline use the planes
  // a, p2, p4, p6, t2, u, v, and the solves the FP64 system lr / li with the
  // transposed block rows tr / ti; structured inputs (A = iG) only the real
  // planes.  They are allocated (zeroed) on first use; the squarings need
  // only x0, x1 and t1, which the constructor allocates.
  bool stage_real = false, stage_cplx = false, stage_system = false;
  void ensure_stage_buffers(bool cplx) {
    const size_t cnt = (size_t)npcap * (size_t)ld;
    if (!stage_real) {
      for (Plane2<W>* p : {&a, &p2, &p4, &p6, &t2, &u, &v}) p->r = dalloc<W>(cnt);
      stage_real = true;
    }
    if (cplx && !stage_cplx) {
      for (Plane2<W>* p : {&a, &p2, &p4, &p6, &t2, &u, &v}) p->i = dalloc<W>(cnt);
      stage_cplx = true;
    }
  }

  void ensure_system() {
    if (stage_system) return;
    // lr and li are one allocation (li = lr + npcap*ldl): the structured
    // solve uses the 4 npcap^2 doubles as one contiguous work area.
    lr = dalloc<double>((size_t)2 * npcap * ldl);
    li = lr + (size_t)npcap * ldl;
    tr = dalloc<double>((size_t)2 * npcap * LU_OUTER);
    ti = dalloc<double>((size_t)2 * npcap * LU_OUTER);
    stage_system = true;
  }

  ~Pipeline() override {
    if (in_arena) return;  // arena memory belongs to the arena
    Plane2<W>* all[] = {&a, &p2, &p4, &p6, &t1, &t2, &u, &v, &x0, &x1};
    for (Plane2<W>* p : all) {
      dfree(p->r);
      dfree(p->i);
    }
    dfree(lr);
    dfree(tr);
    dfree(ti);
    dfree(ipiv);
    dfree(cand);
    dfree(top);
    dfree(inv);
    dfree(ctrs);
    dfree(colpart);
    dfree(colsig);
    dfree(rsum);
    dfree(npart);
    dfree(d_norms);
    dfree(ncmax);
    dfree(off);
    dfree(diag);
    dfree(dsq[0]);
    dfree(dsq[1]);
    for (auto& P : ps) {
      dfree(P.r);
      dfree(P.i);
    }
    dfree(norm0);
    dfree(flags);
    dfree(scal);
    dfree(info_buf);
    for (int b = 0; b < 2; ++b) {
      dfree(oz_s[b]);
      dfree(oz_e[b]);
    }
    dfree(sbuf);
    dfree(crt_e);
    for (int b = 0; b < 2; ++b) dfree(c64_e[b]);
    for (auto& kv : sq_orders) dfree(kv.second);
    for (hipEvent_t e : zev) (void)hipEventDestroy(e);
    if (zst) (void)hipStreamDestroy(zst);
    if (events)
      for (int k = 0; k < 8; ++k) (void)hipEventDestroy(ev[k]);
    for (hipEvent_t e : kev) (void)hipEventDestroy(e);
  }

  bool has_result() const override { return have_result; }

  // Switch to order n_new <= ncap.  The working planes are cleared so that
  // columns beyond the new order are zero again (GEMM k-loops rely on it).
  void set_size(int n_new, hipStream_t st) override {
    if (n_new < 0 || n_new > ncap) throw Error(FEXPM_EINVAL, "order exceeds the engine capacity");
    if (n_new == n) return;
    n = n_new;
    np = round_up(std::max(n, 1), 128);
    // Kernels write only the n x n region and GEMM k-loops read columns up
    // to np of rows < n: zeroing the strip [n, np) x [0, n) of every
    // working plane restores the padding invariant after a change of order.
    if (np > n && n > 0) {
      Plane2<W>* all[] = {&a, &p2, &p4, &p6, &t1, &t2, &u, &v, &x0, &x1};
      PlaneList<W> L;
      L.count = 0;
      for (Plane2<W>* p : all) {  // planes not allocated yet are skipped
        if (p->r) L.p[L.count++] = p->r;
        if (p->i) L.p[L.count++] = p->i;
      }
      zero_strip<W><<<(unsigned)n, 128, 0, st>>>(L, n, (int)np, ld);
      FX_LAUNCH();
    }
    for (auto& P : ps)
      if (P.r) {
        FX_HIP(hipMemsetAsync(P.r, 0, (size_t)ps_rows * ld * sizeof(W), st));
        FX_HIP(hipMemsetAsync(P.i, 0, (size_t)ps_rows * ld * sizeof(W), st));
      }
    have_result = false;
    if (hi) hi->set_size(n_new, st);
  }

  // ---- FP64 products from exact int8 slices (FP64 pipelines, structured path) ----
  // oz_slices (PipelineBase) = 0: FP64 MFMA products; 4..6: the real
  // symmetric products of the Pade stage and of the Cholesky solve (S,
  // X = Z Z^T) run on the int8 MFMA from that many 7-bit slices per value
  // (fexpm_ozaki.hpp).
  int8_t* oz_s[2] = {nullptr, nullptr};
  int* oz_e[2] = {nullptr, nullptr};
  size_t oz_cap = 0;  // bytes per slice buffer
  int oz_rows = 0;

  // Operand row r = s1 * p1[r, 0:k1] followed by s2 * p2[r, 0:nk-k1] (leading dimension ldx).
  struct OzOp {
    const double* p1;
    double s1;
    const double* p2;
    double s2;
    int k1;
    int64_t ldx;
  };

  void oz_reserve(int rows, int kp) {
    const size_t need = (size_t)rows * kp * OZ_MAX_SLICES;
    if (need <= oz_cap && rows <= oz_rows) return;
    for (int b = 0; b < 2; ++b) {
      dfree(oz_s[b]);
      dfree(oz_e[b]);
      oz_s[b] = nullptr;
      oz_e[b] = nullptr;
    }
    oz_cap = 0;
    oz_rows = 0;
    for (int b = 0; b < 2; ++b) {
      oz_s[b] = static_cast<int8_t*>(dmalloc_bytes(need));
      oz_e[b] = static_cast<int*>(dmalloc_bytes((size_t)rows * sizeof(int)));
    }
    oz_cap = need;
    oz_rows = rows;
  }

  template <int S>
  void oz_slice_t(const OzOp& x, int nr, int nk, int rows, int kp, int buf, hipStream_t st) {
    oz_slice<S><<<(unsigned)rows, 256, 0, st>>>(x.p1, x.s1, x.p2, x.s2, x.k1, x.ldx, nr, nk, rows, kp, oz_s[buf],
                                                oz_e[buf]);
    FX_LAUNCH();
  }
  void oz_slice_op(const OzOp& x, int nr, int nk, int rows, int kp, int buf, hipStream_t st) {
    switch (oz_slices) {
      case 4: oz_slice_t<4>(x, nr, nk, rows, kp, buf, st); break;
      case 5: oz_slice_t<5>(x, nr, nk, rows, kp, buf, st); break;
      default: oz_slice_t<6>(x, nr, nk, rows, kp, buf, st); break;
    }
  }
  template <int S>
  void oz_launch_t(int ba, int ra, int bb, int rb, int m, int nn, int kp, const Epilogue<double>& e, bool sym,
                   hipStream_t st) {
    OzArgs<S> g;
    std::memset(&g, 0, sizeof(g));
    g.a = oz_s[ba];
    g.ea = oz_e[ba];
    g.rows_a = ra;
    g.b = oz_s[bb];
    g.eb = oz_e[bb];
    g.rows_b = rb;
    g.m = m;
    g.n = nn;
    g.kpad = kp;
    g.epi = e;
    if (sym) {
      const int T = ra / OZ_BM;
      static_assert(OZ_BM == SQ_BM, "the grouped tile orders are built for SQ_BM tiles");
      g.order = tile_order(T, st);
      oz_gemm_kernel<S, true><<<dim3((unsigned)(T * (T + 1) / 2)), dim3(512), 0, st>>>(g);
    } else {
      oz_gemm_kernel<S, false><<<dim3((unsigned)(rb / OZ_BM), (unsigned)(ra / OZ_BM)), dim3(512), 0, st>>>(g);
    }
    FX_LAUNCH();
  }
  // Operand reuse across consecutive residue products: while c64_keep names
  // a plane, a product whose first operand is that plane alone (scale 1)
  // computes its row exponents and residues once and the following ones
  // with the same order, k and buffer reuse them from the front of the
  // shared residue buffer.  The caller sets c64_keep around such a sequence
  // (KeepOperand), in which no other use of the buffer may come between.
  const void* c64_keep = nullptr;
  bool c64_kept = false;
  int c64_kept_m = 0, c64_kept_k = 0, c64_kept_moduli = 0;
  const void* c64_kept_base = nullptr;
  struct KeepOperand {
    Pipeline* p;
    KeepOperand(Pipeline* q, const void* plane) : p(q) {
      p->c64_keep = plane;
      p->c64_kept = false;
    }
    ~KeepOperand() {
      p->c64_keep = nullptr;
      p->c64_kept = false;
    }
  };

  // Symmetric C (epilogue e, m x m) = A B^T by residue products modulo
  // CRT64_K moduli (fexpm_crt.hpp; B null: B = A).  Returns the FP64-
  // equivalent flops of the product.
  double crt64_product(const OzOp& A, const OzOp* B, int m, int nk, const Epilogue<double>& e, hipStream_t st) {
    const CrtTable64& tab = crt_table64(crt64_moduli);
    const int K = tab.k;
    const int rows = (int)round_up(m, CRT_TILE), kpad = (int)round_up(nk, CRT_KPAD);
    // |C| <= k 2^(2 beta) < M / 2; at most 45 bits (the range of crt64_sres)
    const int beta = std::min(CRT64_BETA_MAX, (int)std::floor((tab.log2M - 1.0 - std::log2((double)kpad)) * 0.5));
    const size_t op_bytes = (size_t)K * rows * kpad;
    const int nops = B ? 2 : 1;
    for (int b = 0; b < nops; ++b)
      if (!c64_e[b]) c64_e[b] = dalloc<int>(round_up(npcap, CRT_TILE));
    // operands of A (and B), then the product residues; sized once for the
    // largest product of the pipeline (one operand with k = 2 npcap, or two
    // with k = npcap) so that the buffer does not grow from one to the next
    const size_t rc = (size_t)round_up(npcap, CRT_TILE);
    const size_t cap_need = (size_t)CRT64_K * rc *
                            (std::max((size_t)round_up(2 * npcap, CRT_KPAD), 2 * (size_t)round_up(npcap, CRT_KPAD)) + rc);
    uint8_t* base = static_cast<uint8_t*>(crt_buf->get(std::max(cap_need, nops * op_bytes +
                                                                 (size_t)CRT64_K * rows * rows)));
    int8_t* c64_op[2] = {reinterpret_cast<int8_t*>(base), reinterpret_cast<int8_t*>(base + op_bytes)};
    uint8_t* c64_res = base + nops * op_bytes;
    const int G = rows / 32, KB = kpad / 32, T = rows / CRT_TILE, tiles_per = T * (T + 1) / 2;
    const int64_t jobs = (int64_t)G * KB;
    const bool keep = c64_keep && A.p1 == c64_keep && !A.p2 && A.s1 == 1.0;
    if (!keep) c64_kept = false;  // the front of the buffer gets another operand
    for (int b = 0; b < nops; ++b) {
      const OzOp& x = b == 0 ? A : *B;
      if (b == 0 && keep) {
        if (c64_kept && c64_kept_m == m && c64_kept_k == nk && c64_kept_moduli == K && c64_kept_base == base)
          continue;
        c64_kept = true;
        c64_kept_m = m;
        c64_kept_k = nk;
        c64_kept_moduli = K;
        c64_kept_base = base;
      }
      crt64_row_exp<<<(m + 3) / 4, 256, 0, st>>>(x.p1, x.s1, x.p2, x.s2, x.k1, x.ldx, m, nk, c64_e[b]);
      FX_LAUNCH();
      const int vec = (x.k1 % 16 == 0 && x.ldx % 2 == 0 && (uintptr_t)x.p1 % 16 == 0 &&
                       (!x.p2 || (uintptr_t)x.p2 % 16 == 0)) ? 1 : 0;
      crt64_residues<<<(unsigned)((jobs + 3) / 4), 256, 0, st>>>(x.p1, x.s1, x.p2, x.s2, x.k1, x.ldx, m, nk,
                                                                  c64_e[b], beta, rows, kpad, vec, tab, c64_op[b]);
      FX_LAUNCH();
    }
    CrtProbs probs;
    std::memset(&probs, 0, sizeof(probs));
    const int64_t op_plane = (int64_t)rows * kpad, res_plane = (int64_t)rows * rows;
    for (int q = 0; q < K; ++q) {
      CrtProb& pq = probs.p[q];
      pq.a = c64_op[0] + q * op_plane;
      pq.b = c64_op[nops - 1] + q * op_plane;
      pq.out = c64_res + q * res_plane;
      pq.m = tab.m[q];
      pq.magic = tab.magic[q];
    }
    crt_gemm_kernel<<<crt_launch_blocks(K, tiles_per, crt_map), 256, 0, st>>>(probs, tiles_per, T, G, KB, crt_map);
    FX_LAUNCH();
    const int T64 = (m + 63) / 64;
    crt64_combine<<<T64 * (T64 + 1) / 2, 256, 0, st>>>(c64_res, rows, c64_e[0], c64_e[nops - 1], beta, m, e, tab);
    FX_LAUNCH();
    return (double)m * m * nk;
  }

  // C (epilogue e) = A B^T from slices (B null: B = A, sliced once).  sym:
  // the product is symmetric (m == n), lower tiles computed and mirrored.
  // With oz_slices == CRT64_K the symmetric products take the residue form
  // (crt64_product) while k fits it, else six slices.  Returns the FP64-
  // equivalent flops of the product.
  double oz_product(const OzOp& A, const OzOp* B, int m, int nn, int nk, const Epilogue<double>& e, bool sym,
                    hipStream_t st) {
    if (oz_slices == CRT64_K && sym && m == nn && round_up(nk, CRT_KPAD) <= CRT64_KMAX)
      return crt64_product(A, B, m, nk, e, st);
    const int kp = (int)round_up(nk, OZ_BK), ra = (int)round_up(m, OZ_BM), rb = (int)round_up(nn, OZ_BM);
    oz_reserve(std::max(ra, rb), kp);
    oz_slice_op(A, m, nk, ra, kp, 0, st);
    if (B) oz_slice_op(*B, nn, nk, rb, kp, 1, st);
    const int bb = B ? 1 : 0, rbb = B ? rb : ra;
    switch (oz_slices) {
      case 4: oz_launch_t<4>(0, ra, bb, rbb, m, nn, kp, e, sym, st); break;
      case 5: oz_launch_t<5>(0, ra, bb, rbb, m, nn, kp, e, sym, st); break;
      default: oz_launch_t<6>(0, ra, bb, rbb, m, nn, kp, e, sym, st); break;
    }
    return (sym ? 1.0 : 2.0) * (double)m * nn * nk;
  }

  // ---- launch helpers ----
  template <int MODE, bool SYM, bool KBLK = false, class C = GemmCfg<W>>
  void gemm(const GemmArgs<W>& g, hipStream_t st) {
    constexpr int NT = (C::BM / C::WM) * (C::BN / C::WN) * 64;
    double tiles;
    if (SYM) {
      const int T = (g.m + C::BM - 1) / C::BM;
      const int nb = T * (T + 1) / 2;
      gemm_nt_kernel<W, MODE, SYM, KBLK, C><<<dim3(nb), dim3(NT), 0, st>>>(g);
      tiles = nb;
    } else {
      dim3 grid((g.n + C::BN - 1) / C::BN, (g.m + C::BM - 1) / C::BM);
      gemm_nt_kernel<W, MODE, SYM, KBLK, C><<<grid, dim3(NT), 0, st>>>(g);
      tiles = (double)grid.x * grid.y;
    }
    FX_LAUNCH();
    const double mult = MODE == MODE_REAL ? 1.0 : (MODE == MODE_C4M ? 4.0 : (MODE == MODE_SQ2 ? 2.0 : 3.0));
    flops += mult * 2.0 * tiles * C::BM * C::BN * (double)g.k;
  }

  static GemmArgs<W> args(const Plane2<W>& A, const Plane2<W>& B, int64_t lda, int64_t ldb, int m, int nn, int k) {
    GemmArgs<W> g;
    std::memset(&g, 0, sizeof(g));
    g.ar = A.r;
    g.ai = A.i;
    g.lda = lda;
    g.br = B.r;
    g.bi = B.i;
    g.ldb = ldb;
    g.m = m;
    g.n = nn;
    g.k = k;
    g.epi.alpha = 1.0;
    return g;
  }

  void set_out(GemmArgs<W>& g, const Plane2<W>& C) {
    g.epi.cr = C.r;
    g.epi.ci = C.i;
    g.epi.ldc = ld;
    g.epi.ld_add = ld;
  }

  void lincomb(const Plane2<W>& out, bool cplx, std::initializer_list<std::pair<double, const Plane2<W>*>> terms,
               double dg, hipStream_t st) {
    double c[4];
    const Plane2<W>* pl[4];
    int k = 0;
    for (auto& t : terms) {
      c[k] = t.first;
      pl[k] = t.second;
      ++k;
    }
    lincomb_n(out, cplx, k, c, pl, dg, st);
  }

  // out = sum_k c[k] in[k] + dg I over the n x n region (at most 4 terms).
  void lincomb_n(const Plane2<W>& out, bool cplx, int nin, const double* c, const Plane2<W>* const* in, double dg,
                 hipStream_t st) {
    Lincomb<W> L;
    std::memset(&L, 0, sizeof(L));
    L.out_r = out.r;
    L.out_i = cplx ? out.i : nullptr;
    for (int k = 0; k < nin; ++k) {
      L.c[k] = c[k];
      L.in_r[k] = in[k]->r;
      L.in_i[k] = in[k]->i;
    }
    L.nin = nin;
    L.diag = dg;
    dim3 grid((n + 255) / 256, n);
    lincomb_kernel<W><<<grid, 256, 0, st>>>(L, n, ld);
    FX_LAUNCH();
  }

  void transpose(const Plane2<W>& in, const Plane2<W>& out, bool cplx, hipStream_t st) {
    dim3 grid((n + 31) / 32, (n + 31) / 32);
    transpose_kernel<W><<<grid, 256, 0, st>>>(in.r, cplx ? in.i : nullptr, out.r, cplx ? out.i : nullptr, n, ld);
    FX_LAUNCH();
  }

  // NT product with symmetric (half-tile) output, real planes:
  // C = alpha A B^T + sum_k c[k] add[k] + dg I (at most 3 terms).
  void sym_real(const Plane2<W>& A, const Plane2<W>& B, const Plane2<W>& C, hipStream_t st,
                std::initializer_list<std::pair<double, const Plane2<W>*>> add = {}, double dg = 0.0,
                double alpha = 1.0) {
    double c[3];
    const Plane2<W>* pl[3];
    int k = 0;
    for (auto& t : add) {
      c[k] = t.first;
      pl[k] = t.second;
      ++k;
    }
    sym_real_n(A, B, C, st, k, c, pl, dg, alpha);
  }

  void sym_real_n(const Plane2<W>& A, const Plane2<W>& B, const Plane2<W>& C, hipStream_t st, int nadd,
                  const double* c, const Plane2<W>* const* add, double dg, double alpha) {
    GemmArgs<W> g = args(A, B, ld, ld, n, n, (int)np);
    set_out(g, C);
    g.epi.alpha = alpha;
    for (int k = 0; k < nadd; ++k) {
      g.epi.addc[k] = c[k];
      g.epi.add_r[k] = add[k]->r;
      g.epi.add_i[k] = add[k]->i;
    }
    g.epi.nadd = nadd;
    g.epi.diag = dg;
    if constexpr (std::is_same<W, double>::value) {
      if (oz_slices > 0 && (oz_scope & 1)) {
        const OzOp a{A.r, 1.0, nullptr, 0.0, (int)np, ld};
        const OzOp b{B.r, 1.0, nullptr, 0.0, (int)np, ld};
        flops += oz_product(a, A.r == B.r ? nullptr : &b, n, n, (int)np, g.epi, true, st);
        return;
      }
    }
    gemm<MODE_REAL, true>(g, st);
  }

  // General complex product C = A * B (B^T supplied as BT), 3M.
  void cplx_mul(const Plane2<W>& A, const Plane2<W>& BT, const Plane2<W>& C, hipStream_t st,
                std::initializer_list<std::pair<double, const Plane2<W>*>> add = {}, double dg = 0.0) {
    GemmArgs<W> g = args(A, BT, ld, ld, n, n, (int)np);
    set_out(g, C);
    int k = 0;
    for (auto& t : add) {
      g.epi.addc[k] = t.first;
      g.epi.add_r[k] = t.second->r;
      g.epi.add_i[k] = t.second->i;
      ++k;
    }
    g.epi.nadd = k;
    g.epi.diag = dg;
    gemm<MODE_C3M, false>(g, st);
  }

  // ---- Pade stage ----
  // Structured: A = i C, C real symmetric in a.r; A^{2k} = (-1)^k C^{2k}.
  // Produces V in v.r and U = i Uq with Uq in u.r.  The stage is split so
  // that the powers of C can be formed before the scaling is known:
  // pade_powers computes C^2, C^4, C^6 (C^8 for m = 9) of the prepared
  // plane, pade_assemble forms U and V at X = f C (f a power of two, so
  // every scaled coefficient b_j f^j and product scale f^k is exact).
  void pade_powers(int m, hipStream_t st) {
    ensure_stage_buffers(false);
    sym_real(a, a, p2, st);  // C2
    if (m >= 5) sym_real(p2, p2, p4, st);
    if (m >= 7) sym_real(p4, p2, p6, st);
    if (m == 9) sym_real(p6, p2, x1, st);  // C8 in x1
  }

  void pade_assemble(int m, double f, hipStream_t st) {
    ensure_stage_buffers(false);
    double b[14];
    for (int j = 0; j <= m; ++j) b[j] = pade_b(m, j);
    const double f2 = f * f, f4 = f2 * f2, f6 = f4 * f2, f8 = f4 * f4;
    if (m == 13) {
      // Z1 = b13 X6 - b11 X4 + b9 X2, Z2 = b12 X6 - b10 X4 + b8 X2
      lincomb(t1, false, {{b[13] * f6, &p6}, {-b[11] * f4, &p4}, {b[9] * f2, &p2}}, 0.0, st);
      lincomb(t2, false, {{b[12] * f6, &p6}, {-b[10] * f4, &p4}, {b[8] * f2, &p2}}, 0.0, st);
      // Wb = X6 Z1 - b7 X6 + b5 X4 - b3 X2 + b1 I   (into x0)
      sym_real(p6, t1, x0, st, {{-b[7] * f6, &p6}, {b[5] * f4, &p4}, {-b[3] * f2, &p2}}, b[1], f6);
      // V = X6 Z2 - b6 X6 + b4 X4 - b2 X2 + b0 I
      sym_real(p6, t2, v, st, {{-b[6] * f6, &p6}, {b[4] * f4, &p4}, {-b[2] * f2, &p2}}, b[0], f6);
      // Uq = X Wb
      sym_real(a, x0, u, st, {}, 0.0, f);
      return;
    }
    // odd = sum_k b_{2k+1} (-1)^k X^{2k}, even = sum_k b_{2k} (-1)^k X^{2k}
    if (m == 3) {
      lincomb(t1, false, {{-b[3] * f2, &p2}}, b[1], st);
      lincomb(v, false, {{-b[2] * f2, &p2}}, b[0], st);
    } else if (m == 5) {
      lincomb(t1, false, {{b[5] * f4, &p4}, {-b[3] * f2, &p2}}, b[1], st);
      lincomb(v, false, {{b[4] * f4, &p4}, {-b[2] * f2, &p2}}, b[0], st);
    } else if (m == 7) {
      lincomb(t1, false, {{-b[7] * f6, &p6}, {b[5] * f4, &p4}, {-b[3] * f2, &p2}}, b[1], st);
      lincomb(v, false, {{-b[6] * f6, &p6}, {b[4] * f4, &p4}, {-b[2] * f2, &p2}}, b[0], st);
    } else {  // 9
      lincomb(t1, false, {{b[9] * f8, &x1}, {-b[7] * f6, &p6}, {b[5] * f4, &p4}, {-b[3] * f2, &p2}}, b[1], st);
      lincomb(v, false, {{b[8] * f8, &x1}, {-b[6] * f6, &p6}, {b[4] * f4, &p4}, {-b[2] * f2, &p2}}, b[0], st);
    }
    sym_real(a, t1, u, st, {}, 0.0, f);
  }

  void pade_structured(int m, hipStream_t st) {
    ensure_stage_buffers(false);
    pade_powers(m, st);
    pade_assemble(m, 1.0, st);
  }

  // ---- Chebyshev stage (structured input, in place of Pade and solve) ----
  // exp(i X) for X = B theta / R, B real symmetric in a.r with spectral
  // radius at most R, from the polynomial of cheb_series(theta):
  //     cos X = sum_j cw[j] Y^j,  sin X = (B / R) sum_j sw[j] Y^j,  Y = C^2 / R^2,
  // C^2k the powers of B (C^2, C^4, C^6 in p2, p4, p6, C^8 in x1, as far as
  // the polynomial reads them).  Spectral plans: B the unscaled matrix,
  // R the spectral bound, theta = R / 2^s (spectral_front formed the
  // powers); 1-norm plans: B = X itself and R = theta = ||A - mu I||_1 / 2^s
  // (front_cheb).  Each polynomial in Y by Horner's rule in
  // Y^4 over blocks of four coefficients (Paterson-Stockmeyer): a block
  // c0 I + c1 Y + c2 Y^2 + c3 Y^3 is a combination of the powers, and a step
  // T <- Y^4 T + block one symmetric product (C^8 times T, scaled by R^-8)
  // with the block in its epilogue; a top block c I makes the first step
  // a combination too (c Y^4 + block).  The factors commute (all are
  // polynomials in B), so every product is symmetric.  No solve.  Returns
  // the planes holding cos X and sin X (real parts).
  std::pair<Plane2<W>*, Plane2<W>*> cheb_structured(double theta, double R, double tail, hipStream_t st) {
    ensure_stage_buffers(false);
    const ChebSeries cs = cheb_series(theta, tail);
    const Plane2<W>* pw[4] = {nullptr, &p2, &p4, &p6};
    const Plane2<W>& mul = x1;
    const double rinv = R > 0.0 ? 1.0 / R : 0.0;
    const double r2 = rinv * rinv;
    const double r8 = r2 * r2 * r2 * r2;
    KeepOperand keep(this, mul.r);  // the Horner steps share their first operand C^8
    auto series = [&](const double* a, int nc, Plane2<W>* cur, Plane2<W>* nxt) {
      double c[4];
      const Plane2<W>* pl[4];
      // block i: a[4i] I + sum_{0 < j < 4, 4i + j < nc} a[4i + j] R^-2j C^2j
      auto block = [&](int i) {
        int q = 0;
        double rj = 1.0;
        for (int j = 1; j < 4 && 4 * i + j < nc; ++j) {
          rj *= r2;
          c[q] = a[4 * i + j] * rj;
          pl[q++] = pw[j];
        }
        return q;
      };
      const int nb = (nc + 3) / 4;
      int i = nb - 1;
      if (nb > 1 && nc % 4 == 1) {
        // top block a[nc - 1] I: Y^4 (a[nc - 1] I) + block(nb - 2) needs no product
        --i;
        int q = block(i);
        c[q] = a[nc - 1] * r8;
        pl[q++] = &mul;
        lincomb_n(*cur, false, q, c, pl, a[4 * i], st);
      } else {
        const int q = block(i);
        lincomb_n(*cur, false, q, c, pl, a[4 * i], st);
      }
      for (--i; i >= 0; --i) {
        const int q = block(i);
        sym_real_n(mul, *cur, *nxt, st, q, c, pl, a[4 * i], r8);
        std::swap(cur, nxt);
      }
      return cur;
    };
    Plane2<W>* cx = series(cs.cw, cs.dc / 2 + 1, &t1, &t2);
    Plane2<W>* sp = series(cs.sw, (cs.ds - 1) / 2 + 1, &u, &v);
    Plane2<W>* sx = sp == &u ? &v : &u;
    sym_real(a, *sp, *sx, st, {}, 0.0, rinv);
    return {cx, sx};
  }

  // ---- spectral scaling (structured inputs) ----
  // A = i G with G real symmetric is normal, so the backward error of the
  // Pade approximant at X = i B / 2^s is that of its eigenvalues: the
  // Higham 2005 condition ||X|| <= theta_13 holds in the 2-norm as soon as
  // rho(B) / 2^s <= theta_13, and s may be chosen from any upper bound of
  // the spectral radius rho(B) instead of the 1-norm.  Here
  //     B = G - mu I + alpha J        (J the all-ones matrix),
  // where alpha != 0 deflates an exact eigenvector 1 of G (all row sums of
  // G equal to c): alpha = (mu - c) / n moves its eigenvalue c - mu to 0, B
  // commutes with J, and exp(A) = e^{i mu} (exp(i B) + (e^{i(c - mu)} - 1) J / n).
  // The bounds rho(B) <= ||B^k||_1^(1/k), k = 1, 2, 4, 6, and
  // rho(B) <= trace(B^2k)^(1/2k), 2k = 4, ..., 12, use the Pade powers,
  // computed here on the unscaled B and reused by pade_assemble with
  // scaled coefficients; with8 (Chebyshev plans) also forms C^8, which the
  // Chebyshev stage needs, and adds the tighter 2k = 14, 16.  For the Laplacian of a dense graph, whose
  // eigenvalue 0 lies far from the rest of the spectrum, this removes 3-4 of
  // 13 squarings at N = 10^4.  Returns the smallest bound (FP64 planes only).
  template <typename Tin>
  double spectral_front(const Tin* d_a, double mu_im, double alpha, bool with8, hipStream_t st) {
    ensure_stage_buffers(false);
    if constexpr (std::is_same<W, double>::value) {
      constexpr int kRows = 64;
      const int nch = (n + kRows - 1) / kRows;
      if (!npart)
        npart = dalloc<double>(std::max((size_t)std::max(1, (ncap + kRows - 1) / kRows) * std::max(ncap, 1),
                                        (size_t)kTraceDots * std::max(ncap, 1)));
      if (!d_norms) d_norms = dalloc<double>(kSpecNorms);
      if (!ncmax) ncmax = dalloc<double>((size_t)(ncap + 255) / 256 + 1);
      dim3 grid((n + 255) / 256, n);
      prepare_planar<Tin, W><<<grid, 256, 0, st>>>(d_a, n, a.r, a.i, ld, 0.0, mu_im, 1.0, 1, alpha);
      FX_LAUNCH();
      pade_powers(13, st);
      if (with8) sym_real(p4, p4, x1, st);  // C^8 in x1, for the Chebyshev stage and two more trace bounds
      const Plane2<W>* pw[4] = {&a, &p2, &p4, &p6};
      for (int k = 0; k < 4; ++k) {
        norm1_part<<<dim3((unsigned)((n + 255) / 256), (unsigned)nch), 256, 0, st>>>(pw[k]->r, ld, n, kRows, npart);
        FX_LAUNCH();
        const int ncb = (n + 255) / 256;
        norm1_cols<<<(unsigned)ncb, 256, 0, st>>>(npart, nch, n, ncmax);
        FX_LAUNCH();
        norm1_final<<<1, 256, 0, st>>>(ncmax, ncb, d_norms + k);
        FX_LAUNCH();
      }
      trace_dots_part<<<(unsigned)n, 256, 0, st>>>(p2.r, p4.r, p6.r, with8 ? x1.r : nullptr, ld, n, npart);
      FX_LAUNCH();
      trace_dots_final<<<1, 256, 0, st>>>(npart, n, with8 ? 1 : 0, d_norms + 4);
      FX_LAUNCH();
      double h[kSpecNorms];
      FX_HIP(hipMemcpyAsync(h, d_norms, sizeof(h), hipMemcpyDeviceToHost, st));
      FX_HIP(hipStreamSynchronize(st));
      return spectral_bound(h);
    } else {
      (void)d_a;
      (void)mu_im;
      (void)alpha;
      (void)with8;
      (void)st;
      throw Error(FEXPM_EINVAL, "spectral scaling runs on FP64 planes");
    }
  }

  // General complex A in a (planar).  Produces U in u, V in v.
  void pade_general(int m, hipStream_t st) {
    ensure_stage_buffers(true);
    double b[14];
    for (int j = 0; j <= m; ++j) b[j] = pade_b(m, j);
    transpose(a, t1, true, st);      // A^T
    cplx_mul(a, t1, p2, st);         // A2
    transpose(p2, t2, true, st);     // A2^T
    if (m >= 5) cplx_mul(p2, t2, p4, st);  // A4
    if (m >= 7) cplx_mul(p4, t2, p6, st);  // A6
    if (m == 13) {
      lincomb(t1, true, {{b[13], &p6}, {b[11], &p4}, {b[9], &p2}}, 0.0, st);
      transpose(t1, x0, true, st);
      cplx_mul(p6, x0, x1, st, {{b[7], &p6}, {b[5], &p4}, {b[3], &p2}}, b[1]);  // bracket of U
      lincomb(t1, true, {{b[12], &p6}, {b[10], &p4}, {b[8], &p2}}, 0.0, st);
      transpose(t1, x0, true, st);
      cplx_mul(p6, x0, v, st, {{b[6], &p6}, {b[4], &p4}, {b[2], &p2}}, b[0]);
      transpose(x1, x0, true, st);
      cplx_mul(a, x0, u, st);
      return;
    }
    if (m == 9) cplx_mul(p6, t2, x1, st);  // A8
    if (m == 3) {
      lincomb(t1, true, {{b[3], &p2}}, b[1], st);
      lincomb(v, true, {{b[2], &p2}}, b[0], st);
    } else if (m == 5) {
      lincomb(t1, true, {{b[5], &p4}, {b[3], &p2}}, b[1], st);
      lincomb(v, true, {{b[4], &p4}, {b[2], &p2}}, b[0], st);
    } else if (m == 7) {
      lincomb(t1, true, {{b[7], &p6}, {b[5], &p4}, {b[3], &p2}}, b[1], st);
      lincomb(v, true, {{b[6], &p6}, {b[4], &p4}, {b[2], &p2}}, b[0], st);
    } else {
      lincomb(t1, true, {{b[9], &x1}, {b[7], &p6}, {b[5], &p4}, {b[3], &p2}}, b[1], st);
      lincomb(v, true, {{b[8], &x1}, {b[6], &p6}, {b[4], &p4}, {b[2], &p2}}, b[0], st);
    }
    transpose(t1, x0, true, st);
    cplx_mul(a, x0, u, st);
  }

  // ---- FP64 solve of (V - U) X = (V + U) ----
  // Returns the index of the split-form diagonal buffer holding diag(X) and
  // leaves Y = X - diag(X) in x0.  With q64 > 0 (structured path of an FP32
  // engine) the first q64 squarings run in FP64 on the free LU buffer first.
  int solve(bool structured, int q64, hipStream_t st) {
    ensure_stage_buffers(!structured);
    ensure_system();
    FX_HIP(hipMemsetAsync(&scal->info, 0, sizeof(int), st));
    if (structured && spd_solve) return solve_spd(q64, st);
    const int inp = (int)np;
    {
      dim3 grid((unsigned)((2 * np + 255) / 256), (unsigned)np);
      form_system<W><<<grid, 256, 0, st>>>(u.r, u.i, v.r, v.i, ld, n, inp, lr, li, ldl, structured ? 1 : 0);
      FX_LAUNCH();
    }
    if (!structured && n <= GJ_NMAX) {
      // small orders: one workgroup eliminates the whole system in LDS
      gj_solve_small<<<1, 1024, 0, st>>>(lr, li, ldl, n, inp, scal);
      FX_LAUNCH();
    } else {
      flops += lu_factor_solve(lr, li, ldl, inp, tr, ti, ipiv, cand, gmax, top, scal, inv, ctrs, st,
                               panel_coop != 0);
    }
    if (q64 <= 0) return solution_tail(lr + inp, li + inp, ldl, lr, li, structured ? 1 : 0, 0, st);
    return solution_tail(lr + inp, li + inp, ldl, lr, li, 1, q64, st);
  }

  // Second stream and events of the overlapped Z updates (created on first use).
  hipStream_t zst = nullptr;
  std::vector<hipEvent_t> zev;
  hipStream_t z_stream(hipStream_t) {
    if (!zst) FX_HIP(hipStreamCreateWithFlags(&zst, hipStreamNonBlocking));
    return zst;
  }
  hipEvent_t z_event(int q) {
    while ((int)zev.size() <= q) {
      hipEvent_t e;
      FX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      zev.push_back(e);
    }
    return zev[(size_t)q];
  }

  // Structured solve by Cholesky (A = iC; V in v.r, U = i Uq with Uq in u.r).
  // M = V - i Uq is normal, S = M^H M = V^2 + Uq^2 is symmetric positive
  // definite, and with S = L L^T, Z = (V + i Uq) L^{-T}:
  //     X = M^{-1} M^H = (V - U)^{-1} (V + U) = Z Z^T.
  // Work area (leading dimension np, contiguous from lr): rows [0, np) Re Z,
  // [np, 2np) Im Z, [2np, 3np) S -> L, [3np, 4np) V V^T; X = Z Z^T is
  // written over S and V V^T.  Blocked right-looking Cholesky with 64-wide
  // panels; the diagonal block and its inverse come from one workgroup,
  // every other step is an FP64 MFMA product, and each panel also applies
  // L^{-T} to the columns of Z (the stacked real and imaginary parts form a
  // single real 2np x np right-hand side).
  int solve_spd(int q64, hipStream_t st) {
    const int64_t L = np;
    const size_t P = (size_t)np * (size_t)np;
    double* zb = lr;           // [Re Z; Im Z]
    double* sm = lr + 2 * P;   // S, then L
    double* vv = lr + 3 * P;   // V V^T
    double* dinv = tr;         // inverses of the diagonal blocks, 64 x 64 each
    const int ip = (int)np;
    const int kpad = (int)round_up(n, 16);
    {
      dim3 grid((unsigned)((np + 255) / 256), (unsigned)(2 * np));
      pack_stacked<W><<<grid, 256, 0, st>>>(v.r, u.r, ld, n, ip, zb, L);
      FX_LAUNCH();
    }
    // S = Uq Uq^T + V V^T (both symmetric: products of commuting polynomials),
    // over the padded order so that the padding rows and columns are zeros.
    // From int8 slices: one product of the concatenated rows [V Uq].
    if (oz_slices > 0 && (oz_scope & 2)) {
      Epilogue<double> e;
      std::memset(&e, 0, sizeof(e));
      e.cr = sm;
      e.ldc = L;
      e.ld_add = L;
      e.alpha = 1.0;
      const OzOp vu{zb, 1.0, zb + P, 1.0, kpad, L};
      flops += oz_product(vu, nullptr, ip, ip, 2 * kpad, e, true, st);
    } else {
    flops += dgemm_real(zb, L, zb, L, vv, L, ip, ip, kpad, 1.0, 0.0, true, st);
    {
      GemmArgs<double> g;
      std::memset(&g, 0, sizeof(g));
      g.ar = zb + P;
      g.lda = L;
      g.br = zb + P;
      g.ldb = L;
      g.m = ip;
      g.n = ip;
      g.k = kpad;
      g.epi.cr = sm;
      g.epi.ldc = L;
      g.epi.ld_add = L;
      g.epi.alpha = 1.0;
      g.epi.nadd = 1;
      g.epi.addc[0] = 1.0;
      g.epi.add_r[0] = vv;
      flops += gemm64<MODE_REAL, true>(g, st);
    }
    }
    // Two levels: 64-wide inner panels update only the columns of their
    // ch_outer-wide outer block; the outer block then updates the trailing
    // S (half tiles) and the remaining columns of Z with K = ch_outer.
    // With z_overlap the Z updates run on a second stream: they read only
    // finished columns of L (the panel's own, after its panel product), so
    // each waits for the event recorded after that panel product, and the
    // latency-bound factorisation of S proceeds meanwhile.
    const bool ov = z_overlap != 0;
    hipStream_t zs = ov ? z_stream(st) : st;
    int nev = 0;
    for (int K0 = 0; K0 < ip; K0 += ch_outer) {
      const int K1 = std::min(ip, K0 + ch_outer);
      if (z_left && K0 > 0) {
        // left-looking Z: bring the block's columns up to date with every
        // finished column at once (K = K0), instead of the right-looking outer
        // update after each block; zs is already ordered after every earlier panel
        flops += dgemm_real(zb, L, sm + (int64_t)K0 * L, L, zb + K0, L, 2 * ip, K1 - K0, K0, -1.0, 1.0, false, zs);
      }
      for (int k0 = K0; k0 < K1; k0 += LU_NB) {
        const int k1 = k0 + LU_NB;
        const int rest = ip - k1;
        double* dk = dinv + (size_t)(k0 / LU_NB) * LU_NB * LU_NB;
        chol_diag<<<1, 256, 0, st>>>(sm, L, k0, n, dk, scal);
        FX_LAUNCH();
        double* panel = sm + (int64_t)k1 * L + k0;  // L[k1:, k0:k1]
        if (rest > 0)
          flops += dgemm_real(panel, L, dk, LU_NB, panel, L, rest, LU_NB, LU_NB, 1.0, 0.0, false, st);
        if (ov) {
          hipEvent_t e = z_event(nev++);
          FX_HIP(hipEventRecord(e, st));
          FX_HIP(hipStreamWaitEvent(zs, e, 0));
        }
        // Z[:, k0:k1] <- Z[:, k0:k1] L_kk^{-T}
        flops += dgemm_real(zb + k0, L, dk, LU_NB, zb + k0, L, 2 * ip, LU_NB, LU_NB, 1.0, 0.0, false, zs);
        if (k1 < K1) {
          // S[k1:, k1:K1] -= L[k1:, k0:k1] L[k1:K1, k0:k1]^T;  Z[:, k1:K1] -= Z[:, k0:k1] L[k1:K1, k0:k1]^T
          flops += dgemm_real(panel, L, panel, L, sm + (int64_t)k1 * L + k1, L, rest, K1 - k1, LU_NB, -1.0, 1.0,
                              false, st);
          flops += dgemm_real(zb + k0, L, panel, L, zb + k1, L, 2 * ip, K1 - k1, LU_NB, -1.0, 1.0, false, zs);
        }
      }
      if (K1 < ip) {
        const int r2 = ip - K1, kk = K1 - K0;
        double* p2 = sm + (int64_t)K1 * L + K0;  // L[K1:, K0:K1]
        flops += dgemm_real(p2, L, p2, L, sm + (int64_t)K1 * L + K1, L, r2, r2, kk, -1.0, 1.0, true, st);
        if (!z_left) flops += dgemm_real(zb + K0, L, p2, L, zb + K1, L, 2 * ip, r2, kk, -1.0, 1.0, false, zs);
      }
    }
    if (ov) {  // X = Z Z^T needs every Z update
      hipEvent_t e = z_event(nev++);
      FX_HIP(hipEventRecord(e, zs));
      FX_HIP(hipStreamWaitEvent(st, e, 0));
    }
    // X = Z Z^T over S and V V^T, from two real symmetric products: with
    // Zr = V L^{-T}, Zi = Uq L^{-T} and V, Uq, S commuting,
    //     Zr Zr^T + Zi Zi^T = (V^2 + Uq^2) S^{-1} = I,
    // so Re X = Zr Zr^T - Zi Zi^T = 2 Zr Zr^T - I and
    //    Im X = Zr Zi^T + Zi Zr^T = 2 Zi Zr^T (= 2 Uq S^{-1} V, symmetric),
    // both to FP64 rounding (the identity holds as far as L L^T = S does).
    // Two thirds of the flops of the 3M complex product.
    double* xr = lr + 2 * P;
    double* xi = lr + 3 * P;
    if (oz_slices > 0 && (oz_scope & 2)) {
      Epilogue<double> e;
      std::memset(&e, 0, sizeof(e));
      e.ldc = L;
      e.ld_add = L;
      e.alpha = 2.0;
      const OzOp zr{zb, 1.0, nullptr, 0.0, kpad, L};
      const OzOp zi{zb + P, 1.0, nullptr, 0.0, kpad, L};
      e.cr = xr;
      e.diag = -1.0;
      flops += oz_product(zr, nullptr, n, n, kpad, e, true, st);
      e.cr = xi;
      e.diag = 0.0;
      flops += oz_product(zi, &zr, n, n, kpad, e, true, st);
    } else {
      GemmArgs<double> g;
      std::memset(&g, 0, sizeof(g));
      g.ar = zb;
      g.lda = L;
      g.br = zb;
      g.ldb = L;
      g.m = n;
      g.n = n;
      g.k = kpad;
      g.epi.cr = xr;
      g.epi.ldc = L;
      g.epi.alpha = 2.0;
      g.epi.diag = -1.0;
      flops += gemm64<MODE_REAL, true>(g, st);
      g.ar = zb + P;
      g.epi.cr = xi;
      g.epi.diag = 0.0;
      flops += gemm64<MODE_REAL, true>(g, st);
    }
    return solution_tail(xr, xi, L, zb, zb + P, 0, q64, st);
  }

  // X (FP64, xr/xi with leading dimension ldx) -> split form for the
  // squarings: Y = X - diag(X) in x0, diag(X) in dsq[0] (index 0 returned).
  // With q64 > 0 the first q64 squarings run in FP64 first: X is split into
  // the free buffers wr/wi (same leading dimension), which ping-pong with
  // xr/xi; the returned index names the diagonal buffer of the last result.
  int solution_tail(double* xr, double* xi, int64_t ldx, double* wr, double* wi, int symmetrize, int q64,
                    hipStream_t st) {
    if (q64 <= 0) {
      dim3 grid((n + 31) / 32, (n + 31) / 32);
      solution_to_work<W><<<grid, 256, 0, st>>>(xr, xi, ldx, x0.r, x0.i, ld, n, symmetrize, dsq[0]);
      FX_LAUNCH();
      return 0;
    }
    const int64_t ldl = ldx;
    const int kpad = (int)round_up(n, 16);
    {
      dim3 grid((kpad + 31) / 32, (n + 31) / 32);
      split_fp64<<<grid, 256, 0, st>>>(xr, xi, ldl, wr, wi, n, kpad, symmetrize, dsq[0]);
      FX_LAUNCH();
    }
    double* cr = wr;
    double* ci = wi;
    double* nr = xr;
    double* ni = xi;
    int dc = 0;
    for (int q = 0; q < q64; ++q) {
      GemmArgs<double> g;
      std::memset(&g, 0, sizeof(g));
      g.ar = cr; g.ai = ci; g.lda = ldl;
      g.br = cr; g.bi = ci; g.ldb = ldl;
      g.m = n; g.n = n; g.k = kpad;
      g.epi.cr = nr; g.epi.ci = ni; g.epi.ldc = ldl; g.epi.alpha = 1.0;
      g.epi.d_in = dsq[dc]; g.epi.d_out = dsq[dc ^ 1];
      using C = GemmCfg<double>;
      constexpr int NT = (C::BM / C::WM) * (C::BN / C::WN) * 64;
      const int T = (n + C::BM - 1) / C::BM;
      gemm_nt_kernel<double, MODE_C3M, true><<<dim3(T * (T + 1) / 2), dim3(NT), 0, st>>>(g);
      FX_LAUNCH();
      flops += 3.0 * 2.0 * (T * (T + 1) / 2) * C::BM * C::BN * (double)kpad;
      std::swap(cr, nr);
      std::swap(ci, ni);
      dc ^= 1;
    }
    {
      dim3 grid((n + 255) / 256, n);
      fp64_to_work<W><<<grid, 256, 0, st>>>(cr, ci, ldl, x0.r, x0.i, ld, n);
      FX_LAUNCH();
    }
    return dc;
  }

  // ---- squarings ----
  // X = diag(d) + Y on entry (x0 holds Y, dsq[0] holds d); each squaring
  // keeps the split form; the diagonal is merged back at the end.
  // s squarings of the split form (Y in x0, d in dsq[dc]); returns the plane
  // holding the last Y and the index of its diagonal buffer.
  // Split-precision squarings (FP32 engines, structured path): each step
  // splits the current Y into 16-bit parts and squares on 16-bit MFMA.
  // Grouped lower-triangular order of T x T tiles of SQ_BM (sym_tile_order,
  // groups of SQ_GROUP), built once per tile count; the copy waits for the
  // host buffer.
  const int* tile_order(int T, hipStream_t st) {
    int*& o = sq_orders[T];
    if (!o) {
      std::vector<int> h((size_t)T * (T + 1) / 2);
      sym_tile_order(T, SQ_GROUP, h.data());
      o = static_cast<int*>(dmalloc_bytes(h.size() * sizeof(int)));
      FX_HIP(hipMemcpyAsync(o, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, st));
      FX_HIP(hipStreamSynchronize(st));
    }
    return o;
  }

  template <class P>
  std::pair<Plane2<W>*, int> square_split16(int s, int dc, hipStream_t st) {
    Plane2<W>* cur = &x0;
    Plane2<W>* nxt = &x1;
    const int rows = (int)np;
    const int kpad = (int)round_up(n, 16);
    const size_t need = (size_t)npcap * (size_t)round_up(ncap, 16) * 6 * 2;  // room for either split
    if (!sbuf || sbuf_bytes < need) {
      dfree(sbuf);
      sbuf = nullptr;
      sbuf = dmalloc_bytes(need);
      sbuf_bytes = need;
    }
    const int T = (n + SQ_BM - 1) / SQ_BM;
    const int nb = T * (T + 1) / 2;
    const int* sq_order = tile_order(T, st);
    for (int q = 0; q < s; ++q) {
      dim3 sg((unsigned)((kpad / 16 + 15) / 16), (unsigned)((rows + 15) / 16));
      split_blocked<P><<<sg, 256, 0, st>>>((const float*)cur->r, (const float*)cur->i, ld, n, n, rows, kpad,
                                          (typename P::H*)sbuf, &scal->info);
      FX_LAUNCH();
      SqArgs<P> a;
      a.a = (const typename P::H*)sbuf;
      a.rows = rows;
      a.yr = (const float*)cur->r;
      a.yi = (const float*)cur->i;
      a.ldy = ld;
      a.cr = (float*)nxt->r;
      a.ci = (float*)nxt->i;
      a.ldc = ld;
      a.d_in = dsq[dc];
      a.d_out = dsq[dc ^ 1];
      a.m = n;
      a.k = kpad;
      a.kblock = kblock > 0 ? kblock : (1 << 30);
      a.order = sq_order;
      kernel_mark(q, false, st);
      sq_split_kernel<P><<<dim3(nb), dim3(sq_threads<P>()), 0, st>>>(a);
      FX_LAUNCH();
      kernel_mark(q, true, st);
      flops += 4.0 * P::NPAIR * 2.0 * nb * (double)SQ_BM * SQ_BM * (double)kpad;
      std::swap(cur, nxt);
      dc ^= 1;
    }
    return {cur, dc};
  }

  // s structured squarings by exact int8 residue products (fexpm_crt.hpp):
  // per squaring the row exponents of Y, the operand residues, one batched
  // launch of the 3 CRT_K residue products over the lower 256 x 256 tiles,
  // and the reconstruction with the diagonal-split epilogue into the other
  // plane.  FP32 engines, n <= CRT_KMAX.
  std::pair<Plane2<W>*, int> square_crt(int s, int dc, hipStream_t st) {
    if constexpr (!std::is_same<W, float>::value) {
      throw Error(FEXPM_EINVAL, "residue squarings apply to FP32 planes");
    } else {
      const int rows = (int)round_up(n, CRT_TILE), kpad = (int)round_up(n, CRT_KPAD);
      const int64_t rows_cap = round_up(ncap, CRT_TILE), kpad_cap = round_up(ncap, CRT_KPAD);
      if (!crt_e) crt_e = dalloc<int>((size_t)npcap);
      const int K = crt_moduli;
      const int nprob = 3 * K;
      // operand residues then product residues of the 3 K products, the
      // buffer sized once for the engine's largest order
      const size_t ops_bytes = (size_t)nprob * rows * kpad;
      uint8_t* base = static_cast<uint8_t*>(crt_buf->get(
          std::max((size_t)nprob * rows_cap * (kpad_cap + rows_cap), ops_bytes + (size_t)nprob * rows * rows)));
      int8_t* crt_ops = reinterpret_cast<int8_t*>(base);
      uint8_t* crt_res = base + ops_bytes;
      const CrtTable& tab = crt_table(K);
      const int beta = crt_beta(tab, n);
      last_crt_beta = beta;
      const int64_t op_plane = (int64_t)rows * kpad, res_plane = (int64_t)rows * rows;
      CrtProbs probs;
      std::memset(&probs, 0, sizeof(probs));
      for (int p = 0; p < nprob; ++p) {
        // problem p: operand p / K (yr', yi', s') modulo moduli p % K; C = X X^T
        CrtProb& q = probs.p[p];
        q.a = crt_ops + p * op_plane;
        q.b = q.a;
        q.out = crt_res + p * res_plane;
        q.m = tab.m[p % K];
        q.magic = tab.magic[p % K];
      }
      const int G = rows / 32, KB = kpad / 32, T = rows / CRT_TILE, tiles_per = T * (T + 1) / 2;
      const int T64 = (n + 63) / 64;
      const int64_t jobs = (int64_t)G * KB;
      Plane2<W>* cur = &x0;
      Plane2<W>* nxt = &x1;
      for (int q = 0; q < s; ++q) {
        crt_row_exp<<<(n + 3) / 4, 256, 0, st>>>(cur->r, cur->i, ld, n, crt_e);
        FX_LAUNCH();
        const unsigned rblocks = (unsigned)((jobs + 3) / 4), cblocks = (unsigned)(T64 * (T64 + 1) / 2);
        if (K == 7) crt_residues_sq<7><<<rblocks, 256, 0, st>>>(cur->r, cur->i, ld, n, crt_e, rows, kpad, beta, tab, crt_ops);
        else if (K == 8) crt_residues_sq<8><<<rblocks, 256, 0, st>>>(cur->r, cur->i, ld, n, crt_e, rows, kpad, beta, tab, crt_ops);
        else crt_residues_sq<9><<<rblocks, 256, 0, st>>>(cur->r, cur->i, ld, n, crt_e, rows, kpad, beta, tab, crt_ops);
        FX_LAUNCH();
        kernel_mark(q, false, st);
        crt_gemm_kernel<<<crt_launch_blocks(nprob, tiles_per, crt_map), 256, 0, st>>>(probs, tiles_per, T, G, KB,
                                                                                     crt_map);
        FX_LAUNCH();
        kernel_mark(q, true, st);
        if (K == 7)
          crt_combine_sq<7><<<cblocks, 256, 0, st>>>(crt_res, rows, crt_e, cur->r, cur->i, ld, n, dsq[dc], dsq[dc ^ 1],
                                                      nxt->r, nxt->i, beta, tab);
        else if (K == 8)
          crt_combine_sq<8><<<cblocks, 256, 0, st>>>(crt_res, rows, crt_e, cur->r, cur->i, ld, n, dsq[dc], dsq[dc ^ 1],
                                                      nxt->r, nxt->i, beta, tab);
        else
          crt_combine_sq<9><<<cblocks, 256, 0, st>>>(crt_res, rows, crt_e, cur->r, cur->i, ld, n, dsq[dc], dsq[dc ^ 1],
                                                      nxt->r, nxt->i, beta, tab);
        FX_LAUNCH();
        flops += 2.0 * nprob * (double)tiles_per * CRT_TILE * CRT_TILE * (double)kpad;  // int8 operations
        std::swap(cur, nxt);
        dc ^= 1;
      }
      return {cur, dc};
    }
  }

  // FP32 structured squaring in 128 x 128 tiles.  When the tile count
  // leaves the last wave of workgroups less than half full, the tiles of
  // the full waves run in 128 x 128 and the rest as 64 x 64 sub-tiles
  // (gemm_sym_tail_kernel), which spread over the whole device.
  template <int MODE>
  void square_big_tail(const GemmArgs<W>& g, hipStream_t st) {
    if constexpr (std::is_same<W, float>::value) {
      using C = FCfg16w;
      const SquarePlan pl = square_plan(g.m, 128, tail_split != 0);
      if (pl.rest > 0) {
        gemm_nt_kernel<float, MODE, true, true, C><<<dim3(pl.full), dim3(C::NT), 0, st>>>(g);
        FX_LAUNCH();
        gemm_sym_tail_kernel<float, MODE, true, FCfgSmall><<<dim3(4 * pl.rest), dim3(FCfgSmall::NT), 0, st>>>(
            g, pl.full);
        FX_LAUNCH();
        const double mult = MODE == MODE_SQ2 ? 2.0 : 3.0;
        flops += mult * 2.0 *
                 ((double)pl.full * C::BM * C::BN + (double)pl.sub * FCfgSmall::BM * FCfgSmall::BN) * (double)g.k;
        return;
      }
      gemm<MODE, true, true, C>(g, st);
    } else {
      gemm<MODE, true, true>(g, st);
    }
  }

  // Form of the structured squarings for the current order: the precision
  // option resolved (FEXPM_PREC_AUTO: residue products for CRT_NMIN <= n <=
  // CRT_KMAX, FP32 MFMA otherwise; residue products need n <= CRT_KMAX).
  int square_mode() const {
    if (!std::is_same<W, float>::value) return FEXPM_PREC_NATIVE;
    if (precision == FEXPM_PREC_AUTO)
      return (n >= CRT_NMIN && n <= CRT_KMAX) ? FEXPM_PREC_INT8_CRT : FEXPM_PREC_NATIVE;
    if (precision == FEXPM_PREC_INT8_CRT && n > CRT_KMAX) return FEXPM_PREC_NATIVE;
    return precision;
  }

  // s squarings of the split form; with structured input the first `two`
  // of them in the two-product form (MODE_SQ2).
  std::pair<Plane2<W>*, int> square_split(int s, bool structured, int dc, hipStream_t st, int two = 0) {
    const int mode = square_mode();
    if (structured && mode == FEXPM_PREC_FP16X2) return square_split16<Fp16x2>(s, dc, st);
    if (structured && mode == FEXPM_PREC_BF16X3) return square_split16<Bf16x3>(s, dc, st);
    if (structured && mode == FEXPM_PREC_INT8_CRT) return square_crt(s, dc, st);
    Plane2<W>* cur = &x0;
    Plane2<W>* nxt = &x1;
    for (int q = 0; q < s; ++q) {
      const bool sq2 = structured && q < two;
      GemmArgs<W> g;
      if (structured) {
        // k over round_up(n, 16): the padding columns of X are zero, so the
        // products beyond it add nothing (same values, 1% fewer MFMAs at n = 10^4)
        g = args(*cur, *cur, ld, ld, n, n, (int)round_up(n, 16));
      } else {
        transpose(*cur, t1, true, st);
        g = args(*cur, t1, ld, ld, n, n, (int)np);
      }
      set_out(g, *nxt);
      g.epi.d_in = dsq[dc];
      g.epi.d_out = dsq[dc ^ 1];
      g.kblock = kblock;
      const bool blk = kblock > 0 && std::is_same<W, float>::value;
      const bool small = std::is_same<W, float>::value && square_tile(n, sq_tile) == 64;
      // (FCfgSmall runs 32x32x2 FP32 MFMAs; FP64 engines never take this branch)
      using SmallC = typename std::conditional<std::is_same<W, float>::value, FCfgSmall, GemmCfg<W>>::type;
      kernel_mark(q, false, st);
      if (structured && sq2) {
        if (blk && small) gemm<MODE_SQ2, true, true, SmallC>(g, st);
        else if (blk) square_big_tail<MODE_SQ2>(g, st);
        else gemm<MODE_SQ2, true>(g, st);
      } else if (structured) {
        if (blk && small) gemm<MODE_C3M, true, true, SmallC>(g, st);
        else if (blk) square_big_tail<MODE_C3M>(g, st);
        else gemm<MODE_C3M, true>(g, st);
      } else {
        if (blk && small) gemm<MODE_C3M, false, true, SmallC>(g, st);
        else if (blk) gemm<MODE_C3M, false, true>(g, st);
        else gemm<MODE_C3M, false>(g, st);
      }
      kernel_mark(q, true, st);
      std::swap(cur, nxt);
      dc ^= 1;
    }
    return {cur, dc};
  }

  Plane2<W>* square(int s, bool structured, int dc, hipStream_t st, int two = 0) {
    auto r = square_split(s, structured, dc, st, two);
    merge_diagonal<W><<<(n + 255) / 256, 256, 0, st>>>(r.first->r, r.first->i, ld, n, dsq[r.second]);
    FX_LAUNCH();
    return r.first;
  }

  // Stages 3-4 in this pipeline's precision: scaled shifted planar A and the
  // Pade U, V.
  template <typename Tin>
  void front_pade(const Tin* d_a, bool structured, int m, double mu_re, double mu_im, double scale,
                  hipStream_t st) {
    ensure_stage_buffers(!structured);
    dim3 grid((n + 255) / 256, n);
    prepare_planar<Tin, W><<<grid, 256, 0, st>>>(d_a, n, a.r, a.i, ld, mu_re, mu_im, scale, structured ? 1 : 0);
    FX_LAUNCH();
    if (structured)
      pade_structured(m, st);
    else
      pade_general(m, st);
  }

  // Chebyshev stage of a 1-norm plan (structured input): the scaled, shifted
  // X = Im(A - mu I) / 2^s and the powers X^2, X^4, X^6, X^8 that the
  // polynomial of cheb_series(theta) reads (p2, p4, p6, x1).
  template <typename Tin>
  void front_cheb(const Tin* d_a, double mu_re, double mu_im, double scale, double theta, double tail,
                  hipStream_t st) {
    ensure_stage_buffers(false);
    dim3 grid((n + 255) / 256, n);
    prepare_planar<Tin, W><<<grid, 256, 0, st>>>(d_a, n, a.r, a.i, ld, mu_re, mu_im, scale, 1);
    FX_LAUNCH();
    const ChebSeries cs = cheb_series(theta, tail);
    const int nc = cs.dc / 2 + 1, ns = (cs.ds - 1) / 2 + 1;
    const int need = (nc > 4 || ns > 4) ? 4 : std::min(3, std::max(nc, ns) - 1);
    if (need >= 1) sym_real(a, a, p2, st);
    if (need >= 2) sym_real(p2, p2, p4, st);
    if (need >= 3) sym_real(p4, p2, p6, st);
    if (need >= 4) sym_real(p4, p4, x1, st);
  }

  // FP64 companion of an FP32 engine: Pade, solve and leading squarings in
  // FP64 (allocated on first use).
  std::unique_ptr<Pipeline<double>> hi;
  Pipeline<double>* companion() {
    if (!hi) {
      hi.reset(new Pipeline<double>(ncap, FEXPM_COMPLEX128, events));
      hi->crt_buf = crt_buf;  // its residue products run before this pipeline's squarings
      hi->panel_coop = panel_coop;
      hi->set_size(n, nullptr);
      FX_HIP(hipDeviceSynchronize());
    }
    return hi.get();
  }
  Scalars* companion_scal() { return hi ? hi->scal : scal; }

  // ---- one run = analysis, host plan, enqueued stages ----
  template <typename Tin>
  void analyze_enqueue(const Tin* d_a, hipStream_t st, Scalars* slot) {
    const int nt = (n + 31) / 32;
    FX_HIP(hipMemsetAsync(flags, 0, sizeof(int), st));
    analyze_tiles<Tin><<<nt * (nt + 1) / 2, 256, 0, st>>>(d_a, n, nt, colpart, diag, flags, colsig);
    FX_LAUNCH();
    reduce_colpart<<<(n + 255) / 256, 256, 0, st>>>(colpart, n, nt, off, colsig, rsum);
    FX_LAUNCH();
    finalize_norms<<<1, 1024, 0, st>>>(off, diag, n, flags, scal, rsum);
    FX_LAUNCH();
    if (slot) FX_HIP(hipMemcpyAsync(slot, scal, sizeof(Scalars), hipMemcpyDeviceToDevice, st));
  }

  struct Plan {
    bool nonfinite, structured, shifted, hybrid, split16;
    bool spectral;  // s from the spectral-radius bound (powers formed before the scaling)
    int two;        // leading squarings in the two-product form
    bool deflate;   // exact eigenvector 1 deflated (all row sums of Im(A) equal)
    double mu_re, mu_im, norm, norm_plain;
    double alpha;   // deflation: B = G - mu I + alpha J
    double c;       // deflation: common row sum of G
    double rho;     // spectral-radius bound of B (spectral plans)
    int m, s, s_do, q64;
    int s_onenorm;  // s of the 1-norm rule
    bool cheb;      // Chebyshev polynomial in place of Pade and solve (structured, FP64 stage)
  };

  // spectral_ok: the caller can wait for the spectral bound on the host
  // (single runs; batch lanes plan without host waits).
  Plan make_plan(const Scalars& hs, int opt_structure, bool spectral_ok) const {
    Plan P;
    std::memset(&P, 0, sizeof(P));
    const bool is_complex = (dtype == FEXPM_COMPLEX64 || dtype == FEXPM_COMPLEX128);
    P.norm_plain = hs.norm_plain;
    P.nonfinite = (hs.flags & FLAG_NONFINITE) || !std::isfinite(hs.norm_plain);
    if (P.nonfinite) {
      P.norm = hs.norm_plain;
      return P;
    }
    P.structured = is_complex && opt_structure && !(hs.flags & (FLAG_NONZERO_RE | FLAG_ASYM));
    P.shifted = hs.norm_shift <= hs.norm_plain;
    P.mu_re = P.shifted ? hs.mu_re : 0.0;
    P.mu_im = P.shifted ? hs.mu_im : 0.0;
    P.norm = P.shifted ? hs.norm_shift : hs.norm_plain;
    choose_degree(P.norm, P.m, P.s);
    P.s_onenorm = P.s;
    P.hybrid = std::is_same<W, float>::value && pade_fp64;
    P.split16 = P.structured && square_mode() != FEXPM_PREC_NATIVE;
    // spectral scaling: structured, degree 13 with squarings, FP64 Pade stage
    P.spectral = spectral_ok && spectral && P.structured && P.m == 13 && P.s > 0 && P.s < 120 &&
                 (P.hybrid || std::is_same<W, double>::value);
    P.deflate = P.spectral && deflate && hs.rowsum_equal && std::isfinite(hs.rowsum_c);
    P.cheb = cheb && P.structured && (P.hybrid || std::is_same<W, double>::value);
    P.c = P.deflate ? hs.rowsum_c : 0.0;
    P.alpha = P.deflate ? (P.mu_im - P.c) / (double)n : 0.0;
    P.rho = 0.0;
    finish_plan(P);
    return P;
  }

  // Stages whose symmetric FP64 products come from int8 slices (bit 0 the
  // Pade products, bit 1 the Cholesky solve's S and X = Z Z^T).  Automatic
  // (oz_scope = 0): the Pade products always; the solve's only for deflated
  // plans of order n >= 4096.  The slices keep about 41 bits of each row's
  // largest value; a Pade term with a dominant rank-one part (the outlier
  // eigenvalue of a Laplacian that is not deflated) loses its small entries
  // to that, which the factorisation amplifies (N = 10^4 weighted ER(0.5)
  // Laplacian, t = 4.5: infidelity 1.5e-11 -> 6.1e-9), while deflated plans
  // keep the FP64 accuracy (t = 4.5: 1.78e-11 -> 1.77e-11, t = 20:
  // 7.7e-11 -> 8.0e-11) and the solve takes 121 -> 87 ms.  Below n = 4096
  // the FP64 MFMA products are the faster ones.
  int slice_scope(const Plan& P) const {
    if (oz_scope != 0) return oz_scope;
    return 1 | ((P.deflate && n >= 4096) ? 2 : 0);
  }

  // Chebyshev plans: theta bounds the spectral radius of X = B / 2^s (the
  // spectral bound, else the 1-norm of the shifted input), and R bounds that
  // of the planes whose powers the stage uses (the unscaled B of spectral
  // plans, the scaled X otherwise).
  static double cheb_theta(const Plan& P) { return (P.spectral ? P.rho : P.norm) * std::ldexp(1.0, -P.s); }
  static double cheb_radius(const Plan& P) { return P.spectral ? P.rho : cheb_theta(P); }

  // Fields that follow from s.
  void finish_plan(Plan& P) const {
    P.s_do = max_squarings >= 0 ? std::min(P.s, max_squarings) : P.s;
    P.q64 = (std::is_same<W, float>::value && (P.hybrid || P.structured)) ? std::min(P.s_do, fp64_squarings) : 0;
    // two-product squarings: structured, not split precision, never the last
    // squaring of the plan (s counts the full plan, so max_squarings cuts
    // never change which squarings use it), and none of the FP64 leading ones
    P.two = (P.structured && !P.split16 && P.q64 == 0) ? std::max(0, std::min(sq2_lead, P.s - 1)) : 0;
    P.two = std::min(P.two, P.s_do);
  }

  // The FP64 pipeline of the Pade stage: this one, or the companion of an FP32 engine.
  Pipeline<double>* fp64_pipe() {
    if constexpr (std::is_same<W, double>::value) return this;
    else return companion();
  }

  // Spectral plans: the powers of the unscaled (deflated) B on the FP64
  // pipeline, their norms read back, and s from the spectral bound.
  template <typename Tin>
  void spectral_plan(const Tin* d_a, Plan& P, hipStream_t st) {
    Pipeline<double>* h = fp64_pipe();
    h->oz_slices = oz_slices;
    h->oz_scope = slice_scope(P);
    h->crt_map = crt_map;
    h->crt64_moduli = crt64_for(P.spectral);
    P.rho = h->template spectral_front<Tin>(d_a, P.mu_im, P.alpha, cheb != 0, st);
    P.s = scaling_for(P.rho);
    finish_plan(P);
  }

  // Stages after the analysis, enqueued on st without host waits (ev[1..5]
  // recorded when `timing`).  With info_slot, the exception bits of this run
  // (singular pivot, barrier time-out, 16-bit range) are copied there.
  template <typename Tin, typename Tout>
  void enqueue_rest(const Tin* d_a, Tout* d_u, const Plan& P, hipStream_t st, bool timing, int* info_slot) {
    flops = 0.0;
    have_result = false;
    if (P.nonfinite) {
      const int64_t cnt = (int64_t)n * n;
      fill_nan<Tout><<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(d_u, cnt);
      FX_LAUNCH();
      if (info_slot) FX_HIP(hipMemsetAsync(info_slot, 0, sizeof(int), st));
      return;
    }
    FX_HIP(hipMemsetAsync(&scal->info, 0, sizeof(int), st));
    const double scale = std::ldexp(1.0, -P.s);
    int dc = 0;
    Pipeline<double>* h = nullptr;
    if (P.hybrid) {
      // FP64 Pade + solve (+ leading squarings), then FP32 squarings
      h = companion();
      h->flops = 0.0;
      h->panel_coop = panel_coop;
      h->spd_solve = spd_solve;
      h->oz_slices = oz_slices;
      h->oz_scope = slice_scope(P);
      h->z_overlap = z_overlap;
      h->ch_outer = ch_outer;
      h->z_left = z_left;
      h->crt_map = crt_map;
      h->crt64_moduli = crt64_for(P.spectral);
      int dh;
      if (P.cheb) {
        // Chebyshev polynomial from the powers formed by spectral_plan (or
        // here, for 1-norm plans); no solve
        const double tail = cheb_tail(P.s, std::is_same<W, float>::value);
        if (!P.spectral) h->front_cheb<Tin>(d_a, P.mu_re, P.mu_im, scale, cheb_theta(P), tail, st);
        auto cs = h->cheb_structured(cheb_theta(P), cheb_radius(P), tail, st);
        if (timing) FX_HIP(hipEventRecord(ev[2], st));
        FX_HIP(hipMemsetAsync(&h->scal->info, 0, sizeof(int), st));
        dh = h->solution_tail(cs.first->r, cs.second->r, h->ld, nullptr, nullptr, 0, 0, st);
      } else {
        if (P.spectral)
          h->pade_assemble(P.m, scale, st);  // powers formed by spectral_plan
        else
          h->front_pade<Tin>(d_a, P.structured, P.m, P.mu_re, P.mu_im, scale, st);
        if (timing) FX_HIP(hipEventRecord(ev[2], st));
        dh = h->solve(P.structured, 0, st);
      }
      auto r = h->square_split(P.q64, P.structured, dh, st);
      dim3 grid((n + 255) / 256, n);
      fp64_to_work<W><<<grid, 256, 0, st>>>(r.first->r, r.first->i, h->ld, x0.r, x0.i, ld, n);
      FX_LAUNCH();
      FX_HIP(hipMemcpyAsync(dsq[0], h->dsq[r.second], (size_t)n * sizeof(double2), hipMemcpyDeviceToDevice, st));
      flops += h->flops;
    } else {
      bool done = false;
      if constexpr (std::is_same<W, double>::value) {
        if (P.cheb) {
          const double tail = cheb_tail(P.s, std::is_same<W, float>::value);
          if (!P.spectral) front_cheb<Tin>(d_a, P.mu_re, P.mu_im, scale, cheb_theta(P), tail, st);
          auto cs = cheb_structured(cheb_theta(P), cheb_radius(P), tail, st);
          if (timing) FX_HIP(hipEventRecord(ev[2], st));
          dc = solution_tail(cs.first->r, cs.second->r, ld, nullptr, nullptr, 0, 0, st);
          done = true;
        }
      }
      if (!done) {
        if (P.spectral)
          pade_assemble(P.m, scale, st);
        else
          front_pade<Tin>(d_a, P.structured, P.m, P.mu_re, P.mu_im, scale, st);
        if (timing) FX_HIP(hipEventRecord(ev[2], st));
        dc = solve(P.structured, P.q64, st);
      }
    }
    if (timing) FX_HIP(hipEventRecord(ev[3], st));
    Plane2<W>* X = square(P.s_do - P.q64, P.structured, dc, st, P.two);
    if (P.deflate) {
      // exp(A) = e^{i mu} (X + delta J), delta = (e^{i(c - mu)} - 1) / n
      const double th = P.c - P.mu_im, sh = std::sin(0.5 * th);
      dim3 grid((n + 255) / 256, n);
      add_constant<W><<<grid, 256, 0, st>>>(X->r, X->i, ld, n, -2.0 * sh * sh / n, std::sin(th) / n);
      FX_LAUNCH();
    }
    if (timing) FX_HIP(hipEventRecord(ev[4], st));
    // phase e^mu
    const double er = std::exp(P.mu_re);
    ph_re = er * std::cos(P.mu_im);
    ph_im = er * std::sin(P.mu_im);
    {
      dim3 grid((n + 255) / 256, n);
      write_output<W, Tout><<<grid, 256, 0, st>>>(X->r, X->i, ld, n, ph_re, ph_im, d_u);
      FX_LAUNCH();
    }
    if (timing) FX_HIP(hipEventRecord(ev[5], st));
    if (info_slot) {
      collect_info<<<1, 1, 0, st>>>(scal, h ? h->scal : nullptr, info_slot);
      FX_LAUNCH();
    }
    res = *X;
    have_result = true;
    last_structured = P.structured;
  }

  // Every buffer that a run would otherwise allocate on first use, at the
  // largest size an order up to the capacity needs (no-allocation pipelines:
  // later runs then take nothing from the arena): the FP64 companion, the
  // int8 slices of the FP64 products, the tile orders of the symmetric
  // slice products and the spectral-bound work buffers.  The Cholesky
  // solve's Z updates stay on the caller's stream.
  void reserve_all() {
    z_overlap = 0;
    if constexpr (std::is_same<W, float>::value) {
      Pipeline<double>* h = companion();
      h->z_overlap = 0;
      h->reserve_all();
    } else {
      oz_reserve((int)npcap, (int)round_up(2 * npcap, OZ_BK));
      for (int T = 1; T * SQ_BM <= npcap; ++T) tile_order(T, nullptr);
      constexpr int kRows = 64;
      npart = dalloc<double>(std::max((size_t)std::max(1, (ncap + kRows - 1) / kRows) * std::max(ncap, 1),
                                      (size_t)kTraceDots * std::max(ncap, 1)));
      d_norms = dalloc<double>(kSpecNorms);
      ncmax = dalloc<double>((size_t)(ncap + 255) / 256 + 1);
    }
  }

  void fill_stats(const Plan& P, fexpm_stats& S) const {
    S.nonfinite = P.nonfinite ? 1 : 0;
    S.norm_plain = P.norm_plain;
    S.norm = P.norm;
    S.m = P.m;
    S.s = P.s;
    S.structured = P.structured ? 1 : 0;
    S.shifted = P.shifted ? 1 : 0;
    S.mu_re = P.mu_re;
    S.mu_im = P.mu_im;
    S.flops_issued = flops;
    S.s_fp64 = P.q64;
    S.pade_fp64 = (P.hybrid && !P.nonfinite) ? 1 : 0;
    S.kblock = kblock;
    S.precision = P.split16 ? square_mode() : FEXPM_PREC_NATIVE;
    S.fp64_slices = (P.hybrid && P.structured && !P.nonfinite) ? oz_slices : 0;
    S.fp64_slices_scope = S.fp64_slices ? slice_scope(P) : 0;
    S.spectral = P.spectral ? 1 : 0;
    S.deflated = P.deflate ? 1 : 0;
    S.s_onenorm = P.nonfinite ? 0 : P.s_onenorm;
    S.rho_bound = P.spectral ? P.rho : 0.0;
    S.s_two_product = P.nonfinite ? 0 : P.two;
    const bool crt = !P.nonfinite && S.precision == FEXPM_PREC_INT8_CRT;
    S.crt_moduli = crt ? crt_moduli : 0;
    S.crt_beta = crt ? last_crt_beta : 0;
    S.fp64_moduli = S.fp64_slices == CRT64_K ? crt64_for(P.spectral) : 0;
    S.reserved3 = 0;
    if (P.cheb && !P.nonfinite) {
      const ChebSeries cs = cheb_series(cheb_theta(P), cheb_tail(P.s, std::is_same<W, float>::value));
      S.chebyshev = std::max(cs.dc, cs.ds);
    } else {
      S.chebyshev = 0;
    }
  }

  static void check_info(int info) {
    if (info & INFO_SPLIT_RANGE)
      throw Error(FEXPM_EINVAL, "split-precision squaring: a scaled operand value left the 16-bit range");
    if (info & 2) throw Error(FEXPM_EHIP, "LU panel grid barrier timed out");
    if (info != 0) throw Error(FEXPM_ESINGULAR, "Pade denominator is exactly singular");
  }

  // Single run with per-phase timing (one host wait for the analysis, one
  // at the end).
  template <typename Tin, typename Tout>
  void run_typed(const Tin* d_a, Tout* d_u, hipStream_t st, fexpm_stats& S, int opt_structure) {
    const bool timing = events;  // pipelines without events (no-allocation path) report no phase times
    time_kernels = timing;
    kev_used = 0;
    if (timing) FX_HIP(hipEventRecord(ev[0], st));
    analyze_enqueue<Tin>(d_a, st, nullptr);
    Scalars hs;
    FX_HIP(hipMemcpyAsync(&hs, scal, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    FX_HIP(hipStreamSynchronize(st));
    if (timing) FX_HIP(hipEventRecord(ev[1], st));
    Plan P = make_plan(hs, opt_structure, true);
    if (P.spectral) spectral_plan<Tin>(d_a, P, st);
    enqueue_rest<Tin, Tout>(d_a, d_u, P, st, timing, info_buf);
    int info = 0;
    FX_HIP(hipMemcpyAsync(&info, info_buf, sizeof(int), hipMemcpyDeviceToHost, st));
    FX_HIP(hipStreamSynchronize(st));
    fill_stats(P, S);
    if (P.nonfinite || !timing) {
      S.ms_total = S.ms_analyze = S.ms_pade = S.ms_solve = S.ms_square = S.ms_output = 0.0;
      S.ms_square_kernel = 0.0;
      S.square_kernel_launches = 0;
      time_kernels = false;
      if (!P.nonfinite) check_info(info);
      return;
    }
    float t01, t12, t23, t34, t45;
    FX_HIP(hipEventElapsedTime(&t01, ev[0], ev[1]));
    FX_HIP(hipEventElapsedTime(&t12, ev[1], ev[2]));
    FX_HIP(hipEventElapsedTime(&t23, ev[2], ev[3]));
    FX_HIP(hipEventElapsedTime(&t34, ev[3], ev[4]));
    FX_HIP(hipEventElapsedTime(&t45, ev[4], ev[5]));
    S.ms_analyze = t01;
    S.ms_pade = t12;
    S.ms_solve = t23;
    S.ms_square = t34;
    S.ms_output = t45;
    S.ms_total = t01 + t12 + t23 + t34 + t45;
    double kms = 0.0;
    for (int q = 0; q < kev_used; ++q) {
      float t = 0.f;
      FX_HIP(hipEventElapsedTime(&t, kev[2 * q], kev[2 * q + 1]));
      kms += t;
    }
    S.ms_square_kernel = kms;
    S.square_kernel_launches = kev_used;
    time_kernels = false;
    check_info(info);
  }

  // Batches: analysis into a device slot; then the rest from the analysis
  // read back by the caller, without host waits.
  void analyze(const void* d_a, Scalars* slot, hipStream_t st) override {
    switch (dtype) {
      case FEXPM_FLOAT32: analyze_enqueue<float>((const float*)d_a, st, slot); break;
      case FEXPM_FLOAT64: analyze_enqueue<double>((const double*)d_a, st, slot); break;
      case FEXPM_COMPLEX64: analyze_enqueue<float2>((const float2*)d_a, st, slot); break;
      default: analyze_enqueue<double2>((const double2*)d_a, st, slot); break;
    }
  }

  void run_planned(const void* d_a, void* d_u, const Scalars& hs, hipStream_t st, fexpm_stats& S,
                   int opt_structure, int* info_slot) override {
    const Plan P = make_plan(hs, opt_structure, false);
    switch (dtype) {
      case FEXPM_FLOAT32:
        enqueue_rest<float, float>((const float*)d_a, (float*)d_u, P, st, false, info_slot);
        break;
      case FEXPM_FLOAT64:
        enqueue_rest<double, double>((const double*)d_a, (double*)d_u, P, st, false, info_slot);
        break;
      case FEXPM_COMPLEX64:
        enqueue_rest<float2, float2>((const float2*)d_a, (float2*)d_u, P, st, false, info_slot);
        break;
      default:
        enqueue_rest<double2, double2>((const double2*)d_a, (double2*)d_u, P, st, false, info_slot);
        break;
    }
    fill_stats(P, S);
  }

  void run(const void* d_a, void* d_u, hipStream_t st, fexpm_stats& S, int opt_structure) override {
    switch (dtype) {
      case FEXPM_FLOAT32:
        run_typed<float, float>((const float*)d_a, (float*)d_u, st, S, opt_structure);
        break;
      case FEXPM_FLOAT64:
        run_typed<double, double>((const double*)d_a, (double*)d_u, st, S, opt_structure);
        break;
      case FEXPM_COMPLEX64:
        run_typed<float2, float2>((const float2*)d_a, (float2*)d_u, st, S, opt_structure);
        break;
      default:
        run_typed<double2, double2>((const double2*)d_a, (double2*)d_u, st, S, opt_structure);
        break;
    }
  }

  void debug_plane(int which, double* re, double* im) override {
    const Plane2<W>* planes[] = {&a, &p2, &p4, &p6, &t1, &t2, &u, &v, &x0, &x1, &res};
    if (which < 0 || which > 10) throw Error(FEXPM_EINVAL, "unknown plane");
    const Plane2<W>& P = *planes[which];
    if (!P.r) throw Error(FEXPM_EINVAL, "plane not available");
    std::vector<W> buf((size_t)n * n);
    FX_HIP(hipDeviceSynchronize());
    FX_HIP(hipMemcpy2D(buf.data(), n * sizeof(W), P.r, ld * sizeof(W), n * sizeof(W), n, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < buf.size(); ++k) re[k] = (double)buf[k];
    if (!P.i) {  // real plane of a structured run
      std::fill(im, im + buf.size(), 0.0);
      return;
    }
    FX_HIP(hipMemcpy2D(buf.data(), n * sizeof(W), P.i, ld * sizeof(W), n * sizeof(W), n, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < buf.size(); ++k) im[k] = (double)buf[k];
  }

  // psi_k = U^k psi0 for k = 1..nsteps and all states at once: Psi (states
  // as rows) <- Psi X^T by the NT GEMM with B = rows of X (no transpose),
  // then refine_states applies the FP64 norm refinement (unitary CTQW
  // propagators) and the phase e^{k mu}.  d_out: nsteps x nstates x n.
  Plane2<W> ps[2];
  double* norm0 = nullptr;
  int ps_rows = 0;
  bool last_structured = false;

  void propagate(const double2* d_psi0, int nstates, int nsteps, double2* d_out, hipStream_t st) override {
    if (nstates > ps_rows) {
      for (auto& P : ps) {
        dfree(P.r);
        dfree(P.i);
        P.r = P.i = nullptr;
      }
      dfree(norm0);
      norm0 = nullptr;
      ps_rows = (int)round_up(nstates, 128);
      for (auto& P : ps) {
        P.r = dalloc<W>((size_t)ps_rows * ld);
        P.i = dalloc<W>((size_t)ps_rows * ld);
      }
      norm0 = dalloc<double>(ps_rows);
    }
    FX_HIP(hipEventRecord(ev[6], st));
    states_to_planes<W><<<nstates, 256, 0, st>>>(d_psi0, n, ps[0].r, ps[0].i, ld, norm0);
    FX_LAUNCH();
    int cur = 0;
    double mre = std::log(std::max(std::hypot(ph_re, ph_im), 1e-300)), mim = std::atan2(ph_im, ph_re);
    for (int k = 1; k <= nsteps; ++k) {
      GemmArgs<W> g = args(ps[cur], res, ld, ld, nstates, n, (int)np);
      g.epi.cr = ps[cur ^ 1].r;
      g.epi.ci = ps[cur ^ 1].i;
      g.epi.ldc = ld;
      g.epi.ld_add = ld;
      g.kblock = kblock > 0 ? kblock : 1 << 20;
      // A few hundred states against n columns make few 128 x 128 tiles
      // (128 x 8192: 64 for 256 CUs): 64 x 64 tiles then (same k order and
      // accumulation blocks per element, so the same values).
      using SmallC = typename std::conditional<std::is_same<W, float>::value, FCfgSmall, GemmCfg<W>>::type;
      const int tiles128 = ((nstates + 127) / 128) * ((n + 127) / 128);
      if (std::is_same<W, float>::value && tiles128 < 512)
        gemm<MODE_C3M, false, true, SmallC>(g, st);
      else if (std::is_same<W, float>::value)
        gemm<MODE_C3M, false, true>(g, st);
      else
        gemm<MODE_C3M, false>(g, st);
      cur ^= 1;
      const double er = std::exp(k * mre);
      refine_states<W><<<nstates, 256, 0, st>>>(ps[cur].r, ps[cur].i, ld, n, norm0, last_structured ? 1 : 0,
                                                er * std::cos(k * mim), er * std::sin(k * mim),
                                                d_out + (int64_t)(k - 1) * nstates * n);
      FX_LAUNCH();
    }
    FX_HIP(hipEventRecord(ev[7], st));
    FX_HIP(hipEventSynchronize(ev[7]));
    float ms = 0.f;
    FX_HIP(hipEventElapsedTime(&ms, ev[6], ev[7]));
    last_prop_ms = ms;
  }

  void apply(const double2* d_psi0, int nstates, double2* d_out, hipStream_t st) override {
    apply_states<W><<<(n + 3) / 4, 256, 0, st>>>(res.r, res.i, ld, n, d_psi0, nstates, ph_re, ph_im, d_out);
    FX_LAUNCH();
    if (last_structured && nstates > 0) {  // unitary U: FP64 norm refinement of every psi
      renorm_states<<<nstates, 256, 0, st>>>(d_psi0, d_out, n);
      FX_LAUNCH();
    }
  }
};

}  // namespace fexpm

#include "fexpm_group.hpp"

// ===========================================================================
// C ABI
using namespace fexpm;

// Copies of n x n matrices between pageable host memory and the device
// through two pinned staging buffers: the host side of each chunk is copied
// by several threads while the DMA engine moves the neighbouring chunk (the
// runtime's own pageable path copies through its staging on one thread).
struct HostStaging {
  static constexpr size_t kChunk = (size_t)32 << 20;
  static constexpr size_t kMin = (size_t)64 << 20;  // smaller copies: plain hipMemcpy
  static constexpr int kThreads = 8;
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ready = false;
  bool failed = false;  // pinned memory unavailable: plain copies from then on
};

struct fexpm_engine {
  int64_t n = 0;
  int dtype = 0;
  bool freed = false;
  std::mutex mtx;
  std::string err;
  fexpm_stats stats;
  hipStream_t stream = nullptr;
  std::unique_ptr<PipelineBase> pipe;
  void* d_in = nullptr;
  void* d_out = nullptr;
  size_t esize = 0;
  int opt_structure = 1;
  int opt_precision = 0;  // FEXPM_PREC_AUTO for complex64 engines (fexpm_create)
  double* d_h = nullptr;  // walk Hamiltonian kept by fexpm_set_graph (FP64, n x n)
  bool have_graph = false;
  HostStaging stg;
  int64_t dev_bytes = 0;  // device memory held (pipelines, I/O buffers, H)
};

static size_t elem_size(int dtype) {
  switch (dtype) {
    case FEXPM_FLOAT32: return 4;
    case FEXPM_FLOAT64: return 8;
    case FEXPM_COMPLEX64: return 8;
    case FEXPM_COMPLEX128: return 16;
  }
  return 0;
}

// Message of the last failure of a call that has no engine (creation,
// testing entry points), per calling thread.
static thread_local std::string g_last_error;

template <typename F> static int guarded(fexpm_engine* e, F&& f) {
  std::string& err = e ? e->err : g_last_error;
  struct AcctScope {  // device allocations of this call count for e
    int64_t* prev;
    explicit AcctScope(int64_t* c) : prev(g_acct) { g_acct = c; }
    ~AcctScope() { g_acct = prev; }
  } acct(e ? &e->dev_bytes : g_acct);
  try {
    f();
    return FEXPM_OK;
  } catch (const Error& x) {
    err = x.what();
    return x.code;
  } catch (const std::bad_alloc&) {
    err = "host allocation failed";
    return FEXPM_ENOMEM;
  } catch (const std::exception& x) {
    err = x.what();
    return FEXPM_EHIP;
  }
}

static void check_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) throw Error(FEXPM_ENODEVICE, "no HIP device available");
}

// Host output buffers are usually fresh memory (a new numpy array): a copy
// into them takes a page fault per 4 KB page on the thread doing the copy
// (about 50 ms per 0.8 GB).  Prefault writes one byte per page on `threads`
// host threads; the buffer is overwritten afterwards, so the values written
// do not matter.  The joins happen in the destructor, so every return path
// waits for the threads before the buffer is copied into or handed back.
class Prefault {
 public:
  Prefault(void* p, size_t bytes, int threads) {
    if (!p || bytes < ((size_t)16 << 20)) return;
    constexpr size_t kPage = 4096;
    const size_t pages = (bytes + kPage - 1) / kPage;
    const size_t per = (pages + threads - 1) / threads;
    char* base = static_cast<char*>(p);
    try {
      for (int t = 0; t < threads; ++t) {
        const size_t lo = (size_t)t * per, hi = std::min(pages, lo + per);
        if (lo >= hi) break;
        th_.emplace_back([base, lo, hi, bytes] {
          volatile char* q = base;
          for (size_t pg = lo; pg < hi; ++pg) q[std::min(pg * kPage, bytes - 1)] = 0;
        });
      }
    } catch (...) {
      // no thread: the copy faults the pages in itself
    }
  }
  ~Prefault() { join(); }
  void join() {
    for (auto& t : th_) t.join();
    th_.clear();
  }

 private:
  std::vector<std::thread> th_;
};

// dst[0, len) = src[0, len) on `threads` host threads (the calling one included)
static void par_copy(void* dst, const void* src, size_t len, int threads) {
  const size_t part = ((len + threads - 1) / threads + 4095) / 4096 * 4096;
  std::vector<std::thread> th;
  size_t off = part;
  try {
    for (; off < len; off += part) {
      const size_t l = std::min(part, len - off);
      th.emplace_back([=] { std::memcpy((char*)dst + off, (const char*)src + off, l); });
    }
  } catch (...) {
    for (; off < len; off += part) std::memcpy((char*)dst + off, (const char*)src + off, std::min(part, len - off));
  }
  std::memcpy(dst, src, std::min(part, len));
  for (auto& t : th) t.join();
}

static void staging_free(HostStaging& h) {
  for (int b = 0; b < 2; ++b) {
    if (h.ev[b]) (void)hipEventSynchronize(h.ev[b]);
    if (h.buf[b]) (void)hipHostFree(h.buf[b]);
    if (h.ev[b]) (void)hipEventDestroy(h.ev[b]);
    h.buf[b] = nullptr;
    h.ev[b] = nullptr;
  }
  h.ready = false;
  h.failed = false;
}

// Allocates the staging buffers on first use; false when pinned host memory
// is not available (e.g. a low locked-memory limit): the caller then copies
// through the runtime's pageable path.
static bool staging_init(HostStaging& h) {
  if (h.ready) return true;
  if (h.failed) return false;
  bool ok = true;
  for (int b = 0; b < 2 && ok; ++b) {
    ok = hipHostMalloc(&h.buf[b], HostStaging::kChunk, hipHostMallocDefault) == hipSuccess &&
         hipEventCreateWithFlags(&h.ev[b], hipEventDisableTiming) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    staging_free(h);
    h.failed = true;
    return false;
  }
  h.ready = true;
  return true;
}

// device d[0, bytes) = host h[0, bytes), enqueued on st (the last chunks may
// still be in flight on return; later work on st is ordered after them)
static void staged_h2d(HostStaging& hs, void* d, const void* h, size_t bytes, hipStream_t st) {
  if (!staging_init(hs)) {
    FX_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    return;
  }
  const size_t C = HostStaging::kChunk, nch = (bytes + C - 1) / C;
  for (size_t c = 0; c < nch; ++c) {
    const int b = (int)(c & 1);
    const size_t off = c * C, len = std::min(C, bytes - off);
    FX_HIP(hipEventSynchronize(hs.ev[b]));  // the DMA that last read buf[b] is done
    par_copy(hs.buf[b], (const char*)h + off, len, HostStaging::kThreads);
    FX_HIP(hipMemcpyAsync((char*)d + off, hs.buf[b], len, hipMemcpyHostToDevice, st));
    FX_HIP(hipEventRecord(hs.ev[b], st));
  }
}

// host h[0, bytes) = device d[0, bytes) after the work enqueued on st; complete on return
static void staged_d2h(HostStaging& hs, void* h, const void* d, size_t bytes, hipStream_t st) {
  if (!staging_init(hs)) {
    FX_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st));
    FX_HIP(hipStreamSynchronize(st));
    return;
  }
  const size_t C = HostStaging::kChunk, nch = (bytes + C - 1) / C;
  auto issue = [&](size_t c) {
    const int b = (int)(c & 1);
    const size_t off = c * C;
    FX_HIP(hipEventSynchronize(hs.ev[b]));
    FX_HIP(hipMemcpyAsync(hs.buf[b], (const char*)d + off, std::min(C, bytes - off), hipMemcpyDeviceToHost, st));
    FX_HIP(hipEventRecord(hs.ev[b], st));
  };
  issue(0);
  if (nch > 1) issue(1);
  for (size_t c = 0; c < nch; ++c) {
    const int b = (int)(c & 1);
    const size_t off = c * C;
    FX_HIP(hipEventSynchronize(hs.ev[b]));
    par_copy((char*)h + off, hs.buf[b], std::min(C, bytes - off), HostStaging::kThreads);
    if (c + 2 < nch) issue(c + 2);
  }
}

static void ensure_io(fexpm_engine* e) {
  const size_t bytes = (size_t)e->n * (size_t)e->n * e->esize;
  if (!e->d_in) e->d_in = dmalloc_bytes(std::max<size_t>(bytes, 16));
  if (!e->d_out) e->d_out = dmalloc_bytes(std::max<size_t>(bytes, 16));
}

static double* ensure_h(fexpm_engine* e) {
  if (!e->d_h) e->d_h = static_cast<double*>(dmalloc_bytes((size_t)e->n * (size_t)e->n * sizeof(double)));
  return e->d_h;
}

static void release(fexpm_engine* e) {
  int64_t* prev = g_acct;
  g_acct = &e->dev_bytes;
  staging_free(e->stg);
  e->pipe.reset();
  dfree(e->d_in);
  dfree(e->d_out);
  e->d_in = e->d_out = nullptr;
  dfree(e->d_h);
  e->d_h = nullptr;
  g_acct = prev;
  e->have_graph = false;
  if (e->stream) (void)hipStreamDestroy(e->stream);
  e->stream = nullptr;
}

extern "C" {

int fexpm_create(int64_t n, int dtype, fexpm_engine** out) {
  if (!out) return FEXPM_EINVAL;
  *out = nullptr;
  if (n < 0 || n > (int64_t)1 << 20 || elem_size(dtype) == 0) return FEXPM_EINVAL;
  fexpm_engine* e = new (std::nothrow) fexpm_engine();
  if (!e) return FEXPM_ENOMEM;
  e->n = n;
  e->dtype = dtype;
  e->esize = elem_size(dtype);
  std::memset(&e->stats, 0, sizeof(e->stats));
  e->stats.n = n;
  e->stats.dtype = dtype;
  const int rc = guarded(e, [&] {
    check_device();
    FX_HIP(hipStreamCreate(&e->stream));
    if (dtype == FEXPM_COMPLEX64) e->opt_precision = FEXPM_PREC_AUTO;
    if (n > 0) {
      if (dtype == FEXPM_FLOAT32 || dtype == FEXPM_COMPLEX64)
        e->pipe.reset(new Pipeline<float>((int)n, dtype));
      else
        e->pipe.reset(new Pipeline<double>((int)n, dtype));
      e->pipe->precision = e->opt_precision;
    }
  });
  if (rc != FEXPM_OK) {
    g_last_error = e->err;  // readable through fexpm_last_error(NULL)
    release(e);
    delete e;
    return rc;
  }
  *out = e;
  return FEXPM_OK;
}

int fexpm_free(fexpm_engine* e) {
  if (!e) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  release(e);
  e->freed = true;
  return FEXPM_OK;
}

int fexpm_destroy(fexpm_engine* e) {
  if (!e) return FEXPM_EINVAL;
  fexpm_free(e);
  delete e;
  return FEXPM_OK;
}

int64_t fexpm_get_n(const fexpm_engine* e) { return e ? e->n : -1; }

static int run_device_locked(fexpm_engine* e, const void* d_a, void* d_u, hipStream_t st) {
  if (e->freed) {
    e->err = "engine memory has been freed";
    return FEXPM_EFREED;
  }
  if (e->n == 0) return FEXPM_OK;
  return guarded(e, [&] {
    e->stats.n = e->n;
    e->stats.dtype = e->dtype;
    e->pipe->run(d_a, d_u, st ? st : e->stream, e->stats, e->opt_structure);
  });
}

int fexpm_run_device(fexpm_engine* e, const void* d_a, void* d_u, void* stream) {
  if (!e || (e->n > 0 && (!d_a || !d_u))) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  return run_device_locked(e, d_a, d_u, (hipStream_t)stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// No-allocation entry point (no counterpart in the reference).  Every work
// buffer of a run of order n <= FEXPM_NOALLOC_NMAX lives in `scratch`, a
// static device array that each GPU holds until the process ends.  Its
// device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and only that address is used on the host; the
// array is cleared once, and the pipelines carved from it keep their own
// zero padding from run to run.  Calls are serialised in the process (one
// mutex) and each returns after its device work is complete; work that a
// caller leaves in flight on the array from another thread would race.
static __device__ __align__(256) unsigned char scratch[FEXPM_NOALLOC_BYTES];

namespace {
struct NoallocCtx {
  bool ready = false;
  Arena arena;                      // over this device's copy of `scratch`
  Pipeline<float>* pf = nullptr;    // float32 / complex64 (arena memory: never destroyed)
  Pipeline<double>* pd = nullptr;   // float64 / complex128
};
std::mutex g_noalloc_mtx;
std::map<int, NoallocCtx> g_noalloc;  // per device

struct ArenaScope {
  Arena* prev;
  explicit ArenaScope(Arena* a) : prev(g_arena) { g_arena = a; }
  ~ArenaScope() { g_arena = prev; }
};
}  // namespace

__global__ __launch_bounds__(256) void zero_words(uint4* __restrict__ p, size_t count) {
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < count; k += (size_t)gridDim.x * 256)
    p[k] = make_uint4(0u, 0u, 0u, 0u);
}

// Device address of `scratch` on the current device, the array cleared, and
// the two pipelines of capacity FEXPM_NOALLOC_NMAX built inside it with every
// buffer a run would allocate on first use.
static void noalloc_setup(NoallocCtx& c) {
  void* ptr = nullptr;
  FX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  c.arena.base = static_cast<char*>(ptr);
  c.arena.cap = FEXPM_NOALLOC_BYTES;
  c.arena.used = 0;
  zero_words<<<1024, 256>>>(reinterpret_cast<uint4*>(c.arena.base), FEXPM_NOALLOC_BYTES / sizeof(uint4));
  FX_LAUNCH();
  FX_HIP(hipDeviceSynchronize());
  ArenaScope scope(&c.arena);
  c.pf = new Pipeline<float>(FEXPM_NOALLOC_NMAX, FEXPM_COMPLEX64, false);
  c.pf->reserve_all();
  c.pd = new Pipeline<double>(FEXPM_NOALLOC_NMAX, FEXPM_COMPLEX128, false);
  c.pd->reserve_all();
  FX_HIP(hipDeviceSynchronize());
  c.ready = true;
}

extern "C" int fexpm_run_device_noalloc(int64_t n, int dtype, const void* d_a, void* d_u, void* stream) {
  if (elem_size(dtype) == 0) {
    g_last_error = "unknown dtype";
    return FEXPM_EINVAL;
  }
  if (n < 0 || n > FEXPM_NOALLOC_NMAX) {
    g_last_error = "fexpm_run_device_noalloc: order n must be in [0, " + std::to_string(FEXPM_NOALLOC_NMAX) +
                   "], the capacity of the static scratch array";
    return FEXPM_EINVAL;
  }
  if (n > 0 && (!d_a || !d_u)) {
    g_last_error = "null device pointer";
    return FEXPM_EINVAL;
  }
  if (n == 0) return FEXPM_OK;
  std::lock_guard<std::mutex> lk(g_noalloc_mtx);
  return guarded(nullptr, [&] {
    int dev = 0;
    FX_HIP(hipGetDevice(&dev));
    NoallocCtx& c = g_noalloc[dev];
    if (!c.ready) {
      try {
        noalloc_setup(c);
      } catch (...) {
        c.ready = false;
        c.arena.used = 0;
        delete c.pf;  // arena pipelines free nothing
        delete c.pd;
        c.pf = nullptr;
        c.pd = nullptr;
        throw;
      }
    }
    ArenaScope scope(&c.arena);  // any allocation now comes from the array, or fails
    hipStream_t st = static_cast<hipStream_t>(stream);
    fexpm_stats S;
    std::memset(&S, 0, sizeof(S));
    S.n = n;
    S.dtype = dtype;
    if (dtype == FEXPM_FLOAT32 || dtype == FEXPM_COMPLEX64) {
      c.pf->dtype = dtype;
      c.pf->set_size((int)n, st);
      c.pf->run(d_a, d_u, st, S, 1);
    } else {
      c.pd->dtype = dtype;
      c.pd->set_size((int)n, st);
      c.pd->run(d_a, d_u, st, S, 1);
    }
  });
}

extern "C" {

int fexpm_upload(fexpm_engine* e, const void* a) {
  if (!e || (e->n > 0 && !a)) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->n == 0) return FEXPM_OK;
  return guarded(e, [&] {
    ensure_io(e);
    const size_t bytes = (size_t)e->n * e->n * e->esize;
    if (bytes >= HostStaging::kMin) {
      staged_h2d(e->stg, e->d_in, a, bytes, e->stream);
      FX_HIP(hipStreamSynchronize(e->stream));
    } else {
      FX_HIP(hipMemcpy(e->d_in, a, bytes, hipMemcpyHostToDevice));
    }
  });
}

int fexpm_run_resident(fexpm_engine* e) {
  if (!e) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->n == 0) return FEXPM_OK;
  if (!e->d_in) {
    e->err = "no resident input: call fexpm_upload first";
    return FEXPM_EINVAL;
  }
  return run_device_locked(e, e->d_in, e->d_out, e->stream);
}

int fexpm_set_graph(fexpm_engine* e, const void* adj, int adj_dtype, double gamma, int laplacian) {
  if (!e || (e->n > 0 && !adj)) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->dtype != FEXPM_COMPLEX64 && e->dtype != FEXPM_COMPLEX128) {
    e->err = "fexpm_set_graph needs a complex engine (A_exp = -i t H is complex)";
    return FEXPM_EINVAL;
  }
  if (adj_dtype != FEXPM_FLOAT32 && adj_dtype != FEXPM_FLOAT64) {
    e->err = "adjacency dtype must be FEXPM_FLOAT32 or FEXPM_FLOAT64";
    return FEXPM_EINVAL;
  }
  if (e->n == 0) return FEXPM_OK;
  return guarded(e, [&] {
    const int n = (int)e->n;
    const size_t cnt = (size_t)n * n, asz = adj_dtype == FEXPM_FLOAT32 ? 4 : 8;
    ensure_io(e);
    ensure_h(e);
    // the adjacency is staged in the input buffer (>= n^2 * 8 bytes for complex
    // engines), which fexpm_run_graph overwrites with A_exp anyway
    FX_HIP(hipMemcpyAsync(e->d_in, adj, cnt * asz, hipMemcpyHostToDevice, e->stream));
    double* rs = nullptr;
    if (laplacian) FX_HIP(hipMalloc(&rs, (size_t)n * sizeof(double)));
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
    if (adj_dtype == FEXPM_FLOAT32) {
      if (rs) graph_row_sums<float><<<n, 256, 0, e->stream>>>((const float*)e->d_in, n, rs);
      graph_hamiltonian<float><<<grid, 256, 0, e->stream>>>((const float*)e->d_in, n, gamma, rs, e->d_h);
    } else {
      if (rs) graph_row_sums<double><<<n, 256, 0, e->stream>>>((const double*)e->d_in, n, rs);
      graph_hamiltonian<double><<<grid, 256, 0, e->stream>>>((const double*)e->d_in, n, gamma, rs, e->d_h);
    }
    FX_LAUNCH();
    FX_HIP(hipStreamSynchronize(e->stream));
    if (rs) (void)hipFree(rs);
    e->have_graph = true;
  });
}

// H = -gamma A or -gamma (D - A) in place, A already in d_h.
static void finish_graph(fexpm_engine* e, double gamma, int laplacian) {
  const int n = (int)e->n;
  double* rs = nullptr;
  if (laplacian) FX_HIP(hipMalloc(&rs, (size_t)n * sizeof(double)));
  try {
    if (rs) {
      graph_row_sums<double><<<n, 256, 0, e->stream>>>(e->d_h, n, rs);
      FX_LAUNCH();
    }
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
    graph_hamiltonian_inplace<<<grid, 256, 0, e->stream>>>(e->d_h, n, gamma, rs);
    FX_LAUNCH();
    FX_HIP(hipStreamSynchronize(e->stream));
  } catch (...) {
    if (rs) (void)hipFree(rs);
    throw;
  }
  if (rs) (void)hipFree(rs);
  e->have_graph = true;
}

static int graph_engine_check(fexpm_engine* e) {
  if (e->freed) return FEXPM_EFREED;
  if (e->dtype != FEXPM_COMPLEX64 && e->dtype != FEXPM_COMPLEX128) {
    e->err = "graph input needs a complex engine (A_exp = -i t H is complex)";
    return FEXPM_EINVAL;
  }
  return FEXPM_OK;
}

int fexpm_set_graph_edges(fexpm_engine* e, int64_t m, const int64_t* src, const int64_t* dst, const double* w,
                          double gamma, int laplacian) {
  if (!e || m < 0 || (m > 0 && (!src || !dst))) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  int rc = graph_engine_check(e);
  if (rc != FEXPM_OK) return rc;
  for (int64_t k = 0; k < m; ++k)
    if (src[k] < 0 || src[k] >= e->n || dst[k] < 0 || dst[k] >= e->n) {
      e->err = "edge " + std::to_string(k) + ": vertex index outside [0, n)";
      return FEXPM_EINVAL;
    }
  if (e->n == 0) return FEXPM_OK;
  return guarded(e, [&] {
    const int n = (int)e->n;
    const int64_t cnt = (int64_t)n * n;
    e->have_graph = false;
    ensure_io(e);  // fexpm_run_graph writes A_exp into the input buffer
    ensure_h(e);
    fill_zero<<<(unsigned)std::min<int64_t>((cnt + 255) / 256, 65536), 256, 0, e->stream>>>(e->d_h, cnt);
    FX_LAUNCH();
    if (m > 0) {
      const size_t ib = (size_t)m * sizeof(int64_t), wb = w ? (size_t)m * sizeof(double) : 0;
      char* buf = nullptr;
      FX_HIP(hipMalloc(&buf, 2 * ib + wb));
      try {
        FX_HIP(hipMemcpyAsync(buf, src, ib, hipMemcpyHostToDevice, e->stream));
        FX_HIP(hipMemcpyAsync(buf + ib, dst, ib, hipMemcpyHostToDevice, e->stream));
        if (w) FX_HIP(hipMemcpyAsync(buf + 2 * ib, w, wb, hipMemcpyHostToDevice, e->stream));
        graph_edges<<<(unsigned)((m + 255) / 256), 256, 0, e->stream>>>(
            e->d_h, n, m, (const int64_t*)buf, (const int64_t*)(buf + ib), w ? (const double*)(buf + 2 * ib) : nullptr);
        FX_LAUNCH();
        FX_HIP(hipStreamSynchronize(e->stream));
      } catch (...) {
        (void)hipStreamSynchronize(e->stream);
        (void)hipFree(buf);
        throw;
      }
      (void)hipFree(buf);
    }
    finish_graph(e, gamma, laplacian);
  });
}

int fexpm_set_graph_er(fexpm_engine* e, double p, uint64_t seed, int weighted, double gamma, int laplacian) {
  if (!e || !(p >= 0.0 && p <= 1.0)) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  int rc = graph_engine_check(e);
  if (rc != FEXPM_OK) return rc;
  if (e->n == 0) return FEXPM_OK;
  return guarded(e, [&] {
    const int n = (int)e->n;
    e->have_graph = false;
    ensure_io(e);  // fexpm_run_graph writes A_exp into the input buffer
    ensure_h(e);
    // hash streams 1 (edges) and 2 (weights), as in fastctqw_amd/graphs.py
    const uint64_t ke = hash_stream_key(seed, 1), kw = hash_stream_key(seed, 2);
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
    graph_er<<<grid, 256, 0, e->stream>>>(e->d_h, n, p, ke, kw, weighted ? 1 : 0);
    FX_LAUNCH();
    finish_graph(e, gamma, laplacian);
  });
}

int fexpm_run_graph(fexpm_engine* e, double t) {
  if (!e) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->n == 0) return FEXPM_OK;
  if (!e->have_graph) {
    e->err = "no graph: call fexpm_set_graph first";
    return FEXPM_EINVAL;
  }
  int rc = guarded(e, [&] {
    ensure_io(e);  // A_exp goes to the input buffer, U to the output buffer
    if (!e->d_in || !e->d_h) throw Error(FEXPM_EINVAL, "graph buffers missing");
    const int64_t cnt = e->n * e->n;
    const unsigned blocks = (unsigned)((cnt + 255) / 256);
    if (e->dtype == FEXPM_COMPLEX64)
      walk_input<float2><<<blocks, 256, 0, e->stream>>>(e->d_h, cnt, t, (float2*)e->d_in);
    else
      walk_input<double2><<<blocks, 256, 0, e->stream>>>(e->d_h, cnt, t, (double2*)e->d_in);
    FX_LAUNCH();
  });
  if (rc != FEXPM_OK) return rc;
  return run_device_locked(e, e->d_in, e->d_out, e->stream);
}

int fexpm_download(fexpm_engine* e, void* u) {
  if (!e || (e->n > 0 && !u)) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->n == 0) return FEXPM_OK;
  if (!e->d_out) return FEXPM_EINVAL;
  return guarded(e, [&] {
    const size_t bytes = (size_t)e->n * e->n * e->esize;
    if (bytes >= HostStaging::kMin) {
      FX_HIP(hipStreamSynchronize(e->stream));
      staged_d2h(e->stg, u, e->d_out, bytes, e->stream);
      return;
    }
    {
      Prefault pf(u, bytes, 8);
    }
    FX_HIP(hipMemcpy(u, e->d_out, bytes, hipMemcpyDeviceToHost));
  });
}

int fexpm_run(fexpm_engine* e, const void* a, void* u) {
  if (!e || (e->n > 0 && (!a || !u))) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) {
    e->err = "engine memory has been freed";
    return FEXPM_EFREED;
  }
  if (e->n == 0) return FEXPM_OK;
  const size_t bytes = (size_t)e->n * e->n * e->esize;
  // the output's pages are faulted in on one host thread while the device
  // works (not when input and output overlap: the input is read meanwhile)
  const char *ca = static_cast<const char*>(a), *cu = static_cast<const char*>(u);
  const bool overlap = ca < cu + bytes && cu < ca + bytes;
  Prefault pf(overlap ? nullptr : u, bytes, 1);
  const bool staged = bytes >= HostStaging::kMin && !overlap;
  int rc = guarded(e, [&] {
    ensure_io(e);
    if (staged) staged_h2d(e->stg, e->d_in, a, bytes, e->stream);
    else FX_HIP(hipMemcpyAsync(e->d_in, a, bytes, hipMemcpyHostToDevice, e->stream));
  });
  if (rc != FEXPM_OK) return rc;
  rc = run_device_locked(e, e->d_in, e->d_out, e->stream);
  if (rc != FEXPM_OK) return rc;
  pf.join();
  return guarded(e, [&] {
    if (staged) {
      staged_d2h(e->stg, u, e->d_out, bytes, e->stream);
      return;
    }
    FX_HIP(hipMemcpyAsync(u, e->d_out, bytes, hipMemcpyDeviceToHost, e->stream));
    FX_HIP(hipStreamSynchronize(e->stream));
  });
}

int fexpm_apply(fexpm_engine* e, const void* psi0, int64_t nstates, void* psi) {
  if (!e || nstates < 0 || (nstates > 0 && (!psi0 || !psi))) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->n == 0 || nstates == 0) return FEXPM_OK;
  if (!e->pipe->has_result()) {
    e->err = "no result to apply: run first";
    return FEXPM_EINVAL;
  }
  return guarded(e, [&] {
    const size_t bytes = (size_t)nstates * e->n * sizeof(double2);
    void* din = nullptr;
    void* dout = nullptr;
    FX_HIP(hipMalloc(&din, bytes));
    hipError_t er = hipMalloc(&dout, bytes);
    if (er != hipSuccess) {
      (void)hipFree(din);
      FX_HIP(er);
    }
    try {
      FX_HIP(hipMemcpyAsync(din, psi0, bytes, hipMemcpyHostToDevice, e->stream));
      e->pipe->apply((const double2*)din, (int)nstates, (double2*)dout, e->stream);
      FX_HIP(hipMemcpyAsync(psi, dout, bytes, hipMemcpyDeviceToHost, e->stream));
      FX_HIP(hipStreamSynchronize(e->stream));
    } catch (...) {
      (void)hipFree(din);
      (void)hipFree(dout);
      throw;
    }
    (void)hipFree(din);
    (void)hipFree(dout);
  });
}

int fexpm_propagate(fexpm_engine* e, const void* psi0, int64_t nstates, int64_t nsteps, void* psi) {
  if (!e || nstates < 0 || nsteps < 0 || (nstates > 0 && nsteps > 0 && (!psi0 || !psi))) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->n == 0 || nstates == 0 || nsteps == 0) return FEXPM_OK;
  if (!e->pipe->has_result()) {
    e->err = "no propagator: run first";
    return FEXPM_EINVAL;
  }
  return guarded(e, [&] {
    const size_t in_bytes = (size_t)nstates * e->n * sizeof(double2);
    const size_t out_bytes = in_bytes * (size_t)nsteps;
    void* din = nullptr;
    void* dout = nullptr;
    FX_HIP(hipMalloc(&din, in_bytes));
    hipError_t er = hipMalloc(&dout, out_bytes);
    if (er != hipSuccess) {
      (void)hipFree(din);
      FX_HIP(er);
    }
    try {
      FX_HIP(hipMemcpyAsync(din, psi0, in_bytes, hipMemcpyHostToDevice, e->stream));
      e->pipe->propagate((const double2*)din, (int)nstates, (int)nsteps, (double2*)dout, e->stream);
      FX_HIP(hipMemcpyAsync(psi, dout, out_bytes, hipMemcpyDeviceToHost, e->stream));
      FX_HIP(hipStreamSynchronize(e->stream));
    } catch (...) {
      (void)hipFree(din);
      (void)hipFree(dout);
      throw;
    }
    (void)hipFree(din);
    (void)hipFree(dout);
  });
}

int fexpm_set_option(fexpm_engine* e, int key, int64_t value) {
  if (!e) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  switch (key) {
    case FEXPM_OPT_PRECISION:
      if (value != FEXPM_PREC_NATIVE && value != FEXPM_PREC_FP16X2 && value != FEXPM_PREC_BF16X3 &&
          value != FEXPM_PREC_INT8_CRT && value != FEXPM_PREC_AUTO) {
        e->err = "unknown precision mode";
        return FEXPM_EINVAL;
      }
      if (value != FEXPM_PREC_NATIVE && e->dtype != FEXPM_COMPLEX64) {
        e->err = "split precision applies to complex64 engines";
        return FEXPM_EINVAL;
      }
      e->opt_precision = (int)value;
      if (e->pipe) e->pipe->precision = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_STRUCTURE:
      e->opt_structure = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_MAX_SQUARINGS:
      if (e->pipe) e->pipe->max_squarings = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_KBLOCK:
      if (value < 0 || value > 1024) {
        e->err = "kblock must be in [0, 1024]";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->kblock = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_PANEL_COOP:
      if (e->pipe) e->pipe->panel_coop = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_SPD_SOLVE:
      if (e->pipe) e->pipe->spd_solve = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_PADE_FP64:
      if (e->pipe) e->pipe->pade_fp64 = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_TAIL_SPLIT:
      if (e->pipe) e->pipe->tail_split = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_Z_OVERLAP:
      if (e->pipe) e->pipe->z_overlap = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_Z_LEFT:
      if (e->pipe) e->pipe->z_left = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_HOST_STAGING:
      staging_free(e->stg);
      e->stg.failed = value == 0;
      return FEXPM_OK;
    case FEXPM_OPT_CRT_MAP:
      if (value != 0 && value != 1) {
        e->err = "residue block map must be 0 or 1";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->crt_map = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_CRT64_MODULI:
      if (value != 0 && (value < CRT64_KMIN || value > CRT64_K)) {
        e->err = "FP64 residue product moduli must be 0 (automatic) or in [8, 13]";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->crt64_opt = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_CRT_MODULI:
      if (value < CRT_KMIN || value > CRT_K) {
        e->err = "residue squaring moduli must be 7, 8 or 9";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->crt_moduli = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_SQUARE_TILE:
      if (value != 0 && value != 64 && value != 128) {
        e->err = "squaring tile must be 0 (auto), 64 or 128";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->sq_tile = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_CHOL_OUTER:
      if (value < 64 || value > 4096 || value % 64 != 0) {
        e->err = "Cholesky outer block must be a multiple of 64 in [64, 4096]";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->ch_outer = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_FP64_SLICES_SCOPE:
      if (value < 0 || value > 3) {
        e->err = "scope must be 0 (automatic), 1 (Pade), 2 (solve) or 3 (both)";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->oz_scope = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_FP64_SLICES:
      if (value != 0 && value != CRT64_K && (value < 4 || value > OZ_MAX_SLICES)) {
        e->err = "int8 slices must be 0, in [4, 6], or 13 (residue products)";
        return FEXPM_EINVAL;
      }
      if (value != 0 && e->dtype != FEXPM_COMPLEX64 && e->dtype != FEXPM_FLOAT32) {
        e->err = "int8-slice FP64 products apply to the FP64 stages of complex64 / float32 engines";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->oz_slices = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_TWO_PRODUCT:
      if (value < 0 || value > 64) {
        e->err = "two-product squarings must be in [0, 64]";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->sq2_lead = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_SPECTRAL:
      if (e->pipe) e->pipe->spectral = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_DEFLATE:
      if (e->pipe) e->pipe->deflate = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_CHEBYSHEV:
      if (e->pipe) e->pipe->cheb = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_FP64_SQUARINGS:
      if (value < 0 || value > 64) {
        e->err = "fp64 squarings must be in [0, 64]";
        return FEXPM_EINVAL;
      }
      if (e->pipe) e->pipe->fp64_squarings = (int)value;
      return FEXPM_OK;
  }
  e->err = "unknown option";
  return FEXPM_EINVAL;
}

int fexpm_get_stats(const fexpm_engine* e, fexpm_stats* out) {
  if (!e || !out) return FEXPM_EINVAL;
  *out = e->stats;
  out->ms_propagate = e->pipe ? e->pipe->last_prop_ms : 0.0;
  out->device_bytes = e->dev_bytes;
  return FEXPM_OK;
}

const char* fexpm_last_error(const fexpm_engine* e) { return e ? e->err.c_str() : g_last_error.c_str(); }

const char* fexpm_status_string(int status) {
  switch (status) {
    case FEXPM_OK: return "ok";
    case FEXPM_EINVAL: return "invalid argument";
    case FEXPM_ENOMEM: return "out of device memory";
    case FEXPM_EHIP: return "HIP runtime error";
    case FEXPM_ESINGULAR: return "singular Pade denominator";
    case FEXPM_EFREED: return "engine memory has been freed";
    case FEXPM_ENODEVICE: return "no HIP device";
  }
  return "unknown status";
}

double fexpm_gpu_memory_gb(void) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return -1.0;
  return (double)fr / (1024.0 * 1024.0 * 1024.0);
}

}  // extern "C"

// ===========================================================================
// Batches of independent expms of uneven order (many walks).

struct fexpm_batch {
  int dtype = 0;
  int64_t nmax = 0;
  int lanes = 1;
  size_t esize = 0;
  std::mutex mtx;
  std::string err;
  std::vector<std::unique_ptr<PipelineBase>> pipes;
  std::vector<hipStream_t> streams;
  std::vector<int64_t> n;       // orders of the uploaded batch
  std::vector<size_t> off;      // element offsets of each matrix in the packed buffers
  std::vector<size_t> voff;     // element offsets of each state
  void* d_in = nullptr;
  void* d_out = nullptr;
  size_t cap_elems = 0;
  double2* d_psi0 = nullptr;
  double2* d_psi = nullptr;
  size_t cap_states = 0;
  std::vector<fexpm_stats> stats;
  double last_ms = 0.0;
  Scalars* d_scal = nullptr;  // per-walk analysis results
  int* d_info = nullptr;      // per-walk exception bits
  size_t cap_walks = 0;
  std::vector<hipEvent_t> ev;  // per-walk start / end
  // grouped launches (complex dtypes)
  std::unique_ptr<GroupRunnerBase> grp;
  int grouped = 1;
  int precision = FEXPM_PREC_NATIVE;
  int spd_solve = 1;
  fexpm_batch_report report{};
  HostStaging stg;  // host copies of 64 MiB and more
  int fail_alloc = 0;  // testing: the next matrix-buffer allocation of an upload fails
};

static void batch_release(fexpm_batch* b) {
  staging_free(b->stg);
  b->grp.reset();
  b->pipes.clear();
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  b->streams.clear();
  if (b->d_in) (void)hipFree(b->d_in);
  if (b->d_out) (void)hipFree(b->d_out);
  if (b->d_psi0) (void)hipFree(b->d_psi0);
  if (b->d_psi) (void)hipFree(b->d_psi);
  b->d_in = b->d_out = nullptr;
  b->d_psi0 = b->d_psi = nullptr;
  if (b->d_scal) (void)hipFree(b->d_scal);
  if (b->d_info) (void)hipFree(b->d_info);
  b->d_scal = nullptr;
  b->d_info = nullptr;
  b->cap_walks = 0;
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
}

template <typename F> static int guarded_b(fexpm_batch* b, F&& f) {
  try {
    f();
    return FEXPM_OK;
  } catch (const Error& x) {
    b->err = x.what();
    return x.code;
  } catch (const std::exception& x) {
    b->err = x.what();
    return FEXPM_EHIP;
  }
}

extern "C" {

int fexpm_batch_create(int dtype, int64_t nmax, int lanes, fexpm_batch** out) {
  if (!out || nmax <= 0 || nmax > (int64_t)1 << 16 || lanes < 1 || lanes > 16 || elem_size(dtype) == 0)
    return FEXPM_EINVAL;
  *out = nullptr;
  fexpm_batch* b = new (std::nothrow) fexpm_batch();
  if (!b) return FEXPM_ENOMEM;
  b->dtype = dtype;
  b->nmax = nmax;
  b->lanes = lanes;
  b->esize = elem_size(dtype);
  int rc = guarded_b(b, [&] {
    check_device();
    for (int l = 0; l < lanes; ++l) {
      hipStream_t st;
      FX_HIP(hipStreamCreate(&st));
      b->streams.push_back(st);
      PipelineBase* p;
      if (dtype == FEXPM_FLOAT32 || dtype == FEXPM_COMPLEX64)
        p = new Pipeline<float>((int)nmax, dtype);
      else
        p = new Pipeline<double>((int)nmax, dtype);
      p->panel_coop = 0;  // lanes run concurrently: no cooperative launches
      p->z_overlap = 0;   // nor second streams (lanes already overlap one another)
      b->pipes.emplace_back(p);
    }
    if (dtype == FEXPM_COMPLEX64)
      b->grp.reset(new GroupRunner<float2, float>());
    else if (dtype == FEXPM_COMPLEX128)
      b->grp.reset(new GroupRunner<double2, double>());
  });
  if (rc != FEXPM_OK) {
    g_last_error = b->err;
    batch_release(b);
    delete b;
    return rc;
  }
  *out = b;
  return FEXPM_OK;
}

int fexpm_batch_destroy(fexpm_batch* b) {
  if (!b) return FEXPM_EINVAL;
  batch_release(b);
  delete b;
  return FEXPM_OK;
}

const char* fexpm_batch_last_error(const fexpm_batch* b) { return b ? b->err.c_str() : g_last_error.c_str(); }

int fexpm_batch_set_option(fexpm_batch* b, int key, int64_t value) {
  if (!b) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(b->mtx);
  switch (key) {
    case FEXPM_OPT_PRECISION:
      if (value != FEXPM_PREC_NATIVE && value != FEXPM_PREC_FP16X2 && value != FEXPM_PREC_BF16X3) {
        b->err = "unknown precision mode";
        return FEXPM_EINVAL;
      }
      if (value != FEXPM_PREC_NATIVE && b->dtype != FEXPM_COMPLEX64) {
        b->err = "split precision applies to complex64 batches";
        return FEXPM_EINVAL;
      }
      for (auto& p : b->pipes) p->precision = (int)value;
      b->precision = (int)value;
      if (b->grp) b->grp->precision = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_KBLOCK:
      if (value < 0 || value > 1024) {
        b->err = "kblock must be in [0, 1024]";
        return FEXPM_EINVAL;
      }
      for (auto& p : b->pipes) p->kblock = (int)value;
      if (b->grp) b->grp->kblock = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_SPD_SOLVE:
      for (auto& p : b->pipes) p->spd_solve = value ? 1 : 0;
      b->spd_solve = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_BATCH_GROUPED:
      b->grouped = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_AUTOTUNE:
      if (b->grp) b->grp->autotune = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_GRAPH:
      if (b->grp) b->grp->use_graph = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_FP64_SLICES:
      if (value != 0 && value != GROUP_OZ_SLICES) {
        b->err = "batches: int8 slices must be 0 or 5";
        return FEXPM_EINVAL;
      }
      if (value != 0 && b->dtype != FEXPM_COMPLEX64) {
        b->err = "int8-slice FP64 products apply to complex64 batches";
        return FEXPM_EINVAL;
      }
      for (auto& p : b->pipes) p->oz_slices = (int)value;
      if (b->grp) b->grp->oz_slices = (int)value;
      return FEXPM_OK;
    case FEXPM_OPT_GROUP_CANDIDATE:
      if (value < -1 || value >= GROUP_CANDIDATES) {
        b->err = "candidate must be in [-1, 5]";
        return FEXPM_EINVAL;
      }
      if (b->grp) b->grp->force_candidate((int)value);
      return FEXPM_OK;
    case FEXPM_OPT_FAIL_ALLOC:
      b->fail_alloc = value ? 1 : 0;
      return FEXPM_OK;
    case FEXPM_OPT_CHEBYSHEV:
      if (b->grp) b->grp->cheb = value ? 1 : 0;
      return FEXPM_OK;
  }
  b->err = "unknown or unsupported batch option";
  return FEXPM_EINVAL;
}

int fexpm_batch_upload(fexpm_batch* b, int64_t count, const int64_t* n, const void* a) {
  if (!b || count < 0 || (count > 0 && (!n || !a))) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(b->mtx);
  for (int64_t i = 0; i < count; ++i)
    if (n[i] < 1 || n[i] > b->nmax) {
      b->err = "order outside [1, nmax]";
      return FEXPM_EINVAL;
    }
  // From here on the batch is empty until every buffer and the copy are in
  // place; each capacity is zero while its buffer is being replaced, so a
  // failed allocation never leaves a capacity larger than the buffer.
  b->n.clear();
  b->off.clear();
  b->voff.clear();
  b->stats.clear();
  return guarded_b(b, [&] {
    std::vector<size_t> off((size_t)count), voff((size_t)count);
    size_t tot = 0, vt = 0;
    for (int64_t i = 0; i < count; ++i) {
      off[i] = tot;
      voff[i] = vt;
      tot += (size_t)n[i] * (size_t)n[i];
      vt += (size_t)n[i];
    }
    if (tot > b->cap_elems) {
      b->cap_elems = 0;
      if (b->d_in) (void)hipFree(b->d_in);
      if (b->d_out) (void)hipFree(b->d_out);
      b->d_in = b->d_out = nullptr;
      if (b->fail_alloc) {  // testing option: fail here, after the old buffers are gone
        b->fail_alloc = 0;
        throw Error(FEXPM_ENOMEM, "injected allocation failure (testing option)");
      }
      FX_HIP(hipMalloc(&b->d_in, tot * b->esize));
      FX_HIP(hipMalloc(&b->d_out, tot * b->esize));
      b->cap_elems = tot;
    }
    if (vt > b->cap_states) {
      b->cap_states = 0;
      if (b->d_psi0) (void)hipFree(b->d_psi0);
      if (b->d_psi) (void)hipFree(b->d_psi);
      b->d_psi0 = b->d_psi = nullptr;
      FX_HIP(hipMalloc(&b->d_psi0, vt * sizeof(double2)));
      FX_HIP(hipMalloc(&b->d_psi, vt * sizeof(double2)));
      b->cap_states = vt;
    }
    if ((size_t)count > b->cap_walks) {
      b->cap_walks = 0;
      if (b->d_scal) (void)hipFree(b->d_scal);
      if (b->d_info) (void)hipFree(b->d_info);
      b->d_scal = nullptr;
      b->d_info = nullptr;
      FX_HIP(hipMalloc(&b->d_scal, (size_t)count * sizeof(Scalars)));
      FX_HIP(hipMalloc(&b->d_info, (size_t)count * sizeof(int)));
      b->cap_walks = (size_t)count;
    }
    while (b->ev.size() < 2 * (size_t)count) {
      hipEvent_t e;
      FX_HIP(hipEventCreate(&e));
      b->ev.push_back(e);
    }
    if (tot * b->esize >= HostStaging::kMin) {
      staged_h2d(b->stg, b->d_in, a, tot * b->esize, nullptr);
      FX_HIP(hipStreamSynchronize(nullptr));
    } else if (tot) {
      FX_HIP(hipMemcpy(b->d_in, a, tot * b->esize, hipMemcpyHostToDevice));
    }
    std::vector<int64_t> ns(n, n + count);
    if (b->grp) b->grp->set_layout(ns, off, b->streams[0]);
    b->n.swap(ns);
    b->off.swap(off);
    b->voff.swap(voff);
    b->stats.assign(count, fexpm_stats());
  });
}

int64_t fexpm_batch_count(const fexpm_batch* b) { return b ? (int64_t)b->n.size() : -1; }

// psi_i = U_i psi0_i on lane pipeline p after its run of walk i; a walk with
// non-finite input has no U (its output is NaN), so psi_i is NaN too.
static void batch_apply(fexpm_batch* b, PipelineBase* p, int64_t i, const fexpm_stats& S, hipStream_t st) {
  if (S.nonfinite) {
    const int64_t cnt = b->n[i];
    fill_nan<double2><<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(b->d_psi + b->voff[i], cnt);
    FX_LAUNCH();
    return;
  }
  p->apply(b->d_psi0 + b->voff[i], 1, b->d_psi + b->voff[i], st);
}

// Grouped run of an uploaded batch of complex walks (see fexpm_batch_report):
// one analysis over all walks and one read-back of the plans; CTQW-form
// walks then run as the grouped program on stream 0, other walks on lane
// pipeline 0 afterwards.
static void batch_run_grouped(fexpm_batch* b, const void* psi0, void* psi) {
  const int64_t count = (int64_t)b->n.size();
  const size_t vt = b->voff.back() + (size_t)b->n.back();
  hipStream_t st = b->streams[0];
  const auto t0 = std::chrono::steady_clock::now();
  if (psi0) FX_HIP(hipMemcpy(b->d_psi0, psi0, vt * sizeof(double2), hipMemcpyHostToDevice));
  b->grp->analyze(b->d_in, b->d_scal, st);
  std::vector<Scalars> hs((size_t)count);
  FX_HIP(hipMemcpyAsync(hs.data(), b->d_scal, (size_t)count * sizeof(Scalars), hipMemcpyDeviceToHost, st));
  FX_HIP(hipStreamSynchronize(st));
  std::vector<GroupPlan> plans;
  std::vector<int64_t> rest;
  const bool is_complex = b->dtype == FEXPM_COMPLEX64 || b->dtype == FEXPM_COMPLEX128;
  for (int64_t i = 0; i < count; ++i) {
    const Scalars& h = hs[(size_t)i];
    const bool nonfinite = (h.flags & FLAG_NONFINITE) || !std::isfinite(h.norm_plain);
    const bool structured = is_complex && !(h.flags & (FLAG_NONZERO_RE | FLAG_ASYM));
    if (nonfinite || !structured) {
      rest.push_back(i);
      continue;
    }
    GroupPlan P;
    std::memset(&P, 0, sizeof(P));
    P.idx = (int)i;
    P.shifted = h.norm_shift <= h.norm_plain ? 1 : 0;
    P.mu_re = P.shifted ? h.mu_re : 0.0;
    P.mu_im = P.shifted ? h.mu_im : 0.0;
    P.norm = P.shifted ? h.norm_shift : h.norm_plain;
    P.norm_plain = h.norm_plain;
    choose_degree(P.norm, P.m, P.s);
    plans.push_back(P);
  }
  const double plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  GroupRunnerBase& G = *b->grp;
  G.run(plans, b->d_in, b->d_out, b->d_scal, psi0 ? b->d_psi0 : nullptr, b->d_psi, b->voff, st);
  // walks outside the CTQW form: lane pipeline 0, after the grouped program
  PipelineBase* p = b->pipes[0].get();
  for (int64_t i : rest) {
    p->set_size((int)b->n[i], st);
    fexpm_stats S;
    std::memset(&S, 0, sizeof(S));
    S.n = b->n[i];
    S.dtype = b->dtype;
    FX_HIP(hipEventRecord(b->ev[2 * i], st));
    p->run_planned((const char*)b->d_in + b->off[i] * b->esize, (char*)b->d_out + b->off[i] * b->esize,
                   hs[(size_t)i], st, S, 1, b->d_info + i);
    if (psi0) batch_apply(b, p, i, S, st);
    FX_HIP(hipEventRecord(b->ev[2 * i + 1], st));
    b->stats[i] = S;
  }
  FX_HIP(hipStreamSynchronize(st));
  // exception bits: Cholesky pivots of grouped walks, lane bits of the rest
  FX_HIP(hipMemcpy(hs.data(), b->d_scal, (size_t)count * sizeof(Scalars), hipMemcpyDeviceToHost));
  int64_t failed = -1, out_of_range = -1;
  for (const GroupPlan& P : plans) {
    const int64_t i = P.idx;
    const int info = hs[(size_t)i].info;
    if ((info & INFO_SPLIT_RANGE) && out_of_range < 0) out_of_range = i;
    if ((info & ~INFO_SPLIT_RANGE) && failed < 0) failed = i;
    fexpm_stats& S = b->stats[i];
    std::memset(&S, 0, sizeof(S));
    S.n = b->n[i];
    S.dtype = b->dtype;
    S.m = P.m;
    S.s = P.s;
    S.structured = 1;
    S.shifted = P.shifted;
    S.mu_re = P.mu_re;
    S.mu_im = P.mu_im;
    S.norm = P.norm;
    S.norm_plain = P.norm_plain;
    S.ms_total = G.device_ms;  // device time of the whole grouped program
    S.flops_issued = (size_t)i < G.walk_flops.size() ? G.walk_flops[(size_t)i] : 0.0;
    S.kblock = G.kblock;
    S.pade_fp64 = b->dtype == FEXPM_COMPLEX64 ? 1 : 0;
    S.precision = b->dtype == FEXPM_COMPLEX64 ? G.precision : FEXPM_PREC_NATIVE;
    S.fp64_slices = b->dtype == FEXPM_COMPLEX64 ? G.oz_slices : 0;
    S.fp64_slices_scope = S.fp64_slices ? 1 : 0;
    if (G.cheb) {
      const ChebSeries cs = cheb_series(P.norm * std::ldexp(1.0, -P.s), cheb_tail(P.s, b->dtype == FEXPM_COMPLEX64));
      S.chebyshev = std::max(cs.dc, cs.ds);
    }
  }
  if (out_of_range >= 0)
    throw Error(FEXPM_EINVAL, "walk " + std::to_string(out_of_range) +
                                  ": split-precision squaring: a scaled operand value left the 16-bit range");
  if (failed >= 0)
    throw Error(FEXPM_ESINGULAR, "walk " + std::to_string(failed) + ": Pade denominator not positive definite");
  for (int64_t i : rest) {
    int info = 0;
    FX_HIP(hipMemcpy(&info, b->d_info + i, sizeof(int), hipMemcpyDeviceToHost));
    float ms = 0.f;
    FX_HIP(hipEventElapsedTime(&ms, b->ev[2 * i], b->ev[2 * i + 1]));
    b->stats[i].ms_total = ms;
    if (info != 0) throw Error(FEXPM_ESINGULAR, "walk " + std::to_string(i) + ": Pade denominator singular");
  }
  if (psi0) FX_HIP(hipMemcpy(psi, b->d_psi, vt * sizeof(double2), hipMemcpyDeviceToHost));
  fexpm_batch_report& R = b->report;
  std::memset(&R, 0, sizeof(R));
  R.grouped = 1;
  R.grouped_walks = (int32_t)plans.size();
  R.fallback_walks = (int32_t)rest.size();
  R.launches = G.launches;
  R.graph = G.graph_used;
  R.graph_built = G.graph_built;
  R.tuned = G.tuned;
  for (int c = 0; c < GC_COUNT; ++c) {
    R.candidate[c] = G.cand[c];
    R.class_ms[c] = G.class_ms[c];
  }
  R.tune_ms = G.tuned ? G.tune_ms : 0.0;
  R.device_ms = G.device_ms;
  R.plan_ms = plan_ms;
}

// U_i = expm(A_i) for the uploaded batch on the device; with psi0 (packed
// complex128 host vectors, length n_i each) also psi_i = U_i psi0_i, copied
// to psi.  Walks are assigned to lanes by decreasing n^3 (greedy least
// loaded); every lane runs its walks on its own stream from its own host
// thread, so small walks of different lanes overlap on the device.
int fexpm_batch_run_resident(fexpm_batch* b, const void* psi0, void* psi) {
  if (!b || (psi0 && !psi)) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(b->mtx);
  const int64_t count = (int64_t)b->n.size();
  if (count == 0) return FEXPM_OK;
  return guarded_b(b, [&] {
    if (b->grp && b->grouped && b->spd_solve) {
      const auto t0 = std::chrono::steady_clock::now();
      batch_run_grouped(b, psi0, psi);
      b->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      return;
    }
    std::memset(&b->report, 0, sizeof(b->report));
    size_t vt = b->voff.back() + (size_t)b->n.back();
    if (psi0) FX_HIP(hipMemcpy(b->d_psi0, psi0, vt * sizeof(double2), hipMemcpyHostToDevice));
    std::vector<int64_t> order(count);
    for (int64_t i = 0; i < count; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return b->n[x] > b->n[y]; });
    std::vector<std::vector<int64_t>> assign(b->lanes);
    std::vector<double> load(b->lanes, 0.0);
    for (int64_t i : order) {
      int best = 0;
      for (int l = 1; l < b->lanes; ++l)
        if (load[l] < load[best]) best = l;
      assign[best].push_back(i);
      load[best] += (double)b->n[i] * b->n[i] * b->n[i];
    }
    std::vector<std::string> errs(b->lanes);
    std::vector<int> codes(b->lanes, FEXPM_OK);
    std::vector<double> lane_enqueue_ms(b->lanes, 0.0), lane_done_ms(b->lanes, 0.0), lane_phase1_ms(b->lanes, 0.0);
    auto t0 = std::chrono::steady_clock::now();
    // Two phases per lane: the analyses of all its walks (one host wait),
    // then every walk's stages enqueued back to back with no host wait in
    // between; exception bits and times are read after the lane drains.
    auto worker = [&](int l) {
      try {
        FX_HIP(hipSetDevice(0));
        PipelineBase* p = b->pipes[l].get();
        hipStream_t st = b->streams[l];
        const std::vector<int64_t>& mine = assign[l];
        for (int64_t i : mine) {
          p->set_size((int)b->n[i], st);
          p->analyze((const char*)b->d_in + b->off[i] * b->esize, b->d_scal + i, st);
        }
        std::vector<Scalars> hs(mine.size());
        for (size_t q = 0; q < mine.size(); ++q)
          FX_HIP(hipMemcpyAsync(&hs[q], b->d_scal + mine[q], sizeof(Scalars), hipMemcpyDeviceToHost, st));
        FX_HIP(hipStreamSynchronize(st));
        const auto t_lane0 = std::chrono::steady_clock::now();
        lane_phase1_ms[l] = std::chrono::duration<double, std::milli>(t_lane0 - t0).count();
        for (size_t q = 0; q < mine.size(); ++q) {
          const int64_t i = mine[q];
          p->set_size((int)b->n[i], st);
          fexpm_stats S;
          std::memset(&S, 0, sizeof(S));
          S.n = b->n[i];
          S.dtype = b->dtype;
          FX_HIP(hipEventRecord(b->ev[2 * i], st));
          p->run_planned((const char*)b->d_in + b->off[i] * b->esize, (char*)b->d_out + b->off[i] * b->esize, hs[q],
                         st, S, 1, b->d_info + i);
          if (psi0) batch_apply(b, p, i, S, st);
          FX_HIP(hipEventRecord(b->ev[2 * i + 1], st));
          b->stats[i] = S;
        }
        const auto t_enq = std::chrono::steady_clock::now();
        FX_HIP(hipStreamSynchronize(st));
        const auto t_done = std::chrono::steady_clock::now();
        lane_enqueue_ms[l] = std::chrono::duration<double, std::milli>(t_enq - t_lane0).count();
        lane_done_ms[l] = std::chrono::duration<double, std::milli>(t_done - t_lane0).count();
        for (int64_t i : mine) {
          int info = 0;
          FX_HIP(hipMemcpy(&info, b->d_info + i, sizeof(int), hipMemcpyDeviceToHost));
          float ms = 0.f;
          FX_HIP(hipEventElapsedTime(&ms, b->ev[2 * i], b->ev[2 * i + 1]));
          b->stats[i].ms_total = ms;
          if (info & INFO_SPLIT_RANGE)
            throw Error(FEXPM_EINVAL, "walk " + std::to_string(i) + ": split-precision operand out of range");
          if (info != 0) throw Error(FEXPM_ESINGULAR, "walk " + std::to_string(i) + ": Pade denominator singular");
        }
      } catch (const Error& x) {
        errs[l] = x.what();
        codes[l] = x.code;
      } catch (const std::exception& x) {
        errs[l] = x.what();
        codes[l] = FEXPM_EHIP;
      }
    };
    std::vector<std::thread> th;
    for (int l = 1; l < b->lanes; ++l) th.emplace_back(worker, l);
    worker(0);
    for (auto& x : th) x.join();
    b->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    // FEXPM_BATCH_TRACE=1: per-lane host timeline (analysis done, stages
    // enqueued, lane drained) on stderr, to tell host-bound from device-bound
    if (std::getenv("FEXPM_BATCH_TRACE"))
      for (int l = 0; l < b->lanes; ++l)
        std::fprintf(stderr, "lane %d: phase 1 done at %.1f ms; phase-2 enqueue %.1f ms, drained %.1f ms\n", l,
                     lane_phase1_ms[l], lane_enqueue_ms[l], lane_done_ms[l]);
    for (int l = 0; l < b->lanes; ++l)
      if (codes[l] != FEXPM_OK) throw Error(codes[l], "lane " + std::to_string(l) + ": " + errs[l]);
    if (psi0) FX_HIP(hipMemcpy(psi, b->d_psi, vt * sizeof(double2), hipMemcpyDeviceToHost));
  });
}

int fexpm_batch_download(fexpm_batch* b, int64_t i, void* u) {
  if (!b || !u) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(b->mtx);
  if (i < 0 || i >= (int64_t)b->n.size()) return FEXPM_EINVAL;
  return guarded_b(b, [&] {
    FX_HIP(hipMemcpy(u, (const char*)b->d_out + b->off[i] * b->esize, (size_t)b->n[i] * b->n[i] * b->esize,
                     hipMemcpyDeviceToHost));
  });
}

int fexpm_batch_get_stats(const fexpm_batch* b, int64_t i, fexpm_stats* out) {
  if (!b || !out || i < 0 || i >= (int64_t)b->stats.size()) return FEXPM_EINVAL;
  *out = b->stats[i];
  return FEXPM_OK;
}

double fexpm_batch_last_ms(const fexpm_batch* b) { return b ? b->last_ms : -1.0; }

int fexpm_batch_get_report(const fexpm_batch* b, fexpm_batch_report* out) {
  if (!b || !out) return FEXPM_EINVAL;
  *out = b->report;
  return FEXPM_OK;
}

}  // extern "C"

// ===========================================================================
// Testing ABI (include/fastexpm_testing.h)
namespace {

template <typename W> struct DevPlane {
  W* p = nullptr;
  explicit DevPlane(size_t cnt) { FX_HIP(hipMalloc(&p, std::max<size_t>(cnt, 1) * sizeof(W))); FX_HIP(hipMemset(p, 0, std::max<size_t>(cnt, 1) * sizeof(W))); }
  ~DevPlane() { if (p) (void)hipFree(p); }
  void put(const double* h, int rows, int cols, int64_t ld) {
    if (!h) return;
    std::vector<W> b((size_t)rows * ld, (W)0);
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) b[(size_t)i * ld + j] = (W)h[(size_t)i * cols + j];
    FX_HIP(hipMemcpy(p, b.data(), b.size() * sizeof(W), hipMemcpyHostToDevice));
  }
  void get(double* h, int rows, int cols, int64_t ld) {
    if (!h) return;
    std::vector<W> b((size_t)rows * ld);
    FX_HIP(hipMemcpy(b.data(), p, b.size() * sizeof(W), hipMemcpyDeviceToHost));
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) h[(size_t)i * cols + j] = (double)b[(size_t)i * ld + j];
  }
};

template <typename W, int MODE, bool SYM>
void test_gemm_t(int m, int n, int k, const double* ar, const double* ai, const double* br, const double* bi,
                 double* cr, double* ci) {
  const int64_t kp = round_up(std::max(k, 1), 16);
  const int64_t ldc = std::max(n, 1);
  DevPlane<W> dar((size_t)m * kp), dai((size_t)m * kp), dbr((size_t)n * kp), dbi((size_t)n * kp);
  DevPlane<W> dcr((size_t)m * ldc), dci((size_t)m * ldc);
  dar.put(ar, m, k, kp);
  dai.put(ai, m, k, kp);
  dbr.put(br, n, k, kp);
  dbi.put(bi, n, k, kp);
  GemmArgs<W> g;
  std::memset(&g, 0, sizeof(g));
  g.ar = dar.p; g.ai = dai.p; g.lda = kp;
  g.br = dbr.p; g.bi = dbi.p; g.ldb = kp;
  g.m = m; g.n = n; g.k = (int)kp;
  g.epi.cr = dcr.p; g.epi.ci = dci.p; g.epi.ldc = ldc; g.epi.alpha = 1.0; g.epi.ld_add = ldc;
  using C = GemmCfg<W>;
  constexpr int NT = (C::BM / C::WM) * (C::BN / C::WN) * 64;
  if (SYM) {
    const int T = (m + C::BM - 1) / C::BM;
    gemm_nt_kernel<W, MODE, true><<<dim3(T * (T + 1) / 2), dim3(NT)>>>(g);
  } else {
    dim3 grid((n + C::BN - 1) / C::BN, (m + C::BM - 1) / C::BM);
    gemm_nt_kernel<W, MODE, false><<<grid, dim3(NT)>>>(g);
  }
  FX_LAUNCH();
  FX_HIP(hipDeviceSynchronize());
  dcr.get(cr, m, n, ldc);
  if (MODE != MODE_REAL) dci.get(ci, m, n, ldc);
}

template <typename W>
void test_gemm_w(int mode, int sym, int m, int n, int k, const double* ar, const double* ai, const double* br,
                 const double* bi, double* cr, double* ci) {
  if (mode == MODE_REAL && !sym) test_gemm_t<W, MODE_REAL, false>(m, n, k, ar, ai, br, bi, cr, ci);
  else if (mode == MODE_REAL) test_gemm_t<W, MODE_REAL, true>(m, n, k, ar, ai, br, bi, cr, ci);
  else if (mode == MODE_C4M && !sym) test_gemm_t<W, MODE_C4M, false>(m, n, k, ar, ai, br, bi, cr, ci);
  else if (mode == MODE_C3M && !sym) test_gemm_t<W, MODE_C3M, false>(m, n, k, ar, ai, br, bi, cr, ci);
  else if (mode == MODE_C3M) test_gemm_t<W, MODE_C3M, true>(m, n, k, ar, ai, br, bi, cr, ci);
  else if (mode == MODE_SQ2 && sym) test_gemm_t<W, MODE_SQ2, true>(m, n, k, ar, ai, br, bi, cr, ci);
  else throw Error(FEXPM_EINVAL, "unsupported gemm variant");
}

// C = A A^T (complex, m x m from m x k) through the split-precision
// squaring kernel with a zero diagonal split; the diagonal of the product
// comes back through d_out.
template <class P>
void test_split_t(int m, int k, const double* ar, const double* ai, double* cr, double* ci) {
  const int kpad = (int)round_up(k, 16), rows = (int)round_up(m, SQ_BM);
  const int64_t ld = std::max(kpad, rows);
  std::vector<float> hr((size_t)rows * ld, 0.0f), hi((size_t)rows * ld, 0.0f);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < k; ++j) {
      hr[(size_t)i * ld + j] = (float)ar[(size_t)i * k + j];
      hi[(size_t)i * ld + j] = (float)ai[(size_t)i * k + j];
    }
  DevPlane<float> yr(hr.size()), yi(hi.size()), outr((size_t)rows * ld), outi((size_t)rows * ld);
  DevPlane<double2> din(rows), dout(rows);
  DevPlane<typename P::H> parts((size_t)rows * kpad * 2 * P::NQ);
  DevPlane<int> info(1);
  FX_HIP(hipMemcpy(yr.p, hr.data(), hr.size() * 4, hipMemcpyHostToDevice));
  FX_HIP(hipMemcpy(yi.p, hi.data(), hi.size() * 4, hipMemcpyHostToDevice));
  FX_HIP(hipMemset(din.p, 0, rows * sizeof(double2)));
  FX_HIP(hipMemset(info.p, 0, sizeof(int)));
  dim3 sg((unsigned)((kpad / 16 + 15) / 16), (unsigned)((rows + 15) / 16));
  split_blocked<P><<<sg, 256>>>(yr.p, yi.p, ld, m, k, rows, kpad, parts.p, info.p);
  FX_LAUNCH();
  SqArgs<P> a;
  a.a = parts.p;
  a.rows = rows;
  a.yr = yr.p;
  a.yi = yi.p;
  a.ldy = ld;
  a.cr = outr.p;
  a.ci = outi.p;
  a.ldc = ld;
  a.d_in = din.p;
  a.d_out = dout.p;
  a.m = m;
  a.k = kpad;
  a.kblock = 8;
  const int T = (m + SQ_BM - 1) / SQ_BM;
  std::vector<int> ho((size_t)T * (T + 1) / 2);
  sym_tile_order(T, SQ_GROUP, ho.data());
  DevPlane<int> order(ho.size());
  FX_HIP(hipMemcpy(order.p, ho.data(), ho.size() * sizeof(int), hipMemcpyHostToDevice));
  a.order = order.p;
  sq_split_kernel<P><<<dim3(T * (T + 1) / 2), dim3(sq_threads<P>())>>>(a);
  FX_LAUNCH();
  FX_HIP(hipDeviceSynchronize());
  int hinfo = 0;
  FX_HIP(hipMemcpy(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost));
  if (hinfo) throw Error(FEXPM_EINVAL, "operand outside the 16-bit range");
  std::vector<float> orr((size_t)rows * ld), oi((size_t)rows * ld);
  std::vector<double2> dd(rows);
  FX_HIP(hipMemcpy(orr.data(), outr.p, orr.size() * 4, hipMemcpyDeviceToHost));
  FX_HIP(hipMemcpy(oi.data(), outi.p, oi.size() * 4, hipMemcpyDeviceToHost));
  FX_HIP(hipMemcpy(dd.data(), dout.p, rows * sizeof(double2), hipMemcpyDeviceToHost));
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      cr[(size_t)i * m + j] = i == j ? dd[i].x : (double)orr[(size_t)i * ld + j];
      ci[(size_t)i * m + j] = i == j ? dd[i].y : (double)oi[(size_t)i * ld + j];
    }
}

}  // namespace

extern "C" {

int fexpm_test_gemm(int wtype, int mode, int sym, int m, int n, int k, const double* ar, const double* ai,
                    const double* br, const double* bi, double* cr, double* ci) {
  if (m <= 0 || n <= 0 || k <= 0 || !ar || !br || !cr || wtype < 0 || wtype > 3) return FEXPM_EINVAL;
  if (mode != MODE_REAL && (!ai || !bi || !ci)) return FEXPM_EINVAL;
  if (sym && m != n) return FEXPM_EINVAL;
  if (wtype >= 2 && (mode != MODE_C3M || !sym)) return FEXPM_EINVAL;
  return guarded(nullptr, [&] {
    check_device();
    if (wtype == 0) test_gemm_w<float>(mode, sym, m, n, k, ar, ai, br, bi, cr, ci);
    else if (wtype == 1) test_gemm_w<double>(mode, sym, m, n, k, ar, ai, br, bi, cr, ci);
    else if (wtype == 2) test_split_t<Fp16x2>(m, k, ar, ai, cr, ci);
    else test_split_t<Bf16x3>(m, k, ar, ai, cr, ci);
  });
}

int fexpm_test_lu_solve(int n, const double* mr, const double* mi, const double* br, const double* bi,
                        double* xr, double* xi, int* info) {
  if (n <= 0 || !mr || !mi || !br || !bi || !xr || !xi) return FEXPM_EINVAL;
  return guarded(nullptr, [&] {
    check_device();
    const int np = (int)round_up(n, 128);
    const int64_t ldl = 2 * (int64_t)np;
    std::vector<double> hr((size_t)np * ldl, 0.0), hi((size_t)np * ldl, 0.0);
    for (int i = 0; i < np; ++i) {
      for (int j = 0; j < np; ++j) {
        if (i < n && j < n) {
          hr[(size_t)i * ldl + j] = mr[(size_t)i * n + j];
          hi[(size_t)i * ldl + j] = mi[(size_t)i * n + j];
          hr[(size_t)i * ldl + np + j] = br[(size_t)i * n + j];
          hi[(size_t)i * ldl + np + j] = bi[(size_t)i * n + j];
        } else if (i == j) {
          hr[(size_t)i * ldl + j] = 1.0;
        }
      }
    }
    DevPlane<double> lr((size_t)np * ldl), li((size_t)np * ldl), tr((size_t)2 * np * LU_OUTER),
        ti((size_t)2 * np * LU_OUTER), top(4 * LU_NB);
    int* ipiv = nullptr;
    PanelCand* cand = nullptr;
    Scalars* scal = nullptr;
    const int gmax = (np + PANEL_R - 1) / PANEL_R;
    FX_HIP(hipMalloc(&ipiv, np * sizeof(int)));
    FX_HIP(hipMalloc(&cand, 2 * gmax * sizeof(PanelCand)));
    FX_HIP(hipMalloc(&scal, sizeof(Scalars)));
    FX_HIP(hipMemset(scal, 0, sizeof(Scalars)));
    FX_HIP(hipMemcpy(lr.p, hr.data(), hr.size() * 8, hipMemcpyHostToDevice));
    FX_HIP(hipMemcpy(li.p, hi.data(), hi.size() * 8, hipMemcpyHostToDevice));
    DevPlane<double> inv(2 * LU_NB * LU_NB);
    unsigned* ctrs = nullptr;
    FX_HIP(hipMalloc(&ctrs, (np / LU_NB) * sizeof(unsigned)));
    lu_factor_solve(lr.p, li.p, ldl, np, tr.p, ti.p, ipiv, cand, gmax, top.p, scal, inv.p, ctrs, nullptr);
    FX_HIP(hipDeviceSynchronize());
    (void)hipFree(ctrs);
    FX_HIP(hipDeviceSynchronize());
    FX_HIP(hipMemcpy(hr.data(), lr.p, hr.size() * 8, hipMemcpyDeviceToHost));
    FX_HIP(hipMemcpy(hi.data(), li.p, hi.size() * 8, hipMemcpyDeviceToHost));
    Scalars hs;
    FX_HIP(hipMemcpy(&hs, scal, sizeof(Scalars), hipMemcpyDeviceToHost));
    (void)hipFree(ipiv);
    (void)hipFree(cand);
    (void)hipFree(scal);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        xr[(size_t)i * n + j] = hr[(size_t)i * ldl + np + j];
        xi[(size_t)i * n + j] = hi[(size_t)i * ldl + np + j];
      }
    if (info) *info = hs.info;
  });
}

int fexpm_debug_hamiltonian(fexpm_engine* e, double* h) {
  if (!e || (e->n > 0 && !h)) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->n == 0) return FEXPM_OK;
  if (!e->have_graph) {
    e->err = "no graph: call fexpm_set_graph first";
    return FEXPM_EINVAL;
  }
  return guarded(e, [&] {
    FX_HIP(hipStreamSynchronize(e->stream));
    FX_HIP(hipMemcpy(h, e->d_h, (size_t)e->n * e->n * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int fexpm_debug_plane(fexpm_engine* e, int which, double* re, double* im) {
  if (!e || !re || !im) return FEXPM_EINVAL;
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->freed) return FEXPM_EFREED;
  if (e->n == 0) return FEXPM_OK;
  return guarded(e, [&] { e->pipe->debug_plane(which, re, im); });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FP64 product by int8 slices (testing ABI): C = A B^T on host data.
namespace {
struct OzBufs {
  double *da = nullptr, *db = nullptr, *dc = nullptr;
  int8_t *sa = nullptr, *sb = nullptr;
  int *ea = nullptr, *eb = nullptr;
  ~OzBufs() {
    for (void* p : {(void*)da, (void*)db, (void*)dc, (void*)sa, (void*)sb, (void*)ea, (void*)eb})
      if (p) (void)hipFree(p);
  }
};

template <int S>
void oz_run_t(int sym, int m, int n, int k, const double* a, const double* b, double* c) {
  const int kpad = (int)round_up(k, OZ_BK), ra = (int)round_up(m, OZ_BM), rb = (int)round_up(n, OZ_BM);
  OzBufs B;
  FX_HIP(hipMalloc(&B.da, (size_t)m * k * sizeof(double)));
  FX_HIP(hipMalloc(&B.db, (size_t)n * k * sizeof(double)));
  FX_HIP(hipMalloc(&B.dc, (size_t)m * n * sizeof(double)));
  FX_HIP(hipMalloc(&B.sa, (size_t)ra * kpad * S));
  FX_HIP(hipMalloc(&B.sb, (size_t)rb * kpad * S));
  FX_HIP(hipMalloc(&B.ea, (size_t)ra * sizeof(int)));
  FX_HIP(hipMalloc(&B.eb, (size_t)rb * sizeof(int)));
  FX_HIP(hipMemcpy(B.da, a, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice));
  FX_HIP(hipMemcpy(B.db, b, (size_t)n * k * sizeof(double), hipMemcpyHostToDevice));
  oz_slice<S><<<ra, 256>>>(B.da, 1.0, nullptr, 0.0, k, k, m, k, ra, kpad, B.sa, B.ea);
  oz_slice<S><<<rb, 256>>>(B.db, 1.0, nullptr, 0.0, k, k, n, k, rb, kpad, B.sb, B.eb);
  FX_LAUNCH();
  OzArgs<S> g;
  std::memset(&g, 0, sizeof(g));
  g.a = B.sa;
  g.ea = B.ea;
  g.rows_a = ra;
  g.b = sym ? B.sa : B.sb;
  g.eb = sym ? B.ea : B.eb;
  g.rows_b = sym ? ra : rb;
  g.m = m;
  g.n = n;
  g.kpad = kpad;
  g.epi.cr = B.dc;
  g.epi.ldc = n;
  g.epi.ld_add = n;
  g.epi.alpha = 1.0;
  if (sym) {
    const int T = ra / OZ_BM;
    oz_gemm_kernel<S, true><<<dim3(T * (T + 1) / 2), dim3(512)>>>(g);
  } else {
    oz_gemm_kernel<S, false><<<dim3(rb / OZ_BM, ra / OZ_BM), dim3(512)>>>(g);
  }
  FX_LAUNCH();
  FX_HIP(hipMemcpy(c, B.dc, (size_t)m * n * sizeof(double), hipMemcpyDeviceToHost));
  FX_HIP(hipDeviceSynchronize());
}
}  // namespace

// The residue form (slices == CRT64_K): the engine's product routine of an
// FP64 pipeline of order m, operands A (and B when b != a) on the device.
static void crt64_run(int m, int k, const double* a, const double* b, double* c) {
  Pipeline<double> P(m, FEXPM_COMPLEX128, false);
  double *da = nullptr, *db = nullptr, *dc = nullptr;
  struct Free {
    double** p[3];
    ~Free() {
      for (double** q : p) (void)hipFree(*q);
    }
  } fr{{&da, &db, &dc}};
  const bool same = a == b;
  FX_HIP(hipMalloc(&da, (size_t)m * k * sizeof(double)));
  if (!same) FX_HIP(hipMalloc(&db, (size_t)m * k * sizeof(double)));
  FX_HIP(hipMalloc(&dc, (size_t)m * m * sizeof(double)));
  FX_HIP(hipMemcpy(da, a, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice));
  if (!same) FX_HIP(hipMemcpy(db, b, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice));
  Epilogue<double> e;
  std::memset(&e, 0, sizeof(e));
  e.cr = dc;
  e.ldc = m;
  e.ld_add = m;
  e.alpha = 1.0;
  using Op = Pipeline<double>::OzOp;
  const Op A{da, 1.0, nullptr, 0.0, k, k};
  const Op B{db, 1.0, nullptr, 0.0, k, k};
  P.crt64_product(A, same ? nullptr : &B, m, k, e, nullptr);
  FX_HIP(hipDeviceSynchronize());
  FX_HIP(hipMemcpy(c, dc, (size_t)m * m * sizeof(double), hipMemcpyDeviceToHost));
}

extern "C" int fexpm_test_ozaki_gemm(int slices, int sym, int m, int n, int k, const double* a, const double* b,
                                     double* c) {
  if (slices == CRT64_K) {
    if (!a || !b || !c || !sym || m <= 0 || m != n || k <= 0 || round_up(k, CRT_KPAD) > CRT64_KMAX)
      return FEXPM_EINVAL;
    return guarded(nullptr, [&] {
      check_device();
      crt64_run(m, k, a, b, c);
    });
  }
  if (!a || !b || !c || m <= 0 || n <= 0 || k <= 0 || (sym && m != n) || slices < 3 || slices > OZ_MAX_SLICES)
    return FEXPM_EINVAL;
  return guarded(nullptr, [&] {
    check_device();
    switch (slices) {
      case 3: oz_run_t<3>(sym, m, n, k, a, b, c); break;
      case 4: oz_run_t<4>(sym, m, n, k, a, b, c); break;
      case 5: oz_run_t<5>(sym, m, n, k, a, b, c); break;
      default: oz_run_t<6>(sym, m, n, k, a, b, c); break;
    }
  });
}

// ---------------------------------------------------------------------------
// Grouped program description on a virtual layout (testing ABI): the same
// builder as the batch runs, no device memory touched.
namespace {
template <typename Tin, typename W>
std::string group_program_text(int precision, int fp64_slices, int64_t count, const int64_t* n, const int32_t* m,
                               const int32_t* s, const int32_t* shifted, const double* mu_re, const double* mu_im,
                               const double* norm, int with_psi) {
  GroupRunner<Tin, W> R(true);
  R.precision = precision;
  R.cheb = norm ? 1 : 0;
  if (fp64_slices >= 0) R.oz_slices = fp64_slices;
  std::vector<int64_t> nv(n, n + count);
  std::vector<size_t> off((size_t)count), voff((size_t)count);
  size_t tot = 0, vt = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (n[i] < 1 || n[i] > (1 << 16)) throw Error(FEXPM_EINVAL, "order outside [1, 65536]");
    if (s[i] < 0 || s[i] > 64 || (m[i] != 3 && m[i] != 5 && m[i] != 7 && m[i] != 9 && m[i] != 13))
      throw Error(FEXPM_EINVAL, "bad plan (m or s)");
    off[(size_t)i] = tot;
    voff[(size_t)i] = vt;
    tot += (size_t)n[i] * n[i];
    vt += (size_t)n[i];
  }
  R.set_layout(nv, off, nullptr);
  std::vector<GroupPlan> plans((size_t)count);
  for (int64_t i = 0; i < count; ++i) {
    GroupPlan& P = plans[(size_t)i];
    std::memset(&P, 0, sizeof(P));
    P.idx = (int)i;
    P.m = m[i];
    P.s = s[i];
    P.shifted = shifted[i] ? 1 : 0;
    P.mu_re = mu_re[i];
    P.mu_im = mu_im[i];
    P.norm = norm ? norm[i] : 0.0;
    if (norm && !(norm[i] >= 0.0 && std::isfinite(norm[i]))) throw Error(FEXPM_EINVAL, "bad plan (norm)");
  }
  R.build(plans, vaddr(VR_IN), vaddr(VR_OUT), reinterpret_cast<Scalars*>(vaddr(VR_SCAL)),
          with_psi ? reinterpret_cast<const double2*>(vaddr(VR_PSI0)) : nullptr,
          reinterpret_cast<double2*>(vaddr(VR_PSI)), voff);
  return R.describe();
}
}  // namespace

extern "C" int fexpm_test_group_program(int dtype, int precision, int fp64_slices, int64_t count, const int64_t* n,
                                        const int32_t* m, const int32_t* s, const int32_t* shifted,
                                        const double* mu_re, const double* mu_im, const double* norm, int with_psi,
                                        char* buf, int64_t cap, int64_t* len) {
  if (count < 1 || !n || !m || !s || !shifted || !mu_re || !mu_im || !len) return FEXPM_EINVAL;
  if (dtype != FEXPM_COMPLEX64 && dtype != FEXPM_COMPLEX128) return FEXPM_EINVAL;
  if (precision < FEXPM_PREC_NATIVE || precision > FEXPM_PREC_BF16X3 ||
      (precision != FEXPM_PREC_NATIVE && dtype != FEXPM_COMPLEX64) ||
      (fp64_slices != -1 && fp64_slices != 0 && (fp64_slices != GROUP_OZ_SLICES || dtype != FEXPM_COMPLEX64)))
    return FEXPM_EINVAL;
  try {
    const std::string t = dtype == FEXPM_COMPLEX64
                              ? group_program_text<float2, float>(precision, fp64_slices, count, n, m, s, shifted,
                                                                  mu_re, mu_im, norm, with_psi)
                              : group_program_text<double2, double>(precision, fp64_slices, count, n, m, s, shifted,
                                                                    mu_re, mu_im, norm, with_psi);
    *len = (int64_t)t.size();
    if (!buf || cap < (int64_t)t.size() + 1) return FEXPM_EINVAL;
    std::memcpy(buf, t.c_str(), t.size() + 1);
    return FEXPM_OK;
  } catch (const Error& x) {
    g_last_error = x.what();
    return x.code;
  } catch (const std::exception& x) {
    g_last_error = x.what();
    return FEXPM_EINVAL;
  }
}

// The tile order of the symmetric GEMM launches, evaluated on the host:
// out[b] = (ti << 16) | tj for b < T(T+1)/2 (no device needed).
extern "C" int fexpm_test_grouped_tile_order(int T, int* out) {
  if (!out || T < 1 || T > 4096) return FEXPM_EINVAL;
  const int nb = T * (T + 1) / 2;
  for (int b = 0; b < nb; ++b) {
    int ti = 0, tj = 0;
    grouped_tri_tile(b, T, ti, tj);
    out[b] = (ti << 16) | tj;
  }
  return FEXPM_OK;
}

// The Chebyshev polynomial of spectral plans, evaluated on the host.
extern "C" int fexpm_test_cheb_series(double theta, int* dc, int* ds, double* cw, double* sw) {
  if (!dc || !ds || !cw || !sw || !(theta >= 0.0) || !(theta <= 16.0)) return FEXPM_EINVAL;
  const ChebSeries c = cheb_series(theta);
  *dc = c.dc;
  *ds = c.ds;
  for (int j = 0; j < kChebMax; ++j) {
    cw[j] = j <= c.dc / 2 ? c.cw[j] : 0.0;
    sw[j] = j <= (c.ds - 1) / 2 ? c.sw[j] : 0.0;
  }
  return FEXPM_OK;
}

// The tile order of the rectangular GEMM launches (nbm x nbn tiles),
// evaluated on the host: out[lin] = (ti << 16) | tj.
extern "C" int fexpm_test_grouped_rect_order(int nbm, int nbn, int* out) {
  if (!out || nbm < 1 || nbn < 1 || nbm > 4096 || nbn > 4096) return FEXPM_EINVAL;
  for (int lin = 0; lin < nbm * nbn; ++lin) {
    int ti = 0, tj = 0;
    grouped_rect_tile(lin, nbm, nbn, ti, tj);
    out[lin] = (ti << 16) | tj;
  }
  return FEXPM_OK;
}

extern "C" int fexpm_test_square_issued(int64_t n, double* flops) {
  if (!flops || n < 1 || n > (int64_t)1 << 20) return FEXPM_EINVAL;
  return guarded(nullptr, [&] {
    check_device();
    const int64_t k = round_up(n, 16);
    const int tile = square_tile(n, 0);
    const SquarePlan p = square_plan((int)n, tile, true);
    *flops = 3.0 * 2.0 * ((double)p.full * tile * tile + (double)p.sub * 64 * 64) * (double)k;
  });
}

