// fastexpm measurement probes: device timing of single kernels, the
// sustained MFMA rates of the device and a plain copy of known size.
//
// A separate library (libfastexpm_probes.so) from the engine: nothing here
// is on the expm path.  The probes time the engine's kernel templates from
// the shared headers, including tilings the engine does not select, on
// operands filled with a deterministic pattern, so that a tiling choice or
// a rate quoted for a kernel can be measured again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fastexpm_probes.h"
#include "fexpm_gemm.hpp"
#include "fexpm_split.hpp"
#include "fexpm_ozaki.hpp"

using namespace fexpm;

namespace {

struct ProbeError : std::runtime_error {
  int code;
  ProbeError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FX_HIP(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess)                                                                          \
      throw ProbeError(e_ == hipErrorOutOfMemory ? FEXPM_ENOMEM : FEXPM_EHIP,                      \
                       std::string(#x) + ": " + hipGetErrorString(e_));                            \
  } while (0)
#define FX_LAUNCH() FX_HIP(hipGetLastError())

thread_local std::string g_probe_error;

template <typename F> int guarded(F&& f) {
  try {
    f();
    return FEXPM_OK;
  } catch (const ProbeError& x) {
    g_probe_error = x.what();
    return x.code;
  } catch (const std::exception& x) {
    g_probe_error = x.what();
    return FEXPM_EHIP;
  }
}

void check_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    throw ProbeError(FEXPM_ENODEVICE, "no HIP device available");
}

inline int64_t round_up(int64_t x, int64_t r) { return (x + r - 1) / r * r; }

// Tilings timed beside the engine's choices (the engine's own are in
// fexpm_gemm.hpp): the FP32 default tiling on 16x16x4 MFMAs, and FP64
// variants of the 64 x 64 tile.
using FCfg16 = TileCfg<128, 128, 16, 4, 64, 32, 16>;
using DCfg8w = TileCfg<64, 64, 16, 2, 32, 16>;           // 8 waves, 32x16 wave tiles
using DCfgTall = TileCfg<128, 64, 16, 2, 32, 32>;        // 8 waves, 128x64 workgroup tile
using DCfgWide = TileCfg<64, 128, 16, 2, 32, 32>;        // 8 waves, 64x128 workgroup tile
using DCfg8wPad4 = TileCfg<64, 64, 16, 4, 32, 16>;       // 4-double LDS row padding
using DCfg8wW4 = TileCfg<64, 64, 16, 2, 32, 16, 16, 4>;  // within 128 VGPRs (4 waves per SIMD)

// The int8-slice product on 64 x 64 tiles of 4 waves (32 x 32 each), 48 KB
// of LDS for S = 6: three workgroups per CU; tiles in 8 x 8 groups.  Slower
// than the engine's 128 x 128 tiles at n = 10^4.
template <int S, bool SYM>
__global__ __launch_bounds__(256, 3) void oz_gemm_kernel64(OzArgs<S> g) {
  int tm, tn;
  if (SYM) {
    grouped_tri_tile(xcd_remap((int)blockIdx.x, (int)gridDim.x), g.rows_a / 64, tm, tn);
  } else {
    const int lin = xcd_remap((int)(blockIdx.y * gridDim.x + blockIdx.x), (int)(gridDim.x * gridDim.y));
    grouped_rect_tile(lin, (int)gridDim.y, (int)gridDim.x, tm, tn);
  }
  oz_tile<S, 64, 2, 2>(g, tm, tn, SYM);
}

template <typename W>
__global__ void fill_pattern(W* p, size_t cnt, uint32_t seed) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  uint32_t h = (uint32_t)(k * 2654435761u) ^ seed;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  p[k] = (W)((double)(h & 0xFFFFFF) / 16777216.0 - 0.5);
}

template <typename W, int MODE, bool SYM, bool KBLK, class CFG>
double time_gemm_t(int m, int n, int k, int reps) {
  const int64_t kp = round_up(std::max(k, 1), 16);
  const size_t ca = (size_t)m * kp, cb = (size_t)n * kp, cc = (size_t)m * n;
  W *ar, *ai, *br, *bi, *cr, *ci;
  FX_HIP(hipMalloc(&ar, ca * sizeof(W)));
  FX_HIP(hipMalloc(&ai, ca * sizeof(W)));
  FX_HIP(hipMalloc(&br, cb * sizeof(W)));
  FX_HIP(hipMalloc(&bi, cb * sizeof(W)));
  FX_HIP(hipMalloc(&cr, cc * sizeof(W)));
  FX_HIP(hipMalloc(&ci, cc * sizeof(W)));
  fill_pattern<W><<<(unsigned)((ca + 255) / 256), 256>>>(ar, ca, 1u);
  fill_pattern<W><<<(unsigned)((ca + 255) / 256), 256>>>(ai, ca, 2u);
  fill_pattern<W><<<(unsigned)((cb + 255) / 256), 256>>>(br, cb, 3u);
  fill_pattern<W><<<(unsigned)((cb + 255) / 256), 256>>>(bi, cb, 4u);
  fill_pattern<W><<<(unsigned)((cc + 255) / 256), 256>>>(cr, cc, 5u);
  fill_pattern<W><<<(unsigned)((cc + 255) / 256), 256>>>(ci, cc, 6u);
  GemmArgs<W> g;
  std::memset(&g, 0, sizeof(g));
  g.ar = ar; g.ai = ai; g.lda = kp; g.br = br; g.bi = bi; g.ldb = kp;
  g.m = m; g.n = n; g.k = (int)kp; g.kblock = 8;
  g.epi.cr = cr; g.epi.ci = ci; g.epi.ldc = n; g.epi.alpha = -1.0; g.epi.beta = SYM ? 0.0 : 1.0;
  dim3 grid, block(CFG::NT);
  if (SYM) {
    const int T = (m + CFG::BM - 1) / CFG::BM;
    grid = dim3(T * (T + 1) / 2);
  } else {
    grid = dim3((n + CFG::BN - 1) / CFG::BN, (m + CFG::BM - 1) / CFG::BM);
  }
  hipEvent_t e0, e1;
  FX_HIP(hipEventCreate(&e0));
  FX_HIP(hipEventCreate(&e1));
  gemm_nt_kernel<W, MODE, SYM, KBLK, CFG><<<grid, block>>>(g);  // warm-up
  FX_LAUNCH();
  FX_HIP(hipEventRecord(e0));
  for (int r = 0; r < reps; ++r) gemm_nt_kernel<W, MODE, SYM, KBLK, CFG><<<grid, block>>>(g);
  FX_HIP(hipEventRecord(e1));
  FX_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (W* p : {ar, ai, br, bi, cr, ci}) (void)hipFree(p);
  return (double)ms / std::max(reps, 1);
}

template <class P>
__global__ void fill_pattern16(typename P::H* p, size_t cnt, uint32_t seed) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  uint32_t h = (uint32_t)(k * 2654435761u) ^ seed;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  p[k] = (typename P::H)((float)(h & 0xFFFFFF) / 16777216.0f - 0.5f);
}

// Split-precision squaring kernel alone (parts already in the k-blocked layout).
template <class P, int WM, int WN, int KPB = 1>
double time_split_t(int m, int k, int reps, int group) {
  const int kpad = (int)round_up(std::max(k, 1), 16), rows = (int)round_up(m, SQ_BM);
  const size_t cp = (size_t)rows * kpad * 2 * P::NQ, cc = (size_t)rows * rows;
  typename P::H* parts;
  float *yr, *yi, *cr, *ci;
  double2 *din, *dout;
  FX_HIP(hipMalloc(&parts, cp * sizeof(typename P::H)));
  FX_HIP(hipMalloc(&yr, cc * sizeof(float)));
  FX_HIP(hipMalloc(&yi, cc * sizeof(float)));
  FX_HIP(hipMalloc(&cr, cc * sizeof(float)));
  FX_HIP(hipMalloc(&ci, cc * sizeof(float)));
  FX_HIP(hipMalloc(&din, rows * sizeof(double2)));
  FX_HIP(hipMalloc(&dout, rows * sizeof(double2)));
  fill_pattern16<P><<<(unsigned)((cp + 255) / 256), 256>>>(parts, cp, 7u);
  fill_pattern<float><<<(unsigned)((cc + 255) / 256), 256>>>(yr, cc, 8u);
  fill_pattern<float><<<(unsigned)((cc + 255) / 256), 256>>>(yi, cc, 9u);
  FX_HIP(hipMemset(din, 0, rows * sizeof(double2)));
  SqArgs<P> a;
  a.a = parts;
  a.rows = rows;
  a.yr = yr;
  a.yi = yi;
  a.ldy = rows;
  a.cr = cr;
  a.ci = ci;
  a.ldc = rows;
  a.d_in = din;
  a.d_out = dout;
  a.m = m;
  a.k = kpad;
  a.kblock = 8;
  const int T = (m + SQ_BM - 1) / SQ_BM;
  int* dorder = nullptr;
  a.order = nullptr;
  if (group > 0) {
    std::vector<int> ho((size_t)T * (T + 1) / 2);
    sym_tile_order(T, group, ho.data());
    FX_HIP(hipMalloc(&dorder, ho.size() * sizeof(int)));
    FX_HIP(hipMemcpy(dorder, ho.data(), ho.size() * sizeof(int), hipMemcpyHostToDevice));
    a.order = dorder;
  }
  dim3 grid(T * (T + 1) / 2), block(64 * WM * WN);
  hipEvent_t e0, e1;
  FX_HIP(hipEventCreate(&e0));
  FX_HIP(hipEventCreate(&e1));
  sq_split_kernel<P, WM, WN, KPB><<<grid, block>>>(a);
  FX_LAUNCH();
  FX_HIP(hipEventRecord(e0));
  for (int r = 0; r < reps; ++r) sq_split_kernel<P, WM, WN, KPB><<<grid, block>>>(a);
  FX_HIP(hipEventRecord(e1));
  FX_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(parts);
  for (float* p : {yr, yi, cr, ci}) (void)hipFree(p);
  (void)hipFree(din);
  (void)hipFree(dout);
  if (dorder) (void)hipFree(dorder);
  return (double)ms / std::max(reps, 1);
}

// The int8-slice FP64 product C = A B^T alone (slices prepared once from
// pattern operands), on the engine's 128 x 128 tiles or on 64 x 64 tiles.
template <int S>
double time_oz_t(int sym, int m, int n, int k, int reps, bool tile64) {
  const int kpad = (int)round_up(k, OZ_BK), ra = (int)round_up(m, OZ_BM), rb = (int)round_up(n, OZ_BM);
  struct Bufs {
    double *da = nullptr, *db = nullptr, *dc = nullptr;
    int8_t *sa = nullptr, *sb = nullptr;
    int *ea = nullptr, *eb = nullptr;
    ~Bufs() {
      for (void* p : {(void*)da, (void*)db, (void*)dc, (void*)sa, (void*)sb, (void*)ea, (void*)eb})
        if (p) (void)hipFree(p);
    }
  } B;
  FX_HIP(hipMalloc(&B.da, (size_t)m * k * sizeof(double)));
  FX_HIP(hipMalloc(&B.db, (size_t)n * k * sizeof(double)));
  FX_HIP(hipMalloc(&B.dc, (size_t)m * n * sizeof(double)));
  FX_HIP(hipMalloc(&B.sa, (size_t)ra * kpad * S));
  FX_HIP(hipMalloc(&B.sb, (size_t)rb * kpad * S));
  FX_HIP(hipMalloc(&B.ea, (size_t)ra * sizeof(int)));
  FX_HIP(hipMalloc(&B.eb, (size_t)rb * sizeof(int)));
  fill_pattern<double><<<(unsigned)(((size_t)m * k + 255) / 256), 256>>>(B.da, (size_t)m * k, 1u);
  fill_pattern<double><<<(unsigned)(((size_t)n * k + 255) / 256), 256>>>(B.db, (size_t)n * k, 2u);
  FX_LAUNCH();
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
  auto launch = [&]() {
    if (tile64) {
      if (sym) {
        const int T = ra / 64;
        oz_gemm_kernel64<S, true><<<dim3(T * (T + 1) / 2), dim3(256)>>>(g);
      } else {
        oz_gemm_kernel64<S, false><<<dim3(rb / 64, ra / 64), dim3(256)>>>(g);
      }
    } else if (sym) {
      const int T = ra / OZ_BM;
      oz_gemm_kernel<S, true><<<dim3(T * (T + 1) / 2), dim3(512)>>>(g);
    } else {
      oz_gemm_kernel<S, false><<<dim3(rb / OZ_BM, ra / OZ_BM), dim3(512)>>>(g);
    }
    FX_LAUNCH();
  };
  launch();  // warm-up
  hipEvent_t e0, e1;
  FX_HIP(hipEventCreate(&e0));
  FX_HIP(hipEventCreate(&e1));
  FX_HIP(hipEventRecord(e0));
  for (int r = 0; r < reps; ++r) launch();
  FX_HIP(hipEventRecord(e1));
  FX_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  FX_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)t / std::max(reps, 1);
}

// Plain copy of known size (HBM traffic calibration for profiler counters).
__global__ __launch_bounds__(256) void copy_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                   size_t cnt) {
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < cnt; k += (size_t)gridDim.x * 256) dst[k] = src[k];
}

// In-register MFMA throughput: every wave issues `iters` rounds of 8
// independent MFMAs (no memory traffic in the loop), so the rate is the
// sustained matrix-core rate of the device at its running clock.
// kind 0: v_mfma_f32_16x16x4f32, 1: v_mfma_f64_16x16x4f64,
// 2: v_mfma_f32_32x32x16_f16, 3: v_mfma_f32_32x32x16_bf16,
// 4: v_mfma_i32_32x32x32_i8, 5: v_mfma_f32_32x32x2f32.
template <int KIND>
__global__ __launch_bounds__(256) void mfma_rate_kernel(float* out, int iters) {
  const float seed = (float)(threadIdx.x & 7) * 1e-3f;
  float keep = 0.0f;
  if constexpr (KIND == 0) {
    f32x4 acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = f32x4{};
    const float a = 1.0f + seed, b = 1.0f - seed;
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 8; ++q) keep += acc[q][0];
  } else if constexpr (KIND == 1) {
    f64x4 acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = f64x4{};
    const double a = 1.0 + seed, b = 1.0 - seed;
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 8; ++q) keep += (float)acc[q][0];
  } else if constexpr (KIND == 4) {
    const int v = 0x01010101 * (1 + (int)(threadIdx.x & 3));
    const i32x4 a = {v, v, v, v}, b = {v, -v, v, -v};
    i32x16 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = i32x16{};
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) keep += (float)acc[q][0];
  } else if constexpr (KIND == 5) {
    f32x16 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = f32x16{};
    const float a = 1.0f + seed, b = 1.0f - seed;
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) keep += acc[q][0];
  } else {
    using V = typename std::conditional<KIND == 2, f16x8, bf16x8>::type;
    using E = typename std::conditional<KIND == 2, _Float16, __bf16>::type;
    V a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      a[e] = (E)(1.0f + seed);
      b[e] = (E)(1.0f - seed);
    }
    f32x16 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = f32x16{};
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if constexpr (KIND == 2) acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[q], 0, 0, 0);
        else acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
      }
#pragma unroll
    for (int q = 0; q < 4; ++q) keep += acc[q][0];
  }
  if (keep == 1234.5f) out[blockIdx.x] = keep;  // keeps the loop alive; never true for these inputs
}

}  // namespace

extern "C" {

const char* fexpm_probe_last_error(void) { return g_probe_error.c_str(); }

int fexpm_probe_mfma_rate_waves(int kind, int waves, int iters, double* tflops, double* ms) {
  if (!tflops || !ms || kind < 0 || kind > 5 || iters <= 0 || iters > (1 << 20) || waves < 1 || waves > 8)
    return FEXPM_EINVAL;
  return guarded([&] {
    check_device();
    int dev = 0;
    hipDeviceProp_t prop;
    FX_HIP(hipGetDevice(&dev));
    FX_HIP(hipGetDeviceProperties(&prop, dev));
    // workgroups of 4 waves, `waves` of them per CU: `waves` waves per SIMD
    const unsigned grid = (unsigned)prop.multiProcessorCount * (unsigned)waves;
    float* out = nullptr;
    FX_HIP(hipMalloc(&out, grid * sizeof(float)));
    hipEvent_t e0, e1;
    FX_HIP(hipEventCreate(&e0));
    FX_HIP(hipEventCreate(&e1));
    auto launch = [&](int n) {
      switch (kind) {
        case 0: mfma_rate_kernel<0><<<grid, 256>>>(out, n); break;
        case 1: mfma_rate_kernel<1><<<grid, 256>>>(out, n); break;
        case 2: mfma_rate_kernel<2><<<grid, 256>>>(out, n); break;
        case 3: mfma_rate_kernel<3><<<grid, 256>>>(out, n); break;
        case 4: mfma_rate_kernel<4><<<grid, 256>>>(out, n); break;
        default: mfma_rate_kernel<5><<<grid, 256>>>(out, n); break;
      }
      FX_LAUNCH();
    };
    launch(std::max(1, iters / 8));
    FX_HIP(hipEventRecord(e0));
    launch(iters);
    FX_HIP(hipEventRecord(e1));
    FX_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    FX_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    // flops per MFMA instruction (one wave): 16x16x4 -> 2048, 32x32x16 -> 32768,
    // 32x32x32 -> 65536, 32x32x2 -> 4096
    const double per = kind < 2 ? 2048.0 * 8 : (kind == 4 ? 65536.0 * 4 : (kind == 5 ? 4096.0 * 4 : 32768.0 * 4));
    const double flops = (double)grid * 4.0 * (double)iters * per;
    *ms = t;
    *tflops = flops / ((double)t * 1e-3) / 1e12;
  });
}

int fexpm_probe_mfma_rate(int kind, int iters, double* tflops, double* ms) {
  return fexpm_probe_mfma_rate_waves(kind, 8, iters, tflops, ms);
}

int fexpm_probe_copy(int64_t bytes, int reps, double* ms) {
  if (!ms || bytes <= 0 || bytes % 16 != 0 || reps <= 0) return FEXPM_EINVAL;
  return guarded([&] {
    check_device();
    const size_t cnt = (size_t)bytes / 16;
    u32x4 *src = nullptr, *dst = nullptr;
    FX_HIP(hipMalloc(&src, (size_t)bytes));
    FX_HIP(hipMalloc(&dst, (size_t)bytes));
    FX_HIP(hipMemset(src, 1, (size_t)bytes));
    const unsigned grid = (unsigned)std::min<size_t>((cnt + 255) / 256, 8192);
    hipEvent_t e0, e1;
    FX_HIP(hipEventCreate(&e0));
    FX_HIP(hipEventCreate(&e1));
    copy_kernel<<<grid, 256>>>(src, dst, cnt);
    FX_LAUNCH();
    FX_HIP(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) copy_kernel<<<grid, 256>>>(src, dst, cnt);
    FX_HIP(hipEventRecord(e1));
    FX_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    FX_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    *ms = (double)t / reps;
  });
}

int fexpm_probe_gemm_time(int wtype, int mode, int sym, int cfg, int m, int n, int k, int reps,
                                    double* ms) {
  if (!ms || m <= 0 || n <= 0 || k <= 0 || reps <= 0 || (sym && m != n)) return FEXPM_EINVAL;
  return guarded([&] {
    check_device();
    double t = -1.0;
    if (wtype == 4) {
      // FP64 product by int8 slices, cfg = number of slices (operands from a fill pattern)
      if (cfg == 5) t = time_oz_t<5>(sym, m, n, k, reps, false);
      else if (cfg == 6) t = time_oz_t<6>(sym, m, n, k, reps, false);
      else if (cfg == 16) t = time_oz_t<6>(sym, m, n, k, reps, true);
      else if (cfg == 4) t = time_oz_t<4>(sym, m, n, k, reps, false);
    } else if (wtype >= 2) {
      // cfg % 10: 0 = 8 waves 2x4, 1 = 4 waves 2x2, 2 = 8 waves 4x2, 3 = 16 waves 4x4, one k-block per LDS
      // stage; 4 / 5 = 16 waves 4x4 (the fp16x2 default) / 8 waves 2x4 with two k-blocks per stage;
      // cfg / 10: tile group of the order table (0 = row-major triangle)
      const int lay = cfg % 10, grp = cfg / 10;
      if (mode == 2 && sym && wtype == 2) {
        if (lay == 0) t = time_split_t<Fp16x2, 2, 4, 1>(m, k, reps, grp);
        else if (lay == 1) t = time_split_t<Fp16x2, 2, 2, 1>(m, k, reps, grp);
        else if (lay == 2) t = time_split_t<Fp16x2, 4, 2, 1>(m, k, reps, grp);
        else if (lay == 3) t = time_split_t<Fp16x2, 4, 4, 1>(m, k, reps, grp);
        else if (lay == 4) t = time_split_t<Fp16x2, 4, 4, 2>(m, k, reps, grp);
        else if (lay == 5) t = time_split_t<Fp16x2, 2, 4, 2>(m, k, reps, grp);
      } else if (mode == 2 && sym && wtype == 3) {
        if (lay == 0) t = time_split_t<Bf16x3, 2, 4>(m, k, reps, grp);
        else if (lay == 1) t = time_split_t<Bf16x3, 2, 2>(m, k, reps, grp);
        else if (lay == 2) t = time_split_t<Bf16x3, 4, 2>(m, k, reps, grp);
      }
    } else if (wtype == 0) {
      // squaring tilings: 0 default (32x32x2 MFMAs), 4 FCfgSmall, 5..8 LDS padding / wave-tile variants on
      // 16x16x4 MFMAs, 9 the default tiling on 16x16x4, 10 / 11 padding / wave-tile variants on 32x32x2,
      // 12 128x128 in 16 waves of 32x32 (four waves per SIMD), 14 / 15 the 16-wave 128x128 / default 64x64
      // tiling with 32-deep k-steps
      if (mode == 2 && sym && cfg == 4) t = time_gemm_t<float, MODE_C3M, true, true, FCfgSmall>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 9) t = time_gemm_t<float, MODE_C3M, true, true, FCfg16>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 10)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<128, 128, 16, 8, 64, 32, 32>>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 11)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<128, 128, 16, 4, 32, 64, 32>>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 12) t = time_gemm_t<float, MODE_C3M, true, true, FCfg16w>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 14)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<128, 128, 32, 4, 32, 32, 32, 4>>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 15)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<64, 64, 32, 4, 32, 32, 32, 4>>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 13)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<64, 64, 16, 4, 32, 16, 16, 4>>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 5)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<128, 128, 16, 8, 64, 32>>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 6)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<128, 128, 16, 4, 32, 64>>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 7)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<128, 128, 16, 12, 64, 32>>(m, n, k, reps);
      else if (mode == 2 && sym && cfg == 8)
        t = time_gemm_t<float, MODE_C3M, true, true, TileCfg<128, 128, 16, 8, 32, 64>>(m, n, k, reps);
      else if (mode == 2 && sym) t = time_gemm_t<float, MODE_C3M, true, true, GemmCfg<float>>(m, n, k, reps);
      else if (mode == 2) t = time_gemm_t<float, MODE_C3M, false, true, GemmCfg<float>>(m, n, k, reps);
      else if (mode == 0 && sym) t = time_gemm_t<float, MODE_REAL, true, false, GemmCfg<float>>(m, n, k, reps);
      else if (mode == 1) t = time_gemm_t<float, MODE_C4M, false, false, GemmCfg<float>>(m, n, k, reps);
    } else {
#define FX_TIME_D(CF)                                                                              \
  do {                                                                                             \
    if (mode == 2 && sym) t = time_gemm_t<double, MODE_C3M, true, false, CF>(m, n, k, reps);       \
    else if (mode == 2) t = time_gemm_t<double, MODE_C3M, false, false, CF>(m, n, k, reps);       \
    else if (mode == 1) t = time_gemm_t<double, MODE_C4M, false, false, CF>(m, n, k, reps);       \
    else if (mode == 0 && sym) t = time_gemm_t<double, MODE_REAL, true, false, CF>(m, n, k, reps); \
    else if (mode == 0) t = time_gemm_t<double, MODE_REAL, false, false, CF>(m, n, k, reps);      \
  } while (0)
      if (cfg == 0) FX_TIME_D(DCfg4w);
      else if (cfg == 1) FX_TIME_D(DCfg8w);
      else if (cfg == 2) FX_TIME_D(DCfgTall);
      else if (cfg == 3) FX_TIME_D(DCfgWide);
      else if (cfg == 4) FX_TIME_D(DCfg8wPad4);
      else if (cfg == 5) FX_TIME_D(DCfg8wW4);
#undef FX_TIME_D
    }
    if (t < 0) throw ProbeError(FEXPM_EINVAL, "unsupported gemm timing variant");
    *ms = t;
  });
}

}  // extern "C"
