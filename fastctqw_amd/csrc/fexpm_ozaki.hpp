// FP64 products on the int8 MFMA of gfx950 (CDNA4), by exact integer slices
// (the splitting of Ozaki et al., "Error-free transformations of matrix
// multiplication by using fast routines of matrix multiplication", 2012,
// with integer slices as in Ootomo, Ozaki & Yokota 2024).
//
// Every row r of an FP64 operand is scaled by the power of two 2^-e_r that
// brings its largest magnitude below 1 and cut into S signed slices,
//     x = 2^e_r * sum_p a_p 2^(-6 - 7p),   a_p integers in [-64, 64],
// exactly except for the part below 2^(e_r - 6 - 7S) (S = 5: 34 bits of the
// row maximum, S = 6: 41 bits).  A product sum_k x_ik y_jk is then
//     2^(e_i + f_j) sum_d 2^(-12 - 7d) sum_{p+q=d} sum_k a_p[ik] b_q[jk],
// whose inner sums are exact in int32 (|a b| <= 2^12, k <= 10^5 terms per
// diagonal d < 6) on v_mfma_i32_32x32x32_i8; only d < S is kept.  The S
// diagonal accumulators are combined in FP64 in the epilogue, which is the
// FP64 GEMM epilogue (alpha, beta, added matrices, diagonal, mirrored
// lower tiles).
//
// Slice layout ("k-blocked", like the 16-bit split parts): for k-block kb
// (32 values of k), row r and slice p, the 32 bytes live at
//     base + ((kb * rows + r) * S + p) * 32,
// so the 128-row stage of one k-block is one contiguous run of 128 * S * 32
// bytes; rows are padded to OZ_BM and k to OZ_BK with zeros.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fexpm_gemm.hpp"

namespace fexpm {

constexpr int OZ_BM = 128;  // square workgroup tile
constexpr int OZ_BK = 32;   // k per stage = k of one MFMA
constexpr int OZ_MAX_SLICES = 6;

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// Row exponents and slices of an operand whose row r is the concatenation
//     x_rk = s1 * x1[r, k] (k < k1),   s2 * x2[r, k - k1] (k1 <= k < nk)
// (x2 may be null: one part; FP64, leading dimension ld, rows < nr valid)
// for rows < rows (padding rows zero) and k < kpad (a multiple of OZ_BK).
// One workgroup per row: the row maximum, then each thread slices whole
// 32-value k-blocks.  ex[r] = e with max |x_rk| < 2^e (0 for a zero row).
// s1, s2 are +-1 (exact).
__device__ __forceinline__ double oz_elem(const double* __restrict__ x1, double s1, const double* __restrict__ x2,
                                          double s2, int k1, int64_t ld, int r, int k) {
  return k < k1 ? s1 * x1[(int64_t)r * ld + k] : s2 * x2[(int64_t)r * ld + (k - k1)];
}

template <int S>
__device__ __forceinline__ void oz_slice_row(const double* __restrict__ x1, double s1,
                                             const double* __restrict__ x2, double s2, int k1, int64_t ld, int nr,
                                             int nk, int rows, int kpad, int8_t* __restrict__ out,
                                             int* __restrict__ ex, int r) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double mx = 0.0;
  if (r < nr)
    for (int k = t; k < nk; k += 256) mx = fmax(mx, fabs(oz_elem(x1, s1, x2, s2, k1, ld, r, k)));
  red[t] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  int e = 0;
  if (red[0] > 0.0) (void)frexp(red[0], &e);  // red[0] = f 2^e, f in [0.5, 1): |x| < 2^e
  if (t == 0) ex[r] = e;
  const double sc = ldexp(1.0, 6 - e);  // first slice: |x 2^-e| < 1, times 2^6
  // 256 consecutive values per pass: thread t reads value t (coalesced),
  // its S slice bytes go to LDS, and 16-byte runs of every slice are
  // written out (8 k-blocks x S slices x 2 halves = 16 S chunks)
  __shared__ __attribute__((aligned(16))) int8_t buf[S][256];
  for (int k0 = 0; k0 < kpad; k0 += 256) {
    const int k = k0 + t;
    double y = (r < nr && k < nk) ? oz_elem(x1, s1, x2, s2, k1, ld, r, k) * sc : 0.0;
#pragma unroll
    for (int p = 0; p < S; ++p) {
      const double a = rint(y);  // |y| <= 64, so a fits; the rest is exact
      buf[p][t] = (int8_t)a;
      y = (y - a) * 128.0;
    }
    __syncthreads();
    if (t < 16 * S) {
      const int kbl = t / (2 * S), w = t % (2 * S), p = w >> 1, h = w & 1;
      const int kb = k0 / OZ_BK + kbl;
      if (kb * OZ_BK < kpad)
        *reinterpret_cast<i32x4*>(out + ((int64_t)kb * rows + r) * S * OZ_BK + p * OZ_BK + 16 * h) =
            *reinterpret_cast<const i32x4*>(&buf[p][kbl * OZ_BK + 16 * h]);
    }
    __syncthreads();
  }
}

// Orders the LDS writes of a wave before its later LDS reads (lanes of one
// wave exchanging data through LDS): wavefront-scope release fence, the
// wave barrier, acquire fence.  No other wave takes part.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The same slicing, four rows per workgroup of 256 threads (one wave per
// row, rows r0 .. r0 + 3 of one operand; each wave syncs only with itself):
// for the small orders of batched walks, where one workgroup per row is
// mostly idle.  rows must be a multiple of 4.
template <int S>
__device__ __forceinline__ void oz_slice_rows4(const double* __restrict__ x1, double s1,
                                               const double* __restrict__ x2, double s2, int k1, int64_t ld, int nr,
                                               int nk, int rows, int kpad, int8_t* __restrict__ out,
                                               int* __restrict__ ex, int r0) {
  constexpr int KV = 256;  // values of the row per pass: four loads per lane in flight
  __shared__ __attribute__((aligned(16))) int8_t buf[4][S][KV];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = r0 + w;
  double mx = 0.0;
  if (r < nr) {
#pragma unroll 4
    for (int k = lane; k < nk; k += 64) mx = fmax(mx, fabs(oz_elem(x1, s1, x2, s2, k1, ld, r, k)));
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  int e = 0;
  if (mx > 0.0) (void)frexp(mx, &e);
  if (lane == 0 && r < rows) ex[r] = e;
  const double sc = ldexp(1.0, 6 - e);
  for (int k0 = 0; k0 < kpad; k0 += KV) {
    double y[KV / 64];
#pragma unroll
    for (int j = 0; j < KV / 64; ++j) {
      const int k = k0 + 64 * j + lane;
      y[j] = (r < nr && k < nk) ? oz_elem(x1, s1, x2, s2, k1, ld, r, k) * sc : 0.0;
    }
#pragma unroll
    for (int j = 0; j < KV / 64; ++j)
#pragma unroll
      for (int p = 0; p < S; ++p) {
        const double a = rint(y[j]);  // |y| <= 64, so a fits; the rest is exact
        buf[w][p][64 * j + lane] = (int8_t)a;
        y[j] = (y[j] - a) * 128.0;
      }
    wave_sync();  // buf[w] is this wave's own: no other wave reads it
    if (r < rows) {
      // KV / OZ_BK k-blocks x S slices x 2 halves of 16 bytes
      for (int c = lane; c < (KV / OZ_BK) * S * 2; c += 64) {
        const int kbl = c / (2 * S), q = c % (2 * S), p = q >> 1, h = q & 1;
        const int kb = k0 / OZ_BK + kbl;
        if (kb * OZ_BK < kpad)
          *reinterpret_cast<i32x4*>(out + ((int64_t)kb * rows + r) * S * OZ_BK + p * OZ_BK + 16 * h) =
              *reinterpret_cast<const i32x4*>(&buf[w][p][kbl * OZ_BK + 16 * h]);
      }
    }
    wave_sync();  // the reads above complete before the next pass writes buf[w]
  }
}

template <int S>
__global__ __launch_bounds__(256) void oz_slice(const double* __restrict__ x1, double s1,
                                                const double* __restrict__ x2, double s2, int k1, int64_t ld, int nr,
                                                int nk, int rows, int kpad, int8_t* __restrict__ out,
                                                int* __restrict__ ex) {
  oz_slice_row<S>(x1, s1, x2, s2, k1, ld, nr, nk, rows, kpad, out, ex, (int)blockIdx.x);
}

template <int S> struct OzArgs {
  const int8_t* a;  // slices of the left operand (rows_a rows in the k-blocked layout)
  const int* ea;
  int rows_a;
  const int8_t* b;  // slices of the right operand (its rows are the columns of C)
  const int* eb;
  int rows_b;
  int m, n, kpad;
  Epilogue<double> epi;
  const int* order;  // optional symmetric tile order (tm << 16 | tn) per logical block, else row-major
};

// Output tile (tm, tn) of C = A B^T from the slices: 8 waves (4 x 2), each
// 32 rows x 64 columns (two 32 x 32 MFMA blocks); every LDS read is
// complete on return.
template <int S, int BMT = OZ_BM, int WAVES_M = 4, int WAVES_N = 2>
__device__ __forceinline__ void oz_tile(const OzArgs<S>& g, int tm, int tn, bool sym) {
  constexpr int NT = 64 * WAVES_M * WAVES_N;
  constexpr int FI = BMT / WAVES_M / 32, FJ = BMT / WAVES_N / 32;  // 32 x 32 blocks per wave
  static_assert(FI >= 1 && FJ >= 1 && OZ_BM % BMT == 0, "tile edge must divide the row padding");
  constexpr int STAGE = BMT * S * OZ_BK;  // bytes of one operand stage in memory
  constexpr int CH = STAGE / 16;            // 16-byte chunks per operand stage
  constexpr int CPT = (CH + NT - 1) / NT;
  // LDS rows of S * 32 bytes, padded by 16 bytes for S < 6: a row stride of
  // 128 or 160 bytes makes the 16-byte reads of consecutive lanes (rows)
  // collide in the same bank groups, the padded stride does not.  For S = 6
  // (192 bytes) the unpadded layout measured faster (9.3 vs 10.4 ms at
  // n = 10^4).
  constexpr int RSTR = S * OZ_BK + (S < 6 ? 16 : 0);
  constexpr int LSTAGE = BMT * RSTR;
  __shared__ __attribute__((aligned(16))) int8_t smem[2][2 * LSTAGE];

  const int row0 = tm * BMT, col0 = tn * BMT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave / WAVES_N) * (32 * FI), wc = (wave % WAVES_N) * (32 * FJ);
  const int l32 = lane & 31, kh = lane >> 5;

  i32x16 acc[S][FI][FJ];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j) acc[d][i][j] = i32x16{};

  const int8_t* baseA = g.a + (int64_t)row0 * S * OZ_BK;
  const int8_t* baseB = g.b + (int64_t)col0 * S * OZ_BK;
  const int64_t strideA = (int64_t)g.rows_a * S * OZ_BK;  // bytes per k-block
  const int64_t strideB = (int64_t)g.rows_b * S * OZ_BK;
  i32x4 ra[CPT], rb[CPT];
  auto load = [&](int kb) {
    const i32x4* pa = reinterpret_cast<const i32x4*>(baseA + kb * strideA);
    const i32x4* pb = reinterpret_cast<const i32x4*>(baseB + kb * strideB);
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int c = threadIdx.x + NT * u;
      if (CH % NT == 0 || c < CH) {  // no guard where the chunks divide evenly
        ra[u] = pa[c];
        rb[u] = pb[c];
      }
    }
  };
  auto store = [&](int8_t* s) {
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int c = threadIdx.x + NT * u;
      if (CH % NT == 0 || c < CH) {
        const int off = (c / (2 * S)) * RSTR + (c % (2 * S)) * 16;
        *reinterpret_cast<i32x4*>(s + off) = ra[u];
        *reinterpret_cast<i32x4*>(s + LSTAGE + off) = rb[u];
      }
    }
  };
  auto compute = [&](const int8_t* s) {
    i32x4 fa[S][FI];
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
      for (int i = 0; i < FI; ++i)
        fa[p][i] = *reinterpret_cast<const i32x4*>(s + (wr + 32 * i + l32) * RSTR + p * OZ_BK + 16 * kh);
#pragma unroll
    for (int q = 0; q < S; ++q) {
      i32x4 fb[FJ];
#pragma unroll
      for (int j = 0; j < FJ; ++j)
        fb[j] = *reinterpret_cast<const i32x4*>(s + LSTAGE + (wc + 32 * j + l32) * RSTR + q * OZ_BK + 16 * kh);
#pragma unroll
      for (int p = 0; p + q < S; ++p)
#pragma unroll
        for (int i = 0; i < FI; ++i)
#pragma unroll
          for (int j = 0; j < FJ; ++j)
            acc[p + q][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[p][i], fb[j], acc[p + q][i][j], 0, 0, 0);
    }
  };

  const int KT = g.kpad / OZ_BK;
  load(0);
  store(smem[0]);
  __syncthreads();
  for (int kt = 0; kt < KT; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < KT) load(kt + 1);
    compute(smem[cur]);
    if (kt + 1 < KT) store(smem[cur ^ 1]);
    __syncthreads();
  }

  // ---- epilogue: combine the diagonals in FP64, then the FP64 GEMM epilogue ----
  const Epilogue<double>& e = g.epi;
  const bool diag_tile = sym && tm == tn;
#pragma unroll
  for (int i = 0; i < FI; ++i) {
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = row0 + wr + 32 * i + 8 * (r >> 2) + 4 * kh + (r & 3);
        const int col = col0 + wc + 32 * j + l32;
        if (row >= g.m || col >= g.n) continue;
        if (diag_tile && row < col) continue;
        double v = 0.0;
#pragma unroll
        for (int d = S - 1; d >= 0; --d) v = v * (1.0 / 128.0) + (double)acc[d][i][j][r];
        v = ldexp(v, g.ea[row] + g.eb[col] - 12);
        const int64_t o = (int64_t)row * e.ldc + col;
        double out = e.alpha * v;
        if (e.beta != 0.0) out += e.beta * e.cr[o];
        if (e.nadd > 0) {
          const int64_t oa = (int64_t)row * e.ld_add + col;
          for (int t = 0; t < e.nadd; ++t) out += e.addc[t] * e.add_r[t][oa];
        }
        if (row == col) out += e.diag;
        e.cr[o] = out;
        if (sym && row != col) e.cr[(int64_t)col * e.ldc + row] = out;
      }
    }
  }
}

template <int S, bool SYM>
__global__ __launch_bounds__(512) void oz_gemm_kernel(OzArgs<S> g) {
  int tm, tn;
  if (SYM) {
    const int lb = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if (g.order) {
      const int code = g.order[lb];
      tm = code >> 16;
      tn = code & 0xFFFF;
    } else {
      tri_tile(lb, tm, tn);
    }
  } else {
    const int nbn = (int)gridDim.x;
    const int lin = xcd_remap((int)(blockIdx.y * gridDim.x + blockIdx.x), (int)(gridDim.x * gridDim.y));
    tm = lin / nbn;
    tn = lin % nbn;
  }
  oz_tile<S>(g, tm, tn, SYM);
}

// Grouped form: many products (one or more per walk of a batch) in one
// persistent launch, with the tile placement of gemm_group_kernel.
template <int S> struct GroupOz {
  OzArgs<S> g;
  int tile0;
  int tiles;
  int tiles_n;
  int sym;
};

template <int S>
__global__ __launch_bounds__(512) void oz_group_kernel(const GroupOz<S>* __restrict__ probs, int nprob, int total,
                                                       int chunk) {
  for (int b = (int)blockIdx.x; b < total; b += (int)gridDim.x) {
    const int t = xcd_chunk_remap(b, total, chunk);
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].tile0 <= t) lo = mid;
      else hi = mid - 1;
    }
    const GroupOz<S>& P = probs[lo];
    const int local = t - P.tile0;
    int tm, tn;
    if (P.sym) grouped_tri_tile(local, P.tiles_n, tm, tn);  // tiles_n = tiles per side
    else grouped_rect_tile(local, P.tiles / P.tiles_n, P.tiles_n, tm, tn);
    oz_tile<S>(P.g, tm, tn, P.sym != 0);
  }
}

}  // namespace fexpm
