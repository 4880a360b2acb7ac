// Split-precision squaring on the 16-bit MFMA of gfx950 (CDNA4).
//
// Every FP32 operand value x is written as a sum of 16-bit parts,
//     SCALE * x = h_0 + h_1 (+ h_2),   h_q = round16(SCALE * x - h_0 - ... - h_{q-1}),
// and a real product sum_k x_ik y_jk is formed from the part products whose
// combined weight is at least 2^-(significand bits * parts):
//   FP16x2  fp16 parts (11-bit significands), 3 products  h0h0 + h0h1 + h1h0,
//           about 22 significant bits; SCALE = 2^14 keeps |x| <= 1 well in
//           the fp16 range and pushes fp16 subnormals below 2^-38 in x units.
//   BF16x3  bf16 parts (8-bit significands), 6 products  h0h0 + h0h1 + h1h0 +
//           h0h2 + h1h1 + h2h0, about 24 significant bits (FP32-like), full
//           FP32 exponent range, SCALE = 1.
// Products accumulate in FP32 on v_mfma_f32_32x32x16_{f16,bf16}.
//
// The kernel computes the complex-symmetric squaring X^2 = X X^T of the
// structured path in split form (X = diag(d) + Y, Y real and imaginary
// planes symmetric, zero diagonal) with four real products per part pair:
//     Re += Yr Yr^T - Yi Yi^T,   Im += Yr Yi^T + Yi Yr^T
// (the minus sign is applied by flipping the sign bits of the Yi fragment).
// Two accumulators per output block keep the register budget within one
// wave per SIMD, which the 3M form (three accumulators) does not.  Lower
// tiles only, mirrored on store, with the blocked accumulation (the
// accumulators are folded into outer sums every kblock k-steps) and the
// FP64 diagonal epilogue of the FP32 kernel.
//
// Operand layout ("k-blocked"): for k-block kb (16 values of k), row r and
// plane/part index w = plane * NQ + part (plane 0 real, 1 imaginary), the 16
// values live at
//     base + ((kb * rows + r) * 2 * NQ + w) * 16,
// so the 128-row stage of one k-block is a single contiguous run of
// 128 * 2 * NQ * 32 bytes.  split_blocked writes this layout from the FP32
// planes (zeros outside the valid rows and columns).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fexpm_gemm.hpp"

namespace fexpm {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

struct Fp16x2 {
  using H = _Float16;
  using V = f16x8;
  static constexpr int NQ = 2;
  static constexpr int NPAIR = 3;
  static constexpr float SCALE = 16384.0f;
  static constexpr float LIMIT = 65504.0f;
  static constexpr int WAVES_M = 4, WAVES_N = 4;  // 16 waves of 32 x 32 (measured fastest)
  static constexpr int KPB = 2;  // k-blocks per LDS stage: 132 KB, 1.6-4.5% faster than one at n = 10^4
  __device__ static constexpr int qa(int p) { return p == 2 ? 1 : 0; }
  __device__ static constexpr int qb(int p) { return p == 1 ? 1 : 0; }
  __device__ static __forceinline__ f32x16 mfma(V a, V b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};

struct Bf16x3 {
  using H = __bf16;
  using V = bf16x8;
  static constexpr int NQ = 3;
  static constexpr int NPAIR = 6;
  static constexpr float SCALE = 1.0f;
  static constexpr float LIMIT = 3.0e38f;
  static constexpr int WAVES_M = 2, WAVES_N = 4;  // 8 waves of 64 x 32 (measured fastest)
  static constexpr int KPB = 1;  // two k-blocks per stage would need 198 KB of LDS
  // pairs (0,0) (0,1) (1,0) (0,2) (1,1) (2,0)
  __device__ static constexpr int qa(int p) { return p == 2 ? 1 : (p == 4 ? 1 : (p == 5 ? 2 : 0)); }
  __device__ static constexpr int qb(int p) { return p == 1 ? 1 : (p == 3 ? 2 : (p == 4 ? 1 : 0)); }
  __device__ static __forceinline__ f32x16 mfma(V a, V b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};

constexpr int SQ_BM = 128;  // workgroup tile (square, lower tiles only)
constexpr int SQ_BK = 16;   // k per stage = k of one MFMA
template <class P> constexpr int sq_threads() { return 64 * P::WAVES_M * P::WAVES_N; }

// Bit set in Scalars::info when a scaled operand value leaves the 16-bit
// range (the split squaring result is then not valid).
constexpr int INFO_SPLIT_RANGE = 4;

// FP32 planes (row-major, ldy; ldy and the plane base 16-byte aligned) ->
// k-blocked split parts of Yr and Yi for rows < rows and k < kpad (a
// multiple of 16); values outside nr x nc are zero.  A workgroup covers 16
// rows x 16 k-blocks: the 16 threads of a row read 16 consecutive 64-byte
// runs, and every thread writes its own 2 * NQ * 32-byte output block.
// Workgroup (bx, by) of the split: rows by*16.., k-blocks bx*16...
template <class P>
__device__ __forceinline__ void split_blocked_block(const float* __restrict__ yr, const float* __restrict__ yi,
                                                    int64_t ldy, int nr, int nc, int rows, int kpad,
                                                    typename P::H* __restrict__ out, int* info, int bx, int by) {
  using H = typename P::H;
  constexpr int NQ = P::NQ;
  const int r = by * 16 + (threadIdx.x >> 4);
  const int kb = bx * 16 + (threadIdx.x & 15);
  if (r >= rows || kb * 16 >= kpad) return;
  float v[2][16];
  const float* src[2] = {yr, yi};
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    if (r < nr && kb * 16 + 16 <= nc) {
      const float4* q = reinterpret_cast<const float4*>(src[p] + (int64_t)r * ldy + kb * 16);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float4 x = q[c];
        v[p][4 * c] = x.x;
        v[p][4 * c + 1] = x.y;
        v[p][4 * c + 2] = x.z;
        v[p][4 * c + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const int k = kb * 16 + c;
        v[p][c] = (r < nr && k < nc) ? src[p][(int64_t)r * ldy + k] : 0.0f;
      }
    }
  }
  H* dst = out + ((int64_t)kb * rows + r) * 2 * NQ * 16;
  bool bad = false;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    H part[NQ][16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      float x = v[p][c] * P::SCALE;
      bad |= !(fabsf(x) <= P::LIMIT);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const H h = (H)x;
        part[q][c] = h;
        x -= (float)h;
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      typename P::V lo, hi;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        lo[e] = part[q][e];
        hi[e] = part[q][8 + e];
      }
      u32x4* d = reinterpret_cast<u32x4*>(dst + (p * NQ + q) * 16);
      d[0] = __builtin_bit_cast(u32x4, lo);
      d[1] = __builtin_bit_cast(u32x4, hi);
    }
  }
  if (bad) atomicOr(info, INFO_SPLIT_RANGE);
}

template <class P>
__global__ __launch_bounds__(256) void split_blocked(const float* __restrict__ yr, const float* __restrict__ yi,
                                                     int64_t ldy, int nr, int nc, int rows, int kpad,
                                                     typename P::H* __restrict__ out, int* info) {
  split_blocked_block<P>(yr, yi, ldy, nr, nc, rows, kpad, out, info, (int)blockIdx.x, (int)blockIdx.y);
}

template <class P> struct SqArgs {
  const typename P::H* a;  // k-blocked split parts
  int rows;                // rows of the k-blocked layout (multiple of 128)
  const float* yr;         // current Y (FP32 planes) for the diagonal epilogue
  const float* yi;
  int64_t ldy;
  float* cr;
  float* ci;
  int64_t ldc;
  const double2* d_in;
  double2* d_out;
  int m, k, kblock;
  const int* order;        // optional tile order (tm << 16 | tn) per logical block, else row-major
};

// Lower-triangular tile order in groups of g x g tiles (row-major over the
// groups, row-major within a group): the blocks that run together on one
// XCD (consecutive logical indices after xcd_remap) then cover a few tile
// rows and columns of one group, so their operand rows are shared in that
// XCD's L2.
inline void sym_tile_order(int T, int g, int* out) {
  int k = 0;
  for (int gi = 0; gi * g < T; ++gi)
    for (int gj = 0; gj <= gi; ++gj)
      for (int tm = gi * g; tm < (gi + 1) * g && tm < T; ++tm)
        for (int tn = gj * g; tn < (gj + 1) * g && tn <= tm; ++tn) out[k++] = (tm << 16) | tn;
}

// LDS position (in 16-bit elements) of 8-value chunk ch (0/1) of a tile
// row: rows of 16 values (32 bytes), the chunk order swapped by bit 2 xor
// bit 3 of the row, so that 8 lanes reading 8 consecutive rows cover the
// 8 16-byte bank groups of 128 bytes and 16 lanes the 16 groups of 256 bytes.
__device__ __forceinline__ int sq_lds(int row, int ch) {
  return row * 16 + 8 * (ch ^ (((row >> 2) ^ (row >> 3)) & 1));
}
// Arrays (plane/part) of a stage are 32 bytes apart in bank position, so
// that the staging stores of consecutive lanes (same row, consecutive
// arrays) do not hit the same banks.
constexpr int SQ_TILE_STRIDE = SQ_BM * SQ_BK + 16;

// Output tile (tm, tn), tm >= tn, of the split squaring; every LDS read is
// complete on return, so a workgroup may call it again for another tile.
template <class P, int WAVES_M, int WAVES_N, int KPB = 1>
__device__ __forceinline__ void sq_split_tile(const SqArgs<P>& g, int tm, int tn) {
  using H = typename P::H;
  using V = typename P::V;
  constexpr int NT = 64 * WAVES_M * WAVES_N;
  constexpr int NQ = P::NQ;
  constexpr int W = 2 * NQ;                   // arrays per operand tile
  constexpr int TILE = SQ_TILE_STRIDE;        // elements per array (with bank padding)
  constexpr int CH = SQ_BM * W * 2;           // 16-byte chunks per operand stage
  constexpr int CPT = CH / NT;                // chunks per thread per operand
  static_assert(CH % NT == 0, "stage chunks");
  static_assert(KPB == 1 || KPB == 2, "one or two k-blocks per barrier");
  // KPB k-blocks per LDS stage (and per barrier), two stages
  __shared__ __attribute__((aligned(16))) H smem[2][KPB][2 * W * TILE];

  const int row0 = tm * SQ_BM, col0 = tn * SQ_BM;
  constexpr int FI = SQ_BM / WAVES_M / 32, FJ = SQ_BM / WAVES_N / 32;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave / WAVES_N) * (32 * FI), wc = (wave % WAVES_N) * (32 * FJ);
  const int l32 = lane & 31, kg = lane >> 5;

  f32x16 acc[2][FI][FJ];  // [re/im][i][j]
  f32x16 outer[2][FI][FJ];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j) {
        acc[p][i][j] = f32x16{};
        outer[p][i][j] = f32x16{};
      }

  // stage staging: chunk c = threadIdx.x + NT * u of an operand stage is
  // row c / (2W), array (c % (2W)) / 2, half (c % 2) of that row.  Two
  // register sets: the loads of k-block kt+2 are issued while kt is being
  // multiplied, so a load has two k-steps to arrive before it is stored.
  u32x4 sa0[CPT], sb0[CPT], sa1[CPT], sb1[CPT];
  const int64_t kstride = (int64_t)g.rows * W * 16;  // elements per k-block
  const H* baseA = g.a + (int64_t)row0 * W * 16;
  const H* baseB = g.a + (int64_t)col0 * W * 16;
  auto load = [&](u32x4* sa, u32x4* sb, int kb) {
    const u32x4* pa = reinterpret_cast<const u32x4*>(baseA + kb * kstride);
    const u32x4* pb = reinterpret_cast<const u32x4*>(baseB + kb * kstride);
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int c = threadIdx.x + NT * u;
      sa[u] = pa[c];
      sb[u] = pb[c];
    }
  };
  auto store = [&](const u32x4* sa, const u32x4* sb, H* s) {
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int c = threadIdx.x + NT * u;
      const int row = c / (2 * W), w = (c % (2 * W)) >> 1, ch = c & 1;
      *reinterpret_cast<u32x4*>(s + w * TILE + sq_lds(row, ch)) = sa[u];
      *reinterpret_cast<u32x4*>(s + (W + w) * TILE + sq_lds(row, ch)) = sb[u];
    }
  };
  auto compute = [&](const H* s) {
    V far_[NQ][FI], fai[NQ][FI], fan[NQ][FI], fbr[NQ][FJ], fbi[NQ][FJ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
#pragma unroll
      for (int i = 0; i < FI; ++i) {
        const int ra = wr + 32 * i + l32;
        far_[q][i] = *reinterpret_cast<const V*>(s + q * TILE + sq_lds(ra, kg));
        const u32x4 vi = *reinterpret_cast<const u32x4*>(s + (NQ + q) * TILE + sq_lds(ra, kg));
        fai[q][i] = __builtin_bit_cast(V, vi);
        fan[q][i] = __builtin_bit_cast(V, vi ^ 0x80008000u);  // -Yi: flip both sign bits of each dword
      }
#pragma unroll
      for (int j = 0; j < FJ; ++j) {
        const int rb = wc + 32 * j + l32;
        fbr[q][j] = *reinterpret_cast<const V*>(s + (W + q) * TILE + sq_lds(rb, kg));
        fbi[q][j] = *reinterpret_cast<const V*>(s + (W + NQ + q) * TILE + sq_lds(rb, kg));
      }
    }
#pragma unroll
    for (int pr = 0; pr < P::NPAIR; ++pr) {
#pragma unroll
      for (int i = 0; i < FI; ++i)
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
          const int qa = P::qa(pr), qb = P::qb(pr);
          acc[0][i][j] = P::mfma(far_[qa][i], fbr[qb][j], acc[0][i][j]);
          acc[0][i][j] = P::mfma(fan[qa][i], fbi[qb][j], acc[0][i][j]);
          acc[1][i][j] = P::mfma(far_[qa][i], fbi[qb][j], acc[1][i][j]);
          acc[1][i][j] = P::mfma(fai[qa][i], fbr[qb][j], acc[1][i][j]);
        }
    }
  };
  const int KT = g.k / SQ_BK;
  auto flush = [&](int kt) {
    if ((kt + 1) % g.kblock == 0 || kt + 1 == KT) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int i = 0; i < FI; ++i)
#pragma unroll
          for (int j = 0; j < FJ; ++j) {
            outer[p][i][j] += acc[p][i][j];
            acc[p][i][j] = f32x16{};
          }
    }
  };

  load(sa0, sb0, 0);
  load(sa1, sb1, min(1, KT - 1));
  store(sa0, sb0, smem[0][0]);
  if constexpr (KPB == 1) {
    __syncthreads();
    // The loads are unconditional (k-block clamped to the last one; the tail
    // re-reads are never stored), so the memory-counter waits before each
    // store only cover the older register set.
    for (int kt = 0; kt < KT; kt += 2) {
      // even step: k-block kt in smem[0]; set 0 is free for kt + 2
      load(sa0, sb0, min(kt + 2, KT - 1));
      compute(smem[0][0]);
      flush(kt);
      if (kt + 1 < KT) store(sa1, sb1, smem[1][0]);
      __syncthreads();
      if (kt + 1 >= KT) break;
      // odd step: k-block kt + 1 in smem[1]; set 1 is free for kt + 3
      load(sa1, sb1, min(kt + 3, KT - 1));
      compute(smem[1][0]);
      flush(kt + 1);
      if (kt + 2 < KT) store(sa0, sb0, smem[0][0]);
      __syncthreads();
    }
  } else {
    // Two k-blocks per stage and per barrier: stage cur holds k-blocks kt
    // and kt + 1; both register sets load kt + 2, kt + 3 (clamped) while
    // they are multiplied and are stored into the other stage, which every
    // wave finished reading before the previous barrier.
    if (KT > 1) store(sa1, sb1, smem[0][1]);
    __syncthreads();
    int cur = 0;
    for (int kt = 0; kt < KT; kt += 2) {
      load(sa0, sb0, min(kt + 2, KT - 1));
      load(sa1, sb1, min(kt + 3, KT - 1));
      compute(smem[cur][0]);
      flush(kt);
      if (kt + 1 < KT) {
        compute(smem[cur][1]);
        flush(kt + 1);
      }
      if (kt + 2 < KT) store(sa0, sb0, smem[cur ^ 1][0]);
      if (kt + 3 < KT) store(sa1, sb1, smem[cur ^ 1][1]);
      __syncthreads();
      cur ^= 1;
    }
  }

  // ---- epilogue (diagonal split, lower tiles mirrored) ----
  const double unscale = 1.0 / ((double)P::SCALE * (double)P::SCALE);
  const bool diag_tile = tm == tn;
#pragma unroll
  for (int i = 0; i < FI; ++i) {
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = row0 + wr + 32 * i + 8 * (r >> 2) + 4 * kg + (r & 3);
        const int col = col0 + wc + 32 * j + l32;
        if (row >= g.m || col >= g.m) continue;
        if (diag_tile && row < col) continue;
        double vr = (double)outer[0][i][j][r] * unscale;
        double vi = (double)outer[1][i][j][r] * unscale;
        if (row == col) {
          const double2 dd = g.d_in[row];
          g.d_out[row] = make_double2(dd.x * dd.x - dd.y * dd.y + vr, 2.0 * dd.x * dd.y + vi);
          vr = 0.0;
          vi = 0.0;
        } else {
          const double2 di = g.d_in[row], dj = g.d_in[col];
          const double sr = di.x + dj.x, si = di.y + dj.y;
          const int64_t oy = (int64_t)row * g.ldy + col;
          const double yr = (double)g.yr[oy], yi = (double)g.yi[oy];
          vr += sr * yr - si * yi;
          vi += sr * yi + si * yr;
        }
        const float outr = (float)vr, outi = (float)vi;
        g.cr[(int64_t)row * g.ldc + col] = outr;
        g.ci[(int64_t)row * g.ldc + col] = outi;
        if (row != col) {
          g.cr[(int64_t)col * g.ldc + row] = outr;
          g.ci[(int64_t)col * g.ldc + row] = outi;
        }
      }
    }
  }
}

template <class P, int WAVES_M = P::WAVES_M, int WAVES_N = P::WAVES_N, int KPB = P::KPB>
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N) void sq_split_kernel(SqArgs<P> g) {
  int tm, tn;
  const int lb = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  if (g.order) {
    const int code = g.order[lb];
    tm = code >> 16;
    tn = code & 0xFFFF;
  } else {
    tri_tile(lb, tm, tn);
  }
  sq_split_tile<P, WAVES_M, WAVES_N, KPB>(g, tm, tn);
}

// Grouped split squarings: one problem per walk, lower tiles of SQ_BM; the
// same tile placement and persistent striding as gemm_group_kernel.
template <class P> struct GroupSq {
  SqArgs<P> g;
  int tile0;
  int tiles;
};

template <class P>
__global__ __launch_bounds__(sq_threads<P>()) void sq_split_group_kernel(const GroupSq<P>* __restrict__ probs,
                                                                         int nprob, int total, int chunk) {
  for (int b = (int)blockIdx.x; b < total; b += (int)gridDim.x) {
    const int t = xcd_chunk_remap(b, total, chunk);
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].tile0 <= t) lo = mid;
      else hi = mid - 1;
    }
    int tm, tn;
    grouped_tri_tile(t - probs[lo].tile0, (probs[lo].g.m + SQ_BM - 1) / SQ_BM, tm, tn);
    sq_split_tile<P, P::WAVES_M, P::WAVES_N, P::KPB>(probs[lo].g, tm, tn);
  }
}

}  // namespace fexpm
