// MFMA GEMM family for gfx950 (CDNA4), row-major "NT" layout:
//
//     C[i][j] = sum_k A[i][k] * B[j][k]
//
// Both operands are read along contiguous rows (k fastest), so a product
// P*Q of row-major matrices is NT(P, Q^T); symmetric right operands need no
// transpose at all.  One kernel template covers every dense contraction of
// the scaling-and-squaring pipeline:
//
//   MODE_REAL  one real accumulator               (structured Pade products)
//   MODE_C4M   complex planar, 4 real products    (LU / solve updates, FP64)
//   MODE_C3M   complex planar, 3 real products    (general squarings)
//              T1 = Ar Br^T, T2 = Ai Bi^T, T3 = (Ar+Ai)(Br+Bi)^T
//              Re = T1 - T2, Im = T3 - T1 - T2
//   SYM        only lower-triangular output tiles are computed; each value
//              is written at (i,j) and (j,i) by the same lane, so the result
//              is exactly symmetric.  Used when the product is symmetric in
//              exact arithmetic (polynomials in one symmetric matrix, and
//              X X^T for complex-symmetric X: the 3M squaring of
//              X = Xr + i Xi with Xr, Xi symmetric costs 3 half products).
//
// Tiles: a workgroup computes BM x BN outputs with (BM/WM)*(BN/WN) waves of
// 64 lanes; every wave owns a WM x WN block made of MF x MF MFMA blocks:
// 16x16 (v_mfma_{f32,f64}_16x16x4, MF = 16) or, for FP32 by default, 32x32
// (v_mfma_f32_32x32x2, MF = 32).  Operand tiles of BK=16 columns are staged
// global -> registers -> LDS with double buffering; rows keep k contiguous
// so one 16-byte LDS read feeds four MFMA k-steps: lane l of a 16x16x4 MFMA
// supplies A[l%16][k] with k = 4*(l/16) + step (of a 32x32x2 one A[l%32][k]
// with k = 4*(l/32) + step, two reads per k-tile), and the same permutation
// of k is used for B, so the products pair correctly.  Accumulator element r
// of lane l holds column l%16 and row acc_row<T>(l, r) of a 16x16 block,
// column l%32 and row acc_row32(l, r) of a 32x32 block.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fexpm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <typename T> struct AccT;
template <> struct AccT<float> { using type = f32x4; };
template <> struct AccT<double> { using type = f64x4; };

template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };

__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// 32x32 FP32 blocks: lane l supplies A[l%32][k] and B^T[l%32][k] with
// k = l/32 of the MFMA's two k values; accumulator element r of lane l holds
// column l%32 and row 8*(r/4) + 4*(l/32) + r%4 of the 32x32 block.  Same
// issue rate per flop as 16x16x4 on paper, measured faster at two waves per
// SIMD (151.9 vs 134.1 TFLOP/s in the register probe).
__device__ __forceinline__ f32x16 mfma32(float a, float b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int acc_row32(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// MODE_SQ2 (structured squarings only, with SYM and the diagonal split):
// X^2 of X = Xr + i Xi whose symmetric parts commute, from two products,
//   Re = (Xr + Xi)(Xr - Xi)^T,  Im = 2 Xr Xi^T
// (equal to Xr^2 - Xi^2 and Xr Xi + Xi Xr when Xr Xi = Xi Xr), two thirds
// of the 3M flops.  Rounding errors that do not commute are not damped by
// it, so at most the first squarings may use it (an engine option, off by
// default: with two MFMAs per k-step the loop is bound by its operand loads
// and saves only about 7% of a squaring's time).
enum GemmMode { MODE_REAL = 0, MODE_C4M = 1, MODE_C3M = 2, MODE_SQ2 = 3 };

// Row (within a 16x16 block) of accumulator element r of a lane; the column
// is lane % 16.  The FP32 16x16x4 MFMA keeps four consecutive rows per lane,
// the FP64 one interleaves them with stride 4.
template <typename T> __device__ __forceinline__ int acc_row(int lane, int r);
template <> __device__ __forceinline__ int acc_row<float>(int lane, int r) { return 4 * (lane >> 4) + r; }
template <> __device__ __forceinline__ int acc_row<double>(int lane, int r) { return (lane >> 4) + 4 * r; }

// Epilogue: C = alpha*acc + beta*C_old + sum_t addc[t]*ADD_t + diag*I.
// ADD_t are planar (re, im) matrices sharing ld_add; for MODE_REAL only the
// re planes are used.  diag is added to the real part only.
template <typename T> struct Epilogue {
  T* cr;
  T* ci;
  int64_t ldc;
  double alpha;
  double beta;
  int nadd;
  const T* add_r[3];
  const T* add_i[3];
  double addc[3];
  int64_t ld_add;
  double diag;
  // Diagonal-split squaring (MODE_C3M only): the operand is Y = X - diag(d)
  // (zero diagonal, the A planes), d in FP64.  Then
  //   X^2 = diag(d^2 + (YY)_ii) + [(d_i + d_j) Y_ij + (YY)_ij]_{i != j},
  // so the MFMA never sums products of an O(1) diagonal with O(eps) terms.
  // Writes d_out and leaves zeros on the output diagonal.
  const double2* d_in;
  double2* d_out;
};

template <typename T> struct GemmArgs {
  const T* ar;
  const T* ai;
  int64_t lda;
  const T* br;
  const T* bi;
  int64_t ldb;
  int m, n, k;  // k must be a multiple of 16; operand columns beyond the
                // logical k are zero in storage
  int kblock;   // KBLK kernels: flush the MFMA accumulators every kblock
                // k-tiles of 16 into the outer sums (>= 1)
  Epilogue<T> epi;
};

// Tile configuration: workgroup tile BM x BN, k-tile BK, LDS row padding
// PAD, wave tile WM x WN, MFMA block edge MF (16: 16x16x4; 32: FP32
// 32x32x2, BK = 16); NT threads.
template <int BM_, int BN_, int BK_, int PAD_, int WM_, int WN_, int MF_ = 16, int MINW_ = 1> struct TileCfg {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, PAD = PAD_, WM = WM_, WN = WN_, MF = MF_;
  static constexpr int NT = (BM / WM) * (BN / WN) * 64;
  static constexpr int MINW = MINW_;  // waves per SIMD the register budget must allow
};
template <typename T> struct GemmCfg;
template <> struct GemmCfg<float> : TileCfg<128, 128, 16, 4, 64, 32, 32> {};
// FP32 squarings of small orders: 4x the workgroups of the default tile
// (a 128-wide tile grid of a few hundred rows leaves most CUs idle); four
// waves per SIMD (128 VGPRs: four workgroups per CU, 18-25% faster at
// n = 1024-4096 than the unconstrained 146 VGPRs).
using FCfgSmall = TileCfg<64, 64, 16, 4, 32, 32, 32, 4>;
// FP32 128 x 128 tiles as 16 waves of 32x32 (the per-wave work of FCfgSmall,
// four waves per SIMD, one workgroup per CU): half the operand loads per
// MFMA of the 64 x 64 tiles; 26.0 ms per n = 10^4 symmetric 3M squaring vs
// 27.2 ms for GemmCfg<float>.
using FCfg16w = TileCfg<128, 128, 16, 4, 32, 32, 32, 4>;
// FP64: 8 waves of 32x16, compiled for four waves per SIMD (at most 128
// VGPRs, two workgroups per CU); measured 63 TFLOP/s (4M, K = 256) and
// 61 TFLOP/s (real symmetric, K = 10^4) on MI355X versus 50 / 57 for 4 waves
// of 32x32.  The 3M kernel needs 134 VGPRs unconstrained (three waves per
// SIMD, so one workgroup per CU): within 128 it runs 57.4 -> 46.9 ms per
// n = 10^4 symmetric product.
template <> struct GemmCfg<double> : TileCfg<64, 64, 16, 2, 32, 16, 16, 4> {};
// FP64 4 waves of 32x32 wave tiles: the grouped (batched) launches' wider
// candidate.
using DCfg4w = TileCfg<64, 64, 16, 2, 32, 32>;

// Staging of a ROWS x BK tile through registers, NT threads.
template <typename T, int ROWS, int BK, int PAD, int NT> struct TileStage {
  using V = typename Vec16<T>::type;
  static constexpr int VEC = 16 / (int)sizeof(T);
  static constexpr int CPR = BK / VEC;
  static constexpr int RPP = NT / CPR;
  static constexpr int PASSES = (ROWS + RPP - 1) / RPP;
  V r[PASSES];

  __device__ __forceinline__ void load(const T* __restrict__ p, int64_t ld, int nvalid, int k0) {
    const int t = threadIdx.x, c = t % CPR, rr = t / CPR;
#pragma unroll
    for (int q = 0; q < PASSES; ++q) {
      const int lr = rr + q * RPP;
      if (lr < ROWS && lr < nvalid)
        r[q] = *reinterpret_cast<const V*>(p + (int64_t)lr * ld + k0 + c * VEC);
      else
        r[q] = V{};
    }
  }
  __device__ __forceinline__ void store(T* lds) const {
    const int t = threadIdx.x, c = t % CPR, rr = t / CPR;
#pragma unroll
    for (int q = 0; q < PASSES; ++q) {
      const int lr = rr + q * RPP;
      if (lr < ROWS) *reinterpret_cast<V*>(lds + lr * (BK + PAD) + c * VEC) = r[q];
    }
  }
};

// Four consecutive k values of one row of an LDS tile.
template <int BK, int PAD>
__device__ __forceinline__ void frag4(const float* lds, int row, int kq, float out[4]) {
  const float4 v = *reinterpret_cast<const float4*>(lds + row * (BK + PAD) + kq);
  out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
}
template <int BK, int PAD>
__device__ __forceinline__ void frag4(const double* lds, int row, int kq, double out[4]) {
  const double2 v0 = *reinterpret_cast<const double2*>(lds + row * (BK + PAD) + kq);
  const double2 v1 = *reinterpret_cast<const double2*>(lds + row * (BK + PAD) + kq + 2);
  out[0] = v0.x; out[1] = v0.y; out[2] = v1.x; out[3] = v1.y;
}

// Lower-triangular tile index: b -> (ti, tj), ti >= tj, row-major over ti.
__device__ __forceinline__ void tri_tile(int b, int& ti, int& tj) {
  int i = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
  while ((i + 1) * (i + 2) / 2 <= b) ++i;
  while (i * (i + 1) / 2 > b) --i;
  ti = i;
  tj = b - i * (i + 1) / 2;
}

// Lower-triangular tile index b -> (ti, tj) in groups of TILE_GROUP x
// TILE_GROUP tiles: group rows gi in order, within one the groups gj <= gi,
// within a group its tiles row-major (tj <= ti in the diagonal group).  The
// same enumeration as sym_tile_order on the host, computed in place.  The
// tiles that run together on one XCD (consecutive b after xcd_remap) then
// share both the row blocks of A and those of B in that XCD's L2.
constexpr int TILE_GROUP = 8;
__host__ __device__ inline void grouped_tri_tile(int b, int T, int& ti, int& tj) {
  constexpr int G = TILE_GROUP;
  int base = 0;
  for (int gi = 0; gi * G < T; ++gi) {
    const int h = T - gi * G < G ? T - gi * G : G;  // tile rows in this group row
    const int full = gi * h * G;                    // tiles of its off-diagonal groups
    const int cnt = full + h * (h + 1) / 2;
    if (b < base + cnt) {
      const int r = b - base;
      if (r < full) {
        const int gj = r / (h * G), w = r % (h * G);
        ti = gi * G + w / G;
        tj = gj * G + w % G;
      } else {
        const int w = r - full;
        int t = 0;
        while ((t + 1) * (t + 2) / 2 <= w) ++t;
        ti = gi * G + t;
        tj = gi * G + (w - t * (t + 1) / 2);
      }
      return;
    }
    base += cnt;
  }
  ti = 0;  // b beyond T(T+1)/2 (never launched)
  tj = 0;
}

// Rectangular tile grid (nbm x nbn) in groups of TILE_GROUP tile rows: a
// group sweeps its columns, visiting its rows for each column, so
// consecutive tiles share the row blocks of B as well as those of A.
__host__ __device__ inline void grouped_rect_tile(int lin, int nbm, int nbn, int& ti, int& tj) {
  constexpr int G = TILE_GROUP;
  const int per = G * nbn;
  const int first = (lin / per) * G;
  const int h = nbm - first < G ? nbm - first : G;
  const int r = lin % per;
  ti = first + r % h;
  tj = r / h;
}

// Spread consecutive tiles over the 8 XCDs so that workgroups that run on
// one XCD (dispatched round-robin: block b on XCD b % 8) work on
// neighbouring tiles and share their operand rows in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  constexpr int NX = 8;
  const int per = nblocks / NX;
  const int full = per * NX;
  if (b >= full) return b;
  return (b % NX) * per + b / NX;
}

// KBLK (MODE_C3M only): blocked accumulation.  The MFMA adds every product
// into its FP32 accumulator one at a time, so a long sequential sum loses
// terms smaller than half an ulp of the running sum.  With KBLK the three
// accumulators cover kblock*16 values of k at a time and are then folded
// into outer Re = T1 - T2 and Im = T3 - T1 - T2 sums, which bounds the
// length of every sequential MFMA sum.
//
// gemm_tile computes output tile (tm, tn) of one problem; `sym` selects the
// lower-triangular form (diagonal tiles keep row >= col, values mirrored).
// It ends with every LDS read complete, so a workgroup may call it again
// for another tile.
template <typename T, int MODE, bool KBLK, class CFG>
__device__ __forceinline__ void gemm_tile(const GemmArgs<T>& g, int tm, int tn, bool sym) {
  using C = CFG;
  constexpr int BM = C::BM, BN = C::BN, BK = C::BK, PAD = C::PAD, WM = C::WM, WN = C::WN, MF = C::MF;
  static_assert(BK % 16 == 0 && (MF == 16 || (MF == 32 && sizeof(T) == 4)), "32x32 blocks: FP32; BK: 16s");
  constexpr int WAVES_N = BN / WN;
  constexpr int NT = (BM / WM) * WAVES_N * 64;
  constexpr int FI = WM / MF, FJ = WN / MF;
  constexpr int NRD = MF == 32 ? BK / 8 : BK / 16;  // 16-byte fragment reads per row and k-tile
  constexpr int KSTEP = MF == 32 ? 8 : 16;            // k covered by one read across the wave
  constexpr int NS = 4 * NRD;            // MFMA steps per k-tile
  constexpr int NEL = MF == 32 ? 16 : 4;  // accumulator elements per lane
  constexpr int NPL = (MODE == MODE_REAL) ? 1 : 2;
  constexpr int NACC = (MODE == MODE_REAL) ? 1 : (MODE == MODE_C3M ? 3 : 2);
  constexpr int TA = BM * (BK + PAD), TB = BN * (BK + PAD);
  using Acc = typename std::conditional<MF == 32, f32x16, typename AccT<T>::type>::type;

  __shared__ __attribute__((aligned(16))) T smem[2][NPL * (TA + TB)];

  const int row0 = tm * BM, col0 = tn * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave / WAVES_N) * WM, wc = (wave % WAVES_N) * WN;
  // 16x16x4: lane l reads k 4*(l/16) .. +3 of row l%16 (four steps);
  // 32x32x2: lane l reads k 4*(l/32) .. +3 and 8 + the same of row l%32
  // (eight steps; lanes below 32 give the first k of each step)
  const int kq = MF == 32 ? 4 * (lane >> 5) : 4 * (lane >> 4);
  const int lrow = lane & (MF - 1);

  Acc acc[NACC][FI][FJ];
#pragma unroll
  for (int a = 0; a < NACC; ++a)
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j) acc[a][i][j] = Acc{};
  static_assert(!KBLK || MODE == MODE_C3M || MODE == MODE_SQ2, "blocked accumulation: squaring kernels");
  constexpr int NOUT = KBLK ? 2 : 1;
  Acc outer[NOUT][FI][FJ];
#pragma unroll
  for (int a = 0; a < NOUT; ++a)
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j) outer[a][i][j] = Acc{};

  const T* pa[2] = {g.ar + (int64_t)row0 * g.lda, (NPL == 2) ? g.ai + (int64_t)row0 * g.lda : nullptr};
  const T* pb[2] = {g.br + (int64_t)col0 * g.ldb, (NPL == 2) ? g.bi + (int64_t)col0 * g.ldb : nullptr};
  const int mvalid = g.m - row0, nvalid = g.n - col0;

  TileStage<T, BM, BK, PAD, NT> sa[NPL];
  TileStage<T, BN, BK, PAD, NT> sb[NPL];

  const int KT = g.k / BK;
#pragma unroll
  for (int p = 0; p < NPL; ++p) {
    sa[p].load(pa[p], g.lda, mvalid, 0);
    sb[p].load(pb[p], g.ldb, nvalid, 0);
  }
#pragma unroll
  for (int p = 0; p < NPL; ++p) {
    sa[p].store(&smem[0][p * (TA + TB)]);
    sb[p].store(&smem[0][p * (TA + TB) + TA]);
  }
  __syncthreads();

  for (int kt = 0; kt < KT; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < KT) {
#pragma unroll
      for (int p = 0; p < NPL; ++p) {
        sa[p].load(pa[p], g.lda, mvalid, (kt + 1) * BK);
        sb[p].load(pb[p], g.ldb, nvalid, (kt + 1) * BK);
      }
    }
    T fa[NPL][FI][NS], fb[NPL][FJ][NS];
#pragma unroll
    for (int p = 0; p < NPL; ++p) {
      const T* la = &smem[cur][p * (TA + TB)];
      const T* lb = la + TA;
#pragma unroll
      for (int q = 0; q < NRD; ++q) {
#pragma unroll
        for (int i = 0; i < FI; ++i) frag4<BK, PAD>(la, wr + MF * i + lrow, kq + KSTEP * q, &fa[p][i][4 * q]);
#pragma unroll
        for (int j = 0; j < FJ; ++j) frag4<BK, PAD>(lb, wc + MF * j + lrow, kq + KSTEP * q, &fb[p][j][4 * q]);
      }
    }
    auto mfma = [](T a, T b, Acc c) -> Acc {
      if constexpr (MF == 32) return mfma32(a, b, c);
      else return mfma16(a, b, c);
    };
#pragma unroll
    for (int s = 0; s < NS; ++s) {
#pragma unroll
      for (int i = 0; i < FI; ++i) {
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
          if constexpr (MODE == MODE_REAL) {
            acc[0][i][j] = mfma(fa[0][i][s], fb[0][j][s], acc[0][i][j]);
          } else if constexpr (MODE == MODE_C4M) {
            acc[0][i][j] = mfma(fa[0][i][s], fb[0][j][s], acc[0][i][j]);
            acc[0][i][j] = mfma(-fa[1][i][s], fb[1][j][s], acc[0][i][j]);
            acc[1][i][j] = mfma(fa[0][i][s], fb[1][j][s], acc[1][i][j]);
            acc[1][i][j] = mfma(fa[1][i][s], fb[0][j][s], acc[1][i][j]);
          } else if constexpr (MODE == MODE_SQ2) {
            acc[0][i][j] = mfma(fa[0][i][s] + fa[1][i][s], fb[0][j][s] - fb[1][j][s], acc[0][i][j]);
            acc[1][i][j] = mfma(fa[0][i][s], fb[1][j][s], acc[1][i][j]);
          } else {
            acc[0][i][j] = mfma(fa[0][i][s], fb[0][j][s], acc[0][i][j]);
            acc[1][i][j] = mfma(fa[1][i][s], fb[1][j][s], acc[1][i][j]);
            acc[2][i][j] = mfma(fa[0][i][s] + fa[1][i][s], fb[0][j][s] + fb[1][j][s], acc[2][i][j]);
          }
        }
      }
    }
    if constexpr (KBLK) {
      if ((kt + 1) % g.kblock == 0 || kt + 1 == KT) {
#pragma unroll
      for (int i = 0; i < FI; ++i) {
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
          if constexpr (MODE == MODE_SQ2) {
            outer[0][i][j] += acc[0][i][j];
            outer[1][i][j] += acc[1][i][j];
          } else {
            outer[0][i][j] += acc[0][i][j] - acc[1][i][j];
            outer[1][i][j] += acc[NACC - 1][i][j] - acc[0][i][j] - acc[1][i][j];
          }
#pragma unroll
          for (int a = 0; a < NACC; ++a) acc[a][i][j] = Acc{};
        }
      }
      }
    }
    if (kt + 1 < KT) {
      const int nxt = cur ^ 1;
#pragma unroll
      for (int p = 0; p < NPL; ++p) {
        sa[p].store(&smem[nxt][p * (TA + TB)]);
        sb[p].store(&smem[nxt][p * (TA + TB) + TA]);
      }
    }
    __syncthreads();
  }

  // ---- epilogue ----
  const Epilogue<T>& e = g.epi;
  const bool diag_tile = sym && (tm == tn);
#pragma unroll
  for (int i = 0; i < FI; ++i) {
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
#pragma unroll
      for (int r = 0; r < NEL; ++r) {
        const int row = row0 + wr + MF * i + (MF == 32 ? acc_row32(lane, r) : acc_row<T>(lane, r));
        const int col = col0 + wc + MF * j + lrow;
        if (row >= g.m || col >= g.n) continue;
        if (diag_tile && row < col) continue;
        double vr, vi = 0.0;
        if constexpr (MODE == MODE_REAL) {
          vr = (double)acc[0][i][j][r];
        } else if constexpr (MODE == MODE_C4M) {
          vr = (double)acc[0][i][j][r];
          vi = (double)acc[1][i][j][r];
        } else if constexpr (KBLK) {
          vr = (double)outer[0][i][j][r];
          vi = (double)outer[1][i][j][r];
        } else if constexpr (MODE == MODE_SQ2) {
          vr = (double)acc[0][i][j][r];
          vi = (double)acc[1][i][j][r];
        } else {
          const T t1 = acc[0][i][j][r], t2 = acc[1][i][j][r], t3 = acc[NACC - 1][i][j][r];
          vr = (double)(t1 - t2);
          vi = (double)(t3 - t1 - t2);
        }
        if constexpr (MODE == MODE_SQ2) vi *= 2.0;
        const int64_t o = (int64_t)row * e.ldc + col;
        if ((MODE == MODE_C3M || MODE == MODE_SQ2) && e.d_in) {
          if (row == col) {
            const double2 dd = e.d_in[row];
            e.d_out[row] = make_double2(dd.x * dd.x - dd.y * dd.y + vr, 2.0 * dd.x * dd.y + vi);
            vr = 0.0;
            vi = 0.0;
          } else {
            const double2 di = e.d_in[row], dj = e.d_in[col];
            const int64_t oy = (int64_t)row * g.lda + col;
            const double yr = (double)g.ar[oy], yi = (double)g.ai[oy];
            if constexpr (MODE == MODE_SQ2) {
              // (D_P Y_Q + Y_P D_Q) with P = Xr + Xi, Q = Xr - Xi; 2 (Dr Yi + Yr Di)
              vr += (di.x + di.y) * (yr - yi) + (yr + yi) * (dj.x - dj.y);
              vi += 2.0 * (di.x * yi + yr * dj.y);
            } else {
              const double sr = di.x + dj.x, si = di.y + dj.y;
              vr += sr * yr - si * yi;
              vi += sr * yi + si * yr;
            }
          }
        }
        T outr = (T)(e.alpha * vr);
        T outi = (T)(e.alpha * vi);
        if (e.beta != 0.0) {
          outr += (T)e.beta * e.cr[o];
          if (NPL == 2) outi += (T)e.beta * e.ci[o];
        }
        if (e.nadd > 0) {
          const int64_t oa = (int64_t)row * e.ld_add + col;
          for (int t = 0; t < e.nadd; ++t) {
            outr += (T)e.addc[t] * e.add_r[t][oa];
            if (NPL == 2) outi += (T)e.addc[t] * e.add_i[t][oa];
          }
        }
        if (row == col) outr += (T)e.diag;
        e.cr[o] = outr;
        if (NPL == 2) e.ci[o] = outi;
        if (sym && row != col) {
          const int64_t om = (int64_t)col * e.ldc + row;
          e.cr[om] = outr;
          if (NPL == 2) e.ci[om] = outi;
        }
      }
    }
  }
}

// One problem per launch: a rectangular tile grid (gridDim.y x gridDim.x) or,
// with SYM, T(T+1)/2 lower-triangular tiles; consecutive tiles are spread
// over the XCDs by xcd_remap.
template <typename T, int MODE, bool SYM, bool KBLK = false, class CFG = GemmCfg<T>>
__global__ __launch_bounds__(CFG::NT, CFG::MINW)
void gemm_nt_kernel(GemmArgs<T> g) {
  int tm, tn;
  if (SYM) {
    const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    grouped_tri_tile(b, (g.m + CFG::BM - 1) / CFG::BM, tm, tn);
  } else {
    const int lin = xcd_remap((int)(blockIdx.y * gridDim.x + blockIdx.x), (int)(gridDim.x * gridDim.y));
    grouped_rect_tile(lin, (int)gridDim.y, (int)gridDim.x, tm, tn);
  }
  gemm_tile<T, MODE, KBLK, CFG>(g, tm, tn, SYM);
}

// Tail of a symmetric product: lower-triangular tiles [t0, t0 + count) of
// edge 2 * CFG::BM, each computed as its sub-tiles of edge CFG::BM (three of
// the four for a diagonal tile).  Pairs with a launch of tiles [0, t0) in the
// larger configuration when those leave a nearly empty last wave of
// workgroups on the device; every output value sees the same k order and
// accumulation blocks as in the larger tile.
template <typename T, int MODE, bool KBLK, class CFG>
__global__ __launch_bounds__(CFG::NT, CFG::MINW)
void gemm_sym_tail_kernel(GemmArgs<T> g, int t0) {
  const int q = (int)blockIdx.x;
  int tm, tn;
  grouped_tri_tile(t0 + q / 4, (g.m + 2 * CFG::BM - 1) / (2 * CFG::BM), tm, tn);
  const int sm = 2 * tm + ((q & 3) >> 1), sn = 2 * tn + (q & 1);
  if (sn > sm) return;  // upper sub-tile of a diagonal tile
  gemm_tile<T, MODE, KBLK, CFG>(g, sm, sn, true);
}

// ---------------------------------------------------------------------------
// Grouped launches: many independent problems (one per walk of a batch, or
// several per walk) in one persistent launch.  Problem p owns the tiles
// [tile0, tile0 + tiles) of the launch; problems are listed by the host in
// decreasing k, so the longest tiles are dispatched first.  Workgroup b
// takes tiles b, b + gridDim.x, ... of the chunk-remapped order.
template <typename T> struct GroupGemm {
  GemmArgs<T> g;
  int tile0;    // first tile of this problem within the launch
  int tiles;    // number of tiles
  int tiles_n;  // column tiles of a rectangular problem
  int sym;      // 1: lower-triangular tiles with mirrored stores
};

// Tile order of a grouped launch: workgroups are dispatched round-robin
// over the 8 XCDs (b % 8), so logical tile t is placed such that every XCD
// receives runs of `chunk` consecutive tiles (sharing operand rows in its
// L2) while chunks alternate between XCDs (balancing uneven tiles).
// A bijection on [0, total): the tail beyond whole 8*chunk groups maps to itself.
__device__ __forceinline__ int xcd_chunk_remap(int b, int total, int chunk) {
  constexpr int NX = 8;
  const int grp = NX * chunk;
  const int full = (total / grp) * grp;
  if (chunk <= 1 || b >= full) return b;
  const int x = b % NX, slot = b / NX;
  return ((slot / chunk) * NX + x) * chunk + slot % chunk;
}

template <typename T, int MODE, bool KBLK, class CFG>
__global__ __launch_bounds__(CFG::NT, CFG::MINW)
void gemm_group_kernel(const GroupGemm<T>* __restrict__ probs, int nprob, int total, int chunk) {
  for (int b = (int)blockIdx.x; b < total; b += (int)gridDim.x) {
    const int t = xcd_chunk_remap(b, total, chunk);
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].tile0 <= t) lo = mid;
      else hi = mid - 1;
    }
    const GroupGemm<T>& P = probs[lo];
    const int local = t - P.tile0;
    int tm, tn;
    if (P.sym) grouped_tri_tile(local, P.tiles_n, tm, tn);  // tiles_n = tiles per side
    else grouped_rect_tile(local, P.tiles / P.tiles_n, P.tiles_n, tm, tn);
    gemm_tile<T, MODE, KBLK, CFG>(P.g, tm, tn, P.sym != 0);
  }
}

}  // namespace fexpm
