// FP32-level products on the int8 MFMA of gfx950 by exact residue
// arithmetic: the integer-modular form of the Ozaki scheme (Ozaki, Uchino &
// Imamura, "Ozaki Scheme II: a GEMM-oriented emulation of floating-point
// matrix multiplication using an integer modular technique", 2024).
//
// Used for the structured squarings X^2 = X X^T of complex64 engines, with
// X = diag(d) + Y (Y = Yr + i Yi symmetric, zero diagonal, FP32 planes):
//
//   1. Every row r of Y is scaled by 2^(BETA - e_r), e_r the exponent with
//      max_k max(|Yr_rk|, |Yi_rk|) < 2^e_r, and rounded to integers
//      yr', yi' (|.| <= 2^BETA); s' = yr' + yi'.
//   2. For each modulus m of a set of pairwise coprime odd moduli (product
//      M), the symmetric residues of yr', yi', s' (int8 in [-127, 127]) are
//      the operands of three int8 products T1 = yr' yr'^T, T2 = yi' yi'^T,
//      T3 = s' s'^T mod m.  Their int32 sums are exact (k 127^2 < 2^31).
//   3. Re = T1 - T2 and Im = T3 - T1 - T2 are formed mod m (exact: the 3M
//      combination costs no accuracy here), and the Chinese remainder theorem
//      gives the integers Re, Im themselves: |Re|, |Im| <= 2^(2 BETA + 1) k
//      < M / 2.  X/M = frac(sum_m s_m / m), s_m = (residue * u_m) mod m with
//      u_m = (M / m)^-1 mod m, evaluated in FP64 (error about 2^-48 of M,
//      i.e. 2^23 integer units against values of about 2^56).
//   4. (Y Y^T)_ij = 2^(e_i + e_j - 2 BETA) (Re + i Im), then the diagonal-split
//      epilogue of the squaring (FP64 diagonal d) and the FP32 planes.
//
// The only rounding is that of step 1 (BETA bits of each row's largest
// element) and of the FP32 output; the products and their sums are exact,
// unlike FP32 MFMA accumulation.  K of the nine moduli of about 2^7.9 are
// used (K = 7, 8 or 9: M of 55.5, 63.4 or 71.2 bits); BETA is the largest
// value <= 27 with k 2^(2 BETA + 1) < M / 2 (crt_beta on the host): 27 for
// K = 9 and k <= 16384, 24 for K = 8 and 2729 < k <= 10915, 20 for K = 7 and
// 3051 < k <= 12203.
//
// Layouts.  Operand residues ("fragment layout"): for a k-block kb (32 values
// of k) and a group g of 32 rows, one 1 KB fragment at ((kb * G + g) * 1024)
// bytes of its plane (G = rows / 32), in which lane l of a wave owns bytes
// 16 l .. 16 l + 15: row 32 g + l % 32, k values 32 kb + 16 (l / 32) .. + 15,
// which is the operand layout of v_mfma_i32_32x32x32_i8; one
// global_load_lds_dwordx4 of a wave moves one fragment into LDS unchanged.
// Product residues: one byte (C mod m in [0, m)) per element of the lower
// tiles, stored transposed, at byte (col * rows + row) of its plane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fexpm_gemm.hpp"

namespace fexpm {

constexpr int CRT_K = 9;         // most moduli of the squaring products (table capacity)
constexpr int CRT_KMIN = 7;      // fewest moduli of the squaring products
constexpr int CRT_BETA = 27;     // most bits kept of each row's largest element
constexpr int CRT_KMAX = 16384;  // longest k (keeps every bound above)
constexpr int CRT_NMIN = 768;    // automatic mode: smallest order for which the residue form is faster
constexpr int CRT_TILE = 256;    // workgroup tile edge of the residue products
constexpr int CRT_KPAD = 64;     // k granularity (one LDS stage: two k-blocks)
constexpr int CRT_NPROB = 3 * CRT_K;

// Moduli and reconstruction constants of K moduli (filled on the host by
// crt_table(K); the first K entries are used).
struct CrtTable {
  uint32_t m[CRT_K];
  uint32_t magic[CRT_K];  // ceil(2^39 / m): floor(u / m) = mulhi(u, magic) >> 7 for 0 <= u < 2^31
  uint32_t inv[CRT_K];    // ((M / m)^-1) mod m
  double rcp[CRT_K];      // 1 / m
  double M;               // product of the K moduli (rounded)
  int k;                  // number of moduli K
};

// u mod m for 0 <= u < 2^31 (exact: the error term of the rounded-up
// reciprocal stays below 1 / m of the quotient's fraction).
__device__ __forceinline__ uint32_t crt_umod(uint32_t u, uint32_t m, uint32_t magic) {
  const uint32_t q = __umulhi(u, magic) >> 7;
  return u - q * m;
}

// Symmetric residue in [-(m-1)/2, (m-1)/2] of an integer |x| <= 2^28 (odd
// m in [227, 255]): u = x + m 2^21 lies in [0, 2^30), since m 2^21 > 2^28.
__device__ __forceinline__ int crt_sres(int x, uint32_t m, uint32_t magic) {
  const uint32_t u = (uint32_t)(x + (int)(m << 21));
  const int r = (int)crt_umod(u, m, magic);
  return r > (int)(m >> 1) ? r - (int)m : r;
}

// Symmetric residue of a + b from symmetric residues a, b (|a + b| <= m - 1).
__device__ __forceinline__ int crt_sadd(int a, int b, uint32_t m) {
  const int h = (int)(m >> 1);
  int r = a + b;
  r = r > h ? r - (int)m : r;
  return r < -h ? r + (int)m : r;
}

// Row exponents of a complex FP32 matrix (planes yr, yi, leading dimension
// ld, padding columns zero): e[r] = e with max_k max(|yr_rk|, |yi_rk|) < 2^e
// (0 for a zero row), rows r < n.  One wave per row, four rows per
// workgroup of 256 threads.
__global__ __launch_bounds__(256) void crt_row_exp(const float* __restrict__ yr, const float* __restrict__ yi,
                                                    int64_t ld, int n, int* __restrict__ e) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = (int)blockIdx.x * 4 + w;
  if (r >= n) return;  // whole waves leave; no workgroup barrier follows
  const int nv = (n + 3) & ~3;  // padding columns up to the next multiple of 4 are zero
  const float* pr = yr + (int64_t)r * ld;
  const float* pi = yi + (int64_t)r * ld;
  float mx = 0.0f;
  for (int k = 4 * lane; k < nv; k += 256) {
    const float4 a = *reinterpret_cast<const float4*>(pr + k);
    const float4 b = *reinterpret_cast<const float4*>(pi + k);
    mx = fmaxf(mx, fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))));
    mx = fmaxf(mx, fmaxf(fmaxf(fabsf(b.x), fabsf(b.y)), fmaxf(fabsf(b.z), fabsf(b.w))));
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  int ex = 0;
  if (mx > 0.0f) (void)frexpf(mx, &ex);  // mx = f 2^ex, f in [0.5, 1)
  if (lane == 0) e[r] = ex;
}

// Operand residues of the structured squaring with K moduli (3 K planes of
// rows x kpad bytes in the fragment layout: plane o * K + i holds operand o
// (0: yr', 1: yi', 2: s') mod moduli i; rows rounded to beta bits).  One wave per fragment
// (k-block kb, row group g), four per workgroup; rows >= n and k >= n are
// zero.  rows is a multiple of CRT_TILE, kpad of CRT_KPAD, kpad <= ld, and
// the padding columns k in [n, ld) of the planes are zero, so a lane reads
// its 16 values (k0 + 16 <= kpad <= ld) with four 16-byte loads per plane.
template <int K>
__global__ __launch_bounds__(256) void crt_residues_sq(const float* __restrict__ yr, const float* __restrict__ yi,
                                                        int64_t ld, int n, const int* __restrict__ e, int rows,
                                                        int kpad, int beta, CrtTable tab, int8_t* __restrict__ out) {
  static_assert(K >= CRT_KMIN && K <= CRT_K, "moduli count");
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int G = rows / 32;
  const int64_t jobs = (int64_t)G * (kpad / 32);
  const int64_t job = (int64_t)blockIdx.x * 4 + w;
  if (job >= jobs) return;
  const int kb = (int)(job / G), g = (int)(job % G);
  const int row = 32 * g + (lane & 31), k0 = 32 * kb + 16 * (lane >> 5);
  int xr[16], xi[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    xr[q] = 0;
    xi[q] = 0;
  }
  if (row < n && k0 < n) {
    const float sc = ldexpf(1.0f, beta - e[row]);  // exact power of two; |y sc| < 2^beta <= 2^27
    const float4* pr = reinterpret_cast<const float4*>(yr + (int64_t)row * ld + k0);
    const float4* pi = reinterpret_cast<const float4*>(yi + (int64_t)row * ld + k0);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float4 a = pr[c], b = pi[c];
      xr[4 * c + 0] = (int)rintf(a.x * sc);
      xr[4 * c + 1] = (int)rintf(a.y * sc);
      xr[4 * c + 2] = (int)rintf(a.z * sc);
      xr[4 * c + 3] = (int)rintf(a.w * sc);
      xi[4 * c + 0] = (int)rintf(b.x * sc);
      xi[4 * c + 1] = (int)rintf(b.y * sc);
      xi[4 * c + 2] = (int)rintf(b.z * sc);
      xi[4 * c + 3] = (int)rintf(b.w * sc);
    }
  }
  const int64_t plane = (int64_t)rows * kpad;
  const int64_t off = ((int64_t)kb * G + g) * 1024 + lane * 16;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const uint32_t m = tab.m[i], mg = tab.magic[i];
    uint32_t pk[3][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      uint32_t a = 0, b = 0, s = 0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int q = 4 * c + t;
        const int ra = crt_sres(xr[q], m, mg), rb = crt_sres(xi[q], m, mg);
        a |= ((uint32_t)ra & 0xFFu) << (8 * t);
        b |= ((uint32_t)rb & 0xFFu) << (8 * t);
        s |= ((uint32_t)crt_sadd(ra, rb, m) & 0xFFu) << (8 * t);
      }
      pk[0][c] = a;
      pk[1][c] = b;
      pk[2][c] = s;
    }
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      uint4 v;
      v.x = pk[o][0];
      v.y = pk[o][1];
      v.z = pk[o][2];
      v.w = pk[o][3];
      *reinterpret_cast<uint4*>(out + (int64_t)(o * K + i) * plane + off) = v;
    }
  }
}

// One product of a batched residue launch: C = A B^T over the lower tiles,
// written as residues mod m.  A launch takes its list by value (kernel
// argument), so it needs no device buffer or host wait.
struct CrtProb {
  const int8_t* a;  // operand planes (fragment layout, `rows` rows)
  const int8_t* b;
  uint8_t* out;     // rows x rows bytes, element (row, col) at col * rows + row
  uint32_t m;
  uint32_t magic;
};
struct CrtProbs {
  CrtProb p[CRT_NPROB];
};

__device__ __forceinline__ void crt_dma16(const int8_t* g, int8_t* l) {
  __builtin_amdgcn_global_load_lds((const void*)g, (__attribute__((address_space(3))) void*)l, 16, 0, 0);
}
// s_waitcnt vmcnt(N) alone (expcnt and lgkmcnt at their maxima), gfx9 encoding.
template <int N> __device__ __forceinline__ void crt_wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Blocks of a batched residue launch over nprob problems of tiles_per lower
// tiles.  map 1 (the engine's default): each problem's tiles padded to a
// multiple of 8, so that the 8 XCDs work on one problem at a time (its
// operands shared in the last-level cache), XCD x on the x-th eighth of its
// grouped tile order; the padding blocks exit at once.  map 0: one block per
// tile, the launch's blocks spread over the XCDs in eighths (xcd_remap: XCD
// x takes the x-th eighth of all tiles of all problems).  At N = 10^4 a
// squaring launch takes 12.21 ms with map 1 and 12.61 ms with map 0.
static inline int crt_launch_blocks(int nprob, int tiles_per, int map) {
  return map ? nprob * ((tiles_per + 7) & ~7) : nprob * tiles_per;
}

// Batched residue products.  Workgroup: a 256 x 256 output tile of one
// problem, 4 waves (one per SIMD) of 128 x 128 (4 x 4 blocks of 32 x 32 on
// v_mfma_i32_32x32x32_i8, 256 accumulator registers).  Operand stages of two
// k-blocks (16 KB each: 8 A fragments then 8 B fragments) move into a ring of
// CRT_RING LDS buffers by direct-to-LDS copies issued CRT_RING - 1 stages
// ahead, interleaved with the MFMAs (each wave copies its share: fragments
// wave and wave + 4 of A and of B).  At the end of step s every wave waits for
// its copies of stage s + 2 and meets the others at a barrier, so stages
// <= s + 2 are complete in LDS and stage s - 1's buffer is free for the copy
// issued in step s + 1; the first fragments of stage s + 1 are read before
// that barrier.  Past the last stage the copies repeat the last k-block into a
// buffer no wave reads, so the loop has no branches around them.
constexpr int CRT_RING = 4;
__global__ __launch_bounds__(256, 1) void crt_gemm_kernel(CrtProbs probs, int tiles_per, int T, int G, int KB,
                                                          int map) {
  constexpr int KBS = 2;
  constexpr int KSTAGE = 16384;
  constexpr int SB = KSTAGE * KBS;
  constexpr int PER = 4 * KBS;  // copies per wave per stage
  __shared__ __attribute__((aligned(1024))) int8_t sm[CRT_RING * SB];
  // Block -> (problem, tile) by crt_launch_blocks' map; blocks are
  // dispatched to the XCDs round-robin (block b on XCD b % 8).
  int p, tile;
  if (map) {
    const int tp8 = (tiles_per + 7) & ~7;
    p = (int)blockIdx.x / tp8;
    const int j = (int)blockIdx.x - p * tp8;
    tile = (j & 7) * (tp8 >> 3) + (j >> 3);
    if (tile >= tiles_per) return;  // padding block (whole workgroup, before any barrier)
  } else {
    const int lin = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    p = lin / tiles_per;
    tile = lin - p * tiles_per;
  }
  int tm, tn;
  grouped_tri_tile(tile, T, tm, tn);  // 8 x 8 groups: 12.48 ms per N = 10^4 squaring, 4 x 4: 12.55, 16 x 16: 12.76
  const CrtProb P = probs.p[p];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  typedef int i32x4v __attribute__((ext_vector_type(4)));
  typedef int i32x16v __attribute__((ext_vector_type(16)));
  i32x16v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = i32x16v{};
  const int8_t* a0 = P.a + (int64_t)tm * 8192 + lane * 16;
  const int8_t* b0 = P.b + (int64_t)tn * 8192 + lane * 16;
  const int64_t kstride = (int64_t)G * 1024;  // bytes per k-block of a plane
  auto issue_kb = [&](int s, int kbl) {
    int8_t* buf = sm + (s % CRT_RING) * SB + kbl * KSTAGE;
    int64_t kb = (int64_t)s * KBS + kbl;
    kb = kb < KB - 1 ? kb : KB - 1;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int gq = wave + 4 * u;  // A group gq and B group gq (wave < 4)
      crt_dma16(a0 + kb * kstride + gq * 1024, buf + gq * 1024);
      crt_dma16(b0 + kb * kstride + gq * 1024, buf + 8192 + gq * 1024);
    }
  };
  i32x4v fa[2][4], fb[2][4];
  auto frags = [&](int s, int kbl, int slot) {
    const int8_t* q = sm + (s % CRT_RING) * SB + kbl * KSTAGE;
#pragma unroll
    for (int i = 0; i < 4; ++i) fa[slot][i] = *reinterpret_cast<const i32x4v*>(q + (wm * 4 + i) * 1024 + lane * 16);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      fb[slot][j] = *reinterpret_cast<const i32x4v*>(q + 8192 + (wn * 4 + j) * 1024 + lane * 16);
  };
  const int NS = KB / KBS;
#pragma unroll
  for (int q = 0; q <= CRT_RING - 2; ++q)
#pragma unroll
    for (int kbl = 0; kbl < KBS; ++kbl) issue_kb(q, kbl);
  crt_wait_vm<(CRT_RING - 3) * PER>();  // stages 0 and 1 landed
  __builtin_amdgcn_s_barrier();
  frags(0, 0, 0);
  int slot = 0;
  for (int s = 0; s < NS; ++s) {
#pragma unroll
    for (int kbl = 0; kbl < KBS; ++kbl) {
      if (kbl + 1 < KBS) frags(s, kbl + 1, slot ^ 1);
      else frags(s + 1, 0, slot ^ 1);  // past the last stage: a stale buffer, values unused
      issue_kb(s + CRT_RING - 1, kbl);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[slot][i], fb[slot][j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // one copy
      }
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
      slot ^= 1;
    }
    crt_wait_vm<(CRT_RING - 3) * PER>();  // this wave's copies of stage s + 2 landed
    __builtin_amdgcn_s_barrier();
  }
  crt_wait_vm<0>();  // no copy in flight when the workgroup's LDS is released
  // epilogue: residues mod m, four consecutive rows per 32-bit store
  const int l32 = lane & 31, kh = lane >> 5;
  const int rows = G * 32;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = tn * CRT_TILE + wn * 128 + j * 32 + l32;
#pragma unroll
      for (int rq = 0; rq < 4; ++rq) {
        uint32_t w4 = 0;
#pragma unroll
        for (int r3 = 0; r3 < 4; ++r3) {
          // |acc| <= k 127^2 <= 2^28.3 for k <= CRT64_KMAX and m 2^22 >= 199 2^22 > 2^29.6:
          // u = acc + m 2^22 in (0, 2^31)
          const uint32_t u = (uint32_t)(acc[i][j][4 * rq + r3] + (int)(P.m << 22));
          w4 |= crt_umod(u, P.m, P.magic) << (8 * r3);
        }
        const int row = tm * CRT_TILE + wm * 128 + i * 32 + 8 * rq + 4 * kh;
        *reinterpret_cast<uint32_t*>(P.out + (int64_t)col * rows + row) = w4;
      }
    }
}

// (residue * u) mod m for residue in [0, 3m): below 2^31 for m <= 255.
__device__ __forceinline__ uint32_t crt_digit(uint32_t r, uint32_t inv, uint32_t m, uint32_t magic) {
  return crt_umod(r * inv, m, magic);
}

// X^2 of the split form from the residues of T1, T2, T3 modulo K moduli
// (planes p, K + p, 2 K + p for modulus p; each rows x rows bytes at
// col * rows + row; operand rows rounded to beta bits),
// for the 64 x 64 lower tiles (TI >= TJ, lower-triangular index blockIdx.x):
//   (Y Y^T)_ij = 2^(e_i + e_j - 2 BETA) (Re + i Im) by CRT, then
//   i != j: X^2_ij = (d_i + d_j) Y_ij + (Y Y^T)_ij (complex),
//   i == j: d_out_i = d_i^2 + (Y Y^T)_ii and X^2_ii = 0 in the planes,
// written at (i, j) and (j, i) of cr / ci (rows and columns < n only).
// Thread t works on rows 4 (t % 16) .. + 3 of column (t / 16) + 16 c.
template <int K>
__global__ __launch_bounds__(256) void crt_combine_sq(const uint8_t* __restrict__ res, int rows,
                                                       const int* __restrict__ e, const float* __restrict__ yr,
                                                       const float* __restrict__ yi, int64_t ld, int n,
                                                       const double2* __restrict__ d_in, double2* __restrict__ d_out,
                                                       float* __restrict__ cr, float* __restrict__ ci, int beta,
                                                       CrtTable tab) {
  static_assert(K >= CRT_KMIN && K <= CRT_K, "moduli count");
  __shared__ float tre[64][65];
  __shared__ float tim[64][65];
  int TI, TJ;
  tri_tile((int)blockIdx.x, TI, TJ);
  const int i0 = TI * 64, j0 = TJ * 64;
  const int t = threadIdx.x;
  const int il = 4 * (t & 15);
  const int64_t plane = (int64_t)rows * rows;
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {
    const int jl = (t >> 4) + 16 * c;
    const int j = j0 + jl;
    const int ib = i0 + il;
    if (j >= n || ib >= n) continue;
    // residues of the 4 rows ib .. ib + 3 in column j, per product and modulus
    uint32_t w[3][K];
#pragma unroll
    for (int o = 0; o < 3; ++o)
#pragma unroll
      for (int q = 0; q < K; ++q)
        w[o][q] = *reinterpret_cast<const uint32_t*>(res + (o * K + q) * plane + (int64_t)j * rows + ib);
    const double2 dj = d_in[j];
    const int ej = e[j];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = ib + r;
      if (i >= n || i < j) continue;  // lower part only (diagonal tiles hold both)
      double yre = 0.0, yim = 0.0;
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const uint32_t m = tab.m[q], mg = tab.magic[q];
        const uint32_t t1 = (w[0][q] >> (8 * r)) & 0xFFu;
        const uint32_t t2 = (w[1][q] >> (8 * r)) & 0xFFu;
        const uint32_t t3 = (w[2][q] >> (8 * r)) & 0xFFu;
        const uint32_t re = t1 + m - t2;          // [0, 2m)
        const uint32_t im = t3 + 2 * m - t1 - t2;  // [0, 3m)
        yre += (double)crt_digit(re, tab.inv[q], m, mg) * tab.rcp[q];
        yim += (double)crt_digit(im, tab.inv[q], m, mg) * tab.rcp[q];
      }
      const double sc = ldexp(tab.M, e[i] + ej - 2 * beta);
      double vr = (yre - rint(yre)) * sc;
      double vi = (yim - rint(yim)) * sc;
      if (i == j) {
        const double2 dd = d_in[i];
        d_out[i] = make_double2(dd.x * dd.x - dd.y * dd.y + vr, 2.0 * dd.x * dd.y + vi);
        vr = 0.0;
        vi = 0.0;
      } else {
        const double2 di = d_in[i];
        const int64_t oy = (int64_t)j * ld + i;  // Y is symmetric: Y_ij = Y_ji
        const double y_r = (double)yr[oy], y_i = (double)yi[oy];
        const double sr = di.x + dj.x, si = di.y + dj.y;
        vr += sr * y_r - si * y_i;
        vi += sr * y_i + si * y_r;
      }
      const float fr = (float)vr, fi = (float)vi;
      cr[(int64_t)j * ld + i] = fr;  // (j, i): consecutive i across the threads of a column
      ci[(int64_t)j * ld + i] = fi;
      tre[il + r][jl] = fr;
      tim[il + r][jl] = fi;
    }
  }
  __syncthreads();
  // (i, j) from the tile in LDS: thread t writes column j0 + (t % 64) of rows i0 + t / 64 + 4 c
  const int jl = t & 63, j = j0 + jl;
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int il2 = (t >> 6) + 4 * c;
    const int i = i0 + il2;
    if (i >= n || j >= n || i < j) continue;
    cr[(int64_t)i * ld + j] = tre[il2][jl];
    ci[(int64_t)i * ld + j] = tim[il2][jl];
  }
}


// ---------------------------------------------------------------------------
// FP64-accurate real products by the same residue arithmetic (the symmetric
// FP64 products of the Pade stage and of the structured solve).  Up to
// thirteen moduli (the nine above and 223, 217, 211, 199; M of 102.2 bits;
// the first 8 .. 12 of them: 63.4, 71.2, 79.1, 86.9 or 94.6 bits); each row
// r of an operand is scaled by 2^(beta - e_r) and rounded to an integer of at
// most 2^beta, beta = min(45, floor((log2 M - 1 - log2 k) / 2)) (43 for k
// from 2^13.2 to 2^14.9), so that |C| <= k 2^(2 beta) < M / 2.  The reconstruction
// X / M = frac(sum_m s_m / m) runs in double-double arithmetic (1 / m as a
// pair of doubles, exact product errors by fma, TwoSum), X = f M to a few
// integer units, and C = 2^(e_i + f_j - 2 beta) X goes through the FP64
// GEMM epilogue (alpha, beta, added matrices, diagonal, mirrored lower tiles).
// Operand rows may be the concatenation s1 x1[r, 0:k1] | s2 x2[r, 0:nk-k1]
// (s1, s2 = +-1), as for S = [V U][V U]^T.
constexpr int CRT64_K = 13;         // most moduli (table capacity)
constexpr int CRT64_KMIN = 8;       // fewest moduli (M of 63.4 bits)
constexpr int CRT64_KMAX = 20480;   // longest k: k 127^2 < 2^28.3 (epilogue offset above)
constexpr int CRT64_BETA_MAX = 45;  // integers of at most 2^45 (crt64_sres)

struct CrtTable64 {
  uint32_t m[CRT64_K];
  uint32_t magic[CRT64_K];
  uint32_t inv[CRT64_K];   // ((M / m)^-1) mod m
  uint32_t c24[CRT64_K];   // 2^24 mod m
  double rcp_hi[CRT64_K];  // 1 / m = rcp_hi + rcp_lo (double-double)
  double rcp_lo[CRT64_K];
  double M_hi, M_lo;       // M = M_hi + M_lo
  double log2M;
  int k;                   // moduli used: the first k of the table
};

// Concatenated operand row r, value k (k < nk).
__device__ __forceinline__ double crt64_elem(const double* __restrict__ x1, double s1, const double* __restrict__ x2,
                                             double s2, int k1, int64_t ld, int r, int k) {
  return k < k1 ? s1 * x1[(int64_t)r * ld + k] : s2 * x2[(int64_t)r * ld + (k - k1)];
}

// e[r] = e with max_k |x_rk| < 2^e (0 for a zero row), rows r < nr.  One
// wave per row, four rows per workgroup.
__global__ __launch_bounds__(256) void crt64_row_exp(const double* __restrict__ x1, double s1,
                                                      const double* __restrict__ x2, double s2, int k1, int64_t ld,
                                                      int nr, int nk, int* __restrict__ e) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = (int)blockIdx.x * 4 + w;
  if (r >= nr) return;  // whole waves leave; no workgroup barrier follows
  double mx = 0.0;
  for (int k = lane; k < nk; k += 64) mx = fmax(mx, fabs(crt64_elem(x1, s1, x2, s2, k1, ld, r, k)));
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  int ex = 0;
  if (mx > 0.0) (void)frexp(mx, &ex);
  if (lane == 0) e[r] = ex;
}

// Split of an integer-valued double |x| <= 2^45: x = h 2^24 + l with
// h = floor(x 2^-24) in [-2^21, 2^21] and l in [0, 2^24) (both exact).
__device__ __forceinline__ void crt64_split(double x, int& h, uint32_t& l) {
  const double hd = floor(x * 0x1p-24);
  h = (int)hd;
  l = (uint32_t)(x - hd * 0x1p24);
}

// Symmetric residue (odd m in [199, 255]) of h 2^24 + l: with c24 = 2^24 mod
// m, u = h c24 + l + m 2^22 lies in (0, 2^31) (|h c24| < 2^29 <= m 2^22,
// l < 2^24), so one exact reduction suffices.
__device__ __forceinline__ int crt64_sres(int h, uint32_t l, uint32_t m, uint32_t magic, uint32_t c24) {
  const uint32_t u = (uint32_t)(h * (int)c24 + (int)(m << 22)) + l;
  const int r = (int)crt_umod(u, m, magic);
  return r > (int)(m >> 1) ? r - (int)m : r;
}

// Operand residues (tab.k planes of rows x kpad bytes, fragment layout) of
// x' = rint(x 2^(beta - e_r)); rows >= nr and k >= nk are zero.  One wave per
// fragment (k-block kb, row group g), four per workgroup.  vec: both parts
// start 16-byte aligned, ld is even and k1 % 16 == 0, so a lane's 16 values
// (k0 % 16 == 0) lie in one part and load as 8 pairs.
__global__ __launch_bounds__(256) void crt64_residues(const double* __restrict__ x1, double s1,
                                                       const double* __restrict__ x2, double s2, int k1, int64_t ld,
                                                       int nr, int nk, const int* __restrict__ e, int beta, int rows,
                                                       int kpad, int vec, CrtTable64 tab, int8_t* __restrict__ out) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int G = rows / 32;
  const int64_t jobs = (int64_t)G * (kpad / 32);
  const int64_t job = (int64_t)blockIdx.x * 4 + w;
  if (job >= jobs) return;
  const int kb = (int)(job / G), g = (int)(job % G);
  const int row = 32 * g + (lane & 31), k0 = 32 * kb + 16 * (lane >> 5);
  double x[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) x[q] = 0.0;
  if (row < nr && k0 < nk) {
    const double sc = ldexp(1.0, beta - e[row]);  // |x sc| < 2^beta
    if (vec && k0 + 16 <= nk) {
      const bool first = k0 < k1;
      const double* src = first ? x1 + (int64_t)row * ld + k0 : x2 + (int64_t)row * ld + (k0 - k1);
      const double sg = (first ? s1 : s2) * sc;  // +-2^(beta - e): exact products
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const double2 v = reinterpret_cast<const double2*>(src)[c];
        x[2 * c] = rint(v.x * sg);
        x[2 * c + 1] = rint(v.y * sg);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = k0 + q;
        if (k < nk) x[q] = rint(crt64_elem(x1, s1, x2, s2, k1, ld, row, k) * sc);
      }
    }
  }
  int h[16];
  uint32_t l[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) crt64_split(x[q], h[q], l[q]);
  const int64_t plane = (int64_t)rows * kpad;
  const int64_t off = ((int64_t)kb * G + g) * 1024 + lane * 16;
#pragma unroll 1
  for (int i = 0; i < tab.k; ++i) {
    const uint32_t m = tab.m[i], mg = tab.magic[i], c = tab.c24[i];
    uint4 v;
    uint32_t* pv = &v.x;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      uint32_t wv = 0;
#pragma unroll
      for (int t = 0; t < 4; ++t)
        wv |= ((uint32_t)crt64_sres(h[4 * cc + t], l[4 * cc + t], m, mg, c) & 0xFFu) << (8 * t);
      pv[cc] = wv;
    }
    *reinterpret_cast<uint4*>(out + (int64_t)i * plane + off) = v;
  }
}

// TwoSum (Knuth): s + err == a + b exactly.  The callers turn off
// floating-point contraction: an add fused with the product before it would
// round differently from the product that the error terms describe.
__device__ __forceinline__ void crt_two_sum(double a, double b, double& s, double& err) {
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  err = (a - (s - bb)) + (b - bb);
}

// C = A B^T from the tab.k residue planes (element (row, col) at byte
// col * rows + row of plane q), 64 x 64 lower tiles (index blockIdx.x), FP64
// epilogue of e (real planes; diag on the diagonal), values written at (i, j)
// and (j, i) for i, j < m.
__global__ __launch_bounds__(256) void crt64_combine(const uint8_t* __restrict__ res, int rows,
                                                      const int* __restrict__ ea, const int* __restrict__ eb,
                                                      int beta, int m, Epilogue<double> ep, CrtTable64 tab) {
  __shared__ double tv[64][65];
  int TI, TJ;
  tri_tile((int)blockIdx.x, TI, TJ);
  const int i0 = TI * 64, j0 = TJ * 64;
  const int t = threadIdx.x;
  const int il = 4 * (t & 15);
  const int64_t plane = (int64_t)rows * rows;
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {
    const int jl = (t >> 4) + 16 * c;
    const int j = j0 + jl;
    const int ib = i0 + il;
    if (j >= m || ib >= m) continue;
    uint32_t w[CRT64_K];
#pragma unroll
    for (int q = 0; q < CRT64_K; ++q)
      w[q] = q < tab.k ? *reinterpret_cast<const uint32_t*>(res + q * plane + (int64_t)j * rows + ib) : 0u;
    const int ej = eb[j];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = ib + r;
      if (i >= m || i < j) continue;
      double hi = 0.0, lo = 0.0;
#pragma unroll
      for (int q = 0; q < CRT64_K; ++q) {
#pragma clang fp contract(off)
        if (q >= tab.k) break;
        const uint32_t d = crt_umod(((w[q] >> (8 * r)) & 0xFFu) * tab.inv[q], tab.m[q], tab.magic[q]);
        const double sd = (double)d;
        const double p = sd * tab.rcp_hi[q];                                // the rounded product TwoSum adds
        const double pe = fma(sd, tab.rcp_hi[q], -p) + sd * tab.rcp_lo[q];  // its exact error, + d rcp_lo
        double s, err;
        crt_two_sum(hi, p, s, err);
        hi = s;
        lo += err + pe;
      }
      const double f = hi - rint(hi);  // exact; the fraction is f + lo
      const double x = f * tab.M_hi + (f * tab.M_lo + lo * tab.M_hi);
      const double v = ldexp(x, ea[i] + ej - 2 * beta);
      const int64_t o = (int64_t)i * ep.ldc + j;
      double out = ep.alpha * v;
      if (ep.beta != 0.0) out += ep.beta * ep.cr[o];
      if (ep.nadd > 0) {
        const int64_t oa = (int64_t)i * ep.ld_add + j;
        for (int a = 0; a < ep.nadd; ++a) out += ep.addc[a] * ep.add_r[a][oa];
      }
      if (i == j) out += ep.diag;
      tv[il + r][jl] = out;
      ep.cr[(int64_t)j * ep.ldc + i] = out;  // (j, i): consecutive i across the threads of a column
    }
  }
  __syncthreads();
  const int jl = t & 63, j = j0 + jl;
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int il2 = (t >> 6) + 4 * c;
    const int i = i0 + il2;
    if (i >= m || j >= m || i < j) continue;
    ep.cr[(int64_t)i * ep.ldc + j] = tv[il2][jl];
  }
}

}  // namespace fexpm
