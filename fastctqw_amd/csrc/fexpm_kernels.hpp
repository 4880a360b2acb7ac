// Bandwidth-bound kernels of the expm pipeline (analysis, conversion,
// linear combinations, transposes, output) and the FP64 blocked LU with
// partial pivoting used for the Pade denominator solve (V - U) X = (V + U).
//
// Matrices inside the engine are "planar": separate row-major re and im
// planes with a padded leading dimension.  Padding columns are zero and are
// never written, so GEMM k-loops may run over the padded extent.
#pragma once
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

namespace fexpm {

// ---------------------------------------------------------------------------
// Input element access: complex64 / complex128 (interleaved) or real.
template <typename Tin> struct InElem;
template <> struct InElem<float2> {
  static __device__ __forceinline__ double re(float2 v) { return v.x; }
  static __device__ __forceinline__ double im(float2 v) { return v.y; }
};
template <> struct InElem<double2> {
  static __device__ __forceinline__ double re(double2 v) { return v.x; }
  static __device__ __forceinline__ double im(double2 v) { return v.y; }
};
template <> struct InElem<float> {
  static __device__ __forceinline__ double re(float v) { return v; }
  static __device__ __forceinline__ double im(float) { return 0.0; }
};
template <> struct InElem<double> {
  static __device__ __forceinline__ double re(double v) { return v; }
  static __device__ __forceinline__ double im(double) { return 0.0; }
};

enum AnalyzeFlags { FLAG_NONZERO_RE = 1, FLAG_ASYM = 2, FLAG_NONFINITE = 4 };

// Double-double accumulation (Knuth's two-sum, additions only, so FMA
// contraction cannot change it): (hi, lo) holds a sum to about 106 bits.
__device__ __forceinline__ void dd_add(double& hi, double& lo, double x) {
  const double s = hi + x;
  const double bb = s - hi;
  const double err = (hi - (s - bb)) + (x - bb);
  hi = s;
  lo += err;
}
// Normalised pair: hi = fl(hi + lo), lo = the exact remainder.
__device__ __forceinline__ double2 dd_norm(double hi, double lo) {
  const double s = hi + lo;
  return make_double2(s, lo - (s - hi));
}

// One 32x32 tile pair (I, J), I >= J, per workgroup of 256 threads:
//  * exact symmetry test a_ij == a_ji (both parts), nonzero real parts,
//    non-finite entries -> atomic OR into flags;
//  * per-row-block partial column sums of |a_ij| over off-diagonal entries,
//    colpart[rowblock * n + col], each entry written exactly once, so the
//    later reduction over row blocks is deterministic;
//  * diagonal entries (re, im) into diag[];
//  * with colsig: per-row-block signed column sums of Im(a_ij), diagonal
//    included, in double-double (colsig[rowblock * n + col] = (hi, lo)).
//    For a symmetric matrix they give its exact row sums.
template <typename Tin>
__device__ __forceinline__ void analyze_tile_pair(const Tin* __restrict__ a, int n, int nt,
                                                  double* __restrict__ colpart, double2* __restrict__ diag,
                                                  int* flags, int b, double2* __restrict__ colsig = nullptr) {
  __shared__ double2 t0[32][33];
  __shared__ double2 t1[32][33];
  __shared__ int sflag;
  int I, J;
  {
    int i = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
    while ((i + 1) * (i + 2) / 2 <= b) ++i;
    while (i * (i + 1) / 2 > b) --i;
    I = i;
    J = b - i * (i + 1) / 2;
  }
  if (I >= nt) return;
  if (threadIdx.x == 0) sflag = 0;
  __syncthreads();
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  int fl = 0;
  for (int r = ty; r < 32; r += 8) {
    const int gi = I * 32 + r, gj = J * 32 + tx;
    double2 v = make_double2(0.0, 0.0);
    if (gi < n && gj < n) {
      const Tin e = a[(int64_t)gi * n + gj];
      v = make_double2(InElem<Tin>::re(e), InElem<Tin>::im(e));
    }
    t0[r][tx] = v;
    const int hi = J * 32 + r, hj = I * 32 + tx;
    double2 w = make_double2(0.0, 0.0);
    if (hi < n && hj < n) {
      const Tin e = a[(int64_t)hi * n + hj];
      w = make_double2(InElem<Tin>::re(e), InElem<Tin>::im(e));
    }
    t1[r][tx] = w;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int gi = I * 32 + r, gj = J * 32 + tx;
    if (gi < n && gj < n) {
      const double2 v = t0[r][tx];
      const double2 w = t1[tx][r];  // element (gj, gi)
      if (v.x != w.x || v.y != w.y) fl |= FLAG_ASYM;
      if (v.x != 0.0 || w.x != 0.0) fl |= FLAG_NONZERO_RE;
      if (!isfinite(v.x) || !isfinite(v.y) || !isfinite(w.x) || !isfinite(w.y)) fl |= FLAG_NONFINITE;
    }
  }
  if (fl) atomicOr(&sflag, fl);
  // column sums: tile (I,J) -> rowblock I, columns of block J;
  // tile (J,I) (if I != J) -> rowblock J, columns of block I.
  if (threadIdx.x < 32) {
    const int c = threadIdx.x;
    double s = 0.0;
    const int gj = J * 32 + c;
    for (int r = 0; r < 32; ++r) {
      const int gi = I * 32 + r;
      if (gi == gj) continue;
      const double2 v = t0[r][c];
      s += sqrt(v.x * v.x + v.y * v.y);
    }
    if (gj < n) colpart[(int64_t)I * n + gj] = s;
    if (I == J) {
      if (gj < n) diag[gj] = t0[c][c];
    }
  } else if (threadIdx.x < 64 && I != J) {
    const int c = threadIdx.x - 32;
    double s = 0.0;
    const int gj = I * 32 + c;
    for (int r = 0; r < 32; ++r) {
      const double2 v = t1[r][c];
      s += sqrt(v.x * v.x + v.y * v.y);
    }
    if (gj < n) colpart[(int64_t)J * n + gj] = s;
  }
  if (colsig) {
    if (threadIdx.x >= 64 && threadIdx.x < 96) {
      const int c = threadIdx.x - 64, gj = J * 32 + c;
      double hi = 0.0, lo = 0.0;
      for (int r = 0; r < 32; ++r) dd_add(hi, lo, t0[r][c].y);
      if (gj < n) colsig[(int64_t)I * n + gj] = dd_norm(hi, lo);
    } else if (threadIdx.x >= 96 && threadIdx.x < 128 && I != J) {
      const int c = threadIdx.x - 96, gj = I * 32 + c;
      double hi = 0.0, lo = 0.0;
      for (int r = 0; r < 32; ++r) dd_add(hi, lo, t1[r][c].y);
      if (gj < n) colsig[(int64_t)J * n + gj] = dd_norm(hi, lo);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && sflag) atomicOr(flags, sflag);
}

template <typename Tin>
__global__ __launch_bounds__(256) void analyze_tiles(const Tin* __restrict__ a, int n, int nt,
                                                     double* __restrict__ colpart,
                                                     double2* __restrict__ diag, int* flags,
                                                     double2* __restrict__ colsig) {
  analyze_tile_pair<Tin>(a, n, nt, colpart, diag, flags, (int)blockIdx.x, colsig);
}

// off[j] = sum over row blocks of colpart (fixed order -> deterministic);
// with colsig also rsum[j] = the double-double sum of colsig over row blocks.
__global__ void reduce_colpart(const double* __restrict__ colpart, int n, int nt, double* __restrict__ off,
                               const double2* __restrict__ colsig = nullptr, double2* __restrict__ rsum = nullptr) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int b = 0; b < nt; ++b) s += colpart[(int64_t)b * n + j];
  off[j] = s;
  if (colsig) {
    double hi = 0.0, lo = 0.0;
    for (int b = 0; b < nt; ++b) {
      const double2 v = colsig[(int64_t)b * n + j];
      dd_add(hi, lo, v.x);
      dd_add(hi, lo, v.y);
    }
    rsum[j] = dd_norm(hi, lo);
  }
}

// Scalars produced on the device and read back by the host once per run.
struct Scalars {
  double tr_re, tr_im;   // trace
  double mu_re, mu_im;   // tr / n
  double norm_plain;     // ||A||_1
  double norm_shift;     // ||A - mu I||_1
  int flags;
  int info;              // LU: nonzero if an exactly zero pivot was met
  int rowsum_equal;      // 1: every row sum of Im(A) equals rowsum_c (double-double sums)
  int pad_;
  double rowsum_c;       // that common row sum
};

// Single workgroup of 1024 threads: trace, mu and both 1-norms.  The trace
// is summed in a fixed tree order.  With rsum (the double-double row sums
// of Im(A)): whether all of them are equal, and their common value.
__device__ __forceinline__ void finalize_norms_block(const double* __restrict__ off,
                                                     const double2* __restrict__ diag, int n,
                                                     const int* flags, Scalars* out,
                                                     const double2* __restrict__ rsum = nullptr) {
  __shared__ double sr[1024], si[1024], s0[1024], s1[1024];
  __shared__ int s_uneq;
  const int t = threadIdx.x;
  if (t == 0) s_uneq = 0;
  __syncthreads();
  if (rsum && n > 0) {
    const double2 r0 = rsum[0];
    int uneq = 0;
    for (int j = t; j < n; j += 1024) {
      const double2 r = rsum[j];
      if (r.x != r0.x || r.y != r0.y) uneq = 1;
    }
    if (uneq) atomicOr(&s_uneq, 1);
  }
  double ar = 0.0, ai = 0.0;
  for (int j = t; j < n; j += 1024) {
    ar += diag[j].x;
    ai += diag[j].y;
  }
  sr[t] = ar;
  si[t] = ai;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) {
      sr[t] += sr[t + w];
      si[t] += si[t + w];
    }
    __syncthreads();
  }
  const double tr_re = sr[0], tr_im = si[0];
  const double mu_re = tr_re / n, mu_im = tr_im / n;
  double m0 = 0.0, m1 = 0.0;
  for (int j = t; j < n; j += 1024) {
    const double2 d = diag[j];
    const double c0 = off[j] + sqrt(d.x * d.x + d.y * d.y);
    const double dr = d.x - mu_re, di = d.y - mu_im;
    const double c1 = off[j] + sqrt(dr * dr + di * di);
    m0 = fmax(m0, c0);
    m1 = fmax(m1, c1);
    if (c0 != c0) m0 = c0;  // propagate NaN
    if (c1 != c1) m1 = c1;
  }
  __syncthreads();
  s0[t] = m0;
  s1[t] = m1;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) {
      const double x0 = s0[t + w], x1 = s1[t + w];
      s0[t] = (x0 != x0) ? x0 : fmax(s0[t], x0);
      s1[t] = (x1 != x1) ? x1 : fmax(s1[t], x1);
    }
    __syncthreads();
  }
  if (t == 0) {
    out->tr_re = tr_re;
    out->tr_im = tr_im;
    out->mu_re = mu_re;
    out->mu_im = mu_im;
    out->norm_plain = s0[0];
    out->norm_shift = s1[0];
    out->flags = *flags;
    out->info = 0;
    out->rowsum_equal = (rsum && n > 0 && !s_uneq) ? 1 : 0;
    out->pad_ = 0;
    out->rowsum_c = (rsum && n > 0) ? rsum[0].x + rsum[0].y : 0.0;
  }
}

__global__ __launch_bounds__(1024) void finalize_norms(const double* __restrict__ off,
                                                       const double2* __restrict__ diag, int n,
                                                       const int* flags, Scalars* out,
                                                       const double2* __restrict__ rsum) {
  finalize_norms_block(off, diag, n, flags, out, rsum);
}

// ---------------------------------------------------------------------------
// Input -> planar working matrix: w = (A - mu I) * scale.
// structured: A = i*G with G real symmetric; only
// wr = (G - Im(mu) I + alpha J) * scale, J the all-ones matrix (alpha != 0
// deflates an exact eigenvector 1 of G, see Pipeline::spectral_front).
template <typename Tin, typename W>
__global__ void prepare_planar(const Tin* __restrict__ a, int n, W* __restrict__ wr, W* __restrict__ wi,
                               int64_t ld, double mu_re, double mu_im, double scale, int structured,
                               double alpha = 0.0) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const Tin e = a[(int64_t)i * n + j];
  double re = InElem<Tin>::re(e), im = InElem<Tin>::im(e);
  if (i == j) {
    re -= mu_re;
    im -= mu_im;
  }
  if (structured) {
    wr[(int64_t)i * ld + j] = (W)((im + alpha) * scale);
  } else {
    wr[(int64_t)i * ld + j] = (W)(re * scale);
    wi[(int64_t)i * ld + j] = (W)(im * scale);
  }
}

// Small general systems: Gauss-Jordan elimination with partial pivoting of
// the complex n x 2n system [M | B] (n <= GJ_NMAX) held in LDS by one
// workgroup, replacing the blocked LU (a cooperative launch and about 30
// launches per 64 columns, 1.3 ms at n <= 128) for the orders where that is
// all launch overhead.  Input and output are the planar FP64 system of the
// LU route: M in columns 0..n-1, B in columns inp..inp+n-1 of rows 0..n-1
// (leading dimension ldl); on return the B columns hold X = M^-1 B.  Pivot:
// the largest |re| + |im| in the column (lowest row on ties); an exactly
// zero pivot sets info bit 1 and leaves X undefined, as the LU does.
constexpr int GJ_NMAX = 64;
__global__ __launch_bounds__(1024) void gj_solve_small(double* __restrict__ lr, double* __restrict__ li,
                                                       int64_t ldl, int n, int inp, Scalars* __restrict__ sc) {
  constexpr int RS = 2 * GJ_NMAX + 1;  // LDS row stride (complex elements)
  __shared__ double2 s[GJ_NMAX * RS];
  __shared__ int piv;
  __shared__ int singular;
  const int t = threadIdx.x, w = 2 * n;
  if (t == 0) singular = 0;
  for (int q = t; q < n * w; q += 1024) {
    const int r = q / w, c = q % w;
    const int64_t o = (int64_t)r * ldl + (c < n ? c : inp + (c - n));
    s[r * RS + c] = make_double2(lr[o], li[o]);
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (t < 64) {  // wave 0: pivot row, argmax of |re| + |im| over rows k..n-1 (n <= 64: one row per lane)
      const int r = k + t;
      double v = -1.0;
      if (r < n) {
        const double2 x = s[r * RS + k];
        v = fabs(x.x) + fabs(x.y);
      }
      int best = r;
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int ob = __shfl_xor(best, off);
        if (ov > v || (ov == v && ob < best)) {
          v = ov;
          best = ob;
        }
      }
      if (t == 0) {
        piv = best;
        if (!(v > 0.0)) singular = 1;
      }
    }
    __syncthreads();
    const int p = piv;
    if (p != k) {  // columns < k of rows k and p are zero
      for (int c = k + t; c < w; c += 1024) {
        const double2 a = s[k * RS + c];
        s[k * RS + c] = s[p * RS + c];
        s[p * RS + c] = a;
      }
    }
    __syncthreads();
    const double2 d = s[k * RS + k];
    const double dd = d.x * d.x + d.y * d.y;
    const double2 inv = dd > 0.0 ? make_double2(d.x / dd, -d.y / dd) : make_double2(0.0, 0.0);
    __syncthreads();  // every thread has read the pivot before row k changes
    for (int c = k + 1 + t; c < w; c += 1024) {
      const double2 a = s[k * RS + c];
      s[k * RS + c] = make_double2(a.x * inv.x - a.y * inv.y, a.x * inv.y + a.y * inv.x);
    }
    __syncthreads();
    // eliminate column k from every other row (columns > k; column k itself is
    // never read again)
    const int span = w - k - 1;
    for (int q = t; q < n * span; q += 1024) {
      const int r = q / span, c = k + 1 + q % span;
      if (r == k) continue;
      const double2 f = s[r * RS + k];
      const double2 b = s[k * RS + c];
      double2 x = s[r * RS + c];
      x.x -= f.x * b.x - f.y * b.y;
      x.y -= f.x * b.y + f.y * b.x;
      s[r * RS + c] = x;
    }
    __syncthreads();
  }
  for (int q = t; q < n * n; q += 1024) {
    const int r = q / n, c = q % n;
    const double2 x = s[r * RS + n + c];
    const int64_t o = (int64_t)r * ldl + inp + c;
    lr[o] = x.x;
    li[o] = x.y;
  }
  if (t == 0 && singular) sc->info |= 1;
}

// out = sum_t c_t P_t + diag*I on an n x n planar region (im plane optional).
template <typename W> struct Lincomb {
  W* out_r;
  W* out_i;
  const W* in_r[4];
  const W* in_i[4];
  double c[4];
  int nin;
  double diag;
};

template <typename W>
__global__ void lincomb_kernel(Lincomb<W> L, int n, int64_t ld) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const int64_t o = (int64_t)i * ld + j;
  W sr = (i == j) ? (W)L.diag : (W)0;
  W si = (W)0;
  for (int t = 0; t < L.nin; ++t) {
    sr += (W)L.c[t] * L.in_r[t][o];
    if (L.out_i) si += (W)L.c[t] * L.in_i[t][o];
  }
  L.out_r[o] = sr;
  if (L.out_i) L.out_i[o] = si;
}

// out = in^T on an n x n planar region, 32x32 LDS tiles (im plane optional).
template <typename W>
__global__ __launch_bounds__(256) void transpose_kernel(const W* __restrict__ ir, const W* __restrict__ ii,
                                                        W* __restrict__ orr, W* __restrict__ oi, int n,
                                                        int64_t ld) {
  __shared__ W tr[32][33];
  __shared__ W ti[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, j = bx + tx;
    if (i < n && j < n) {
      tr[r][tx] = ir[(int64_t)i * ld + j];
      if (ii) ti[r][tx] = ii[(int64_t)i * ld + j];
    }
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = bx + r, j = by + tx;  // output (i, j) = input (j, i)
    if (i < n && j < n) {
      orr[(int64_t)i * ld + j] = tr[tx][r];
      if (ii) oi[(int64_t)i * ld + j] = ti[tx][r];
    }
  }
}

// Augmented FP64 system [V - U | V + U] over the padded np x 2np extent.
// Padding rows/columns of the left block are the identity, of the right
// block zero, so the padded LU reduces to the unpadded one.
// structured: V real, U = i*Uq (Uq real) -> V - U = V - i Uq, V + U = V + i Uq.
template <typename W>
__global__ void form_system(const W* __restrict__ ur, const W* __restrict__ ui, const W* __restrict__ vr,
                            const W* __restrict__ vi, int64_t ldw, int n, int np, double* __restrict__ lr,
                            double* __restrict__ li, int64_t ldl, int structured) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;  // 0 .. 2np
  const int i = blockIdx.y;
  if (j >= 2 * np) return;
  const int jj = j < np ? j : j - np;
  const bool left = j < np;
  double re = 0.0, im = 0.0;
  if (i < n && jj < n) {
    const int64_t o = (int64_t)i * ldw + jj;
    if (structured) {
      const double v = (double)vr[o], u = (double)ur[o];
      re = v;
      im = left ? -u : u;
    } else {
      const double v0 = (double)vr[o], v1 = (double)vi[o];
      const double u0 = (double)ur[o], u1 = (double)ui[o];
      re = left ? v0 - u0 : v0 + u0;
      im = left ? v1 - u1 : v1 + u1;
    }
  } else if (left && i == jj) {
    re = 1.0;
  }
  lr[(int64_t)i * ldl + j] = re;
  li[(int64_t)i * ldl + j] = im;
}

// X (FP64, right block of the system) -> working precision, optionally
// symmetrised: x = (X + X^T)/2.  The diagonal goes to d (FP64) and the
// planes keep a zero diagonal (split form X = diag(d) + Y for squaring).
template <typename W>
__global__ __launch_bounds__(256) void solution_to_work(const double* __restrict__ xr, const double* __restrict__ xi,
                                                        int64_t ldx, W* __restrict__ wr, W* __restrict__ wi,
                                                        int64_t ldw, int n, int symmetrize, double2* __restrict__ d) {
  __shared__ double tr[32][33];
  __shared__ double ti[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (symmetrize) {
    for (int r = ty; r < 32; r += 8) {  // load transposed block (bx.., by..)
      const int i = bx + r, j = by + tx;
      if (i < n && j < n) {
        tr[r][tx] = xr[(int64_t)i * ldx + j];
        ti[r][tx] = xi[(int64_t)i * ldx + j];
      }
    }
    __syncthreads();
  }
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, j = bx + tx;
    if (i < n && j < n) {
      double vr = xr[(int64_t)i * ldx + j], vi = xi[(int64_t)i * ldx + j];
      if (symmetrize) {
        vr = 0.5 * (vr + tr[tx][r]);
        vi = 0.5 * (vi + ti[tx][r]);
      }
      if (i == j) {
        d[i] = make_double2(vr, vi);
        vr = 0.0;
        vi = 0.0;
      }
      wr[(int64_t)i * ldw + j] = (W)vr;
      wi[(int64_t)i * ldw + j] = (W)vi;
    }
  }
}

// FP64 solution (right block) -> split form in FP64 for FP64 squarings:
// dst = sym(X) with zero diagonal (columns up to kpad, zeros beyond n),
// d = diag(X).
__global__ __launch_bounds__(256) void split_fp64(const double* __restrict__ xr, const double* __restrict__ xi,
                                                  int64_t ld, double* __restrict__ yr, double* __restrict__ yi, int n,
                                                  int kpad, int symmetrize, double2* __restrict__ d) {
  __shared__ double tr[32][33];
  __shared__ double ti[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (symmetrize) {
    for (int r = ty; r < 32; r += 8) {
      const int i = bx + r, j = by + tx;
      if (i < n && j < n) {
        tr[r][tx] = xr[(int64_t)i * ld + j];
        ti[r][tx] = xi[(int64_t)i * ld + j];
      }
    }
    __syncthreads();
  }
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, j = bx + tx;
    if (i >= n || j >= kpad) continue;
    double vr = 0.0, vi = 0.0;
    if (j < n) {
      vr = xr[(int64_t)i * ld + j];
      vi = xi[(int64_t)i * ld + j];
      if (symmetrize) {
        vr = 0.5 * (vr + tr[tx][r]);
        vi = 0.5 * (vi + ti[tx][r]);
      }
      if (i == j) {
        d[i] = make_double2(vr, vi);
        vr = 0.0;
        vi = 0.0;
      }
    }
    yr[(int64_t)i * ld + j] = vr;
    yi[(int64_t)i * ld + j] = vi;
  }
}

// FP64 planes (split form, zero diagonal) -> working precision planes.
template <typename W>
__global__ void fp64_to_work(const double* __restrict__ xr, const double* __restrict__ xi, int64_t ldx,
                             W* __restrict__ wr, W* __restrict__ wi, int64_t ldw, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  wr[(int64_t)i * ldw + j] = (W)xr[(int64_t)i * ldx + j];
  wi[(int64_t)i * ldw + j] = (W)xi[(int64_t)i * ldx + j];
}

// Split form back to a plain matrix: X_ii = d_i (rounded to W).
template <typename W>
__global__ void merge_diagonal(W* __restrict__ wr, W* __restrict__ wi, int64_t ldw, int n,
                               const double2* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  wr[(int64_t)i * ldw + i] = (W)d[i].x;
  wi[(int64_t)i * ldw + i] = (W)d[i].y;
}

// x_ij += c for every i, j < n (both planes): the rank-one term c * J of a
// deflated exponential.
template <typename W>
__global__ void add_constant(W* __restrict__ wr, W* __restrict__ wi, int64_t ldw, int n, double cr, double ci) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const int64_t o = (int64_t)i * ldw + j;
  wr[o] = (W)((double)wr[o] + cr);
  wi[o] = (W)((double)wi[o] + ci);
}

// 1-norm of a real n x n FP64 plane in three deterministic steps: partial
// column sums of |x_ij| over row chunks of `rows` rows (part[chunk * n + j]),
// the chunks summed per column with the maximum of each 256 columns, and one
// workgroup's maximum of those.
__global__ __launch_bounds__(256) void norm1_part(const double* __restrict__ x, int64_t ld, int n, int rows,
                                                  double* __restrict__ part) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (j >= n) return;
  const int r1 = min(n, (c + 1) * rows);
  double s = 0.0;
  for (int i = c * rows; i < r1; ++i) s += fabs(x[(int64_t)i * ld + j]);
  part[(int64_t)c * n + j] = s;
}

// Column sums over the chunks (one thread per column), the maximum of each
// workgroup's 256 columns into cmax[blockIdx.x] (NaN propagates).
__global__ __launch_bounds__(256) void norm1_cols(const double* __restrict__ part, int nchunks, int n,
                                                  double* __restrict__ cmax) {
  __shared__ double sm[256];
  const int j = blockIdx.x * 256 + threadIdx.x;
  double m = 0.0;
  if (j < n) {
    for (int c = 0; c < nchunks; ++c) m += part[(int64_t)c * n + j];
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double y = sm[threadIdx.x + w];
      sm[threadIdx.x] = (y != y) ? y : fmax(sm[threadIdx.x], y);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) cmax[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void norm1_final(const double* __restrict__ cmax, int nblocks,
                                                   double* __restrict__ out) {
  __shared__ double sm[256];
  double m = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) {
    const double y = cmax[b];
    m = (y != y) ? y : fmax(m, y);
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double y = sm[threadIdx.x + w];
      sm[threadIdx.x] = (y != y) ? y : fmax(sm[threadIdx.x], y);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// Traces of even powers of a real symmetric B from its powers B^2, B^4,
// B^6 and, when b8 is given, B^8 (FP64 planes, n x n region):
// trace(B^(2j+2k)) = sum_il (B^2j)_il (B^2k)_il.  One workgroup per row
// writes part[row * kTraceDots + q] for q = 0..6: trace B^4, B^6, B^8, B^10,
// B^12, B^14, B^16 (the last two 0 without b8); trace_dots_final sums the
// rows in order (deterministic) and reports -1 for the last two without b8.
constexpr int kTraceDots = 7;
__global__ __launch_bounds__(256) void trace_dots_part(const double* __restrict__ b2, const double* __restrict__ b4,
                                                       const double* __restrict__ b6, const double* __restrict__ b8,
                                                       int64_t ld, int n, double* __restrict__ part) {
  __shared__ double sm[kTraceDots][256];
  const int i = blockIdx.x;
  double acc[kTraceDots] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < n; j += 256) {
    const int64_t o = (int64_t)i * ld + j;
    const double x2 = b2[o], x4 = b4[o], x6 = b6[o];
    acc[0] += x2 * x2;
    acc[1] += x2 * x4;
    acc[2] += x4 * x4;
    acc[3] += x4 * x6;
    acc[4] += x6 * x6;
    if (b8) {
      const double x8 = b8[o];
      acc[5] += x6 * x8;
      acc[6] += x8 * x8;
    }
  }
#pragma unroll
  for (int q = 0; q < kTraceDots; ++q) sm[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < kTraceDots; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kTraceDots) part[(int64_t)i * kTraceDots + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void trace_dots_final(const double* __restrict__ part, int n, int with8,
                                                        double* __restrict__ out) {
  __shared__ double sm[kTraceDots][256];
  double acc[kTraceDots] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += 256)
#pragma unroll
    for (int q = 0; q < kTraceDots; ++q) acc[q] += part[(int64_t)i * kTraceDots + q];
#pragma unroll
  for (int q = 0; q < kTraceDots; ++q) sm[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < kTraceDots; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kTraceDots) out[threadIdx.x] = (threadIdx.x >= 5 && !with8) ? -1.0 : sm[threadIdx.x][0];
}

// Output: out[i][j] = e^mu * (xr + i xi), interleaved complex or real.
template <typename Tout> struct OutElem;
template <> struct OutElem<float2> {
  static __device__ __forceinline__ float2 make(double re, double im) { return make_float2((float)re, (float)im); }
};
template <> struct OutElem<double2> {
  static __device__ __forceinline__ double2 make(double re, double im) { return make_double2(re, im); }
};
template <> struct OutElem<float> {
  static __device__ __forceinline__ float make(double re, double) { return (float)re; }
};
template <> struct OutElem<double> {
  static __device__ __forceinline__ double make(double re, double) { return re; }
};

template <typename W, typename Tout>
__global__ void write_output(const W* __restrict__ xr, const W* __restrict__ xi, int64_t ld, int n,
                             double ph_re, double ph_im, Tout* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const int64_t o = (int64_t)i * ld + j;
  const double a = (double)xr[o], b = (double)xi[o];
  out[(int64_t)i * n + j] = OutElem<Tout>::make(ph_re * a - ph_im * b, ph_re * b + ph_im * a);
}

// Fill an n x n output with NaN (non-finite input).
template <typename Tout>
__global__ void fill_nan(Tout* out, int64_t count) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) out[k] = OutElem<Tout>::make(__builtin_nan(""), __builtin_nan(""));
}

// ---------------------------------------------------------------------------
// FP64 blocked LU with partial pivoting on the augmented planar system
// S = [M | B] (np rows, 2np columns, ld), panel width LU_NB.
//
// Panel factorisation runs one launch per column.  Workgroup w owns rows
// [k0 + w*PANEL_R, k0 + (w+1)*PANEL_R) of the panel and writes only those
// rows, so no two workgroups write the same address.  Launch j:
//   1. every workgroup reduces the candidates published by launch j-1
//      (its local argmax of |re|+|im| in column k0+j and that row's panel
//      values) to the global pivot p (ties -> smaller row index);
//   2. the owner of row c = k0+j stores the pivot row there, the owner of
//      row p stores the old row c (published as "top") there;
//   3. every row below c is scaled and rank-1 updated within the panel;
//   4. the local argmax of column c+1 and the new row c+1 are published for
//      launch j+1 (buffers alternate by parity of j).
constexpr int LU_NB = 64;
constexpr int PANEL_R = 64;

struct PanelCand {
  double val;
  int idx;
  int pad;
  double re[LU_NB];
  double im[LU_NB];
};

__device__ __forceinline__ double cabs1(double r, double i) { return fabs(r) + fabs(i); }

// Workgroup argmax over (val, idx): larger val wins, ties -> smaller idx.
__device__ __forceinline__ void argmax_merge(double& v, int& k, double v2, int k2) {
  if (v2 > v || (v2 == v && k2 >= 0 && (k < 0 || k2 < k))) {
    v = v2;
    k = k2;
  }
}

__global__ __launch_bounds__(256) void lu_panel_init(const double* __restrict__ sr, const double* __restrict__ si,
                                                     int64_t ld, int np, int k0, PanelCand* cand,
                                                     double* top) {
  __shared__ double sv[256];
  __shared__ int sk[256];
  const int w = blockIdx.x;
  const int r0 = k0 + w * PANEL_R;
  const int t = threadIdx.x;
  double v = -1.0;
  int k = -1;
  if (t < PANEL_R) {
    const int i = r0 + t;
    if (i < np) {
      v = cabs1(sr[(int64_t)i * ld + k0], si[(int64_t)i * ld + k0]);
      k = i;
      if (!(v >= 0.0)) v = -1.0, k = -1;  // NaN never wins
    }
  }
  sv[t] = v;
  sk[t] = k;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      double a = sv[t];
      int ka = sk[t];
      argmax_merge(a, ka, sv[t + s], sk[t + s]);
      sv[t] = a;
      sk[t] = ka;
    }
    __syncthreads();
  }
  const int best = sk[0];
  if (t == 0) {
    cand[w].val = sv[0];
    cand[w].idx = best;
  }
  if (best >= 0 && t < LU_NB) {
    cand[w].re[t] = sr[(int64_t)best * ld + k0 + t];
    cand[w].im[t] = si[(int64_t)best * ld + k0 + t];
  }
  if (r0 <= k0 && k0 < r0 + PANEL_R && t < LU_NB) {
    top[t] = sr[(int64_t)k0 * ld + k0 + t];
    top[LU_NB + t] = si[(int64_t)k0 * ld + k0 + t];
  }
}

__global__ __launch_bounds__(256) void lu_panel_step(double* __restrict__ sr, double* __restrict__ si, int64_t ld,
                                                     int np, int k0, int j, int G, const PanelCand* __restrict__ cin,
                                                     PanelCand* __restrict__ cout, const double* __restrict__ tin,
                                                     double* __restrict__ tout, int* __restrict__ ipiv,
                                                     Scalars* sc) {
  __shared__ double rr[PANEL_R][LU_NB + 1];
  __shared__ double ri[PANEL_R][LU_NB + 1];
  __shared__ double pr[LU_NB], pi[LU_NB];
  __shared__ double sv[256];
  __shared__ int sk[256];
  __shared__ int swin;
  const int t = threadIdx.x;
  const int w = blockIdx.x;
  const int c = k0 + j;
  const int r0 = k0 + w * PANEL_R;

  // 1. global pivot: reduce (val, row, publishing workgroup) triplets
  __shared__ int sw[256];
  {
    double v = -1.0;
    int k = -1, q0 = -1;
    for (int q = t; q < G; q += 256) {
      const double v2 = cin[q].val;
      const int k2 = cin[q].idx;
      if (k2 >= 0 && (v2 > v || (v2 == v && (k < 0 || k2 < k)))) {
        v = v2;
        k = k2;
        q0 = q;
      }
    }
    sv[t] = v;
    sk[t] = k;
    sw[t] = q0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
        const double v2 = sv[t + s];
        const int k2 = sk[t + s];
        if (k2 >= 0 && (v2 > sv[t] || (v2 == sv[t] && (sk[t] < 0 || k2 < sk[t])))) {
          sv[t] = v2;
          sk[t] = k2;
          sw[t] = sw[t + s];
        }
      }
      __syncthreads();
    }
    if (t == 0) swin = sw[0];
    __syncthreads();
  }
  const int p = sk[0] >= 0 ? sk[0] : c;
  const double pval = sv[0];
  const int win = swin;
  if (t < LU_NB) {
    if (win >= 0) {
      pr[t] = cin[win].re[t];
      pi[t] = cin[win].im[t];
    } else {
      pr[t] = sr[(int64_t)c * ld + k0 + t];
      pi[t] = si[(int64_t)c * ld + k0 + t];
    }
  }
  if (w == 0 && t == 0) {
    ipiv[c] = p;
    if (!(pval > 0.0)) atomicOr(&sc->info, 1);
  }
  __syncthreads();

  // 2./3. my rows
  const int lo = max(r0, c), hi = min(r0 + PANEL_R, np);
  if (lo < hi) {
    // load rows lo..hi-1 (entire panel width)
    for (int e = t; e < (hi - lo) * LU_NB; e += 256) {
      const int lr = e / LU_NB, col = e % LU_NB;
      const int i = lo + lr;
      double vr, vi;
      if (i == c) {
        vr = pr[col];
        vi = pi[col];
      } else if (i == p) {
        vr = tin[col];
        vi = tin[LU_NB + col];
      } else {
        vr = sr[(int64_t)i * ld + k0 + col];
        vi = si[(int64_t)i * ld + k0 + col];
      }
      rr[i - r0][col] = vr;
      ri[i - r0][col] = vi;
    }
    __syncthreads();
    // reciprocal of the pivot
    const double dr = pr[j], di = pi[j];
    const double den = dr * dr + di * di;
    const bool ok = den > 0.0;
    const double qr = ok ? dr / den : 0.0, qi = ok ? -di / den : 0.0;
    // thread -> (row, column group of 16); column j is left untouched
    // until every thread of the row has read it
    const int lr = t >> 2, cg = t & 3;
    const int i = r0 + lr;
    const bool upd = i > c && i < hi;
    if (upd) {
      const double ar = rr[lr][j], ai = ri[lr][j];
      const double lre = ar * qr - ai * qi, lim = ar * qi + ai * qr;
      for (int col = cg * 16; col < cg * 16 + 16; ++col) {
        if (col > j) {
          rr[lr][col] -= lre * pr[col] - lim * pi[col];
          ri[lr][col] -= lre * pi[col] + lim * pr[col];
        }
      }
    }
    __syncthreads();
    if (upd && cg == 0) {
      const double ar = rr[lr][j], ai = ri[lr][j];
      rr[lr][j] = ar * qr - ai * qi;
      ri[lr][j] = ar * qi + ai * qr;
    }
    __syncthreads();
    // write back
    for (int e = t; e < (hi - lo) * LU_NB; e += 256) {
      const int q = e / LU_NB, col = e % LU_NB;
      const int ii = lo + q;
      sr[(int64_t)ii * ld + k0 + col] = rr[ii - r0][col];
      si[(int64_t)ii * ld + k0 + col] = ri[ii - r0][col];
    }
  }
  // 4. candidates for column j+1
  if (j + 1 < LU_NB) {
    const int jn = j + 1, cn = c + 1;
    double v = -1.0;
    int k = -1;
    if (t < PANEL_R) {
      const int ii = r0 + t;
      if (ii >= cn && ii < np && ii >= lo && ii < hi) {
        v = cabs1(rr[t][jn], ri[t][jn]);
        k = ii;
        if (!(v >= 0.0)) v = -1.0, k = -1;
      }
    }
    __syncthreads();
    sv[t] = v;
    sk[t] = k;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
        double a = sv[t];
        int ka = sk[t];
        argmax_merge(a, ka, sv[t + s], sk[t + s]);
        sv[t] = a;
        sk[t] = ka;
      }
      __syncthreads();
    }
    const int best = sk[0];
    if (t == 0) {
      cout[w].val = sv[0];
      cout[w].idx = best;
    }
    if (best >= 0 && t < LU_NB) {
      cout[w].re[t] = rr[best - r0][t];
      cout[w].im[t] = ri[best - r0][t];
    }
    if (r0 <= cn && cn < r0 + PANEL_R && cn < np && t < LU_NB) {
      tout[t] = rr[cn - r0][t];
      tout[LU_NB + t] = ri[cn - r0][t];
    }
  }
}

// Whole-panel factorisation in one cooperative launch (G workgroups, all
// co-resident by the cooperative launch contract; ctr is a zeroed counter).  Workgroup w keeps rows
// [k0 + 64w, k0 + 64w + 64) of the 64-wide panel in LDS for all 64 columns
// and exchanges, per column, its argmax candidate (value, row, row values)
// and -- for the owner of row c -- row c itself through global buffers,
// separated by one grid barrier.  Buffers alternate by column parity, so a
// workgroup running one column ahead never overwrites what another still
// reads (a workgroup can be at most one barrier ahead of any other).  Exchanged values are stored and loaded with agent-scope atomics,
// which keeps them coherent across the XCDs' separate L2 caches.
__device__ __forceinline__ void st_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier over G co-resident workgroups: one arrival counter that is
// never reset within a launch (barrier b waits for G * (b + 1) arrivals).
// The wait is bounded in time: after 2^28 ticks of the constant-rate device
// clock (about 2.7 s at its 100 MHz) it records an error in sc->info (bit 1)
// and returns.  The error is sticky: once bit 1 is set every later barrier
// returns at once and later panel launches exit at their start, so a lost
// workgroup ends the factorisation with an error instead of a long spin.
constexpr uint64_t kBarrierTicks = (uint64_t)1 << 28;
__device__ __forceinline__ void grid_barrier(unsigned* ctr, unsigned target, Scalars* sc) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (ld_agent(&sc->info) & 2) break;  // an earlier barrier timed out
      __builtin_amdgcn_s_sleep(1);
      if (wall_clock64() - t0 > kBarrierTicks) {
        atomicOr(&sc->info, 2);
        break;
      }
    }
    __threadfence();
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void lu_panel_coop(double* __restrict__ sr, double* __restrict__ si, int64_t ld,
                                                     int np, int k0, int G, PanelCand* cand, double* top,
                                                     int* __restrict__ ipiv, Scalars* sc, unsigned* ctr) {
  __shared__ double rr[PANEL_R][LU_NB + 1];
  __shared__ double ri[PANEL_R][LU_NB + 1];
  __shared__ double pr[LU_NB], pi[LU_NB];
  __shared__ double sv[256];
  __shared__ int sk[256];
  __shared__ int sw[256];
  __shared__ int s_abort;
  const int t = threadIdx.x;
  // an earlier panel's barrier timed out: the factorisation has failed
  if (t == 0) s_abort = ld_agent(&sc->info) & 2;
  __syncthreads();
  if (s_abort) return;
  const int w = blockIdx.x;
  const int r0 = k0 + w * PANEL_R;
  const int hi = min(r0 + PANEL_R, np);
  const int nrow = max(hi - r0, 0);
  for (int e = t; e < nrow * LU_NB; e += 256) {
    const int lr = e / LU_NB, col = e % LU_NB;
    rr[lr][col] = sr[(int64_t)(r0 + lr) * ld + k0 + col];
    ri[lr][col] = si[(int64_t)(r0 + lr) * ld + k0 + col];
  }
  __syncthreads();
  for (int j = 0; j < LU_NB; ++j) {
    const int c = k0 + j;
    const int par = j & 1;
    PanelCand* cb = cand + par * G;
    double* tb = top + par * 2 * LU_NB;
    // 1. local candidate of column j among rows >= c (wave 0, shuffles)
    if (t < 64) {
      double v = -1.0;
      int k = -1;
      if (t < nrow) {
        const int i = r0 + t;
        if (i >= c) {
          v = cabs1(rr[t][j], ri[t][j]);
          k = i;
          if (!(v >= 0.0)) v = -1.0, k = -1;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off);
        const int k2 = __shfl_xor(k, off);
        argmax_merge(v, k, v2, k2);
      }
      const int best = k;  // identical in every lane after the butterfly
      if (t == 0) {
        st_agent(&cb[w].val, v);
        st_agent(&cb[w].idx, best);
      }
      if (best >= 0) {
        st_agent(&cb[w].re[t], rr[best - r0][t]);
        st_agent(&cb[w].im[t], ri[best - r0][t]);
      }
      if (c >= r0 && c < hi) {
        st_agent(&tb[t], rr[c - r0][t]);
        st_agent(&tb[LU_NB + t], ri[c - r0][t]);
      }
    }
    grid_barrier(ctr, (unsigned)G * (unsigned)(j + 1), sc);
    // 2. global pivot (wave 0, shuffles; ties -> lower row, then lower workgroup)
    if (t < 64) {
      double v = -1.0;
      int k = -1, q0 = -1;
      for (int q = t; q < G; q += 64) {
        const double v2 = ld_agent(&cb[q].val);
        const int k2 = ld_agent(&cb[q].idx);
        if (k2 >= 0 && (v2 > v || (v2 == v && (k < 0 || k2 < k)))) {
          v = v2;
          k = k2;
          q0 = q;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off);
        const int k2 = __shfl_xor(k, off);
        const int q2 = __shfl_xor(q0, off);
        if (k2 >= 0 && (v2 > v || (v2 == v && (k < 0 || k2 < k)))) {
          v = v2;
          k = k2;
          q0 = q2;
        }
      }
      if (t == 0) {
        sv[0] = v;
        sk[0] = k;
        sw[0] = q0;
      }
    }
    __syncthreads();
    const int p = sk[0] >= 0 ? sk[0] : c;
    const double pval = sv[0];
    const int win = sw[0];
    if (t < LU_NB) {
      if (win >= 0) {
        pr[t] = ld_agent(&cb[win].re[t]);
        pi[t] = ld_agent(&cb[win].im[t]);
      } else {
        pr[t] = ld_agent(&tb[t]);
        pi[t] = ld_agent(&tb[LU_NB + t]);
      }
    }
    if (w == 0 && t == 0) {
      ipiv[c] = p;
      if (!(pval > 0.0)) atomicOr(&sc->info, 1);
    }
    __syncthreads();
    // 3. interchange rows c and p inside LDS (each side by its owner)
    if (p != c && p >= r0 && p < hi && t < LU_NB) {
      rr[p - r0][t] = ld_agent(&tb[t]);
      ri[p - r0][t] = ld_agent(&tb[LU_NB + t]);
    }
    if (c >= r0 && c < hi && t < LU_NB) {
      rr[c - r0][t] = pr[t];
      ri[c - r0][t] = pi[t];
    }
    __syncthreads();
    // 4. scale and rank-1 update of my rows below c
    const double dr = pr[j], di = pi[j];
    const double den = dr * dr + di * di;
    const bool ok = den > 0.0;
    const double qr = ok ? dr / den : 0.0, qi = ok ? -di / den : 0.0;
    const int lr = t >> 2, cg4 = t & 3;
    const int i = r0 + lr;
    const bool upd = lr < nrow && i > c;
    if (upd) {
      const double ar = rr[lr][j], ai = ri[lr][j];
      const double lre = ar * qr - ai * qi, lim = ar * qi + ai * qr;
      for (int col = cg4 * 16; col < cg4 * 16 + 16; ++col) {
        if (col > j) {
          rr[lr][col] -= lre * pr[col] - lim * pi[col];
          ri[lr][col] -= lre * pi[col] + lim * pr[col];
        }
      }
    }
    __syncthreads();
    if (upd && cg4 == 0) {
      const double ar = rr[lr][j], ai = ri[lr][j];
      rr[lr][j] = ar * qr - ai * qi;
      ri[lr][j] = ar * qi + ai * qr;
    }
    __syncthreads();
  }
  for (int e = t; e < nrow * LU_NB; e += 256) {
    const int lr = e / LU_NB, col = e % LU_NB;
    sr[(int64_t)(r0 + lr) * ld + k0 + col] = rr[lr][col];
    si[(int64_t)(r0 + lr) * ld + k0 + col] = ri[lr][col];
  }
}

// The panel of lu_panel_coop when its rows fit one workgroup (np - k0 <=
// PANEL_R1 = 128 rows, 133 KB of LDS): the same pivot rule (largest
// |re| + |im| among rows >= c, lowest row on ties), interchanges, scaling
// and rank-1 updates, element by element with the same operations, but
// with workgroup barriers instead of grid barriers (64 of those cost about
// 0.6 ms per panel at any order).  One ordinary launch, usable by
// concurrent lanes as well.
constexpr int PANEL_R1 = 128;
__global__ __launch_bounds__(256) void lu_panel_single(double* __restrict__ sr, double* __restrict__ si, int64_t ld,
                                                       int np, int k0, int* __restrict__ ipiv, Scalars* sc) {
  __shared__ double rr[PANEL_R1][LU_NB + 1];
  __shared__ double ri[PANEL_R1][LU_NB + 1];
  __shared__ double pr[LU_NB], pi[LU_NB];
  __shared__ int s_piv;
  const int t = threadIdx.x;
  const int nrow = min(np - k0, PANEL_R1);
  for (int e = t; e < nrow * LU_NB; e += 256) {
    const int lr = e / LU_NB, col = e % LU_NB;
    rr[lr][col] = sr[(int64_t)(k0 + lr) * ld + k0 + col];
    ri[lr][col] = si[(int64_t)(k0 + lr) * ld + k0 + col];
  }
  __syncthreads();
  for (int j = 0; j < LU_NB; ++j) {
    const int c = k0 + j;
    // 1. pivot among rows >= c (wave 0: rows t and t + 64, then shuffles)
    if (t < 64) {
      double v = -1.0;
      int k = -1;
      for (int lr = t; lr < nrow; lr += 64) {
        if (k0 + lr >= c) {
          double x = cabs1(rr[lr][j], ri[lr][j]);
          int kx = k0 + lr;
          if (!(x >= 0.0)) x = -1.0, kx = -1;
          argmax_merge(v, k, x, kx);
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off);
        const int k2 = __shfl_xor(k, off);
        argmax_merge(v, k, v2, k2);
      }
      if (t == 0) {
        const int p = k >= 0 ? k : c;
        s_piv = p;
        ipiv[c] = p;
        if (!(v > 0.0)) atomicOr(&sc->info, 1);
      }
    }
    __syncthreads();
    const int p = s_piv;
    // 2. interchange rows c and p of the panel; pr / pi = the pivot row
    if (t < LU_NB) {
      const double ar = rr[p - k0][t], ai = ri[p - k0][t];
      if (p != c) {
        rr[p - k0][t] = rr[c - k0][t];
        ri[p - k0][t] = ri[c - k0][t];
        rr[c - k0][t] = ar;
        ri[c - k0][t] = ai;
      }
      pr[t] = ar;
      pi[t] = ai;
    }
    __syncthreads();
    // 3. scale and rank-1 update of the rows below c
    const double dr = pr[j], di = pi[j];
    const double den = dr * dr + di * di;
    const bool ok = den > 0.0;
    const double qr = ok ? dr / den : 0.0, qi = ok ? -di / den : 0.0;
    const int lr = t >> 1, half = t & 1;
    const bool upd = lr < nrow && k0 + lr > c;
    if (upd) {
      const double ar = rr[lr][j], ai = ri[lr][j];
      const double lre = ar * qr - ai * qi, lim = ar * qi + ai * qr;
      for (int col = half * 32; col < half * 32 + 32; ++col) {
        if (col > j) {
          rr[lr][col] -= lre * pr[col] - lim * pi[col];
          ri[lr][col] -= lre * pi[col] + lim * pr[col];
        }
      }
    }
    __syncthreads();
    if (upd && half == 0) {
      const double ar = rr[lr][j], ai = ri[lr][j];
      rr[lr][j] = ar * qr - ai * qi;
      ri[lr][j] = ar * qi + ai * qr;
    }
    __syncthreads();
  }
  for (int e = t; e < nrow * LU_NB; e += 256) {
    const int lr = e / LU_NB, col = e % LU_NB;
    sr[(int64_t)(k0 + lr) * ld + k0 + col] = rr[lr][col];
    si[(int64_t)(k0 + lr) * ld + k0 + col] = ri[lr][col];
  }
}

// Apply the panel's row interchanges to columns [c0, c1).
__global__ void lu_laswp(double* __restrict__ sr, double* __restrict__ si, int64_t ld, int k0, int nb,
                         const int* __restrict__ ipiv, int c0, int c1) {
  const int col = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= c1) return;
  for (int q = 0; q < nb; ++q) {
    const int r = k0 + q, p = ipiv[r];
    if (p != r) {
      const int64_t a = (int64_t)r * ld + col, b = (int64_t)p * ld + col;
      const double xr = sr[a], xi = si[a];
      sr[a] = sr[b];
      si[a] = si[b];
      sr[b] = xr;
      si[b] = xi;
    }
  }
}

// Block-row triangular solve, 64 columns per workgroup of 64 threads:
//   lower (unit):   S[k0:k0+64, cols] = L11^{-1} S[k0:k0+64, cols]
//   upper (non-unit): S[k0:k0+64, cols] = U11^{-1} S[k0:k0+64, cols]
// The solved block is also written transposed (row = column index - c0,
// 64 k-values) to tr/ti as the right operand of the following GEMM.
// The transposed copy goes to tr[(col - c0) * ldt + toff + r].
template <bool UPPER>
__global__ __launch_bounds__(64) void lu_trsm_block_row(double* __restrict__ sr, double* __restrict__ si, int64_t ld,
                                                        int k0, int c0, int c1, double* __restrict__ tr,
                                                        double* __restrict__ ti, int ldt, int toff) {
  __shared__ double lr[LU_NB][LU_NB + 1];
  __shared__ double li[LU_NB][LU_NB + 1];
  __shared__ double xr[LU_NB][64 + 1];
  __shared__ double xi[LU_NB][64 + 1];
  __shared__ double dr[LU_NB], di[LU_NB];
  const int t = threadIdx.x;
  const int cb = c0 + blockIdx.x * 64;
  for (int e = t; e < LU_NB * LU_NB; e += 64) {
    const int r = e / LU_NB, q = e % LU_NB;
    lr[r][q] = sr[(int64_t)(k0 + r) * ld + k0 + q];
    li[r][q] = si[(int64_t)(k0 + r) * ld + k0 + q];
  }
  const int col = cb + t;
  const bool active = col < c1;
  for (int r = 0; r < LU_NB; ++r) {
    xr[r][t] = active ? sr[(int64_t)(k0 + r) * ld + col] : 0.0;
    xi[r][t] = active ? si[(int64_t)(k0 + r) * ld + col] : 0.0;
  }
  __syncthreads();
  if (UPPER) {
    if (t < LU_NB) {
      const double a = lr[t][t], b = li[t][t];
      const double den = a * a + b * b;
      dr[t] = a / den;
      di[t] = -b / den;
    }
    __syncthreads();
    for (int r = LU_NB - 1; r >= 0; --r) {
      double ar = xr[r][t], ai = xi[r][t];
      for (int q = r + 1; q < LU_NB; ++q) {
        const double ur = lr[r][q], ui = li[r][q];
        const double br = xr[q][t], bi = xi[q][t];
        ar -= ur * br - ui * bi;
        ai -= ur * bi + ui * br;
      }
      const double zr = ar * dr[r] - ai * di[r], zi = ar * di[r] + ai * dr[r];
      xr[r][t] = zr;
      xi[r][t] = zi;
    }
  } else {
    for (int r = 1; r < LU_NB; ++r) {
      double ar = xr[r][t], ai = xi[r][t];
      for (int q = 0; q < r; ++q) {
        const double ur = lr[r][q], ui = li[r][q];
        const double br = xr[q][t], bi = xi[q][t];
        ar -= ur * br - ui * bi;
        ai -= ur * bi + ui * br;
      }
      xr[r][t] = ar;
      xi[r][t] = ai;
    }
  }
  if (active) {
    for (int r = 0; r < LU_NB; ++r) {
      sr[(int64_t)(k0 + r) * ld + col] = xr[r][t];
      si[(int64_t)(k0 + r) * ld + col] = xi[r][t];
      tr[(int64_t)(col - c0) * ldt + toff + r] = xr[r][t];
      ti[(int64_t)(col - c0) * ldt + toff + r] = xi[r][t];
    }
  }
}

// Inverse of the 64x64 diagonal block at (k0, k0): unit lower (L) or
// non-unit upper (U) triangle.  Right-looking elimination on X = I with 256
// threads: step q finalises row q of X and updates the remaining rows in
// parallel (64 synchronised steps).  Output: inv (64 x 64 complex, planar,
// row-major, ld 64).
template <bool UPPER>
__global__ __launch_bounds__(256) void lu_invert_diag_block(const double* __restrict__ sr,
                                                            const double* __restrict__ si, int64_t ld, int k0,
                                                            double* __restrict__ invr, double* __restrict__ invi) {
  __shared__ double lr[LU_NB][LU_NB + 1];
  __shared__ double li[LU_NB][LU_NB + 1];
  __shared__ double xr[LU_NB][LU_NB + 1];
  __shared__ double xi[LU_NB][LU_NB + 1];
  const int t = threadIdx.x;
  for (int e = t; e < LU_NB * LU_NB; e += 256) {
    const int r = e / LU_NB, q = e % LU_NB;
    lr[r][q] = sr[(int64_t)(k0 + r) * ld + k0 + q];
    li[r][q] = si[(int64_t)(k0 + r) * ld + k0 + q];
    xr[r][q] = (r == q) ? 1.0 : 0.0;
    xi[r][q] = 0.0;
  }
  __syncthreads();
  if (UPPER) {
    for (int q = LU_NB - 1; q >= 0; --q) {
      // row q of X: divide by U[q][q] (its columns c >= q are the nonzero ones)
      if (t < LU_NB && t >= q) {
        const double dr = lr[q][q], di = li[q][q];
        const double den = dr * dr + di * di;
        const double qr = dr / den, qi = -di / den;
        const double ar = xr[q][t], ai = xi[q][t];
        xr[q][t] = ar * qr - ai * qi;
        xi[q][t] = ar * qi + ai * qr;
      }
      __syncthreads();
      // rows r < q: X[r][c] -= U[r][q] X[q][c] for c >= q
      for (int e = t; e < q * LU_NB; e += 256) {
        const int r = e / LU_NB, c = e % LU_NB;
        if (c >= q) {
          const double ur = lr[r][q], ui = li[r][q], br = xr[q][c], bi = xi[q][c];
          xr[r][c] -= ur * br - ui * bi;
          xi[r][c] -= ur * bi + ui * br;
        }
      }
      __syncthreads();
    }
  } else {
    for (int q = 0; q < LU_NB - 1; ++q) {
      // rows r > q: X[r][c] -= L[r][q] X[q][c] for c <= q
      for (int e = t; e < (LU_NB - 1 - q) * LU_NB; e += 256) {
        const int r = q + 1 + e / LU_NB, c = e % LU_NB;
        if (c <= q) {
          const double ur = lr[r][q], ui = li[r][q], br = xr[q][c], bi = xi[q][c];
          xr[r][c] -= ur * br - ui * bi;
          xi[r][c] -= ur * bi + ui * br;
        }
      }
      __syncthreads();
    }
  }
  for (int e = t; e < LU_NB * LU_NB; e += 256) {
    invr[e] = xr[e / LU_NB][e % LU_NB];
    invi[e] = xi[e / LU_NB][e % LU_NB];
  }
}

// Block-row solve by an explicit triangular inverse: for the 64 rows at k0
// and columns [c0, c1), S[k0:k0+64, col] <- inv * S[k0:k0+64, col], and the
// result transposed to tr[(col - c0) * ldt + toff + r].  256 threads per
// workgroup handle 64 columns: thread (c = t % 64, g = t / 64) forms rows
// 16g..16g+15 of column c; the transposed copy is written from LDS with
// consecutive lanes on consecutive r.
__global__ __launch_bounds__(256) void lu_apply_inverse(double* __restrict__ sr, double* __restrict__ si, int64_t ld,
                                                        int k0, int c0, int c1, const double* __restrict__ invr,
                                                        const double* __restrict__ invi, double* __restrict__ tr,
                                                        double* __restrict__ ti, int ldt, int toff) {
  __shared__ double ir[LU_NB][LU_NB + 1];
  __shared__ double ii[LU_NB][LU_NB + 1];
  __shared__ double br[LU_NB][64 + 1];
  __shared__ double bi[LU_NB][64 + 1];
  const int t = threadIdx.x;
  const int cb = c0 + blockIdx.x * 64;
  for (int e = t; e < LU_NB * LU_NB; e += 256) {
    const int r = e / LU_NB, q = e % LU_NB;
    ir[r][q] = invr[e];
    ii[r][q] = invi[e];
  }
  for (int e = t; e < LU_NB * 64; e += 256) {
    const int r = e / 64, c = e % 64;
    const int col = cb + c;
    const bool ok = col < c1;
    br[r][c] = ok ? sr[(int64_t)(k0 + r) * ld + col] : 0.0;
    bi[r][c] = ok ? si[(int64_t)(k0 + r) * ld + col] : 0.0;
  }
  __syncthreads();
  const int c = t & 63, g = t >> 6;
  double ar[16], ai[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) ar[k] = ai[k] = 0.0;
  for (int q = 0; q < LU_NB; ++q) {
    const double xr = br[q][c], xi = bi[q][c];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const double vr = ir[16 * g + k][q], vi = ii[16 * g + k][q];
      ar[k] += vr * xr - vi * xi;
      ai[k] += vr * xi + vi * xr;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    br[16 * g + k][c] = ar[k];
    bi[16 * g + k][c] = ai[k];
  }
  __syncthreads();
  for (int e = t; e < LU_NB * 64; e += 256) {
    const int r = e / 64, cc = e % 64;
    const int col = cb + cc;
    if (col < c1) {
      sr[(int64_t)(k0 + r) * ld + col] = br[r][cc];
      si[(int64_t)(k0 + r) * ld + col] = bi[r][cc];
    }
  }
  for (int e = t; e < LU_NB * 64; e += 256) {
    const int cc = e / 64, r = e % 64;
    const int col = cb + cc;
    if (col < c1) {
      tr[(int64_t)(col - c0) * ldt + toff + r] = br[r][cc];
      ti[(int64_t)(col - c0) * ldt + toff + r] = bi[r][cc];
    }
  }
}

// ---------------------------------------------------------------------------
// Propagation of many states: Psi (states as rows, planar working
// precision) <-> complex128 host layout, with an FP64 refinement per step.

// Psi planes <- psi0 (nstates x n complex128); norm0[s] = ||psi0_s||_2.
template <typename W>
__global__ __launch_bounds__(256) void states_to_planes(const double2* __restrict__ psi0, int n, W* __restrict__ pr,
                                                        W* __restrict__ pi, int64_t ld, double* __restrict__ norm0) {
  __shared__ double red[256];
  const int s = blockIdx.x, t = threadIdx.x;
  double acc = 0.0;
  for (int j = t; j < n; j += 256) {
    const double2 v = psi0[(int64_t)s * n + j];
    pr[(int64_t)s * ld + j] = (W)v.x;
    pi[(int64_t)s * ld + j] = (W)v.y;
    acc += v.x * v.x + v.y * v.y;
  }
  red[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) norm0[s] = sqrt(red[0]);
}

// One refinement per step: FP64 norm of each propagated state; for unitary
// propagators (renorm != 0) the state is rescaled to its initial norm in
// place; out[s][j] = e^{k mu} * state (complex128).
template <typename W>
__global__ __launch_bounds__(256) void refine_states(W* __restrict__ pr, W* __restrict__ pi, int64_t ld, int n,
                                                     const double* __restrict__ norm0, int renorm, double ph_re,
                                                     double ph_im, double2* __restrict__ out) {
  __shared__ double red[256];
  const int s = blockIdx.x, t = threadIdx.x;
  double acc = 0.0;
  for (int j = t; j < n; j += 256) {
    const double a = (double)pr[(int64_t)s * ld + j], b = (double)pi[(int64_t)s * ld + j];
    acc += a * a + b * b;
  }
  red[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double nrm = sqrt(red[0]);
  const double f = (renorm && nrm > 0.0) ? norm0[s] / nrm : 1.0;
  for (int j = t; j < n; j += 256) {
    const double a = f * (double)pr[(int64_t)s * ld + j], b = f * (double)pi[(int64_t)s * ld + j];
    if (renorm) {
      pr[(int64_t)s * ld + j] = (W)a;
      pi[(int64_t)s * ld + j] = (W)b;
    }
    out[(int64_t)s * n + j] = make_double2(ph_re * a - ph_im * b, ph_re * b + ph_im * a);
  }
}

// ---------------------------------------------------------------------------
// psi_out[s][i] = e^mu sum_j X[i][j] psi0[s][j], FP64 accumulation; one
// wavefront per row, states looped (nstates small).
template <typename W>
__global__ __launch_bounds__(256) void apply_states(const W* __restrict__ xr, const W* __restrict__ xi, int64_t ld,
                                                    int n, const double2* __restrict__ psi0, int nstates,
                                                    double ph_re, double ph_im, double2* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  for (int s = 0; s < nstates; ++s) {
    const double2* p = psi0 + (int64_t)s * n;
    double ar = 0.0, ai = 0.0;
    for (int j = lane; j < n; j += 64) {
      const double a = (double)xr[(int64_t)row * ld + j], b = (double)xi[(int64_t)row * ld + j];
      const double2 v = p[j];
      ar += a * v.x - b * v.y;
      ai += a * v.y + b * v.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      ar += __shfl_xor(ar, off);
      ai += __shfl_xor(ai, off);
    }
    if (lane == 0) out[(int64_t)s * n + row] = make_double2(ph_re * ar - ph_im * ai, ph_re * ai + ph_im * ar);
  }
}

// One-step FP64 refinement of psi = U psi0 for a unitary U (structured
// input, |e^mu| = 1): each state is rescaled to the norm of its psi0, which
// removes the radial part of the error of the FP32 U.  One workgroup per
// state; zero states are left as they are.
__device__ __forceinline__ void renorm_state_block(const double2* __restrict__ p, double2* __restrict__ q, int n) {
  __shared__ double r0[256], r1[256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int j = t; j < n; j += 256) {
    a += p[j].x * p[j].x + p[j].y * p[j].y;
    b += q[j].x * q[j].x + q[j].y * q[j].y;
  }
  r0[t] = a;
  r1[t] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      r0[t] += r0[t + w];
      r1[t] += r1[t + w];
    }
    __syncthreads();
  }
  if (!(r0[0] > 0.0) || !(r1[0] > 0.0)) return;
  const double f = sqrt(r0[0] / r1[0]);
  for (int j = t; j < n; j += 256) q[j] = make_double2(f * q[j].x, f * q[j].y);
}
__global__ __launch_bounds__(256) void renorm_states(const double2* __restrict__ psi0, double2* __restrict__ out,
                                                     int n) {
  renorm_state_block(psi0 + (int64_t)blockIdx.x * n, out + (int64_t)blockIdx.x * n, n);
}

// ---------------------------------------------------------------------------
// Structured solve (A = iC, C real symmetric).  With V and U = i Uq from the
// real Pade products, M = V - i Uq is normal (V, Uq commute) and
// S = M^H M = V^2 + Uq^2 is real symmetric positive definite.  For
// S = L L^T and Z = (V + i Uq) L^{-T}:  X = M^{-1} M^H = Z Z^T.
// The stacked real matrix [Re Z; Im Z] (2np x np) is solved in place.

// big[r][j] = V[r][j] for r < np, Uq[r - np][j] for r >= np (zero padding).
template <typename W>
__global__ __launch_bounds__(256) void pack_stacked(const W* __restrict__ vr, const W* __restrict__ ur, int64_t ldw,
                                                    int n, int np, double* __restrict__ big, int64_t ld) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (j >= np) return;
  const int i = r < np ? r : r - np;
  const W* src = r < np ? vr : ur;
  big[(int64_t)r * ld + j] = (i < n && j < n) ? (double)src[(int64_t)i * ldw + j] : 0.0;
}

// Cholesky factor of the 64 x 64 diagonal block at (k0, k0) of s (FP64,
// lower triangle, in place) and its inverse into inv (64 x 64 row-major,
// zero above the diagonal).  Rows and columns at or beyond n are padding
// and are factored as the identity.  A pivot that is not positive sets bit
// 0 of sc->info and is replaced by 1.
__device__ __forceinline__ void chol_diag_block(double* __restrict__ s, int64_t ld, int k0, int n,
                                                double* __restrict__ inv, Scalars* sc) {
  __shared__ double a[64][65];   // scaled factor (input of the inverse)
  __shared__ double col[2][64];  // column j of step j (unscaled), double-buffered by parity
  __shared__ double ldg[64], rd[64];
  const int t = threadIdx.x;
  // thread (ti, tk) keeps the 16 elements (ti + 16p, tk + 16q) in registers
  const int ti = t >> 4, tk = t & 15;
  double e[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = ti + 16 * p, k = tk + 16 * q;
      e[p][q] = (k0 + i >= n || k0 + k >= n) ? (i == k ? 1.0 : 0.0) : s[(int64_t)(k0 + i) * ld + k0 + k];
    }
  // Right-looking on unscaled columns: step j publishes column j (its
  // owners hold tk == j % 16), then every element (i, k) with j < k <= i
  // loses a_ij a_kj / a_jj; the scaling by 1/sqrt(a_jj) is applied at the
  // end.  One barrier per step.
#pragma unroll
  for (int jq = 0; jq < 4; ++jq)
#pragma unroll 1
  for (int jb = 0; jb < 16; ++jb) {
    const int j = 16 * jq + jb;
    double* cb = col[j & 1];
    if (tk == jb) {
#pragma unroll
      for (int p = 0; p < 4; ++p) cb[ti + 16 * p] = e[p][jq];
    }
    __syncthreads();
    double d = cb[j];
    if (!(d > 0.0)) {
      if (t == 0) atomicOr(&sc->info, 1);
      d = 1.0;
    }
    if (t == 0) {
      ldg[j] = sqrt(d);
      rd[j] = 1.0 / ldg[j];
    }
    const double r2 = 1.0 / d;
    double ci[4], ck[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      ci[p] = cb[ti + 16 * p] * r2;
      ck[p] = cb[tk + 16 * p];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = ti + 16 * p, k = tk + 16 * q;
        if (k > j && k <= i) e[p][q] -= ci[p] * ck[q];
      }
  }
  __syncthreads();
  // L[i][k] = a_ik / sqrt(d_k) below the diagonal, sqrt(d_i) on it
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = ti + 16 * p, k = tk + 16 * q;
      double v = 0.0;
      if (k < i) v = e[p][q] * rd[k];
      else if (k == i) v = ldg[i];
      a[i][k] = v;
      if (k <= i && k0 + i < n && k0 + k < n) s[(int64_t)(k0 + i) * ld + k0 + k] = v;
    }
  __syncthreads();
  // inverse by forward substitution, column c = t / 4; the four lanes of a
  // column split every inner product by k mod 4 and combine with shuffles.
  const int c = t >> 2, q = t & 3;
  double x[16];
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) x[kk] = 0.0;
#pragma unroll
  for (int i = 0; i < 64; ++i) {
    double p = 0.0;
#pragma unroll
    for (int kk = 0; kk < (i + 3) / 4; ++kk) {
      const int k = 4 * kk + q;
      if (k < i) p += a[i][k] * x[kk];
    }
    p += __shfl_xor(p, 1);
    p += __shfl_xor(p, 2);
    const double xi = ((i == c ? 1.0 : 0.0) - p) * rd[i];
    if ((i & 3) == q) x[i >> 2] = xi;
    if (q == 0) inv[i * 64 + c] = xi;
  }
}

__global__ __launch_bounds__(256) void chol_diag(double* __restrict__ s, int64_t ld, int k0, int n,
                                                 double* __restrict__ inv, Scalars* sc) {
  chol_diag_block(s, ld, k0, n, inv, sc);
}


// ---------------------------------------------------------------------------
// Walk Hamiltonian on the device (the reference builds it on the host,
// FastCTQW/Operator.py:49-56): H = -gamma * A, or -gamma * (D - A) with
// D = diag(row sums of A); then A_exp = -i t H in the engine's input type.

// One workgroup per row: FP64 row sum of a real adjacency (fixed tree order).
template <typename Ta>
__global__ __launch_bounds__(256) void graph_row_sums(const Ta* __restrict__ adj, int n, double* __restrict__ rs) {
  __shared__ double part[256];
  const int i = blockIdx.x;
  double acc = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) acc += (double)adj[(int64_t)i * n + j];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) rs[i] = part[0];
}

template <typename Ta>
__global__ __launch_bounds__(256) void graph_hamiltonian(const Ta* __restrict__ adj, int n, double gamma,
                                                         const double* __restrict__ rs, double* __restrict__ h) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const double a = (double)adj[(int64_t)i * n + j];
  double v;
  if (rs)
    v = (i == j) ? rs[i] - a : -a;  // Laplacian D - A
  else
    v = a;
  h[(int64_t)i * n + j] = -gamma * v;
}

// In place: h = -gamma * A or -gamma * (D - A), A held in h (rs: its row sums).
__global__ __launch_bounds__(256) void graph_hamiltonian_inplace(double* __restrict__ h, int n, double gamma,
                                                                 const double* __restrict__ rs) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const int64_t o = (int64_t)i * n + j;
  const double a = h[o];
  const double v = rs ? ((i == j) ? rs[i] - a : -a) : a;
  h[o] = -gamma * v;
}

// Counter-based hash of the seeded graph generators (splitmix64; the same
// bits as fastctqw_amd/graphs.py): u = 53 hash bits of counter ^ key in [0, 1),
// key = splitmix64(splitmix64(seed) ^ stream).
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t hash_stream_key(uint64_t seed, uint64_t stream) {
  return splitmix64(splitmix64(seed) ^ stream);
}
__device__ __forceinline__ double hash_uniform(uint64_t key, uint64_t ctr) {
  return (double)(splitmix64(ctr ^ key) >> 11) * (1.0 / 9007199254740992.0);
}

// Dense adjacency of the seeded G(n, p): edge {i, j}, i < j, iff
// u(edge key, i n + j) < p, weight 1 or 1 - u(weight key, i n + j).
__global__ __launch_bounds__(256) void graph_er(double* __restrict__ a, int n, double p, uint64_t key_e,
                                                uint64_t key_w, int weighted) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  double v = 0.0;
  if (i != j) {
    const uint64_t lo = (uint64_t)min(i, j), hi = (uint64_t)max(i, j);
    const uint64_t ctr = lo * (uint64_t)n + hi;
    if (hash_uniform(key_e, ctr) < p) v = weighted ? 1.0 - hash_uniform(key_w, ctr) : 1.0;
  }
  a[(int64_t)i * n + j] = v;
}

__global__ __launch_bounds__(256) void fill_zero(double* __restrict__ a, int64_t cnt) {
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < cnt; k += (int64_t)gridDim.x * 256) a[k] = 0.0;
}

// Undirected edge list -> dense adjacency (zeroed before): a_ij = a_ji = w,
// a self-loop on the diagonal once; repeated pairs add their weights.
__global__ __launch_bounds__(256) void graph_edges(double* __restrict__ a, int n, int64_t m,
                                                   const int64_t* __restrict__ src, const int64_t* __restrict__ dst,
                                                   const double* __restrict__ w) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= m) return;
  const int64_t i = src[k], j = dst[k];
  const double v = w ? w[k] : 1.0;
  atomicAdd(&a[i * n + j], v);
  if (i != j) atomicAdd(&a[j * n + i], v);
}

// A_exp = -i t H (real H) into the engine input: complex64 / complex128.
template <typename Tc>
__global__ __launch_bounds__(256) void walk_input(const double* __restrict__ h, int64_t cnt, double t,
                                                  Tc* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= cnt) return;
  Tc v;
  v.x = 0;
  v.y = -t * h[k];
  out[k] = v;
}

// Zero columns [n, np) of rows [0, n) in up to 24 planes of leading
// dimension ld (the padding strip that GEMM k-loops read); one workgroup
// per row, np - n < 128.
template <typename W> struct PlaneList {
  W* p[24];
  int count;
};

template <typename W>
__global__ __launch_bounds__(128) void zero_strip(PlaneList<W> L, int n, int np, int64_t ld) {
  const int j = n + (int)threadIdx.x;
  if (j >= np) return;
  const int64_t o = (int64_t)blockIdx.x * ld + j;
  for (int q = 0; q < L.count; ++q) L.p[q][o] = (W)0;
}

// Exception bits of one run: its own scalars and those of its FP64
// companion (LU / Cholesky pivots), gathered into one device int.
__global__ void collect_info(const Scalars* a, const Scalars* b, int* out) {
  out[0] = a->info | (b ? b->info : 0);
}

}  // namespace fexpm
