// Grouped execution of a batch of CTQW walks of uneven order and time.
//
// Every stage of the structured pipeline (A = i G, G real symmetric) is one
// launch over all walks that need it: the analysis, the scaled input, each
// Pade product, the Cholesky panels of the FP64 solve, and each squaring
// round r (walks with s_i > r).  A launch is described by a device array of
// problems; workgroup b finds its problem by a binary search over the
// problems' first blocks (element-wise stages) or first tiles (GEMM
// stages).  GEMM stages list their problems by decreasing k, so the longest
// tiles start first.
//
// The host turns the per-walk plans (Pade degree m_i, squarings s_i, trace
// shift) into a program of stages, "materialises" it for one choice of tile
// shapes (descriptor arrays in one device buffer), and enqueues it.  The
// sequence after the analysis is captured once into a hipGraph and
// replayed while the plans stay the same.  Tile shapes are chosen by an
// autotuner: the program is run once per candidate, every GEMM launch is
// timed with events, and each class of launches (squarings, symmetric FP64
// products, solve updates) keeps its fastest candidate.
//
// Memory per walk (np = order padded to 64, P = np^2): FP64 planes
// a, u, v, p2, p4, p6, t1, t2, x0 (9P) and the inverses of the 64 x 64
// diagonal blocks of the Cholesky factor (64 np); squaring planes in the
// working precision W (4P); analysis and diagonal buffers.  p2..t1 are
// contiguous and become the solve's work area [Re Z; Im Z], S -> L, V V^T
// once the Pade products are done.  Padding columns of every plane stay
// zero: kernels write only the n x n region (or zeros), and the arenas are
// cleared when the batch's layout changes.
#pragma once
#include <deque>

namespace fexpm {

// ---------------------------------------------------------------------------
// Device side

// Problem p with start[p] <= b < start[p + 1] (start ascending, start[0] = 0).
__device__ __forceinline__ int group_find(const int* __restrict__ start, int nprob, int b) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class Op>
__global__ __launch_bounds__(Op::NT) void group_op_kernel(const typename Op::Prob* __restrict__ probs,
                                                          const int* __restrict__ start, int nprob) {
  const int b = (int)blockIdx.x;
  const int p = group_find(start, nprob, b);
  Op::run(probs[p], b - start[p]);
}

__global__ void zero_ints(int* __restrict__ p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

// Analysis of one walk's input, 32 x 32 tile pairs (see analyze_tiles).
template <typename Tin> struct OpAnalyze {
  static constexpr int NT = 256;
  struct Prob {
    const Tin* a;
    int n, nt;
    double* colpart;
    double2* diag;
    int* flags;
  };
  static int blocks(const Prob& p) { return p.nt * (p.nt + 1) / 2; }
  static __device__ __forceinline__ void run(const Prob& p, int b) {
    analyze_tile_pair<Tin>(p.a, p.n, p.nt, p.colpart, p.diag, p.flags, b);
  }
};

// off[j] = sum over row blocks of colpart (fixed order).
struct OpReduceCols {
  static constexpr int NT = 256;
  struct Prob {
    const double* colpart;
    int n, nt;
    double* off;
  };
  static int blocks(const Prob& p) { return (p.n + 255) / 256; }
  static __device__ __forceinline__ void run(const Prob& p, int b) {
    const int j = b * 256 + (int)threadIdx.x;
    if (j >= p.n) return;
    double s = 0.0;
    for (int q = 0; q < p.nt; ++q) s += p.colpart[(int64_t)q * p.n + j];
    p.off[j] = s;
  }
};

// Trace, mu and the two 1-norms of one walk (one workgroup).
struct OpFinalize {
  static constexpr int NT = 1024;
  struct Prob {
    const double* off;
    const double2* diag;
    int n;
    const int* flags;
    Scalars* out;
  };
  static int blocks(const Prob&) { return 1; }
  static __device__ __forceinline__ void run(const Prob& p, int) {
    finalize_norms_block(p.off, p.diag, p.n, p.flags, p.out);
  }
};

// g = (Im(A) - Im(mu) I) * scale into an FP64 plane (structured walks).
template <typename Tin> struct OpPrepare {
  static constexpr int NT = 256;
  struct Prob {
    const Tin* a;
    int n;
    int64_t ld;
    double* g;
    double mu_im, scale;
  };
  static int blocks(const Prob& p) { return p.n; }
  static __device__ __forceinline__ void run(const Prob& p, int i) {
    for (int j = (int)threadIdx.x; j < p.n; j += NT) {
      double im = InElem<Tin>::im(p.a[(int64_t)i * p.n + j]);
      if (i == j) im -= p.mu_im;
      p.g[(int64_t)i * p.ld + j] = im * p.scale;
    }
  }
};

// out = sum_t c_t in_t + diag I on the n x n region (FP64, one row per block).
struct OpLincomb {
  static constexpr int NT = 256;
  struct Prob {
    double* out;
    const double* in[4];
    double c[4];
    int nin;
    double diag;
    int n;
    int64_t ld;
  };
  static int blocks(const Prob& p) { return p.n; }
  static __device__ __forceinline__ void run(const Prob& p, int i) {
    for (int j = (int)threadIdx.x; j < p.n; j += NT) {
      const int64_t o = (int64_t)i * p.ld + j;
      double s = (i == j) ? p.diag : 0.0;
      for (int t = 0; t < p.nin; ++t) s += p.c[t] * p.in[t][o];
      p.out[o] = s;
    }
  }
};

// Stacked [V; Uq] (2np x np, zero padding) for the Cholesky solve.
struct OpPack {
  static constexpr int NT = 256;
  struct Prob {
    const double* v;
    const double* u;
    int n, np;
    int64_t ld;
    double* zb;
  };
  static int blocks(const Prob& p) { return 2 * p.np; }
  static __device__ __forceinline__ void run(const Prob& p, int r) {
    const int i = r < p.np ? r : r - p.np;
    const double* src = r < p.np ? p.v : p.u;
    for (int j = (int)threadIdx.x; j < p.np; j += NT)
      p.zb[(int64_t)r * p.ld + j] = (i < p.n && j < p.n) ? src[(int64_t)i * p.ld + j] : 0.0;
  }
};

// FP64 solution X -> split form in W: Y = X - diag(X), d = diag(X).
template <typename W> struct OpSolution {
  static constexpr int NT = 256;
  struct Prob {
    const double* xr;
    const double* xi;
    int64_t ldx;
    W* wr;
    W* wi;
    int64_t ldw;
    int n;
    double2* d;
  };
  static int blocks(const Prob& p) { return p.n; }
  static __device__ __forceinline__ void run(const Prob& p, int i) {
    for (int j = (int)threadIdx.x; j < p.n; j += NT) {
      double vr = p.xr[(int64_t)i * p.ldx + j], vi = p.xi[(int64_t)i * p.ldx + j];
      if (i == j) {
        p.d[i] = make_double2(vr, vi);
        vr = 0.0;
        vi = 0.0;
      }
      p.wr[(int64_t)i * p.ldw + j] = (W)vr;
      p.wi[(int64_t)i * p.ldw + j] = (W)vi;
    }
  }
};

// U = e^mu (Y + diag(d)) in the caller's interleaved layout.
template <typename W, typename Tout> struct OpOutput {
  static constexpr int NT = 256;
  struct Prob {
    const W* yr;
    const W* yi;
    int64_t ld;
    int n;
    const double2* d;
    double ph_re, ph_im;
    Tout* out;
  };
  static int blocks(const Prob& p) { return p.n; }
  static __device__ __forceinline__ void run(const Prob& p, int i) {
    for (int j = (int)threadIdx.x; j < p.n; j += NT) {
      double a, b;
      if (i == j) {
        a = p.d[i].x;
        b = p.d[i].y;
      } else {
        a = (double)p.yr[(int64_t)i * p.ld + j];
        b = (double)p.yi[(int64_t)i * p.ld + j];
      }
      p.out[(int64_t)i * p.n + j] = OutElem<Tout>::make(p.ph_re * a - p.ph_im * b, p.ph_re * b + p.ph_im * a);
    }
  }
};

// psi = e^mu (Y + diag(d)) psi0 with FP64 accumulation; one wavefront per row.
template <typename W> struct OpApply {
  static constexpr int NT = 256;
  struct Prob {
    const W* yr;
    const W* yi;
    int64_t ld;
    int n;
    const double2* d;
    const double2* psi0;
    double ph_re, ph_im;
    double2* out;
  };
  static int blocks(const Prob& p) { return (p.n + 3) / 4; }
  static __device__ __forceinline__ void run(const Prob& p, int b) {
    const int lane = (int)threadIdx.x & 63;
    const int row = b * 4 + ((int)threadIdx.x >> 6);
    if (row >= p.n) return;
    double ar = 0.0, ai = 0.0;
    for (int j = lane; j < p.n; j += 64) {
      const double a = (double)p.yr[(int64_t)row * p.ld + j], c = (double)p.yi[(int64_t)row * p.ld + j];
      const double2 v = p.psi0[j];
      ar += a * v.x - c * v.y;
      ai += a * v.y + c * v.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      ar += __shfl_xor(ar, off);
      ai += __shfl_xor(ai, off);
    }
    if (lane == 0) {
      const double2 dd = p.d[row], v = p.psi0[row];
      ar += dd.x * v.x - dd.y * v.y;
      ai += dd.x * v.y + dd.y * v.x;
      p.out[row] = make_double2(p.ph_re * ar - p.ph_im * ai, p.ph_re * ai + p.ph_im * ar);
    }
  }
};

// FP64 norm refinement of psi = U psi0 (unitary U): psi rescaled to the
// norm of psi0 (renorm_state_block), one workgroup per walk.
struct OpRenorm {
  static constexpr int NT = 256;
  struct Prob {
    const double2* psi0;
    double2* out;
    int n;
  };
  static int blocks(const Prob&) { return 1; }
  static __device__ __forceinline__ void run(const Prob& p, int) { renorm_state_block(p.psi0, p.out, p.n); }
};

// Diagonal 64 x 64 block of the Cholesky factor and its inverse.
struct OpChol {
  static constexpr int NT = 256;
  struct Prob {
    double* s;
    int64_t ld;
    int k0, n;
    double* inv;
    Scalars* sc;
  };
  static int blocks(const Prob&) { return 1; }
  static __device__ __forceinline__ void run(const Prob& p, int) { chol_diag_block(p.s, p.ld, p.k0, p.n, p.inv, p.sc); }
};

// Split of one walk's squaring operand into k-blocked 16-bit parts.
template <class P> struct OpSplit {
  static constexpr int NT = 256;
  struct Prob {
    const float* yr;
    const float* yi;
    int64_t ldy;
    int n, rows, kpad;
    typename P::H* out;
    int* info;
  };
  static int gx(const Prob& p) { return (p.kpad / 16 + 15) / 16; }
  static int blocks(const Prob& p) { return gx(p) * ((p.rows + 15) / 16); }
  static __device__ __forceinline__ void run(const Prob& p, int b) {
    const int g = (p.kpad / 16 + 15) / 16;
    split_blocked_block<P>(p.yr, p.yi, p.ldy, p.n, p.n, p.rows, p.kpad, p.out, p.info, b % g, b / g);
  }
};

// Exact int8 slices of one FP64 operand (rows of a walk's plane, optionally
// the concatenation of two planes), four rows per workgroup.  Five slices
// (34 bits of each row's largest element): on the 256-walk batch with
// uneven orders 64..2048 the same worst infidelity as six (2.0e-11 /
// 2.5e-11) with the FP64 class 67 -> 53 ms; four slices left 2.2e-8.
constexpr int GROUP_OZ_SLICES = 5;
struct OpOzSlice {
  static constexpr int NT = 256;
  struct Prob {
    const double* x1;
    double s1;
    const double* x2;
    double s2;
    int k1;
    int64_t ld;
    int nr, nk, rows, kpad;
    int8_t* out;
    int* ex;
  };
  static int blocks(const Prob& p) { return p.rows / 4; }  // rows: a multiple of OZ_BM
  static __device__ __forceinline__ void run(const Prob& p, int b) {
    oz_slice_rows4<GROUP_OZ_SLICES>(p.x1, p.s1, p.x2, p.s2, p.k1, p.ld, p.nr, p.nk, p.rows, p.kpad, p.out, p.ex,
                                    4 * b);
  }
};

// ---------------------------------------------------------------------------
// Host side

enum GroupKind {
  GK_ANALYZE,
  GK_REDUCE,
  GK_FINALIZE,
  GK_PREPARE,
  GK_LINCOMB,
  GK_PACK,
  GK_SOLUTION,
  GK_OUTPUT,
  GK_APPLY,
  GK_CHOL,
  GK_GEMM,
  GK_SPLIT_F16,   // split of the squaring operand (fp16x2 parts)
  GK_SPLIT_BF16,  // (bf16x3 parts)
  GK_SQ_F16,      // split squaring on 16-bit MFMA, tiles of SQ_BM (GEMM-like stage)
  GK_SQ_BF16,
  GK_OZSLICE,     // int8 slices of FP64 operands
  GK_OZGEMM,      // FP64 products from int8 slices, tiles of OZ_BM (GEMM-like stage)
  GK_RENORM       // FP64 norm refinement of the states psi_i
};

static inline bool gk_tiled(int kind) {
  return kind == GK_GEMM || kind == GK_SQ_F16 || kind == GK_SQ_BF16 || kind == GK_OZGEMM;
}

// Tuning classes of GEMM launches.
enum GroupClass { GC_SQUARE = 0, GC_SYM64 = 1, GC_SOLVE = 2, GC_COUNT = 3 };

// Autotuner candidates: tile configuration x dispatch.  Dispatch 0: one
// workgroup per tile in launch order; 1: one workgroup per tile, runs of 4
// tiles per XCD; 2: persistent grid (as many workgroups as fit on the
// device at once), runs of 4 tiles per XCD.
constexpr int GROUP_CANDIDATES = 6;
static inline int cand_cfg(int c) { return c / 3; }
static inline int cand_dispatch(int c) { return c % 3; }

struct GroupStage {
  int kind = 0;
  int cls = -1;      // GEMM: tuning class
  int mode = 0;      // GEMM: MODE_REAL or MODE_C3M
  bool f32 = false;  // GEMM: FP32 operands (blocked-accumulation squarings)
  size_t psize = 0;
  std::vector<unsigned char> probs;
  std::vector<int> blocks;  // element-wise: workgroups per problem
  std::vector<char> sym;    // GEMM: lower-triangular problem
  std::vector<int> walk;    // GEMM: owning walk (flop accounting)
  std::vector<int> inplace; // GEMM: output overwrites operand A (needs one column tile)

  template <class P> void add(const P& p, int nb) {
    psize = sizeof(P);
    const size_t o = probs.size();
    probs.resize(o + sizeof(P));
    std::memcpy(probs.data() + o, &p, sizeof(P));
    blocks.push_back(nb);
  }
  template <typename T> void add_gemm(const GemmArgs<T>& g, bool is_sym, int w, bool in_place = false) {
    add(g, 0);
    sym.push_back(is_sym ? 1 : 0);
    walk.push_back(w);
    inplace.push_back(in_place ? 1 : 0);
  }
  int count() const { return psize ? (int)(probs.size() / psize) : 0; }
};

struct GroupLaunch {
  int kind, cls, mode;
  bool f32;
  int cfg, dispatch;
  int nprob, total, grid, chunk;
  size_t off_probs, off_start;
};

// Tile edge of a GEMM configuration (all grouped configurations are square).
static inline int group_tile(bool f32, int cfg) {
  if (f32) return cfg == 0 ? FCfgSmall::BM : FCfg16w::BM;
  return cfg == 0 ? GemmCfg<double>::BM : DCfg4w::BM;
}

template <typename KernelT> static int group_capacity(KernelT kernel, int nt) {
  static std::map<const void*, int> cache;
  static std::mutex mtx;
  std::lock_guard<std::mutex> lk(mtx);
  auto it = cache.find((const void*)kernel);
  if (it != cache.end()) return it->second;
  int dev = 0, per_cu = 0, cap = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kernel, nt, 0) == hipSuccess)
    cap = per_cu * prop.multiProcessorCount;
  cap = std::max(cap, 8);
  cache[(const void*)kernel] = cap;
  return cap;
}

template <typename T, int MODE, bool KBLK, class CFG> static int group_gemm_capacity() {
  return group_capacity(gemm_group_kernel<T, MODE, KBLK, CFG>, CFG::NT);
}

// A materialised program: descriptor arrays in one device buffer.
struct GroupProgram {
  std::vector<GroupLaunch> launches;
  unsigned char* d_arena = nullptr;
  size_t cap = 0;
  std::vector<double> walk_flops;
  ~GroupProgram() {
    if (d_arena) (void)hipFree(d_arena);
  }

  // Split squaring stage: problems are SqArgs<P>, lower tiles of SQ_BM.
  template <class P>
  void materialize_sq(const GroupStage& S, GroupLaunch& L, std::vector<unsigned char>& host, const int* cand) {
    const int np_ = S.count();
    const int c = cand[S.cls];
    L.cfg = 0;
    L.dispatch = cand_dispatch(c);
    std::vector<int> order((size_t)np_);
    for (int p = 0; p < np_; ++p) order[p] = p;
    auto arg = [&](int p) -> const SqArgs<P>& {
      return *reinterpret_cast<const SqArgs<P>*>(S.probs.data() + (size_t)p * S.psize);
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return arg(x).k > arg(y).k; });
    std::vector<GroupSq<P>> buf((size_t)np_);
    int tot = 0;
    for (int q = 0; q < np_; ++q) {
      const SqArgs<P>& g = arg(order[q]);
      if (g.m <= 0 || g.k <= 0 || g.k % SQ_BK != 0 || g.rows % SQ_BM != 0 || g.rows < g.m)
        throw Error(FEXPM_EINVAL, "grouped split squaring: bad problem shape");
      const int T = (g.m + SQ_BM - 1) / SQ_BM;
      const int tiles = T * (T + 1) / 2;
      walk_flops[(size_t)S.walk[order[q]]] += 4.0 * P::NPAIR * 2.0 * tiles * (double)SQ_BM * SQ_BM * (double)g.k;
      std::memset(&buf[q], 0, sizeof(GroupSq<P>));
      buf[q].g = g;
      buf[q].tile0 = tot;
      buf[q].tiles = tiles;
      tot += tiles;
    }
    L.total = tot;
    L.chunk = L.dispatch == 0 ? 1 : 4;
    int capw = tot;
    if (L.dispatch == 2)
      capw = std::max(8, group_capacity(sq_split_group_kernel<P>, sq_threads<P>()) / 8 * 8);
    L.grid = std::min(tot, capw);
    host.resize((host.size() + 255) / 256 * 256);
    L.off_probs = host.size();
    const size_t bytes = buf.size() * sizeof(GroupSq<P>);
    host.resize(host.size() + bytes);
    std::memcpy(host.data() + L.off_probs, buf.data(), bytes);
  }

  // FP64 products from int8 slices: problems are OzArgs, tiles of OZ_BM.
  void materialize_oz(const GroupStage& S, GroupLaunch& L, std::vector<unsigned char>& host, const int* cand) {
    using A = OzArgs<GROUP_OZ_SLICES>;
    const int np_ = S.count();
    L.cfg = 0;
    L.dispatch = cand_dispatch(cand[S.cls]);
    std::vector<int> order((size_t)np_);
    for (int p = 0; p < np_; ++p) order[p] = p;
    auto arg = [&](int p) -> const A& { return *reinterpret_cast<const A*>(S.probs.data() + (size_t)p * S.psize); };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return arg(x).kpad > arg(y).kpad; });
    std::vector<GroupOz<GROUP_OZ_SLICES>> buf((size_t)np_);
    int tot = 0;
    for (int q = 0; q < np_; ++q) {
      const int p = order[q];
      const A& g = arg(p);
      const bool sy = S.sym[p] != 0;
      if (g.m <= 0 || g.n <= 0 || g.kpad <= 0 || g.kpad % OZ_BK != 0 || g.rows_a < g.m || g.rows_b < g.n ||
          g.rows_a % OZ_BM != 0 || g.rows_b % OZ_BM != 0 || (sy && g.m != g.n))
        throw Error(FEXPM_EINVAL, "grouped int8-slice product: bad problem shape");
      const int tm = (g.m + OZ_BM - 1) / OZ_BM, tn = (g.n + OZ_BM - 1) / OZ_BM;
      const int tiles = sy ? tm * (tm + 1) / 2 : tm * tn;
      walk_flops[(size_t)S.walk[p]] += 2.0 * tiles * (double)OZ_BM * OZ_BM * (double)g.kpad;  // FP64-equivalent
      std::memset(&buf[q], 0, sizeof(buf[q]));
      buf[q].g = g;
      buf[q].tile0 = tot;
      buf[q].tiles = tiles;
      buf[q].tiles_n = tn;
      buf[q].sym = sy ? 1 : 0;
      tot += tiles;
    }
    L.total = tot;
    L.chunk = L.dispatch == 0 ? 1 : 4;
    int capw = tot;
    if (L.dispatch == 2) capw = std::max(8, group_capacity(oz_group_kernel<GROUP_OZ_SLICES>, 512) / 8 * 8);
    L.grid = std::min(tot, capw);
    host.resize((host.size() + 255) / 256 * 256);
    L.off_probs = host.size();
    const size_t bytes = buf.size() * sizeof(buf[0]);
    host.resize(host.size() + bytes);
    std::memcpy(host.data() + L.off_probs, buf.data(), bytes);
  }

  void materialize(const std::deque<GroupStage>& stages, const int* cand, int nwalks) {
    std::vector<unsigned char> host;
    auto align = [&]() { host.resize((host.size() + 255) / 256 * 256); };
    launches.clear();
    walk_flops.assign((size_t)nwalks, 0.0);
    for (const GroupStage& S : stages) {
      const int np_ = S.count();
      if (np_ == 0) continue;
      GroupLaunch L;
      std::memset(&L, 0, sizeof(L));
      L.kind = S.kind;
      L.cls = S.cls;
      L.mode = S.mode;
      L.f32 = S.f32;
      L.nprob = np_;
      L.chunk = 1;
      if (S.kind == GK_OZGEMM) {
        materialize_oz(S, L, host, cand);
      } else if (S.kind == GK_SQ_F16) {
        materialize_sq<Fp16x2>(S, L, host, cand);
      } else if (S.kind == GK_SQ_BF16) {
        materialize_sq<Bf16x3>(S, L, host, cand);
      } else if (S.kind != GK_GEMM) {
        std::vector<int> start((size_t)np_);
        int tot = 0;
        for (int p = 0; p < np_; ++p) {
          start[p] = tot;
          tot += S.blocks[p];
        }
        L.total = tot;
        L.grid = tot;
        align();
        L.off_probs = host.size();
        host.insert(host.end(), S.probs.begin(), S.probs.end());
        align();
        L.off_start = host.size();
        host.resize(host.size() + start.size() * sizeof(int));
        std::memcpy(host.data() + L.off_start, start.data(), start.size() * sizeof(int));
      } else {
        const int c = cand[S.cls];
        L.cfg = cand_cfg(c);
        L.dispatch = cand_dispatch(c);
        const int tile = group_tile(S.f32, L.cfg);
        // problems by decreasing k (stable)
        std::vector<int> order((size_t)np_);
        for (int p = 0; p < np_; ++p) order[p] = p;
        auto kof = [&](int p) {
          return S.f32 ? reinterpret_cast<const GemmArgs<float>*>(S.probs.data() + (size_t)p * S.psize)->k
                       : reinterpret_cast<const GemmArgs<double>*>(S.probs.data() + (size_t)p * S.psize)->k;
        };
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return kof(x) > kof(y); });
        const double mult = S.mode == MODE_REAL ? 1.0 : (S.mode == MODE_C4M ? 4.0 : 3.0);
        const size_t gsize = S.f32 ? sizeof(GroupGemm<float>) : sizeof(GroupGemm<double>);
        std::vector<unsigned char> buf((size_t)np_ * gsize);
        int tot = 0;
        for (int q = 0; q < np_; ++q) {
          const int p = order[q];
          int m, n, k;
          if (S.f32) {
            const GemmArgs<float>& g = *reinterpret_cast<const GemmArgs<float>*>(S.probs.data() + (size_t)p * S.psize);
            m = g.m; n = g.n; k = g.k;
          } else {
            const GemmArgs<double>& g = *reinterpret_cast<const GemmArgs<double>*>(S.probs.data() + (size_t)p * S.psize);
            m = g.m; n = g.n; k = g.k;
          }
          if (m <= 0 || n <= 0 || k <= 0 || k % 16 != 0) throw Error(FEXPM_EINVAL, "grouped GEMM: bad problem shape");
          const int tm = (m + tile - 1) / tile, tn = (n + tile - 1) / tile;
          if (S.inplace[p] && tn != 1) throw Error(FEXPM_EINVAL, "grouped GEMM: in-place problem wider than a tile");
          const bool sy = S.sym[p] != 0;
          if (sy && m != n) throw Error(FEXPM_EINVAL, "grouped GEMM: symmetric problem must be square");
          const int tiles = sy ? tm * (tm + 1) / 2 : tm * tn;
          walk_flops[(size_t)S.walk[p]] += mult * 2.0 * tiles * (double)tile * tile * (double)k;
          if (S.f32) {
            GroupGemm<float> G;
            std::memset(&G, 0, sizeof(G));
            G.g = *reinterpret_cast<const GemmArgs<float>*>(S.probs.data() + (size_t)p * S.psize);
            G.tile0 = tot;
            G.tiles = tiles;
            G.tiles_n = tn;
            G.sym = sy ? 1 : 0;
            std::memcpy(buf.data() + (size_t)q * gsize, &G, sizeof(G));
          } else {
            GroupGemm<double> G;
            std::memset(&G, 0, sizeof(G));
            G.g = *reinterpret_cast<const GemmArgs<double>*>(S.probs.data() + (size_t)p * S.psize);
            G.tile0 = tot;
            G.tiles = tiles;
            G.tiles_n = tn;
            G.sym = sy ? 1 : 0;
            std::memcpy(buf.data() + (size_t)q * gsize, &G, sizeof(G));
          }
          tot += tiles;
        }
        L.total = tot;
        L.chunk = L.dispatch == 0 ? 1 : 4;
        int capw = tot;
        if (L.dispatch == 2) {
          if (S.f32)
            capw = L.cfg == 0 ? group_gemm_capacity<float, MODE_C3M, true, FCfgSmall>()
                              : group_gemm_capacity<float, MODE_C3M, true, FCfg16w>();
          else if (S.mode == MODE_REAL)
            capw = L.cfg == 0 ? group_gemm_capacity<double, MODE_REAL, false, GemmCfg<double>>()
                              : group_gemm_capacity<double, MODE_REAL, false, DCfg4w>();
          else
            capw = L.cfg == 0 ? group_gemm_capacity<double, MODE_C3M, false, GemmCfg<double>>()
                              : group_gemm_capacity<double, MODE_C3M, false, DCfg4w>();
          capw = std::max(8, capw / 8 * 8);  // a multiple of the XCD count keeps each workgroup on its runs
        }
        L.grid = std::min(tot, capw);
        align();
        L.off_probs = host.size();
        host.insert(host.end(), buf.begin(), buf.end());
      }
      launches.push_back(L);
    }
    if (host.size() > cap) {
      if (d_arena) (void)hipFree(d_arena);
      d_arena = nullptr;
      cap = 0;
      FX_HIP(hipMalloc(&d_arena, host.size()));
      cap = host.size();
    }
    if (!host.empty()) FX_HIP(hipMemcpy(d_arena, host.data(), host.size(), hipMemcpyHostToDevice));
  }
};

// Launch of one materialised stage.
template <typename Tin, typename W>
static void group_enqueue_one(const GroupLaunch& L, unsigned char* base, hipStream_t st) {
  using OutOp = OpOutput<W, Tin>;
  const void* pp = base + L.off_probs;
  const int* ps = reinterpret_cast<const int*>(base + L.off_start);
  const unsigned grid = (unsigned)L.grid;
#define FX_OP(OP)                                                                                          \
  group_op_kernel<OP><<<grid, OP::NT, 0, st>>>(reinterpret_cast<const typename OP::Prob*>(pp), ps, L.nprob); \
  break;
  switch (L.kind) {
    case GK_ANALYZE: FX_OP(OpAnalyze<Tin>)
    case GK_REDUCE: FX_OP(OpReduceCols)
    case GK_FINALIZE: FX_OP(OpFinalize)
    case GK_PREPARE: FX_OP(OpPrepare<Tin>)
    case GK_LINCOMB: FX_OP(OpLincomb)
    case GK_PACK: FX_OP(OpPack)
    case GK_SOLUTION: FX_OP(OpSolution<W>)
    case GK_OUTPUT: FX_OP(OutOp)
    case GK_APPLY: FX_OP(OpApply<W>)
    case GK_RENORM: FX_OP(OpRenorm)
    case GK_CHOL: FX_OP(OpChol)
    case GK_OZSLICE: FX_OP(OpOzSlice)
    case GK_OZGEMM:
      oz_group_kernel<GROUP_OZ_SLICES><<<grid, 512, 0, st>>>(
          reinterpret_cast<const GroupOz<GROUP_OZ_SLICES>*>(pp), L.nprob, L.total, L.chunk);
      break;
    case GK_SPLIT_F16: FX_OP(OpSplit<Fp16x2>)
    case GK_SPLIT_BF16: FX_OP(OpSplit<Bf16x3>)
    case GK_SQ_F16:
      sq_split_group_kernel<Fp16x2><<<grid, sq_threads<Fp16x2>(), 0, st>>>(
          reinterpret_cast<const GroupSq<Fp16x2>*>(pp), L.nprob, L.total, L.chunk);
      break;
    case GK_SQ_BF16:
      sq_split_group_kernel<Bf16x3><<<grid, sq_threads<Bf16x3>(), 0, st>>>(
          reinterpret_cast<const GroupSq<Bf16x3>*>(pp), L.nprob, L.total, L.chunk);
      break;
    case GK_GEMM: {
#define FX_GG(T, MODE, KB, CF)                                                                             \
  gemm_group_kernel<T, MODE, KB, CF><<<grid, CF::NT, 0, st>>>(reinterpret_cast<const GroupGemm<T>*>(pp), \
                                                               L.nprob, L.total, L.chunk)
      if (L.f32) {
        if (L.cfg == 0) FX_GG(float, MODE_C3M, true, FCfgSmall);
        else FX_GG(float, MODE_C3M, true, FCfg16w);
      } else if (L.mode == MODE_REAL) {
        if (L.cfg == 0) FX_GG(double, MODE_REAL, false, GemmCfg<double>);
        else FX_GG(double, MODE_REAL, false, DCfg4w);
      } else {
        if (L.cfg == 0) FX_GG(double, MODE_C3M, false, GemmCfg<double>);
        else FX_GG(double, MODE_C3M, false, DCfg4w);
      }
#undef FX_GG
      break;
    }
    default: throw Error(FEXPM_EINVAL, "grouped launch: unknown stage");
  }
#undef FX_OP
  FX_LAUNCH();
}

// Per-walk plan of a grouped (structured, finite) walk.
struct GroupPlan {
  int idx;  // index in the batch
  int m, s;
  int shifted;
  double mu_re, mu_im, norm, norm_plain;
};

// Virtual address regions of a host-only program description.
enum VRegion { VR_F64 = 1, VR_W, VR_SMALL, VR_SPLIT, VR_IN, VR_OUT, VR_PSI0, VR_PSI, VR_SCAL, VR_OZ };
constexpr int VR_SHIFT = 40;
inline void* vaddr(int region) { return reinterpret_cast<void*>((uintptr_t)region << VR_SHIFT); }

// "region:byte offset" of a virtual address (0:0 for null).
inline std::string vfmt(const void* p) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  char b[64];
  std::snprintf(b, sizeof(b), "%llu:%llu", (unsigned long long)(a >> VR_SHIFT),
                (unsigned long long)(a & (((uintptr_t)1 << VR_SHIFT) - 1)));
  return b;
}
inline std::string dfmt(double v) {
  char b[40];
  std::snprintf(b, sizeof(b), "%.17g", v);
  return b;
}

// One text line per problem of a stage (testing ABI).
template <typename T> static std::string describe_gemm(const GroupStage& S, int p) {
  const GemmArgs<T>& g = *reinterpret_cast<const GemmArgs<T>*>(S.probs.data() + (size_t)p * S.psize);
  std::string o = "GEMM cls=" + std::to_string(S.cls) + " mode=" + std::to_string(S.mode) +
                  " sym=" + std::to_string((int)S.sym[p]) + " inplace=" + std::to_string(S.inplace[p]) +
                  " ar=" + vfmt(g.ar) + " ai=" + vfmt(g.ai) + " lda=" + std::to_string(g.lda) +
                  " br=" + vfmt(g.br) + " bi=" + vfmt(g.bi) + " ldb=" + std::to_string(g.ldb) +
                  " m=" + std::to_string(g.m) + " n=" + std::to_string(g.n) + " k=" + std::to_string(g.k) +
                  " cr=" + vfmt(g.epi.cr) + " ci=" + vfmt(g.epi.ci) + " ldc=" + std::to_string(g.epi.ldc) +
                  " alpha=" + dfmt(g.epi.alpha) + " beta=" + dfmt(g.epi.beta) +
                  " nadd=" + std::to_string(g.epi.nadd) + " ld_add=" + std::to_string(g.epi.ld_add) +
                  " diag=" + dfmt(g.epi.diag) + " din=" + vfmt(g.epi.d_in) + " dout=" + vfmt(g.epi.d_out);
  for (int t = 0; t < g.epi.nadd; ++t)
    o += " add" + std::to_string(t) + "=" + vfmt(g.epi.add_r[t]) + " c" + std::to_string(t) + "=" +
         dfmt(g.epi.addc[t]);
  return o;
}

template <class P> static std::string describe_sq(const GroupStage& S, int p) {
  const SqArgs<P>& a = *reinterpret_cast<const SqArgs<P>*>(S.probs.data() + (size_t)p * S.psize);
  return "SQ a=" + vfmt(a.a) + " rows=" + std::to_string(a.rows) + " yr=" + vfmt(a.yr) + " yi=" + vfmt(a.yi) +
         " ldy=" + std::to_string(a.ldy) + " cr=" + vfmt(a.cr) + " ci=" + vfmt(a.ci) +
         " ldc=" + std::to_string(a.ldc) + " din=" + vfmt(a.d_in) + " dout=" + vfmt(a.d_out) +
         " m=" + std::to_string(a.m) + " k=" + std::to_string(a.k);
}

template <class P> static std::string describe_split(const GroupStage& S, int p) {
  const typename OpSplit<P>::Prob& q =
      *reinterpret_cast<const typename OpSplit<P>::Prob*>(S.probs.data() + (size_t)p * S.psize);
  return "SPLIT yr=" + vfmt(q.yr) + " yi=" + vfmt(q.yi) + " ldy=" + std::to_string(q.ldy) +
         " n=" + std::to_string(q.n) + " rows=" + std::to_string(q.rows) + " kpad=" + std::to_string(q.kpad) +
         " out=" + vfmt(q.out) + " info=" + vfmt(q.info);
}

template <typename Tin, typename W> static std::string describe_stage(const GroupStage& S) {
  std::string out;
  const int cnt = S.count();
  out += "STAGE kind=" + std::to_string(S.kind) + " problems=" + std::to_string(cnt) + "\n";
  for (int p = 0; p < cnt; ++p) {
    const unsigned char* b = S.probs.data() + (size_t)p * S.psize;
    std::string l;
    switch (S.kind) {
      case GK_PREPARE: {
        const auto& q = *reinterpret_cast<const typename OpPrepare<Tin>::Prob*>(b);
        l = "PREPARE a=" + vfmt(q.a) + " n=" + std::to_string(q.n) + " ld=" + std::to_string(q.ld) +
            " g=" + vfmt(q.g) + " mu_im=" + dfmt(q.mu_im) + " scale=" + dfmt(q.scale);
        break;
      }
      case GK_LINCOMB: {
        const auto& q = *reinterpret_cast<const OpLincomb::Prob*>(b);
        l = "LINCOMB out=" + vfmt(q.out) + " nin=" + std::to_string(q.nin) + " diag=" + dfmt(q.diag) +
            " n=" + std::to_string(q.n) + " ld=" + std::to_string(q.ld);
        for (int t = 0; t < q.nin; ++t)
          l += " in" + std::to_string(t) + "=" + vfmt(q.in[t]) + " c" + std::to_string(t) + "=" + dfmt(q.c[t]);
        break;
      }
      case GK_PACK: {
        const auto& q = *reinterpret_cast<const OpPack::Prob*>(b);
        l = "PACK v=" + vfmt(q.v) + " u=" + vfmt(q.u) + " n=" + std::to_string(q.n) + " np=" + std::to_string(q.np) +
            " ld=" + std::to_string(q.ld) + " zb=" + vfmt(q.zb);
        break;
      }
      case GK_CHOL: {
        const auto& q = *reinterpret_cast<const OpChol::Prob*>(b);
        l = "CHOL s=" + vfmt(q.s) + " ld=" + std::to_string(q.ld) + " k0=" + std::to_string(q.k0) +
            " n=" + std::to_string(q.n) + " inv=" + vfmt(q.inv) + " sc=" + vfmt(q.sc);
        break;
      }
      case GK_SOLUTION: {
        const auto& q = *reinterpret_cast<const typename OpSolution<W>::Prob*>(b);
        l = "SOLUTION xr=" + vfmt(q.xr) + " xi=" + vfmt(q.xi) + " ldx=" + std::to_string(q.ldx) +
            " wr=" + vfmt(q.wr) + " wi=" + vfmt(q.wi) + " ldw=" + std::to_string(q.ldw) +
            " n=" + std::to_string(q.n) + " d=" + vfmt(q.d);
        break;
      }
      case GK_OUTPUT: {
        const auto& q = *reinterpret_cast<const typename OpOutput<W, Tin>::Prob*>(b);
        l = "OUTPUT yr=" + vfmt(q.yr) + " yi=" + vfmt(q.yi) + " ld=" + std::to_string(q.ld) +
            " n=" + std::to_string(q.n) + " d=" + vfmt(q.d) + " phr=" + dfmt(q.ph_re) + " phi=" + dfmt(q.ph_im) +
            " out=" + vfmt(q.out);
        break;
      }
      case GK_APPLY: {
        const auto& q = *reinterpret_cast<const typename OpApply<W>::Prob*>(b);
        l = "APPLY yr=" + vfmt(q.yr) + " yi=" + vfmt(q.yi) + " ld=" + std::to_string(q.ld) +
            " n=" + std::to_string(q.n) + " d=" + vfmt(q.d) + " psi0=" + vfmt(q.psi0) +
            " phr=" + dfmt(q.ph_re) + " phi=" + dfmt(q.ph_im) + " out=" + vfmt(q.out);
        break;
      }
      case GK_RENORM: {
        const auto& q = *reinterpret_cast<const OpRenorm::Prob*>(b);
        l = "RENORM psi0=" + vfmt(q.psi0) + " out=" + vfmt(q.out) + " n=" + std::to_string(q.n);
        break;
      }
      case GK_GEMM: l = S.f32 ? describe_gemm<float>(S, p) : describe_gemm<double>(S, p); break;
      case GK_OZSLICE: {
        const auto& q = *reinterpret_cast<const OpOzSlice::Prob*>(b);
        l = "OZSLICE x1=" + vfmt(q.x1) + " s1=" + dfmt(q.s1) + " x2=" + vfmt(q.x2) + " s2=" + dfmt(q.s2) +
            " k1=" + std::to_string(q.k1) + " ld=" + std::to_string(q.ld) + " nr=" + std::to_string(q.nr) +
            " nk=" + std::to_string(q.nk) + " rows=" + std::to_string(q.rows) + " kpad=" + std::to_string(q.kpad) +
            " out=" + vfmt(q.out) + " ex=" + vfmt(q.ex);
        break;
      }
      case GK_OZGEMM: {
        const auto& g = *reinterpret_cast<const OzArgs<GROUP_OZ_SLICES>*>(b);
        l = "OZGEMM sym=" + std::to_string((int)S.sym[p]) + " a=" + vfmt(g.a) + " rows_a=" + std::to_string(g.rows_a) +
            " b=" + vfmt(g.b) + " rows_b=" + std::to_string(g.rows_b) + " m=" + std::to_string(g.m) +
            " n=" + std::to_string(g.n) + " kpad=" + std::to_string(g.kpad) + " cr=" + vfmt(g.epi.cr) +
            " ldc=" + std::to_string(g.epi.ldc) + " alpha=" + dfmt(g.epi.alpha) + " beta=" + dfmt(g.epi.beta) +
            " nadd=" + std::to_string(g.epi.nadd) + " ld_add=" + std::to_string(g.epi.ld_add) +
            " diag=" + dfmt(g.epi.diag);
        for (int t = 0; t < g.epi.nadd; ++t)
          l += " add" + std::to_string(t) + "=" + vfmt(g.epi.add_r[t]) + " c" + std::to_string(t) + "=" +
               dfmt(g.epi.addc[t]);
        break;
      }
      case GK_SPLIT_F16: l = describe_split<Fp16x2>(S, p); break;
      case GK_SPLIT_BF16: l = describe_split<Bf16x3>(S, p); break;
      case GK_SQ_F16: l = describe_sq<Fp16x2>(S, p); break;
      case GK_SQ_BF16: l = describe_sq<Bf16x3>(S, p); break;
      default: l = "UNKNOWN"; break;
    }
    out += l + "\n";
  }
  return out;
}

// Runner for one batch dtype: Tin = float2 (complex64, FP32 squarings) or
// double2 (complex128, FP64 squarings); Pade products and solve in FP64.
struct GroupRunnerBase {
  virtual ~GroupRunnerBase() {}
  virtual void set_layout(const std::vector<int64_t>& n, const std::vector<size_t>& off, hipStream_t st) = 0;
  virtual void analyze(const void* d_in, Scalars* d_scal, hipStream_t st) = 0;
  // Builds (or reuses) the program for these plans, tunes it when asked,
  // and runs it on st; psi0/psi are packed device states (or null).
  virtual void run(const std::vector<GroupPlan>& plans, const void* d_in, void* d_out, Scalars* d_scal,
                   const double2* d_psi0, double2* d_psi, const std::vector<size_t>& voff, hipStream_t st) = 0;
  // testing: every class runs candidate c (no tuning); -1 = autotuner
  virtual void force_candidate(int c) = 0;
  int kblock = 8;
  int autotune = 1;
  int use_graph = 1;
  int forced = -1;
  int precision = FEXPM_PREC_NATIVE;  // squarings of complex64 batches: split 16-bit forms
  int oz_slices = 0;                  // Pade products from int8 slices (complex64 batches: 6), 0 = FP64 MFMA
  int cheb = 1;                       // Chebyshev polynomial in place of the Pade products and solve
  // report of the last run
  int launches = 0, graph_used = 0, graph_built = 0, tuned = 0;
  int cand[GC_COUNT] = {0, 0, 0};
  double class_ms[GC_COUNT] = {0, 0, 0};
  double tune_ms = 0.0, device_ms = 0.0;
  std::vector<double> walk_flops;  // by batch index
};

template <typename Tin, typename W> struct GroupRunner : GroupRunnerBase {
  struct WalkMem {
    int n = 0, np = 0;
    double* f64 = nullptr;  // 9 planes of np^2, then 64 np inverse blocks
    W* w = nullptr;         // 4 planes of np^2
    double* colpart = nullptr;
    double* off = nullptr;
    double2* diag = nullptr;
    double2* dsq = nullptr;  // 2 x n
    int* flags = nullptr;
    size_t in_off = 0;       // element offset of the walk's input/output
  };
  std::vector<WalkMem> wm;
  double* f64_arena = nullptr;
  W* w_arena = nullptr;
  unsigned char* small_arena = nullptr;
  size_t f64_cap = 0, w_cap = 0, small_cap = 0;
  int* d_flags = nullptr;  // all walks' flags, contiguous (one clear launch)
  int nwalk = 0;
  // k-blocked 16-bit split parts of the squaring operand (split precision),
  // rows padded to SQ_BM and k to 16, room for three parts per plane
  void* split_arena = nullptr;
  size_t split_cap = 0;
  std::vector<size_t> split_off;  // bytes, per walk
  bool split_ready = false;
  // int8 slices of the Pade products: three slice buffers and their row
  // exponents per walk (rows padded to OZ_BM, k = np)
  void* oz_arena = nullptr;
  size_t oz_cap = 0;
  std::vector<size_t> oz_off;  // bytes, per walk
  bool oz_ready = false;

  GroupProgram analysis;   // analysis program (per layout)
  const void* analysis_in = nullptr;
  Scalars* analysis_scal = nullptr;
  GroupProgram prog;       // program of the current plans
  std::deque<GroupStage> stages;
  std::vector<double> plan_key;
  const void* key_in = nullptr;
  void* key_out = nullptr;
  const void* key_psi0 = nullptr;
  bool tuned_prec[3] = {false, false, false};  // the autotuner ran for this layout, per precision
  int cand_prec[3][GC_COUNT] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<hipEvent_t> evpool;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // virt: no device memory; arenas get the addresses region << 40 (regions
  // of vaddr) so that the program can be built and described on the host.
  bool virt = false;

  explicit GroupRunner(bool virtual_layout = false) : virt(virtual_layout) {
    // complex64 batches: Pade products from 6 int8 slices (FP64 accuracy)
    if (std::is_same<W, float>::value) oz_slices = GROUP_OZ_SLICES;
  }
  ~GroupRunner() override {
    drop_graph();
    if (!virt) {
      if (f64_arena) (void)hipFree(f64_arena);
      if (w_arena) (void)hipFree(w_arena);
      if (small_arena) (void)hipFree(small_arena);
      if (split_arena) (void)hipFree(split_arena);
      if (oz_arena) (void)hipFree(oz_arena);
    }
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }

  void ensure_events() {
    if (!ev0) FX_HIP(hipEventCreate(&ev0));
    if (!ev1) FX_HIP(hipEventCreate(&ev1));
  }

  void force_candidate(int c) override {
    forced = c;
    drop_graph();
    stages.clear();
    plan_key.clear();
    for (bool& t : tuned_prec) t = false;
    for (int k = 0; k < GC_COUNT; ++k) cand[k] = c >= 0 ? c : 0;
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  static int64_t pad64(int64_t x) { return (std::max<int64_t>(x, 1) + 63) / 64 * 64; }

  void set_layout(const std::vector<int64_t>& n, const std::vector<size_t>& off, hipStream_t) override {
    const int cnt = (int)n.size();
    bool same = cnt == nwalk;
    for (int i = 0; same && i < cnt; ++i) same = wm[i].n == (int)n[i] && wm[i].in_off == off[i];
    if (same) return;
    drop_graph();
    stages.clear();
    plan_key.clear();
    analysis_in = nullptr;
    split_ready = false;
    oz_ready = false;
    for (bool& t : tuned_prec) t = false;
    nwalk = cnt;
    wm.assign(cnt, WalkMem());
    size_t f64_tot = 0, w_tot = 0, small_tot = 0;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    std::vector<size_t> fo(cnt), wo(cnt), so(cnt);
    for (int i = 0; i < cnt; ++i) {
      const int64_t np_ = pad64(n[i]);
      const size_t P = (size_t)np_ * (size_t)np_;
      fo[i] = f64_tot;
      f64_tot += al((9 * P + 64 * (size_t)np_) * sizeof(double));
      wo[i] = w_tot;
      w_tot += al(4 * P * sizeof(W));
      so[i] = small_tot;
      const int nt = (int)((n[i] + 31) / 32);
      small_tot += al((size_t)nt * n[i] * sizeof(double)) + al(n[i] * sizeof(double)) +
                   al(n[i] * sizeof(double2)) + al(2 * n[i] * sizeof(double2));
    }
    const size_t flags_off = small_tot;
    small_tot += al((size_t)std::max(cnt, 1) * sizeof(int));
    auto grow = [](auto*& p, size_t& capb, size_t need) {
      if (need > capb) {
        if (p) (void)hipFree(p);
        p = nullptr;
        capb = 0;
        FX_HIP(hipMalloc((void**)&p, need));
        capb = need;
      }
    };
    if (virt) {
      f64_arena = reinterpret_cast<double*>(vaddr(VR_F64));
      w_arena = reinterpret_cast<W*>(vaddr(VR_W));
      small_arena = reinterpret_cast<unsigned char*>(vaddr(VR_SMALL));
      f64_cap = f64_tot;
      w_cap = w_tot;
      small_cap = small_tot;
    } else {
      grow(f64_arena, f64_cap, std::max<size_t>(f64_tot, 256));
      grow(w_arena, w_cap, std::max<size_t>(w_tot, 256));
      grow(small_arena, small_cap, std::max<size_t>(small_tot, 256));
      // padding strips must read as zero: clear everything once per layout
      FX_HIP(hipMemset(f64_arena, 0, f64_tot));
      FX_HIP(hipMemset(w_arena, 0, w_tot));
      FX_HIP(hipMemset(small_arena, 0, small_tot));
    }
    d_flags = reinterpret_cast<int*>(small_arena + flags_off);
    for (int i = 0; i < cnt; ++i) {
      WalkMem& M = wm[i];
      M.n = (int)n[i];
      M.np = (int)pad64(n[i]);
      M.in_off = off[i];
      M.f64 = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(f64_arena) + fo[i]);
      M.w = reinterpret_cast<W*>(reinterpret_cast<unsigned char*>(w_arena) + wo[i]);
      unsigned char* sp = small_arena + so[i];
      const int nt = (M.n + 31) / 32;
      M.colpart = reinterpret_cast<double*>(sp);
      sp += al((size_t)nt * M.n * sizeof(double));
      M.off = reinterpret_cast<double*>(sp);
      sp += al(M.n * sizeof(double));
      M.diag = reinterpret_cast<double2*>(sp);
      sp += al(M.n * sizeof(double2));
      M.dsq = reinterpret_cast<double2*>(sp);
      M.flags = d_flags + i;
    }
  }

  // FP64 plane q of walk i: 0 a, 1 u, 2 v, 3 p2, 4 p4, 5 p6, 6 t1, 7 t2, 8 x0.
  double* fp(int i, int q) const { return wm[i].f64 + (size_t)q * wm[i].np * wm[i].np; }
  double* dinv(int i) const { return wm[i].f64 + (size_t)9 * wm[i].np * wm[i].np; }
  // squaring plane of walk i: buffer b in {0, 1}, part 0 re / 1 im
  W* wp(int i, int b, int part) const { return wm[i].w + (size_t)(2 * b + part) * wm[i].np * wm[i].np; }

  static int split_rows(int n) { return (int)round_up(std::max(n, 1), SQ_BM); }
  static int split_kpad(int n) { return (int)round_up(std::max(n, 1), 16); }

  // Split-part buffers of the current layout (allocated on first use).
  void ensure_split() {
    if (split_ready) return;
    split_off.assign((size_t)nwalk, 0);
    size_t tot = 0;
    for (int i = 0; i < nwalk; ++i) {
      split_off[i] = tot;
      tot += ((size_t)split_rows(wm[i].n) * split_kpad(wm[i].n) * 2 * 3 * sizeof(uint16_t) + 255) / 256 * 256;
    }
    if (virt) {
      split_arena = vaddr(VR_SPLIT);
      split_cap = tot;
    } else if (tot > split_cap) {
      if (split_arena) (void)hipFree(split_arena);
      split_arena = nullptr;
      split_cap = 0;
      FX_HIP(hipMalloc(&split_arena, tot));
      split_cap = tot;
    }
    split_ready = true;
  }
  void* split_buf(int i) const { return static_cast<unsigned char*>(split_arena) + split_off[i]; }

  static int oz_rows(int n) { return (int)round_up(std::max(n, 1), OZ_BM); }
  static int oz_kp(int n) { return (int)round_up(std::max(n, 1), OZ_BK); }
  static size_t oz_buf_bytes(int n) {
    return ((size_t)oz_rows(n) * oz_kp(n) * GROUP_OZ_SLICES + 255) / 256 * 256;
  }
  static size_t oz_ex_bytes(int n) { return ((size_t)oz_rows(n) * sizeof(int) + 255) / 256 * 256; }
  // Slice buffers of the current layout (allocated on first use): per walk
  // three slice buffers, then their three exponent arrays.
  void ensure_oz() {
    if (oz_ready) return;
    oz_off.assign((size_t)nwalk, 0);
    size_t tot = 0;
    for (int i = 0; i < nwalk; ++i) {
      oz_off[i] = tot;
      tot += 3 * (oz_buf_bytes(wm[i].n) + oz_ex_bytes(wm[i].n));
    }
    if (virt) {
      oz_arena = vaddr(VR_OZ);
      oz_cap = tot;
    } else if (tot > oz_cap) {
      if (oz_arena) (void)hipFree(oz_arena);
      oz_arena = nullptr;
      oz_cap = 0;
      FX_HIP(hipMalloc(&oz_arena, tot));
      oz_cap = tot;
    }
    oz_ready = true;
  }
  int8_t* ozb(int i, int b) const {
    return static_cast<int8_t*>(oz_arena) + oz_off[i] + (size_t)b * oz_buf_bytes(wm[i].n);
  }
  int* oze(int i, int b) const {
    return reinterpret_cast<int*>(static_cast<unsigned char*>(oz_arena) + oz_off[i] +
                                  3 * oz_buf_bytes(wm[i].n) + (size_t)b * oz_ex_bytes(wm[i].n));
  }
  void oz_slice_prob(GroupStage& S, int i, const double* x, int b) {
    const int n = wm[i].n;
    OpOzSlice::Prob p{x, 1.0, nullptr, 0.0, n, wm[i].np, n, n, oz_rows(n), oz_kp(n), ozb(i, b), oze(i, b)};
    S.add(p, OpOzSlice::blocks(p));
  }
  // C = (slices ba)(slices bb)^T with the Pade epilogue (symmetric product).
  void oz_prod(GroupStage& S, int i, int ba, int bb, double* c,
               std::initializer_list<std::pair<double, const double*>> add = {}, double dg = 0.0,
               double alpha = 1.0) {
    const int n = wm[i].n;
    OzArgs<GROUP_OZ_SLICES> g;
    std::memset(&g, 0, sizeof(g));
    g.a = ozb(i, ba);
    g.ea = oze(i, ba);
    g.rows_a = oz_rows(n);
    g.b = ozb(i, bb);
    g.eb = oze(i, bb);
    g.rows_b = oz_rows(n);
    g.m = n;
    g.n = n;
    g.kpad = oz_kp(n);
    g.epi.cr = c;
    g.epi.ldc = wm[i].np;
    g.epi.ld_add = wm[i].np;
    g.epi.alpha = alpha;
    int k = 0;
    for (auto& t : add) {
      g.epi.addc[k] = t.first;
      g.epi.add_r[k] = t.second;
      ++k;
    }
    g.epi.nadd = k;
    g.epi.diag = dg;
    S.add(g, 0);
    S.sym.push_back(1);
    S.walk.push_back(i);
    S.inplace.push_back(0);
  }

  template <class P>
  void split_round(GroupStage& sp, GroupStage& sq, const GroupPlan& P_, int r, Scalars* d_scal) {
    const int i = P_.idx, n = wm[i].n, np_ = wm[i].np;
    const int cb = r & 1, nb = cb ^ 1;
    typename OpSplit<P>::Prob p{reinterpret_cast<const float*>(wp(i, cb, 0)),
                                reinterpret_cast<const float*>(wp(i, cb, 1)),
                                np_, n, split_rows(n), split_kpad(n),
                                static_cast<typename P::H*>(split_buf(i)), &d_scal[i].info};
    sp.add(p, OpSplit<P>::blocks(p));
    SqArgs<P> a;
    std::memset(&a, 0, sizeof(a));
    a.a = static_cast<const typename P::H*>(split_buf(i));
    a.rows = split_rows(n);
    a.yr = reinterpret_cast<const float*>(wp(i, cb, 0));
    a.yi = reinterpret_cast<const float*>(wp(i, cb, 1));
    a.ldy = np_;
    a.cr = reinterpret_cast<float*>(wp(i, nb, 0));
    a.ci = reinterpret_cast<float*>(wp(i, nb, 1));
    a.ldc = np_;
    a.d_in = wm[i].dsq + (size_t)cb * n;
    a.d_out = wm[i].dsq + (size_t)nb * n;
    a.m = n;
    a.k = split_kpad(n);
    a.kblock = kblock > 0 ? kblock : (1 << 30);
    a.order = nullptr;
    sq.add(a, 0);
    sq.walk.push_back(i);
  }

  void analyze(const void* d_in, Scalars* d_scal, hipStream_t st) override {
    if (nwalk == 0) return;
    if (analysis_in != d_in || analysis_scal != d_scal) {
      std::deque<GroupStage> S(3);
      S[0].kind = GK_ANALYZE;
      S[1].kind = GK_REDUCE;
      S[2].kind = GK_FINALIZE;
      for (int i = 0; i < nwalk; ++i) {
        const WalkMem& M = wm[i];
        const int nt = (M.n + 31) / 32;
        typename OpAnalyze<Tin>::Prob a{reinterpret_cast<const Tin*>(d_in) + M.in_off, M.n, nt, M.colpart, M.diag,
                                        M.flags};
        S[0].add(a, OpAnalyze<Tin>::blocks(a));
        OpReduceCols::Prob r{M.colpart, M.n, nt, M.off};
        S[1].add(r, OpReduceCols::blocks(r));
        OpFinalize::Prob f{M.off, M.diag, M.n, M.flags, d_scal + i};
        S[2].add(f, 1);
      }
      const int none[GC_COUNT] = {0, 0, 0};
      analysis.materialize(S, none, nwalk);
      analysis_in = d_in;
      analysis_scal = d_scal;
    }
    zero_ints<<<(nwalk + 255) / 256, 256, 0, st>>>(d_flags, nwalk);
    FX_LAUNCH();
    for (const GroupLaunch& L : analysis.launches) group_enqueue_one<Tin, W>(L, analysis.d_arena, st);
  }

  static GemmArgs<double> g64(const double* a, const double* b, int64_t ld, int m, int n, int k, double* c) {
    GemmArgs<double> g;
    std::memset(&g, 0, sizeof(g));
    g.ar = a;
    g.br = b;
    g.lda = ld;
    g.ldb = ld;
    g.m = m;
    g.n = n;
    g.k = k;
    g.epi.cr = c;
    g.epi.ldc = ld;
    g.epi.ld_add = ld;
    g.epi.alpha = 1.0;
    return g;
  }

  // The program of the structured pipeline for the given plans.
  // The Pade lincombs (odd and even polynomial sums) of every walk.
  void pade_lincombs(const std::vector<GroupPlan>& plans, GroupStage& lc) {
    for (const GroupPlan& P : plans) {
      const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
      double b[14];
      for (int j = 0; j <= P.m; ++j) b[j] = pade_b(P.m, j);
      const double* p2 = fp(i, 3);
      const double* p4 = fp(i, 4);
      const double* p6 = fp(i, 5);
      const double* c8 = fp(i, 7);
      auto add = [&](double* out, std::initializer_list<std::pair<double, const double*>> terms, double dg) {
        OpLincomb::Prob L;
        std::memset(&L, 0, sizeof(L));
        L.out = out;
        int k = 0;
        for (auto& t : terms) {
          L.c[k] = t.first;
          L.in[k] = t.second;
          ++k;
        }
        L.nin = k;
        L.diag = dg;
        L.n = n;
        L.ld = np_;
        lc.add(L, OpLincomb::blocks(L));
      };
      if (P.m == 13) {
        add(fp(i, 6), {{b[13], p6}, {-b[11], p4}, {b[9], p2}}, 0.0);  // Z1 in t1
        add(fp(i, 7), {{b[12], p6}, {-b[10], p4}, {b[8], p2}}, 0.0);  // Z2 in t2
      } else if (P.m == 3) {
        add(fp(i, 6), {{-b[3], p2}}, b[1]);
        add(fp(i, 2), {{-b[2], p2}}, b[0]);
      } else if (P.m == 5) {
        add(fp(i, 6), {{b[5], p4}, {-b[3], p2}}, b[1]);
        add(fp(i, 2), {{b[4], p4}, {-b[2], p2}}, b[0]);
      } else if (P.m == 7) {
        add(fp(i, 6), {{-b[7], p6}, {b[5], p4}, {-b[3], p2}}, b[1]);
        add(fp(i, 2), {{-b[6], p6}, {b[4], p4}, {-b[2], p2}}, b[0]);
      } else {  // 9
        add(fp(i, 6), {{b[9], c8}, {-b[7], p6}, {b[5], p4}, {-b[3], p2}}, b[1]);
        add(fp(i, 2), {{b[8], c8}, {-b[6], p6}, {b[4], p4}, {-b[2], p2}}, b[0]);
      }
    }
  }

  // Pade products from int8 slices (three slice buffers per walk; each
  // operand is sliced in the stage before the product that reads it):
  //   C2 = a a (0,0); C4 = C2 C2 (1,1); C6 = C4 C2 (2,1); C8 = C6 C2 (2,1);
  //   Wb = C6 Z1 + ..., V = C6 Z2 + ... (2,0), (2,1); Uq = a W (0,1).
  template <class F> void build_pade_oz(const std::vector<GroupPlan>& plans, F& stage) {
    ensure_oz();
    {
      GroupStage& z = stage(GK_OZSLICE);
      for (const GroupPlan& P : plans) oz_slice_prob(z, P.idx, fp(P.idx, 0), 0);
      GroupStage& g = stage(GK_OZGEMM, GC_SYM64);
      for (const GroupPlan& P : plans) oz_prod(g, P.idx, 0, 0, fp(P.idx, 3));  // C2
    }
    {
      GroupStage& z = stage(GK_OZSLICE);
      for (const GroupPlan& P : plans)
        if (P.m >= 5) oz_slice_prob(z, P.idx, fp(P.idx, 3), 1);
      GroupStage& g = stage(GK_OZGEMM, GC_SYM64);
      for (const GroupPlan& P : plans)
        if (P.m >= 5) oz_prod(g, P.idx, 1, 1, fp(P.idx, 4));  // C4
    }
    {
      GroupStage& z = stage(GK_OZSLICE);
      for (const GroupPlan& P : plans)
        if (P.m >= 7) oz_slice_prob(z, P.idx, fp(P.idx, 4), 2);
      GroupStage& g = stage(GK_OZGEMM, GC_SYM64);
      for (const GroupPlan& P : plans)
        if (P.m >= 7) oz_prod(g, P.idx, 2, 1, fp(P.idx, 5));  // C6
    }
    {
      GroupStage& z = stage(GK_OZSLICE);
      for (const GroupPlan& P : plans)
        if (P.m >= 9) oz_slice_prob(z, P.idx, fp(P.idx, 5), 2);  // C6 (C4's slices are no longer read)
      GroupStage& g = stage(GK_OZGEMM, GC_SYM64);
      for (const GroupPlan& P : plans)
        if (P.m == 9) oz_prod(g, P.idx, 2, 1, fp(P.idx, 7));  // C8 in t2
    }
    pade_lincombs(plans, stage(GK_LINCOMB));
    {
      GroupStage& z = stage(GK_OZSLICE);
      for (const GroupPlan& P : plans)
        if (P.m == 13) {
          oz_slice_prob(z, P.idx, fp(P.idx, 6), 0);  // Z1
          oz_slice_prob(z, P.idx, fp(P.idx, 7), 1);  // Z2
        }
      GroupStage& g = stage(GK_OZGEMM, GC_SYM64);
      for (const GroupPlan& P : plans)
        if (P.m == 13) {
          const int i = P.idx;
          double b[14];
          for (int j = 0; j <= 13; ++j) b[j] = pade_b(13, j);
          oz_prod(g, i, 2, 0, fp(i, 8), {{-b[7], fp(i, 5)}, {b[5], fp(i, 4)}, {-b[3], fp(i, 3)}}, b[1]);  // Wb
          oz_prod(g, i, 2, 1, fp(i, 2), {{-b[6], fp(i, 5)}, {b[4], fp(i, 4)}, {-b[2], fp(i, 3)}}, b[0]);  // V
        }
    }
    {
      GroupStage& z = stage(GK_OZSLICE);
      for (const GroupPlan& P : plans) {
        oz_slice_prob(z, P.idx, fp(P.idx, 0), 0);
        oz_slice_prob(z, P.idx, fp(P.idx, P.m == 13 ? 8 : 6), 1);
      }
      GroupStage& g = stage(GK_OZGEMM, GC_SYM64);
      for (const GroupPlan& P : plans) oz_prod(g, P.idx, 0, 1, fp(P.idx, 1));  // Uq
    }
  }

  // Chebyshev polynomial of exp(i X_i) in place of the Pade products and the
  // solve: X_i = the prepared plane, whose spectral radius is at most
  // theta_i = ||A_i - mu_i I||_1 / 2^s_i (the 1-norm plan's own bound), and
  // cheb_series(theta_i) gives cos X = sum_j cw[j] w^j and
  // sin X = (X / theta) sum_j sw[j] w^j, w = X^2 / theta^2.  Per walk: the
  // powers X^2, X^4, X^6, X^8 its blocks need (planes 3, 4, 5, 7), the top
  // blocks as combinations (cos in plane 1, sin in plane 6; a top block of one
  // coefficient folds the first Horner step in), Horner steps
  // T <- theta^-8 X^8 T + block as symmetric products with the block in the
  // epilogue (cos ping-pongs in planes 1 / 2, sin in 6 / 8; the rounds are
  // aligned so that every chain ends in the last one), X times the sin
  // polynomial scaled by 1 / theta, and the split form of cos X + i sin X.
  // complex64 batches take the products from int8 slices (buffer 0: X, then
  // the cos iterates; 1: X^2, then the sin iterates; 2: X^4, then X^8).
  template <class F> void build_cheb(const std::vector<GroupPlan>& plans, F& stage) {
    const bool oz = oz_slices > 0 && std::is_same<W, float>::value;
    if (oz) ensure_oz();
    struct Chain {
      const double* a = nullptr;  // coefficients in w
      int nc = 0;                 // their count
      int top = 0;                // block index of the top combination
      int cur = 0;                // plane of the current iterate
      int alt = 0;                // the other plane of the ping-pong
    };
    struct Walk {
      ChebSeries cs;
      Chain c[2];  // 0 cos, 1 sin
      double r2 = 0.0, r8 = 0.0, rinv = 0.0;
      int need = 0;  // highest power of X^2 the blocks read (0 .. 3), 4 with Horner steps
    };
    std::vector<Walk> wk(plans.size());
    int H = 0;
    for (size_t k = 0; k < plans.size(); ++k) {
      const GroupPlan& P = plans[k];
      Walk& w = wk[k];
      const double theta = P.norm * std::ldexp(1.0, -P.s);
      w.cs = cheb_series(theta, cheb_tail(P.s, std::is_same<W, float>::value));
      w.rinv = theta > 0.0 ? 1.0 / theta : 0.0;
      w.r2 = w.rinv * w.rinv;
      w.r8 = w.r2 * w.r2 * w.r2 * w.r2;
      for (int q = 0; q < 2; ++q) {
        Chain& c = w.c[q];
        c.a = q == 0 ? w.cs.cw : w.cs.sw;
        c.nc = q == 0 ? w.cs.dc / 2 + 1 : (w.cs.ds - 1) / 2 + 1;
        const int nb = (c.nc + 3) / 4;
        c.top = (nb > 1 && c.nc % 4 == 1) ? nb - 2 : nb - 1;
        c.cur = q == 0 ? 1 : 6;
        c.alt = q == 0 ? 2 : 8;
        w.need = std::max(w.need, std::min(3, c.nc - 1));
        if (nb > 1) w.need = 4;
        H = std::max(H, c.top);
      }
    }
    auto block = [&](const Walk& w, const Chain& c, int i, const double** pl, double* cf) {
      int q = 0;
      double rj = 1.0;
      for (int j = 1; j < 4 && 4 * i + j < c.nc; ++j) {
        rj *= w.r2;
        cf[q] = c.a[4 * i + j] * rj;
        pl[q++] = nullptr;  // filled by the caller: the power X^2j of the walk
      }
      return q;
    };
    auto power = [&](int i, int j) -> double* { return fp(i, j == 4 ? 7 : 2 + j); };  // X^2j
    // powers
    if (oz) {
      GroupStage& z0 = stage(GK_OZSLICE);
      for (const GroupPlan& P : plans) oz_slice_prob(z0, P.idx, fp(P.idx, 0), 0);
      GroupStage& g2 = stage(GK_OZGEMM, GC_SYM64);
      for (size_t k = 0; k < plans.size(); ++k)
        if (wk[k].need >= 1) oz_prod(g2, plans[k].idx, 0, 0, power(plans[k].idx, 1));
      GroupStage& z1 = stage(GK_OZSLICE);
      for (size_t k = 0; k < plans.size(); ++k)
        if (wk[k].need >= 2) oz_slice_prob(z1, plans[k].idx, power(plans[k].idx, 1), 1);
      GroupStage& g4 = stage(GK_OZGEMM, GC_SYM64);
      for (size_t k = 0; k < plans.size(); ++k)
        if (wk[k].need >= 2) oz_prod(g4, plans[k].idx, 1, 1, power(plans[k].idx, 2));
      GroupStage& z2 = stage(GK_OZSLICE);
      for (size_t k = 0; k < plans.size(); ++k)
        if (wk[k].need >= 3) oz_slice_prob(z2, plans[k].idx, power(plans[k].idx, 2), 2);
      GroupStage& g6 = stage(GK_OZGEMM, GC_SYM64);
      for (size_t k = 0; k < plans.size(); ++k) {
        const int i = plans[k].idx;
        if (wk[k].need >= 3) oz_prod(g6, i, 2, 1, power(i, 3));
        if (wk[k].need >= 4) oz_prod(g6, i, 2, 2, power(i, 4));
      }
    } else {
      GroupStage& g2 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      GroupStage& g4 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      GroupStage& g6 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (size_t k = 0; k < plans.size(); ++k) {
        const int i = plans[k].idx, n = wm[i].n, np_ = wm[i].np;
        if (wk[k].need >= 1) g2.add_gemm(g64(fp(i, 0), fp(i, 0), np_, n, n, np_, power(i, 1)), true, i);
        if (wk[k].need >= 2) g4.add_gemm(g64(power(i, 1), power(i, 1), np_, n, n, np_, power(i, 2)), true, i);
        if (wk[k].need >= 3) g6.add_gemm(g64(power(i, 2), power(i, 1), np_, n, n, np_, power(i, 3)), true, i);
        if (wk[k].need >= 4) g6.add_gemm(g64(power(i, 2), power(i, 2), np_, n, n, np_, power(i, 4)), true, i);
      }
    }
    // top blocks
    {
      GroupStage& lc = stage(GK_LINCOMB);
      for (size_t k = 0; k < plans.size(); ++k) {
        const int i = plans[k].idx;
        const Walk& w = wk[k];
        for (int q = 0; q < 2; ++q) {
          const Chain& c = w.c[q];
          OpLincomb::Prob L;
          std::memset(&L, 0, sizeof(L));
          const double* pl[4];
          double cf[4];
          int t = block(w, c, c.top, pl, cf);
          for (int j = 0; j < t; ++j) pl[j] = power(i, j + 1);
          if (c.top < (c.nc + 3) / 4 - 1) {  // the folded step: a[nc - 1] theta^-8 X^8
            cf[t] = c.a[c.nc - 1] * w.r8;
            pl[t++] = power(i, 4);
          }
          L.out = fp(i, c.cur);
          for (int j = 0; j < t; ++j) {
            L.in[j] = pl[j];
            L.c[j] = cf[j];
          }
          L.nin = t;
          L.diag = c.a[4 * c.top];
          L.n = wm[i].n;
          L.ld = wm[i].np;
          lc.add(L, OpLincomb::blocks(L));
        }
      }
    }
    // Horner rounds
    if (oz && H > 0) {
      GroupStage& z = stage(GK_OZSLICE);
      for (size_t k = 0; k < plans.size(); ++k)
        if (wk[k].need >= 4) oz_slice_prob(z, plans[k].idx, power(plans[k].idx, 4), 2);
    }
    for (int r = 0; r < H; ++r) {
      GroupStage* z = oz ? &stage(GK_OZSLICE) : nullptr;
      GroupStage& g = oz ? stage(GK_OZGEMM, GC_SYM64) : stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (size_t k = 0; k < plans.size(); ++k) {
        const int i = plans[k].idx, n = wm[i].n, np_ = wm[i].np;
        Walk& w = wk[k];
        for (int q = 0; q < 2; ++q) {
          Chain& c = w.c[q];
          if (r < H - c.top) continue;  // this chain's steps end with the last round
          const int bi = H - 1 - r;     // block added by this step
          const double* pl[4];
          double cf[4];
          const int t = block(w, c, bi, pl, cf);
          if (oz) {
            oz_slice_prob(*z, i, fp(i, c.cur), q);
            switch (t) {
              case 0: oz_prod(g, i, 2, q, fp(i, c.alt), {}, c.a[4 * bi], w.r8); break;
              case 1: oz_prod(g, i, 2, q, fp(i, c.alt), {{cf[0], power(i, 1)}}, c.a[4 * bi], w.r8); break;
              case 2:
                oz_prod(g, i, 2, q, fp(i, c.alt), {{cf[0], power(i, 1)}, {cf[1], power(i, 2)}}, c.a[4 * bi], w.r8);
                break;
              default:
                oz_prod(g, i, 2, q, fp(i, c.alt), {{cf[0], power(i, 1)}, {cf[1], power(i, 2)}, {cf[2], power(i, 3)}},
                        c.a[4 * bi], w.r8);
                break;
            }
          } else {
            GemmArgs<double> a = g64(power(i, 4), fp(i, c.cur), np_, n, n, np_, fp(i, c.alt));
            a.epi.alpha = w.r8;
            a.epi.nadd = t;
            for (int j = 0; j < t; ++j) {
              a.epi.addc[j] = cf[j];
              a.epi.add_r[j] = power(i, j + 1);
            }
            a.epi.diag = c.a[4 * bi];
            g.add_gemm(a, true, i);
          }
          std::swap(c.cur, c.alt);
        }
      }
    }
    // sin X = (X / theta) S, into the free sin plane
    if (oz) {
      GroupStage& z = stage(GK_OZSLICE);
      for (size_t k = 0; k < plans.size(); ++k) {
        oz_slice_prob(z, plans[k].idx, fp(plans[k].idx, 0), 0);
        oz_slice_prob(z, plans[k].idx, fp(plans[k].idx, wk[k].c[1].cur), 1);
      }
      GroupStage& g = stage(GK_OZGEMM, GC_SYM64);
      for (size_t k = 0; k < plans.size(); ++k)
        oz_prod(g, plans[k].idx, 0, 1, fp(plans[k].idx, wk[k].c[1].alt), {}, 0.0, wk[k].rinv);
    } else {
      GroupStage& g = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (size_t k = 0; k < plans.size(); ++k) {
        const int i = plans[k].idx, n = wm[i].n, np_ = wm[i].np;
        GemmArgs<double> a = g64(fp(i, 0), fp(i, wk[k].c[1].cur), np_, n, n, np_, fp(i, wk[k].c[1].alt));
        a.epi.alpha = wk[k].rinv;
        g.add_gemm(a, true, i);
      }
    }
    GroupStage& so = stage(GK_SOLUTION);
    for (size_t k = 0; k < plans.size(); ++k) {
      const int i = plans[k].idx, np_ = wm[i].np;
      typename OpSolution<W>::Prob p{fp(i, wk[k].c[0].cur), fp(i, wk[k].c[1].alt), np_, wp(i, 0, 0), wp(i, 0, 1),
                                     np_, wm[i].n, wm[i].dsq};
      so.add(p, OpSolution<W>::blocks(p));
    }
  }

  // Stages are held in a deque: references returned by stage() stay valid
  // while later stages are appended.
  void build(const std::vector<GroupPlan>& plans, const void* d_in, void* d_out, Scalars* d_scal,
             const double2* d_psi0, double2* d_psi, const std::vector<size_t>& voff) {
    stages.clear();
    auto stage = [&](int kind, int cls = -1, int mode = 0, bool f32 = false) -> GroupStage& {
      stages.emplace_back();
      GroupStage& S = stages.back();
      S.kind = kind;
      S.cls = cls;
      S.mode = mode;
      S.f32 = f32;
      return S;
    };
    const int G = (int)plans.size();
    // 1. scaled, shifted G_i = (Im A_i - Im(mu_i) I) / 2^s_i
    {
      GroupStage& S = stage(GK_PREPARE);
      for (const GroupPlan& P : plans) {
        const int i = P.idx;
        typename OpPrepare<Tin>::Prob p{reinterpret_cast<const Tin*>(d_in) + wm[i].in_off, wm[i].n, wm[i].np,
                                        fp(i, 0), P.shifted ? P.mu_im : 0.0, std::ldexp(1.0, -P.s)};
        S.add(p, OpPrepare<Tin>::blocks(p));
      }
    }
    if (cheb) build_cheb(plans, stage);
    else {
    // 2. Pade products (real symmetric, half tiles); A^{2k} = (-1)^k G^{2k}
    const bool oz = oz_slices > 0 && std::is_same<W, float>::value;
    if (oz) build_pade_oz(plans, stage);
    else {
      GroupStage& s1 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans) {
        const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
        s1.add_gemm(g64(fp(i, 0), fp(i, 0), np_, n, n, np_, fp(i, 3)), true, i);  // C2
      }
      GroupStage& s2 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans)
        if (P.m >= 5) {
          const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
          s2.add_gemm(g64(fp(i, 3), fp(i, 3), np_, n, n, np_, fp(i, 4)), true, i);  // C4
        }
      GroupStage& s3 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans)
        if (P.m >= 7) {
          const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
          s3.add_gemm(g64(fp(i, 4), fp(i, 3), np_, n, n, np_, fp(i, 5)), true, i);  // C6
        }
      GroupStage& s4 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans)
        if (P.m == 9) {
          const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
          s4.add_gemm(g64(fp(i, 5), fp(i, 3), np_, n, n, np_, fp(i, 7)), true, i);  // C8 in t2
        }
      GroupStage& lc = stage(GK_LINCOMB);
      for (const GroupPlan& P : plans) {
        const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
        double b[14];
        for (int j = 0; j <= P.m; ++j) b[j] = pade_b(P.m, j);
        const double* p2 = fp(i, 3);
        const double* p4 = fp(i, 4);
        const double* p6 = fp(i, 5);
        const double* c8 = fp(i, 7);
        auto add = [&](double* out, std::initializer_list<std::pair<double, const double*>> terms, double dg) {
          OpLincomb::Prob L;
          std::memset(&L, 0, sizeof(L));
          L.out = out;
          int k = 0;
          for (auto& t : terms) {
            L.c[k] = t.first;
            L.in[k] = t.second;
            ++k;
          }
          L.nin = k;
          L.diag = dg;
          L.n = n;
          L.ld = np_;
          lc.add(L, OpLincomb::blocks(L));
        };
        if (P.m == 13) {
          add(fp(i, 6), {{b[13], p6}, {-b[11], p4}, {b[9], p2}}, 0.0);  // Z1 in t1
          add(fp(i, 7), {{b[12], p6}, {-b[10], p4}, {b[8], p2}}, 0.0);  // Z2 in t2
        } else if (P.m == 3) {
          add(fp(i, 6), {{-b[3], p2}}, b[1]);
          add(fp(i, 2), {{-b[2], p2}}, b[0]);
        } else if (P.m == 5) {
          add(fp(i, 6), {{b[5], p4}, {-b[3], p2}}, b[1]);
          add(fp(i, 2), {{b[4], p4}, {-b[2], p2}}, b[0]);
        } else if (P.m == 7) {
          add(fp(i, 6), {{-b[7], p6}, {b[5], p4}, {-b[3], p2}}, b[1]);
          add(fp(i, 2), {{-b[6], p6}, {b[4], p4}, {-b[2], p2}}, b[0]);
        } else {  // 9
          add(fp(i, 6), {{b[9], c8}, {-b[7], p6}, {b[5], p4}, {-b[3], p2}}, b[1]);
          add(fp(i, 2), {{b[8], c8}, {-b[6], p6}, {b[4], p4}, {-b[2], p2}}, b[0]);
        }
      }
      // m = 13: Wb = C6 Z1 - b7 C6 + b5 C4 - b3 C2 + b1 I (x0); V = C6 Z2 - b6 C6 + b4 C4 - b2 C2 + b0 I
      GroupStage& s5 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans)
        if (P.m == 13) {
          const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
          double b[14];
          for (int j = 0; j <= 13; ++j) b[j] = pade_b(13, j);
          GemmArgs<double> g = g64(fp(i, 5), fp(i, 6), np_, n, n, np_, fp(i, 8));
          g.epi.nadd = 3;
          g.epi.addc[0] = -b[7];
          g.epi.add_r[0] = fp(i, 5);
          g.epi.addc[1] = b[5];
          g.epi.add_r[1] = fp(i, 4);
          g.epi.addc[2] = -b[3];
          g.epi.add_r[2] = fp(i, 3);
          g.epi.diag = b[1];
          s5.add_gemm(g, true, i);
          GemmArgs<double> h = g64(fp(i, 5), fp(i, 7), np_, n, n, np_, fp(i, 2));
          h.epi.nadd = 3;
          h.epi.addc[0] = -b[6];
          h.epi.add_r[0] = fp(i, 5);
          h.epi.addc[1] = b[4];
          h.epi.add_r[1] = fp(i, 4);
          h.epi.addc[2] = -b[2];
          h.epi.add_r[2] = fp(i, 3);
          h.epi.diag = b[0];
          s5.add_gemm(h, true, i);
        }
      // Uq = G * (Wb or odd polynomial)
      GroupStage& s6 = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans) {
        const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
        s6.add_gemm(g64(fp(i, 0), fp(i, P.m == 13 ? 8 : 6), np_, n, n, np_, fp(i, 1)), true, i);
      }
    }
    // 3. Cholesky solve: S = V V^T + Uq Uq^T = L L^T, Z = [V; Uq] L^{-T}, X = Z Z^T
    {
      GroupStage& pk = stage(GK_PACK);
      for (const GroupPlan& P : plans) {
        const int i = P.idx;
        OpPack::Prob p{fp(i, 2), fp(i, 1), wm[i].n, wm[i].np, wm[i].np, fp(i, 3)};
        pk.add(p, OpPack::blocks(p));
      }
      GroupStage& vv = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans) {
        const int i = P.idx, np_ = wm[i].np, kpad = (int)round_up(wm[i].n, 16);
        const size_t Q = (size_t)np_ * np_;
        double* zb = fp(i, 3);
        vv.add_gemm(g64(zb, zb, np_, np_, np_, kpad, zb + 3 * Q), true, i);
      }
      GroupStage& sm = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans) {
        const int i = P.idx, np_ = wm[i].np, kpad = (int)round_up(wm[i].n, 16);
        const size_t Q = (size_t)np_ * np_;
        double* zb = fp(i, 3);
        GemmArgs<double> g = g64(zb + Q, zb + Q, np_, np_, np_, kpad, zb + 2 * Q);
        g.epi.nadd = 1;
        g.epi.addc[0] = 1.0;
        g.epi.add_r[0] = zb + 3 * Q;
        sm.add_gemm(g, true, i);
      }
      int max_np = 0;
      for (const GroupPlan& P : plans) max_np = std::max(max_np, wm[P.idx].np);
      for (int k0 = 0; k0 < max_np; k0 += LU_NB) {
        GroupStage& ch = stage(GK_CHOL);
        GroupStage& pn = stage(GK_GEMM, GC_SOLVE, MODE_REAL);
        GroupStage& up = stage(GK_GEMM, GC_SOLVE, MODE_REAL);
        for (const GroupPlan& P : plans) {
          const int i = P.idx, n = wm[i].n, ip = wm[i].np;
          if (k0 >= ip) continue;
          const size_t Q = (size_t)ip * ip;
          const int64_t L = ip;
          double* zb = fp(i, 3);
          double* s = zb + 2 * Q;
          double* dk = dinv(i) + (size_t)(k0 / LU_NB) * LU_NB * LU_NB;
          const int K0 = k0 / CH_OUTER * CH_OUTER;
          const int K1 = std::min(ip, K0 + CH_OUTER);
          const int k1 = k0 + LU_NB;
          const int rest = ip - k1;
          OpChol::Prob c{s, L, k0, n, dk, d_scal + i};
          ch.add(c, 1);
          double* panel = s + (int64_t)k1 * L + k0;  // L[k1:, k0:k1]
          if (rest > 0) {
            GemmArgs<double> g = g64(panel, dk, L, rest, LU_NB, LU_NB, panel);
            g.ldb = LU_NB;
            g.epi.alpha = 1.0;
            pn.add_gemm(g, false, i, true);
          }
          {
            GemmArgs<double> g = g64(zb + k0, dk, L, 2 * ip, LU_NB, LU_NB, zb + k0);
            g.ldb = LU_NB;
            pn.add_gemm(g, false, i, true);
          }
          if (k1 < K1) {
            GemmArgs<double> g = g64(panel, panel, L, rest, K1 - k1, LU_NB, s + (int64_t)k1 * L + k1);
            g.epi.alpha = -1.0;
            g.epi.beta = 1.0;
            up.add_gemm(g, false, i);
            GemmArgs<double> h = g64(zb + k0, panel, L, 2 * ip, K1 - k1, LU_NB, zb + k1);
            h.epi.alpha = -1.0;
            h.epi.beta = 1.0;
            up.add_gemm(h, false, i);
          } else if (K1 < ip) {
            const int r2 = ip - K1, kk = K1 - K0;
            double* p2 = s + (int64_t)K1 * L + K0;  // L[K1:, K0:K1]
            GemmArgs<double> g = g64(p2, p2, L, r2, r2, kk, s + (int64_t)K1 * L + K1);
            g.epi.alpha = -1.0;
            g.epi.beta = 1.0;
            up.add_gemm(g, true, i);
            GemmArgs<double> h = g64(zb + K0, p2, L, 2 * ip, r2, kk, zb + K1);
            h.epi.alpha = -1.0;
            h.epi.beta = 1.0;
            up.add_gemm(h, false, i);
          }
        }
      }
      // X = Z Z^T as Re X = 2 Zr Zr^T - I, Im X = 2 Zi Zr^T (Zr Zr^T + Zi Zi^T
      // = I because L L^T = V^2 + Uq^2): two real symmetric products.
      GroupStage& xs = stage(GK_GEMM, GC_SYM64, MODE_REAL);
      for (const GroupPlan& P : plans) {
        const int i = P.idx, n = wm[i].n, np_ = wm[i].np, kpad = (int)round_up(n, 16);
        const size_t Q = (size_t)np_ * np_;
        double* zb = fp(i, 3);
        GemmArgs<double> g = g64(zb, zb, np_, n, n, kpad, zb + 2 * Q);
        g.epi.alpha = 2.0;
        g.epi.diag = -1.0;
        xs.add_gemm(g, true, i);
        GemmArgs<double> h = g64(zb + Q, zb, np_, n, n, kpad, zb + 3 * Q);
        h.epi.alpha = 2.0;
        xs.add_gemm(h, true, i);
      }
      GroupStage& so = stage(GK_SOLUTION);
      for (const GroupPlan& P : plans) {
        const int i = P.idx, np_ = wm[i].np;
        const size_t Q = (size_t)np_ * np_;
        double* zb = fp(i, 3);
        typename OpSolution<W>::Prob p{zb + 2 * Q, zb + 3 * Q, np_, wp(i, 0, 0), wp(i, 0, 1), np_, wm[i].n,
                                       wm[i].dsq};
        so.add(p, OpSolution<W>::blocks(p));
      }
    }
    }  // Pade route
    // 4. squaring rounds: round r squares every walk with s_i > r
    int smax = 0;
    for (const GroupPlan& P : plans) smax = std::max(smax, P.s);
    const bool f32 = std::is_same<W, float>::value;
    const int split = f32 ? precision : FEXPM_PREC_NATIVE;
    if (split != FEXPM_PREC_NATIVE) ensure_split();
    for (int r = 0; r < smax && split != FEXPM_PREC_NATIVE; ++r) {
      const bool h = split == FEXPM_PREC_FP16X2;
      GroupStage& sp = stage(h ? GK_SPLIT_F16 : GK_SPLIT_BF16);
      GroupStage& sq = stage(h ? GK_SQ_F16 : GK_SQ_BF16, GC_SQUARE);
      for (const GroupPlan& P : plans)
        if (P.s > r) {
          if (h) split_round<Fp16x2>(sp, sq, P, r, d_scal);
          else split_round<Bf16x3>(sp, sq, P, r, d_scal);
        }
    }
    for (int r = 0; r < smax && split == FEXPM_PREC_NATIVE; ++r) {
      GroupStage& sq = stage(GK_GEMM, GC_SQUARE, MODE_C3M, f32);
      for (const GroupPlan& P : plans) {
        if (P.s <= r) continue;
        const int i = P.idx, n = wm[i].n, np_ = wm[i].np;
        const int cb = r & 1, nb = cb ^ 1;
        GemmArgs<W> g;
        std::memset(&g, 0, sizeof(g));
        g.ar = wp(i, cb, 0);
        g.ai = wp(i, cb, 1);
        g.lda = np_;
        g.br = wp(i, cb, 0);
        g.bi = wp(i, cb, 1);
        g.ldb = np_;
        g.m = n;
        g.n = n;
        g.k = (int)round_up(n, 16);  // the padding columns are zero
        g.kblock = kblock > 0 ? kblock : (1 << 30);
        g.epi.cr = wp(i, nb, 0);
        g.epi.ci = wp(i, nb, 1);
        g.epi.ldc = np_;
        g.epi.ld_add = np_;
        g.epi.alpha = 1.0;
        g.epi.d_in = wm[i].dsq + (size_t)cb * wm[i].n;
        g.epi.d_out = wm[i].dsq + (size_t)nb * wm[i].n;
        sq.add_gemm(g, true, i);
      }
    }
    // 5. U_i = e^mu_i X_i and psi_i = U_i psi0_i
    {
      GroupStage& out = stage(GK_OUTPUT);
      GroupStage& ap = stage(GK_APPLY);
      GroupStage& rn = stage(GK_RENORM);  // every grouped walk is a unitary CTQW propagator
      for (const GroupPlan& P : plans) {
        const int i = P.idx, b = P.s & 1;
        const double er = std::exp(P.shifted ? P.mu_re : 0.0);
        const double th = P.shifted ? P.mu_im : 0.0;
        const double phr = er * std::cos(th), phi = er * std::sin(th);
        const double2* d = wm[i].dsq + (size_t)b * wm[i].n;
        typename OpOutput<W, Tin>::Prob o{wp(i, b, 0), wp(i, b, 1), wm[i].np, wm[i].n, d, phr, phi,
                                          reinterpret_cast<Tin*>(d_out) + wm[i].in_off};
        out.add(o, OpOutput<W, Tin>::blocks(o));
        if (d_psi0) {
          typename OpApply<W>::Prob a{wp(i, b, 0), wp(i, b, 1), wm[i].np, wm[i].n, d, d_psi0 + voff[i], phr, phi,
                                      d_psi + voff[i]};
          ap.add(a, OpApply<W>::blocks(a));
          OpRenorm::Prob r{d_psi0 + voff[i], d_psi + voff[i], wm[i].n};
          rn.add(r, OpRenorm::blocks(r));
        }
      }
    }
    (void)G;
  }

  // Text description of the program (testing ABI): region sizes, then every
  // stage and problem.
  std::string describe() const {
    std::string o = "REGIONS f64=" + std::to_string(f64_cap) + " w=" + std::to_string(w_cap) +
                    " small=" + std::to_string(small_cap) + " split=" + std::to_string(split_cap) +
                    " oz=" + std::to_string(oz_cap) +
                    " wsize=" + std::to_string(sizeof(W)) + " insize=" + std::to_string(sizeof(Tin)) + "\n";
    for (const GroupStage& S : stages) o += describe_stage<Tin, W>(S);
    return o;
  }

  void enqueue(hipStream_t st) {
    for (const GroupLaunch& L : prog.launches) group_enqueue_one<Tin, W>(L, prog.d_arena, st);
  }

  // One eager run with every GEMM launch timed; adds per-class device time.
  void timed_run(hipStream_t st, double* cls_ms) {
    size_t ng = 0;
    for (const GroupLaunch& L : prog.launches) ng += gk_tiled(L.kind);
    while (evpool.size() < 2 * ng) {
      hipEvent_t e;
      FX_HIP(hipEventCreate(&e));
      evpool.push_back(e);
    }
    size_t q = 0;
    std::vector<int> cls;
    for (const GroupLaunch& L : prog.launches) {
      if (gk_tiled(L.kind)) FX_HIP(hipEventRecord(evpool[2 * q], st));
      group_enqueue_one<Tin, W>(L, prog.d_arena, st);
      if (gk_tiled(L.kind)) {
        FX_HIP(hipEventRecord(evpool[2 * q + 1], st));
        cls.push_back(L.cls);
        ++q;
      }
    }
    FX_HIP(hipStreamSynchronize(st));
    for (size_t k = 0; k < q; ++k) {
      float ms = 0.f;
      FX_HIP(hipEventElapsedTime(&ms, evpool[2 * k], evpool[2 * k + 1]));
      cls_ms[cls[k]] += ms;
    }
  }

  void tune(hipStream_t st, int nw) {
    const auto t0 = std::chrono::steady_clock::now();
    double best[GC_COUNT];
    int arg[GC_COUNT];
    for (int c = 0; c < GC_COUNT; ++c) {
      best[c] = 1e300;
      arg[c] = 0;
    }
    for (int k = 0; k < GROUP_CANDIDATES; ++k) {
      const int ck[GC_COUNT] = {k, k, k};
      prog.materialize(stages, ck, nw);
      double ms[GC_COUNT] = {0, 0, 0};
      timed_run(st, ms);
      for (int c = 0; c < GC_COUNT; ++c)
        if (ms[c] < best[c]) {
          best[c] = ms[c];
          arg[c] = k;
        }
    }
    for (int c = 0; c < GC_COUNT; ++c) {
      cand[c] = arg[c];
      class_ms[c] = best[c];
    }
    tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  void run(const std::vector<GroupPlan>& plans, const void* d_in, void* d_out, Scalars* d_scal,
           const double2* d_psi0, double2* d_psi, const std::vector<size_t>& voff, hipStream_t st) override {
    graph_built = 0;
    tuned = 0;
    graph_used = 0;
    if (plans.empty()) {
      launches = 0;
      device_ms = 0.0;
      return;
    }
    // key: everything the program depends on
    std::vector<double> key;
    key.reserve(plans.size() * 6 + 1);
    key.push_back((double)kblock);
    key.push_back((double)precision);
    key.push_back((double)oz_slices);
    key.push_back((double)cheb);
    for (const GroupPlan& P : plans) {
      key.push_back(P.norm);
      key.push_back(P.idx);
      key.push_back(P.m);
      key.push_back(P.s);
      key.push_back(P.shifted);
      key.push_back(P.mu_re);
      key.push_back(P.mu_im);
    }
    const bool same = key == plan_key && key_in == d_in && key_out == d_out && key_psi0 == (const void*)d_psi0 &&
                      !stages.empty();
    if (!same) {
      drop_graph();
      build(plans, d_in, d_out, d_scal, d_psi0, d_psi, voff);
      plan_key = key;
      key_in = d_in;
      key_out = d_out;
      key_psi0 = d_psi0;
      const int pi = std::min(std::max(precision, 0), 2);
      if (forced >= 0) {
        for (int c = 0; c < GC_COUNT; ++c) cand[c] = forced;
      } else if (autotune && !tuned_prec[pi]) {
        tune(st, nwalk);
        tuned = 1;
        tuned_prec[pi] = true;
        for (int c = 0; c < GC_COUNT; ++c) cand_prec[pi][c] = cand[c];
      } else if (tuned_prec[pi]) {
        for (int c = 0; c < GC_COUNT; ++c) cand[c] = cand_prec[pi][c];
      }
      prog.materialize(stages, cand, nwalk);
      walk_flops = prog.walk_flops;
    }
    launches = (int)prog.launches.size();
    ensure_events();
    FX_HIP(hipEventRecord(ev0, st));
    if (use_graph) {
      if (!gexec) {
        FX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
          enqueue(st);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        FX_HIP(hipStreamEndCapture(st, &graph));
        FX_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        graph_built = 1;
        // the capture recorded nothing on the stream: ev0 is re-recorded
        FX_HIP(hipEventRecord(ev0, st));
      }
      FX_HIP(hipGraphLaunch(gexec, st));
      graph_used = 1;
    } else {
      enqueue(st);
    }
    FX_HIP(hipEventRecord(ev1, st));
    FX_HIP(hipEventSynchronize(ev1));
    float ms = 0.f;
    FX_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    device_ms = ms;
  }
};

}  // namespace fexpm
