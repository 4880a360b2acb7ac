// pybind11 module `_fastexpm_core`: the same four classes, constructor,
// method and keyword names as the reference binding (src/main.cu:8-64):
//   ExpMatFloat32 / ExpMatFloat64 / ExpMatComplex64 / ExpMatComplex128
//   __init__(n), run(arr_a) -> ndarray, free()
// Every call goes through the C ABI declared in include/fastexpm.h; no HIP
// or device type appears here.  The GIL is released while the device works
// (the engine serialises calls on one instance itself).
//
// Extra methods beyond the reference (used by Operator/Ctqw and the tests):
//   upload(arr_a), run_resident(), download(), run_device(ptr_a, ptr_u),
//   apply(psi0) -> psi (U of the last run times states, FP64 accumulation),
//   stats() -> dict, set_option(key, value), n (property).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fastexpm.h"
#include "../../include/fastexpm_testing.h"

namespace py = pybind11;

static void raise_status(int rc, const fexpm_engine* e) {
  if (rc == FEXPM_OK) return;
  std::string msg = std::string("fastexpm: ") + fexpm_status_string(rc);
  if (e) {
    const char* d = fexpm_last_error(e);
    if (d && *d) msg += std::string(": ") + d;
  }
  switch (rc) {
    case FEXPM_ENOMEM:
      throw std::bad_alloc();
    case FEXPM_EINVAL:
      throw std::invalid_argument(msg);
    default:
      throw std::runtime_error(msg);
  }
}

static py::dict stats_dict(const fexpm_stats& s) {
  py::dict d;
  d["n"] = s.n;
  d["dtype"] = s.dtype;
  d["m"] = s.m;
  d["s"] = s.s;
  d["structured"] = (bool)s.structured;
  d["shifted"] = (bool)s.shifted;
  d["nonfinite"] = (bool)s.nonfinite;
  d["mu"] = std::complex<double>(s.mu_re, s.mu_im);
  d["norm"] = s.norm;
  d["norm_plain"] = s.norm_plain;
  d["ms_total"] = s.ms_total;
  d["ms_analyze"] = s.ms_analyze;
  d["ms_pade"] = s.ms_pade;
  d["ms_solve"] = s.ms_solve;
  d["ms_square"] = s.ms_square;
  d["ms_output"] = s.ms_output;
  d["flops_issued"] = s.flops_issued;
  d["s_fp64"] = s.s_fp64;
  d["kblock"] = s.kblock;
  d["pade_fp64"] = (bool)s.pade_fp64;
  d["precision"] = s.precision == FEXPM_PREC_FP16X2    ? "fp16x2"
                   : s.precision == FEXPM_PREC_BF16X3  ? "bf16x3"
                   : s.precision == FEXPM_PREC_INT8_CRT ? "int8crt"
                                                        : "native";
  d["fp64_slices"] = s.fp64_slices;
  d["fp64_slices_scope"] = s.fp64_slices_scope;
  d["spectral"] = s.spectral;
  d["deflated"] = s.deflated;
  d["s_onenorm"] = s.s_onenorm;
  d["rho_bound"] = s.rho_bound;
  d["s_two_product"] = s.s_two_product;
  d["chebyshev"] = s.chebyshev;
  d["device_bytes"] = s.device_bytes;
  d["ms_square_kernel"] = s.ms_square_kernel;
  d["square_kernel_launches"] = s.square_kernel_launches;
  d["crt_moduli"] = s.crt_moduli;
  d["crt_beta"] = s.crt_beta;
  d["fp64_moduli"] = s.fp64_moduli;
  d["ms_propagate"] = s.ms_propagate;
  return d;
}

template <typename T, int DT> class ExpMat {
 public:
  explicit ExpMat(int64_t n) : n_(n) {
    if (n < 0) throw std::invalid_argument("fastexpm: n must be non-negative");
    fexpm_engine* e = nullptr;
    int rc = fexpm_create(n, DT, &e);
    raise_status(rc, nullptr);
    e_ = e;
  }
  ~ExpMat() {
    if (e_) fexpm_destroy(e_);
  }
  ExpMat(const ExpMat&) = delete;
  ExpMat& operator=(const ExpMat&) = delete;

  py::array_t<T> run(py::array_t<T, py::array::c_style | py::array::forcecast> a) {
    check_shape(a);
    py::array_t<T> out({n_, n_});
    const T* src = a.data();
    T* dst = out.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_run(e_, src, dst);
    }
    raise_status(rc, e_);
    return out;
  }

  void upload(py::array_t<T, py::array::c_style | py::array::forcecast> a) {
    check_shape(a);
    const T* src = a.data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_upload(e_, src);
    }
    raise_status(rc, e_);
  }

  void run_resident() {
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_run_resident(e_);
    }
    raise_status(rc, e_);
  }

  void set_graph(py::array adj, double gamma, bool laplacian) {
    if (adj.ndim() != 2 || adj.shape(0) != n_ || adj.shape(1) != n_)
      throw std::invalid_argument("fastexpm.set_graph: adjacency must have shape (n, n)");
    int adt;
    py::array a;
    if (adj.dtype().is(py::dtype::of<float>())) {
      a = py::array_t<float, py::array::c_style | py::array::forcecast>(adj);
      adt = FEXPM_FLOAT32;
    } else {
      a = py::array_t<double, py::array::c_style | py::array::forcecast>(adj);
      adt = FEXPM_FLOAT64;
    }
    const void* src = a.data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_set_graph(e_, src, adt, gamma, laplacian ? 1 : 0);
    }
    raise_status(rc, e_);
  }

  void set_graph_edges(py::array_t<int64_t, py::array::c_style | py::array::forcecast> src,
                       py::array_t<int64_t, py::array::c_style | py::array::forcecast> dst, py::object w,
                       double gamma, bool laplacian) {
    if (src.ndim() != 1 || dst.ndim() != 1 || src.shape(0) != dst.shape(0))
      throw std::invalid_argument("fastexpm.set_graph_edges: src and dst must be 1-D of equal length");
    const int64_t m = src.shape(0);
    py::array_t<double, py::array::c_style | py::array::forcecast> wa;
    const double* wp = nullptr;
    if (!w.is_none()) {
      wa = py::array_t<double, py::array::c_style | py::array::forcecast>(w);
      if (wa.ndim() != 1 || wa.shape(0) != m)
        throw std::invalid_argument("fastexpm.set_graph_edges: one weight per edge expected");
      wp = wa.data();
    }
    const int64_t* ps = src.data();
    const int64_t* pd = dst.data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_set_graph_edges(e_, m, ps, pd, wp, gamma, laplacian ? 1 : 0);
    }
    raise_status(rc, e_);
  }

  void set_graph_er(double p, uint64_t seed, bool weighted, double gamma, bool laplacian) {
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_set_graph_er(e_, p, seed, weighted ? 1 : 0, gamma, laplacian ? 1 : 0);
    }
    raise_status(rc, e_);
  }

  py::array_t<double> debug_hamiltonian() {
    py::array_t<double> out({n_, n_});
    double* p = out.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_debug_hamiltonian(e_, p);
    }
    raise_status(rc, e_);
    return out;
  }

  void run_graph(double t) {
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_run_graph(e_, t);
    }
    raise_status(rc, e_);
  }

  py::array_t<T> download() {
    py::array_t<T> out({n_, n_});
    T* dst = out.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_download(e_, dst);
    }
    raise_status(rc, e_);
    return out;
  }

  void run_device(uintptr_t d_a, uintptr_t d_u, uintptr_t stream) {
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_run_device(e_, reinterpret_cast<const void*>(d_a), reinterpret_cast<void*>(d_u),
                            reinterpret_cast<void*>(stream));
    }
    raise_status(rc, e_);
  }

  py::array_t<std::complex<double>> apply(
      py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> psi0) {
    int64_t ns;
    if (psi0.ndim() == 1) {
      if (psi0.shape(0) != n_) throw std::invalid_argument("fastexpm.apply: state length must be n");
      ns = 1;
    } else if (psi0.ndim() == 2) {
      if (psi0.shape(1) != n_) throw std::invalid_argument("fastexpm.apply: states must have shape (k, n)");
      ns = psi0.shape(0);
    } else {
      throw std::invalid_argument("fastexpm.apply: expected a 1-D or 2-D array");
    }
    std::vector<ssize_t> shape(psi0.shape(), psi0.shape() + psi0.ndim());
    py::array_t<std::complex<double>> out(shape);
    const void* src = psi0.data();
    void* dst = out.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_apply(e_, src, ns, dst);
    }
    raise_status(rc, e_);
    return out;
  }

  py::array_t<std::complex<double>> propagate(
      py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> psi0, int64_t nsteps) {
    if (psi0.ndim() != 2 || psi0.shape(1) != n_)
      throw std::invalid_argument("fastexpm.propagate: states must have shape (k, n)");
    if (nsteps < 0) throw std::invalid_argument("fastexpm.propagate: nsteps must be >= 0");
    const int64_t ns = psi0.shape(0);
    py::array_t<std::complex<double>> out({nsteps, ns, n_});
    const void* src = psi0.data();
    void* dst = out.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_propagate(e_, src, ns, nsteps, dst);
    }
    raise_status(rc, e_);
    return out;
  }

  void free() {
    int rc = fexpm_free(e_);
    raise_status(rc, e_);
  }

  void set_option(int key, int64_t value) { raise_status(fexpm_set_option(e_, key, value), e_); }

  py::dict stats() const {
    fexpm_stats s;
    raise_status(fexpm_get_stats(e_, &s), e_);
    return stats_dict(s);
  }

  int64_t n() const { return n_; }

  // Internal working plane after a run (testing ABI, include/fastexpm_testing.h).
  py::array_t<std::complex<double>> debug_plane(int which) {
    py::array_t<double> re({n_, n_}), im({n_, n_});
    raise_status(fexpm_debug_plane(e_, which, re.mutable_data(), im.mutable_data()), e_);
    py::array_t<std::complex<double>> out({n_, n_});
    auto o = out.mutable_unchecked<2>();
    auto r = re.unchecked<2>();
    auto i = im.unchecked<2>();
    for (ssize_t a = 0; a < n_; ++a)
      for (ssize_t b = 0; b < n_; ++b) o(a, b) = std::complex<double>(r(a, b), i(a, b));
    return out;
  }

 private:
  template <typename A> void check_shape(const A& a) const {
    if (a.ndim() != 2 || a.shape(0) != n_ || a.shape(1) != n_)
      throw std::invalid_argument("fastexpm: input matrix must have shape (n, n)");
  }
  int64_t n_;
  fexpm_engine* e_ = nullptr;
};

// Batch of independent expms of uneven order (C ABI fexpm_batch_*).
class ExpMatBatch {
 public:
  ExpMatBatch(const std::string& dtype, int64_t nmax, int lanes) {
    if (dtype == "complex64") dt_ = FEXPM_COMPLEX64, es_ = 8;
    else if (dtype == "complex128") dt_ = FEXPM_COMPLEX128, es_ = 16;
    else throw std::invalid_argument("ExpMatBatch: dtype must be complex64 or complex128");
    fexpm_batch* b = nullptr;
    int rc = fexpm_batch_create(dt_, nmax, lanes, &b);
    if (rc != FEXPM_OK)
      throw std::runtime_error(std::string("fastexpm batch: ") + fexpm_status_string(rc) + ": " +
                               fexpm_batch_last_error(nullptr));
    b_ = b;
  }
  ~ExpMatBatch() {
    if (b_) fexpm_batch_destroy(b_);
  }
  void check(int rc) {
    if (rc != FEXPM_OK)
      throw std::runtime_error(std::string("fastexpm batch: ") + fexpm_status_string(rc) + ": " +
                               fexpm_batch_last_error(b_));
  }
  void set_option(int key, int64_t value) { check(fexpm_batch_set_option(b_, key, value)); }
  int64_t count() const { return fexpm_batch_count(b_); }

  // mats: list of square complex arrays (cast to the batch dtype), packed
  // into one host buffer (allocated once, filled by several threads)
  void upload(py::list mats) {
    std::vector<int64_t> ns;
    std::vector<py::array> cs;
    std::vector<size_t> offs;
    size_t total = 0;
    for (auto h : mats) {
      py::array a = py::array::ensure(h, py::array::c_style | py::array::forcecast);
      if (a.ndim() != 2 || a.shape(0) != a.shape(1)) throw std::invalid_argument("batch: square matrices expected");
      py::array c = dt_ == FEXPM_COMPLEX64
                        ? (py::array)py::array_t<std::complex<float>, py::array::c_style | py::array::forcecast>(a)
                        : (py::array)py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>(a);
      offs.push_back(total);
      total += (size_t)c.nbytes();
      ns.push_back(a.shape(0));
      cs.push_back(c);
    }
    std::unique_ptr<char[]> host(new char[std::max<size_t>(total, 1)]);
    std::vector<const void*> src;
    std::vector<size_t> len;
    for (const py::array& c : cs) {
      src.push_back(c.data());
      len.push_back((size_t)c.nbytes());
    }
    int rc;
    {
      py::gil_scoped_release nogil;
      // matrix i goes to thread i % nth (the arrays stay alive in cs)
      const int nth = total >= ((size_t)64 << 20) ? 8 : 1;
      auto work = [&](int t) {
        for (size_t i = (size_t)t; i < src.size(); i += (size_t)nth) std::memcpy(host.get() + offs[i], src[i], len[i]);
      };
      std::vector<std::thread> th;
      int t0 = 0;
      try {
        for (int t = 1; t < nth; ++t) {
          th.emplace_back(work, t);
          t0 = t;
        }
      } catch (...) {
      }
      for (int t = t0 + 1; t < nth; ++t) work(t);  // threads that could not be started
      work(0);
      for (auto& t : th) t.join();
      rc = fexpm_batch_upload(b_, (int64_t)ns.size(), ns.data(), host.get());
    }
    // the sizes change only with the engine's: a rejected upload leaves the
    // previous batch (or, after a failed allocation, an empty one)
    if (rc == FEXPM_OK) n_ = ns;
    else if (fexpm_batch_count(b_) == 0) n_.clear();
    check(rc);
  }
  // Runs the uploaded batch; with psi0 (list of vectors) returns the list of psi.
  py::object run_resident(py::object psi0) {
    if (psi0.is_none()) {
      int rc;
      {
        py::gil_scoped_release nogil;
        rc = fexpm_batch_run_resident(b_, nullptr, nullptr);
      }
      check(rc);
      return py::none();
    }
    py::list lst = psi0;
    if ((size_t)py::len(lst) != n_.size()) throw std::invalid_argument("batch: one state per matrix expected");
    size_t tot = 0;
    for (int64_t v : n_) tot += (size_t)v;
    std::vector<std::complex<double>> in(tot), out(tot);
    size_t off = 0;
    for (size_t i = 0; i < n_.size(); ++i) {
      auto v = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>(lst[i]);
      if (v.ndim() != 1 || v.shape(0) != n_[i]) throw std::invalid_argument("batch: state length must match n_i");
      std::memcpy(in.data() + off, v.data(), (size_t)n_[i] * sizeof(std::complex<double>));
      off += (size_t)n_[i];
    }
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = fexpm_batch_run_resident(b_, in.data(), out.data());
    }
    check(rc);
    py::list res;
    off = 0;
    for (size_t i = 0; i < n_.size(); ++i) {
      py::array_t<std::complex<double>> v(n_[i]);
      std::memcpy(v.mutable_data(), out.data() + off, (size_t)n_[i] * sizeof(std::complex<double>));
      res.append(v);
      off += (size_t)n_[i];
    }
    return res;
  }
  py::array download(int64_t i) {
    if (i < 0 || i >= (int64_t)n_.size()) throw std::invalid_argument("batch: index out of range");
    py::array out = dt_ == FEXPM_COMPLEX64 ? (py::array)py::array_t<std::complex<float>>({n_[i], n_[i]})
                                           : (py::array)py::array_t<std::complex<double>>({n_[i], n_[i]});
    check(fexpm_batch_download(b_, i, out.mutable_data()));
    return out;
  }
  py::dict stats(int64_t i) {
    fexpm_stats s;
    check(fexpm_batch_get_stats(b_, i, &s));
    return stats_dict(s);
  }
  double last_ms() const { return fexpm_batch_last_ms(b_); }
  py::dict report() {
    fexpm_batch_report r;
    check(fexpm_batch_get_report(b_, &r));
    py::dict d;
    d["grouped"] = r.grouped;
    d["grouped_walks"] = r.grouped_walks;
    d["fallback_walks"] = r.fallback_walks;
    d["launches"] = r.launches;
    d["graph"] = r.graph;
    d["graph_built"] = r.graph_built;
    d["tuned"] = r.tuned;
    py::list cand, cms;
    for (int c = 0; c < 3; ++c) {
      cand.append(r.candidate[c]);
      cms.append(r.class_ms[c]);
    }
    d["candidate"] = cand;
    d["class_ms"] = cms;
    d["tune_ms"] = r.tune_ms;
    d["device_ms"] = r.device_ms;
    d["plan_ms"] = r.plan_ms;
    return d;
  }

 private:
  fexpm_batch* b_ = nullptr;
  int dt_ = 0;
  size_t es_ = 0;
  std::vector<int64_t> n_;
};

template <typename T, int DT> static void bind(py::module_& m, const char* name) {
  using C = ExpMat<T, DT>;
  py::class_<C>(m, name)
      .def(py::init<int64_t>(), py::arg("n"),
           "Initializes the MatrixExpCalculator for n x n matrices, allocating GPU memory.")
      .def("run", &C::run, py::arg("arr_a"),
           "Performs matrix exponential calculation on the given NumPy array.")
      .def("free", &C::free, "Free memory")
      .def("upload", &C::upload, py::arg("arr_a"))
      .def("run_resident", &C::run_resident)
      .def("set_graph", &C::set_graph, py::arg("adj"), py::arg("gamma") = 1.0, py::arg("laplacian") = false)
      .def("run_graph", &C::run_graph, py::arg("t"))
      .def("set_graph_edges", &C::set_graph_edges, py::arg("src"), py::arg("dst"), py::arg("w") = py::none(),
           py::arg("gamma") = 1.0, py::arg("laplacian") = false)
      .def("set_graph_er", &C::set_graph_er, py::arg("p"), py::arg("seed"), py::arg("weighted") = false,
           py::arg("gamma") = 1.0, py::arg("laplacian") = false)
      .def("debug_hamiltonian", &C::debug_hamiltonian)
      .def("download", &C::download)
      .def("run_device", &C::run_device, py::arg("d_a"), py::arg("d_u"), py::arg("stream") = 0)
      .def("apply", &C::apply, py::arg("psi0"))
      .def("propagate", &C::propagate, py::arg("psi0"), py::arg("nsteps"))
      .def("set_option", &C::set_option, py::arg("key"), py::arg("value"))
      .def("stats", &C::stats)
      .def("debug_plane", &C::debug_plane, py::arg("which"))
      .def_property_readonly("n", &C::n);
}

PYBIND11_MODULE(_fastexpm_core, m) {
  m.doc() = "MI355X matrix exponential engine (C ABI: include/fastexpm.h)";
  bind<float, FEXPM_FLOAT32>(m, "ExpMatFloat32");
  bind<double, FEXPM_FLOAT64>(m, "ExpMatFloat64");
  bind<std::complex<float>, FEXPM_COMPLEX64>(m, "ExpMatComplex64");
  bind<std::complex<double>, FEXPM_COMPLEX128>(m, "ExpMatComplex128");
  m.attr("OPT_PRECISION") = (int)FEXPM_OPT_PRECISION;
  m.attr("OPT_STRUCTURE") = (int)FEXPM_OPT_STRUCTURE;
  m.attr("OPT_KBLOCK") = (int)FEXPM_OPT_KBLOCK;
  m.attr("OPT_FP64_SQUARINGS") = (int)FEXPM_OPT_FP64_SQUARINGS;
  m.attr("OPT_PADE_FP64") = (int)FEXPM_OPT_PADE_FP64;
  m.attr("OPT_PANEL_COOP") = (int)FEXPM_OPT_PANEL_COOP;
  m.attr("OPT_SPD_SOLVE") = (int)FEXPM_OPT_SPD_SOLVE;
  m.attr("OPT_MAX_SQUARINGS") = (int)FEXPM_OPT_MAX_SQUARINGS;
  m.attr("OPT_GROUP_CANDIDATE") = (int)FEXPM_OPT_GROUP_CANDIDATE;
  m.attr("OPT_FP64_SLICES_SCOPE") = (int)FEXPM_OPT_FP64_SLICES_SCOPE;
  m.attr("OPT_TAIL_SPLIT") = (int)FEXPM_OPT_TAIL_SPLIT;
  m.attr("OPT_Z_OVERLAP") = (int)FEXPM_OPT_Z_OVERLAP;
  m.attr("OPT_CHOL_OUTER") = (int)FEXPM_OPT_CHOL_OUTER;
  m.attr("OPT_SQUARE_TILE") = (int)FEXPM_OPT_SQUARE_TILE;
  m.attr("OPT_Z_LEFT") = (int)FEXPM_OPT_Z_LEFT;
  m.attr("OPT_HOST_STAGING") = (int)FEXPM_OPT_HOST_STAGING;
  m.attr("OPT_BATCH_GROUPED") = (int)FEXPM_OPT_BATCH_GROUPED;
  m.attr("OPT_AUTOTUNE") = (int)FEXPM_OPT_AUTOTUNE;
  m.attr("OPT_GRAPH") = (int)FEXPM_OPT_GRAPH;
  m.attr("OPT_FP64_SLICES") = (int)FEXPM_OPT_FP64_SLICES;
  m.attr("OPT_SPECTRAL") = (int)FEXPM_OPT_SPECTRAL;
  m.attr("OPT_FAIL_ALLOC") = (int)FEXPM_OPT_FAIL_ALLOC;
  m.attr("OPT_CRT_MAP") = (int)FEXPM_OPT_CRT_MAP;
  m.attr("OPT_DEFLATE") = (int)FEXPM_OPT_DEFLATE;
  m.attr("OPT_TWO_PRODUCT") = (int)FEXPM_OPT_TWO_PRODUCT;
  m.attr("OPT_CHEBYSHEV") = (int)FEXPM_OPT_CHEBYSHEV;
  m.attr("OPT_CRT_MODULI") = (int)FEXPM_OPT_CRT_MODULI;
  m.attr("OPT_CRT64_MODULI") = (int)FEXPM_OPT_CRT64_MODULI;
  m.attr("PREC_NATIVE") = (int)FEXPM_PREC_NATIVE;
  m.attr("PREC_FP16X2") = (int)FEXPM_PREC_FP16X2;
  m.attr("PREC_BF16X3") = (int)FEXPM_PREC_BF16X3;
  m.attr("PREC_INT8_CRT") = (int)FEXPM_PREC_INT8_CRT;
  m.attr("PREC_AUTO") = (int)FEXPM_PREC_AUTO;
  py::class_<ExpMatBatch>(m, "ExpMatBatch")
      .def(py::init<const std::string&, int64_t, int>(), py::arg("dtype"), py::arg("nmax"), py::arg("lanes") = 4)
      .def("upload", &ExpMatBatch::upload, py::arg("mats"))
      .def("run_resident", &ExpMatBatch::run_resident, py::arg("psi0") = py::none())
      .def("set_option", &ExpMatBatch::set_option, py::arg("key"), py::arg("value"))
      .def("download", &ExpMatBatch::download, py::arg("i"))
      .def("stats", &ExpMatBatch::stats, py::arg("i"))
      .def("last_ms", &ExpMatBatch::last_ms)
      .def("count", &ExpMatBatch::count)
      .def("report", &ExpMatBatch::report);
  m.def("gpu_memory_gb", &fexpm_gpu_memory_gb);
}
