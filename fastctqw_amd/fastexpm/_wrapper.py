"""Python wrappers over the pybind11 module ``_fastexpm_core``.

Behaviour follows the reference wrappers (FastCTQW/fastexpm/_wrapper.py:9-130):
``run`` raises TypeError for a non-ndarray (:92) and ValueError for a shape
other than (N, N) (:94), returns ``backend.run(A).reshape(N, N)`` (:96);
``__exit__`` frees and turns a failing ``free`` into a warning without
swallowing the original exception (:79-88).  ``free`` is idempotent.

The native module is required: there is no CPU fallback.  If it has not been
built, importing this package raises ImportError naming the build command.
"""
from __future__ import annotations

import warnings

import numpy as np

try:
    from . import _fastexpm_core
except ImportError as exc:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "fastexpm native module is not built: run `python -m fastctqw_amd.build` "
        f"(gfx950 HIP library + pybind11 layer). Original error: {exc}"
    ) from exc


class _ExpMatBase:
    _core_cls = None

    def __init__(self, N: int):
        self.N = N
        self._backend = self._core_cls(N)

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        try:
            self.free()
        except Exception as cleanup_error:
            warnings.warn(f"[fastexpm] Failed to free resources in __exit__: {cleanup_error}")
        return False

    def run(self, A) -> np.ndarray:
        if not isinstance(A, np.ndarray):
            raise TypeError("Input must be a NumPy array.")
        if A.shape != (self.N, self.N):
            raise ValueError(f"Input matrix must have shape ({self.N}, {self.N}).")
        return self._backend.run(A).reshape(self.N, self.N)

    def free(self):
        self._backend.free()

    # ---- device-resident extensions (no counterpart in the reference) ----
    def upload(self, A: np.ndarray) -> None:
        """Copy A to the engine's device buffer."""
        if not isinstance(A, np.ndarray):
            raise TypeError("Input must be a NumPy array.")
        if A.shape != (self.N, self.N):
            raise ValueError(f"Input matrix must have shape ({self.N}, {self.N}).")
        self._backend.upload(A)

    def run_resident(self) -> None:
        """U = expm(A) for the uploaded A; U stays on the device."""
        self._backend.run_resident()

    def download(self) -> np.ndarray:
        return self._backend.download()

    def set_graph(self, adj, gamma: float = 1.0, laplacian: bool = False) -> None:
        """Keep H = -gamma*A (or -gamma*(D - A)) on the device, formed in FP64
        from a real adjacency (float32 inputs stay float32 on the wire)."""
        adj = np.asarray(adj)
        if np.iscomplexobj(adj):
            raise TypeError("set_graph expects a real adjacency matrix")
        if adj.dtype != np.float32:
            adj = adj.astype(np.float64, copy=False)
        self._backend.set_graph(adj, float(gamma), bool(laplacian))

    def set_graph_edges(self, src, dst, w=None, gamma: float = 1.0, laplacian: bool = False) -> None:
        """Walk Hamiltonian H = -gamma*A or -gamma*(D - A) from an undirected
        edge list (vertex indices src[k], dst[k], optional weights w[k]),
        built in FP64 on the device."""
        self._backend.set_graph_edges(np.asarray(src, dtype=np.int64), np.asarray(dst, dtype=np.int64),
                                      None if w is None else np.asarray(w, dtype=np.float64),
                                      float(gamma), bool(laplacian))

    def set_graph_er(self, p: float, seed: int, weighted: bool = False, gamma: float = 1.0,
                     laplacian: bool = False) -> None:
        """Walk Hamiltonian of the seeded G(N, p) generated on the device: the
        graph of fastctqw_amd.graphs.erdos_renyi(N, p, seed, weighted), bit for bit."""
        self._backend.set_graph_er(float(p), int(seed), bool(weighted), float(gamma), bool(laplacian))

    def run_graph(self, t: float) -> None:
        """U = expm(-i t H) for the graph of set_graph, resident on the device."""
        self._backend.run_graph(float(t))

    def apply(self, psi0) -> np.ndarray:
        """psi = U psi0 (complex128) for the U of the last run, FP64 accumulation."""
        return self._backend.apply(np.asarray(psi0, dtype=np.complex128))

    def propagate(self, psi0, nsteps: int) -> np.ndarray:
        """psi[k-1] = U^k psi0 for k = 1..nsteps (complex128, shape
        (nsteps, nstates, N)) with the U of the last run."""
        psi0 = np.atleast_2d(np.asarray(psi0, dtype=np.complex128))
        return self._backend.propagate(psi0, int(nsteps))

    def stats(self) -> dict:
        return self._backend.stats()

    def set_structure_detection(self, enabled: bool) -> None:
        self._backend.set_option(_fastexpm_core.OPT_STRUCTURE, 1 if enabled else 0)

    _PRECISIONS = {"native": "PREC_NATIVE", "fp16x2": "PREC_FP16X2", "bf16x3": "PREC_BF16X3",
                   "int8crt": "PREC_INT8_CRT", "auto": "PREC_AUTO"}

    def set_precision(self, mode: str) -> None:
        """Squaring precision of a complex64 engine on structured (CTQW)
        inputs: "native" (FP32 MFMA), "fp16x2" or "bf16x3" (split 16-bit
        products with FP32 accumulation), "int8crt" (rows rounded to 27 bits,
        exact int8 products modulo 9 moduli, Chinese-remainder reconstruction;
        orders up to 16384) or "auto" (the complex64 default: "int8crt" for
        768 <= n <= 16384, "native" otherwise; stats()["precision"] names the
        form a run used).  No counterpart in the reference."""
        if mode not in self._PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(self._PRECISIONS)}")
        self._backend.set_option(_fastexpm_core.OPT_PRECISION, getattr(_fastexpm_core, self._PRECISIONS[mode]))


class ExpMatFloat32(_ExpMatBase):
    _core_cls = _fastexpm_core.ExpMatFloat32


class ExpMatFloat64(_ExpMatBase):
    _core_cls = _fastexpm_core.ExpMatFloat64


class ExpMatComplex64(_ExpMatBase):
    _core_cls = _fastexpm_core.ExpMatComplex64


class ExpMatComplex128(_ExpMatBase):
    _core_cls = _fastexpm_core.ExpMatComplex128


class ExpMatBatch:
    """Many independent exponentials of uneven order on the GPU (no
    counterpart in the reference).  ``run(mats, psi0=None)`` returns the list
    of U (and of psi = U psi0 when states are given)."""

    def __init__(self, nmax: int, dtype=np.complex64, lanes: int = 4):
        self.nmax = nmax
        name = np.dtype(dtype).name
        self._backend = _fastexpm_core.ExpMatBatch(name, int(nmax), int(lanes))

    def set_precision(self, mode: str) -> None:
        """Squaring precision for every lane ("native", "fp16x2", "bf16x3")."""
        if mode not in _ExpMatBase._PRECISIONS or mode in ("int8crt", "auto"):
            raise ValueError("precision must be one of ['bf16x3', 'fp16x2', 'native']")
        self._backend.set_option(_fastexpm_core.OPT_PRECISION,
                                 getattr(_fastexpm_core, _ExpMatBase._PRECISIONS[mode]))

    def upload(self, mats):
        self._backend.upload([np.asarray(m) for m in mats])
        self._count = len(mats)

    def run_resident(self, psi0=None):
        return self._backend.run_resident(None if psi0 is None else [np.asarray(v) for v in psi0])

    def download(self, i: int) -> np.ndarray:
        return self._backend.download(int(i))

    def count(self) -> int:
        """Number of matrices of the uploaded batch."""
        return int(self._backend.count())

    def run(self, mats, psi0=None):
        self.upload(mats)
        psi = self.run_resident(psi0)
        us = [self.download(i) for i in range(len(mats))]
        return us if psi is None else (us, psi)

    def stats(self, i: int) -> dict:
        return self._backend.stats(int(i))

    def last_ms(self) -> float:
        return self._backend.last_ms()

    def set_grouped(self, on: bool = True, autotune: bool = True, graph: bool = True) -> None:
        """Grouped launches over all walks (default) or concurrent lanes; the
        autotuner and hipGraph replay of grouped runs."""
        self._backend.set_option(_fastexpm_core.OPT_BATCH_GROUPED, int(bool(on)))
        self._backend.set_option(_fastexpm_core.OPT_AUTOTUNE, int(bool(autotune)))
        self._backend.set_option(_fastexpm_core.OPT_GRAPH, int(bool(graph)))

    def report(self) -> dict:
        """How the last run executed: grouped or lanes, walks per path, launch
        count, graph replay, autotuner choices and device time."""
        return self._backend.report()
