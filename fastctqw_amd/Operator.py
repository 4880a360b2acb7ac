"""Walk operator U(t) = expm(-i H t) (mirror of FastCTQW/Operator.py).

Conventions kept from the reference:
* H = -gamma*A or -gamma*(D - A) (Operator.py:49-53), A_exp = -1j*H*t
  (:55-56), dtype complex64 for float32/complex64 data and complex128 for
  float64/complex128 (:27-32), TypeError otherwise;
* ValueError for a non-square / non-symmetric A (:80-90) and a non-Hermitian
  H (:92-103);
* device "cpu" -> scipy.linalg.expm (:74-76).

Deliberate differences (drop-in supersets):
* device "cuda" (also "gpu", "hip") runs the gfx950 engine with the engine
  class matching the operator dtype; the reference always picked
  ExpMatComplex64 (:40), truncating complex128 input;
* on the GPU, a real adjacency is sent once and H = -gamma*A or
  -gamma*(D - A) is formed on the device in FP64 (Operator.py:49-53); each
  ``buildExpmOperator(t)`` forms A_exp = -i t H there too, so no N x N
  complex array is built on the host or crosses PCIe.  The host H is built
  only if something reads ``_hamiltonian``; ``setHamiltonian`` switches back
  to uploading the host A_exp.  U stays in device memory and ``getOperator``
  copies it to the host only when asked (:165-166); ``applyTo`` computes
  U psi on the device with FP64 accumulation.  A complex adjacency whose
  imaginary part is zero (``nx.to_numpy_array(G, dtype=np.complex64)``, the
  README call) takes the same device route with its real part.
* ``from_adjacency_list`` / ``reset_adjList`` (empty stubs in the reference,
  Operator.py:114-116, 147-148) take an edge list; on the GPU only the edges
  cross PCIe and H is formed on the device.  ``from_erdos_renyi`` generates
  the seeded G(N, p) on the device (no counterpart in the reference).
"""
from __future__ import annotations

import numpy as np
from numpy.typing import NDArray

_GPU_DEVICES = ("cuda", "gpu", "hip")


def _operator_dtype(adjacency_dtype):
    """Operator dtype of an adjacency dtype (the map of Operator.py:27-32)."""
    dt = np.dtype(adjacency_dtype)
    if dt in (np.float32, np.complex64):
        return np.complex64
    if dt in (np.float64, np.complex128):
        return np.complex128
    raise TypeError(f"A.dtype is error : {dt}")


def _expm_cpu(a):
    from scipy.linalg import expm

    return expm(a)


class Operator:
    def __init__(self, data, gamma: float = 1, laplacian: bool = False, device: str = "cpu",
                 dtype=np.complex64) -> None:
        self._expMator = None
        self._expmMat = None
        self._on_device = False
        self._h = None
        self._adj = None
        self._edges = None  # (n, src, dst, w, adjacency dtype) of an edge-list graph
        self._er = None     # (n, p, seed, weighted, adjacency dtype) of a device-generated graph
        self._graph_on_device = False
        if isinstance(data, np.ndarray):
            if not self._is_adjacency_matrix(data):
                raise ValueError("A is not a adj matrix")
            self._n = data.shape[0]
            self._gamma = gamma
            self._laplacian = laplacian
            self._device = device
            if data.dtype in (np.float32, np.complex64):
                self._dtype = np.complex64
            elif data.dtype in (np.float64, np.complex128):
                self._dtype = np.complex128
            else:
                raise TypeError(f"A.dtype is error : {data.dtype}")
            self._set_graph(data)
        elif isinstance(data, (int, np.integer)):
            self._n = int(data)
            self._device = device
            self._dtype = np.dtype(dtype).type
            self._expMator = self._initExpMator(self._n, self._dtype)
        else:
            raise TypeError(f"data must be an ndarray or int, got {type(data)!r}")

    def free(self):
        if self._expMator is not None:
            self._expMator.free()

    def _set_graph(self, A: np.ndarray) -> None:
        """H for adjacency A: on the device for a real A on a GPU operator
        (a real symmetric A gives a real symmetric, hence Hermitian, H),
        otherwise on the host with the reference's Hermitian check."""
        self._edges = self._er = None
        real = not np.iscomplexobj(A) or not np.any(A.imag)
        if self._device in _GPU_DEVICES and real:
            if self._expMator is None:
                self._expMator = self._initExpMator(self._n, self._dtype)
            ra = A.real if np.iscomplexobj(A) else A
            self._expMator.set_graph(np.ascontiguousarray(ra), self._gamma, self._laplacian)
            self._adj, self._h, self._graph_on_device = A, None, True
            return
        self._adj, self._graph_on_device = None, False
        self._h = self._buildHamiltonian(A, self._gamma)
        if not self._is_hermitian(self._h):
            raise ValueError("_hamiltonian is not a hamiltonian matrix")
        if self._expMator is None:
            self._expMator = self._initExpMator(self._n, self._dtype)

    @property
    def _hamiltonian(self):
        if self._h is None and self._adj is None and self._edges is not None:
            from .graphs import dense_from_edges

            n, src, dst, w, adt = self._edges
            self._adj = dense_from_edges(n, src, dst, w, dtype=adt)
        if self._h is None and self._adj is None and self._er is not None:
            from .graphs import erdos_renyi

            n, p, seed, weighted, adt = self._er
            self._adj = erdos_renyi(n, p, seed, weighted=weighted).astype(adt)
        if self._h is None and self._adj is not None:
            self._h = self._buildHamiltonian(self._adj, self._gamma)
        return self._h

    @_hamiltonian.setter
    def _hamiltonian(self, value) -> None:
        self._h = value
        self._graph_on_device = False

    def _buildHamiltonian(self, A: np.ndarray, gamma: float) -> NDArray:
        if self._laplacian:
            return -gamma * np.asarray(np.diag(A.sum(axis=1)) - A, dtype=self._dtype)
        return -gamma * np.asarray(A, self._dtype)

    def _buildUnitory(self, time):
        return np.asarray(-1j * self._hamiltonian * time, dtype=self._dtype)

    def _initExpMator(self, n: int, dtype=np.complex64):
        if self._device == "cpu":
            return None
        if self._device in _GPU_DEVICES:
            from .fastexpm import ExpMatComplex64, ExpMatComplex128

            if dtype == np.complex64:
                return ExpMatComplex64(n)
            if dtype == np.complex128:
                return ExpMatComplex128(n)
            raise TypeError(f"self._phaseFactor.dtype: {dtype} is error")
        raise ValueError(f"unknown device {self._device!r}")

    def buildExpmOperator(self, time: float) -> None:
        """U = e^{-iHt}."""
        if self._device == "cpu":
            self._expmMat = _expm_cpu(self._buildUnitory(time))
            self._on_device = False
        elif self._graph_on_device:
            self._expMator.run_graph(time)
            self._expmMat = None
            self._on_device = True
        else:
            self._expMator.upload(self._buildUnitory(time))
            self._expMator.run_resident()
            self._expmMat = None
            self._on_device = True

    def applyTo(self, psi0: np.ndarray) -> np.ndarray:
        """U psi0 in the operator dtype; on the GPU with FP64 accumulation."""
        if self._on_device:
            return self._expMator.apply(psi0).astype(self._dtype, copy=False)
        return self.getOperator() @ psi0

    def lastRunStats(self) -> dict:
        return self._expMator.stats() if self._expMator is not None else {}

    def _is_adjacency_matrix(self, A: np.ndarray) -> bool:
        if A.ndim != 2:
            print(f"Error: Matrix is not 2D. Dimensions: {A.ndim}")
            return False
        if A.shape[0] != A.shape[1]:
            print(f"Error: Matrix is not square. Shape: {A.shape}")
            return False
        if not np.array_equal(A, A.T):
            print("Error: Matrix is not symmetric (expected for undirected graph).")
            return False
        return True

    def _is_hermitian(self, hamiltonian) -> bool:
        return np.allclose(hamiltonian, hamiltonian.conj().T)

    @classmethod
    def from_networkx_graph(cls, nx_graph, gamma: float = 1, laplacian: bool = False,
                            device: str = "cpu") -> "Operator":
        import networkx as nx

        return cls(nx.to_numpy_array(nx_graph), gamma, laplacian, device)

    @classmethod
    def from_numpy_array(cls, A: np.ndarray, gamma: float = 1, laplacian: bool = False,
                         device: str = "cpu") -> "Operator":
        return cls(A, gamma, laplacian, device)

    @classmethod
    def from_adjacency_list(cls, adj_list, gamma: float = 1, laplacian: bool = False, device: str = "cpu",
                            num_nodes=None, dtype=np.float32) -> "Operator":
        """Operator of an undirected graph given as an adjacency list (see
        fastctqw_amd.graphs.edge_list for the accepted forms); ``dtype`` is
        the adjacency dtype (float32 -> complex64 operator, float64 ->
        complex128).  The reference declares this constructor with an empty
        body (Operator.py:114-116)."""
        from .graphs import edge_list

        n, src, dst, w = edge_list(adj_list, num_nodes)
        op = cls(n, device=device, dtype=_operator_dtype(dtype))
        op._set_edges(n, src, dst, w, gamma, laplacian, dtype)
        return op

    @classmethod
    def from_erdos_renyi(cls, n: int, p: float, seed: int, weighted: bool = False, gamma: float = 1,
                         laplacian: bool = False, device: str = "cpu", dtype=np.float32) -> "Operator":
        """Operator of the seeded G(n, p) of fastctqw_amd.graphs.erdos_renyi;
        on the GPU the graph is generated in device memory (same bits)."""
        op = cls(n, device=device, dtype=_operator_dtype(dtype))
        op._gamma, op._laplacian = gamma, laplacian
        if device in _GPU_DEVICES:
            op._expMator.set_graph_er(p, seed, weighted, gamma, laplacian)
            op._adj, op._h, op._graph_on_device = None, None, True
            op._er = (n, p, seed, weighted, dtype)
        else:
            from .graphs import erdos_renyi

            op._set_graph(erdos_renyi(n, p, seed, weighted=weighted).astype(dtype))
        return op

    def _set_edges(self, n, src, dst, w, gamma, laplacian, adt) -> None:
        self._gamma, self._laplacian = gamma, laplacian
        if self._device in _GPU_DEVICES:
            self._expMator.set_graph_edges(src, dst, w, gamma, laplacian)
            self._adj, self._h, self._graph_on_device = None, None, True
            self._edges, self._er = (n, src, dst, w, adt), None
        else:
            from .graphs import dense_from_edges

            self._set_graph(dense_from_edges(n, src, dst, w, dtype=adt))
        self._expmMat = None
        self._on_device = False

    def reset_adjList(self, adj_list, gamma: float = 1, time=None, laplacian: bool = False,
                      device=None, num_nodes=None, dtype=np.float32) -> None:
        """Replace the graph by an adjacency list (empty in the reference,
        Operator.py:147-148; ``time`` is accepted for its signature and
        unused: pass the time to buildExpmOperator)."""
        from .graphs import edge_list

        n, src, dst, w = edge_list(adj_list, num_nodes)
        if device is not None and device != self._device:
            self.free()
            self._expMator, self._device = None, device
        odt = _operator_dtype(dtype)
        if n != self._n or odt != self._dtype or self._expMator is None:
            if self._expMator is not None:
                self._expMator.free()
            self._n, self._dtype = n, odt
            self._expMator = self._initExpMator(n, odt)
        self._set_edges(n, src, dst, w, gamma, laplacian, dtype)

    @classmethod
    def from_num_nodes(cls, n: int, dtype=np.complex64, device: str = "cpu"):
        return cls(n, device=device, dtype=dtype)

    def reset_numpy_array(self, A: np.ndarray, gamma: float = 1, laplacian: bool = False):
        if A.dtype in (np.float32, np.complex64):
            dtype = np.complex64
        elif A.dtype in (np.float64, np.complex128):
            dtype = np.complex128
        else:
            raise TypeError(f"A.dtype is error : {A.dtype}")
        if not self._is_adjacency_matrix(A):
            raise ValueError("A is not a adj matrix")
        changed = dtype != getattr(self, "_dtype", None) or A.shape[0] != self._n
        self._dtype = dtype
        self._n = A.shape[0]
        self._gamma = gamma
        self._laplacian = laplacian
        if self._device != "cpu" and changed and self._expMator is not None:
            self._expMator.free()
            self._expMator = None
        self._set_graph(A)
        self._expmMat = None
        self._on_device = False

    def reset_networkx_graph(self, nx_graph, gamma: float = 1, laplacian: bool = False):
        import networkx as nx

        self.reset_numpy_array(nx.to_numpy_array(nx_graph), gamma, laplacian)

    def setHamiltonian(self, hamiltonian: NDArray):
        self._hamiltonian = hamiltonian  # later walks upload -1j * H * t from the host

    def setexpmOperator(self, newexpmMat):
        self._expmMat = newexpmMat
        self._on_device = False

    def getOperator(self):
        if self._expmMat is None and self._on_device:
            self._expmMat = self._expMator.download()
        return self._expmMat
