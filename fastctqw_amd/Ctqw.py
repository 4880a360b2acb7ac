"""Continuous-time quantum walk (mirror of FastCTQW/Ctqw.py).

``runWalk(time)`` builds U = expm(-i H t) (Ctqw.py:90-94) and sets the final
state to U psi0.  Deliberate differences (drop-in supersets):
* ``from_networkx_graph`` and ``from_numpy_array`` accept ``device=`` (the
  README usage, README.md:64) and pass ``ctqwId`` by keyword (the reference
  shifted it into the dtype slot, Ctqw.py:57);
* on the GPU the product U psi0 runs on the device with FP64 accumulation
  instead of a host matvec;
* ``from_adjacency_list`` (an empty stub in the reference, Ctqw.py:60-71)
  builds the walk from an edge list, and ``from_erdos_renyi`` from the
  seeded G(N, p), generated on the device for device="cuda".
"""
from __future__ import annotations

import numpy as np

from .Operator import Operator
from .State import State


class Ctqw:
    def __init__(self, data, initState: State, gamma: float = 1, laplacian: bool = False,
                 dtype=np.complex64, device: str = "cpu", ctqwId: str = "Undef") -> None:
        if isinstance(data, np.ndarray):
            self._n = data.shape[0]
        elif isinstance(data, (int, np.integer)):
            self._n = int(data)
        else:
            raise TypeError(f"data must be an ndarray or int, got {type(data)!r}")
        self._device = device
        self.ctqwId = ctqwId
        self._initStator = initState
        self._finalStator = State.getSingleNodeState(self._n, 0)
        if isinstance(data, np.ndarray):
            self._Operator = Operator.from_numpy_array(data, gamma, laplacian, device)
        else:
            self._Operator = Operator.from_num_nodes(self._n, dtype, device)

    @classmethod
    def from_numpy_array(cls, A: np.ndarray, initState: State, gamma: float = 1,
                         laplacian: bool = False, ctqwId: str = "Undef",
                         device: str = "cpu") -> "Ctqw":
        return cls(A, initState, gamma, laplacian, device=device, ctqwId=ctqwId)

    @classmethod
    def from_networkx_graph(cls, graph, initState: State, gamma: float = 1,
                            laplacian: bool = False, ctqwId: str = "Undef",
                            dtype=np.float32, device: str = "cpu") -> "Ctqw":
        import networkx as nx

        if graph.is_directed():
            raise TypeError("graph is a directed graph, but Ctqw recive a undirected graph")
        return cls(nx.to_numpy_array(graph, dtype=dtype), initState, gamma, laplacian,
                   device=device, ctqwId=ctqwId)

    @classmethod
    def _from_operator(cls, op: Operator, initState: State, device: str, ctqwId: str) -> "Ctqw":
        self = cls.__new__(cls)
        self._n = op._n
        self._device = device
        self.ctqwId = ctqwId
        self._initStator = initState
        self._finalStator = State.getSingleNodeState(self._n, 0)
        self._Operator = op
        return self

    @classmethod
    def from_adjacency_list(cls, adj_list, initState: State, time: float = 1, gamma: float = 1,
                            laplacian: bool = False, ctqwId: str = "Undef", dtype=np.float32,
                            device: str = "cpu", num_nodes=None) -> "Ctqw":
        """Walk on an undirected graph given as an adjacency list (the forms
        of fastctqw_amd.graphs.edge_list); ``dtype`` is the adjacency dtype.
        ``time`` is accepted for the reference's signature (Ctqw.py:60-68)
        and unused: pass the time to runWalk."""
        op = Operator.from_adjacency_list(adj_list, gamma, laplacian, device, num_nodes=num_nodes, dtype=dtype)
        return cls._from_operator(op, initState, device, ctqwId)

    @classmethod
    def from_erdos_renyi(cls, n: int, p: float, seed: int, initState: State, weighted: bool = False,
                         gamma: float = 1, laplacian: bool = False, ctqwId: str = "Undef",
                         dtype=np.float32, device: str = "cpu") -> "Ctqw":
        """Walk on the seeded G(n, p) of fastctqw_amd.graphs.erdos_renyi
        (generated in device memory for device="cuda")."""
        op = Operator.from_erdos_renyi(n, p, seed, weighted, gamma, laplacian, device, dtype)
        return cls._from_operator(op, initState, device, ctqwId)

    @classmethod
    def from_num_nodes(cls, n: int, dtype=np.complex64, device: str = "cpu") -> "Ctqw":
        return cls(n, State.getSingleNodeState(n, 0), dtype=dtype, device=device)

    def reset_numpy_array(self, A: np.ndarray, initState: State, gamma: float = 1,
                          laplacian: bool = False):
        self._n = A.shape[0]
        self._Operator.reset_numpy_array(A, gamma, laplacian)
        self._initStator = initState
        self._finalStator = State.getSingleNodeState(self._n, 0)

    def reset_networkx_graph(self, nx_graph, initState: State, gamma: float = 1,
                             laplacian: bool = False):
        import networkx as nx

        self.reset_numpy_array(nx.to_numpy_array(nx_graph), initState, gamma, laplacian)

    def runWalk(self, time: float = 0):
        self._Operator.buildExpmOperator(time)
        self._finalStator.setState(self._Operator.applyTo(self._initStator.getState()))

    def getFinalState(self) -> State:
        return self._finalStator

    def getOperator(self) -> Operator:
        return self._Operator
