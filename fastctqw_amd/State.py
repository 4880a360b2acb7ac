"""Quantum state vectors (mirror of FastCTQW/State.py:5-59).

Constructors return complex64 vectors like the reference
(getUniformSuperposition :27-32, getSingleNodeState :35-41,
getLocalSuperposition :44-49).  A non-1-D array raises ValueError (:9-10,
:54-55); an unnormalised state only warns (:14-25).
"""
from __future__ import annotations

import warnings

import numpy as np


class State:
    def __init__(self, initState: np.ndarray):
        if not self._isState(initState):
            raise ValueError("initState is not a quantum state")
        self._state = initState

    def _isState(self, state: np.ndarray, tol: float = 1e-5) -> bool:
        state = np.asarray(state)
        if state.ndim != 1:
            warnings.warn(f"input is not a 1-D array (ndim = {state.ndim})")
            return False
        sum_of_squares = float(np.sum(np.abs(state) ** 2))
        if not np.isclose(sum_of_squares, 1.0, atol=tol):
            warnings.warn(f"state is not normalised: sum |psi|^2 = {sum_of_squares}")
        return True

    @classmethod
    def getUniformSuperposition(cls, n: int) -> "State":
        return cls(np.full(n, 1.0 / np.sqrt(n), dtype=np.complex64))

    @classmethod
    def getSingleNodeState(cls, n: int, m: int) -> "State":
        s = np.zeros(n, dtype=np.complex64)
        s[m] = 1.0
        return cls(s)

    @classmethod
    def getLocalSuperposition(cls, n: int, markNodes: list) -> "State":
        s = np.zeros(n, dtype=np.complex64)
        for i in markNodes:
            s[i] = 1.0
        s = s / np.linalg.norm(s)
        return cls(s)

    def getProbabilities(self) -> np.ndarray:
        return np.abs(self._state) ** 2

    def setState(self, initState: np.ndarray) -> None:
        if not self._isState(initState):
            raise ValueError("initState is not a quantum state")
        self._state = initState

    def getState(self) -> np.ndarray:
        return self._state
