"""Closed forms of exp(i theta G) for graphs whose spectrum is known.

Each is derived from the spectral decomposition and is independent of any
Pade or Taylor code, so it pins the oracle and the engine alike:

* complete graph K_n, G = J - I:  J has eigenvalues n (vector 1) and 0, so
  exp(i th (J - I)) = e^{-i th} (I + (e^{i th n} - 1)/n J).
* cycle C_n: circulant with eigenvalues 2 cos(2 pi k / n) on the Fourier
  basis, so exp(i th C)_{jl} = (1/n) sum_k e^{i th 2cos(2 pi k/n)} w^{k(j-l)}.
* hypercube Q_d: G = sum_k sigma_x on bit k, so
  exp(i th G) = (x)_k (cos th I + i sin th sigma_x).
* diagonal D: exp(D) = diag(exp(d_j)).
"""
from __future__ import annotations

import numpy as np


def complete_graph(n: int) -> np.ndarray:
    return np.ones((n, n)) - np.eye(n)


def expm_complete(n: int, theta: float) -> np.ndarray:
    """exp(1j*theta*(J - I)) for the n-vertex complete graph."""
    j = np.ones((n, n), dtype=np.complex128)
    return np.exp(-1j * theta) * (np.eye(n) + (np.exp(1j * theta * n) - 1.0) / n * j)


def cycle_graph(n: int) -> np.ndarray:
    a = np.zeros((n, n))
    idx = np.arange(n)
    a[idx, (idx + 1) % n] = 1.0
    a[(idx + 1) % n, idx] = 1.0
    return a


def expm_cycle(n: int, theta: float) -> np.ndarray:
    """exp(1j*theta*A) for the n-cycle adjacency A (n >= 3)."""
    k = np.arange(n)
    lam = 2.0 * np.cos(2.0 * np.pi * k / n)
    # first row of the circulant: c_d = (1/n) sum_k e^{i th lam_k} w^{k d}
    d = np.arange(n)
    w = np.exp(2j * np.pi * np.outer(k, d) / n)
    c = (np.exp(1j * theta * lam)[:, None] * w).sum(axis=0) / n
    out = np.empty((n, n), dtype=np.complex128)
    for r in range(n):
        out[r] = np.roll(c, r)
    return out


def hypercube_graph(dim: int) -> np.ndarray:
    n = 1 << dim
    a = np.zeros((n, n))
    for v in range(n):
        for b in range(dim):
            a[v, v ^ (1 << b)] = 1.0
    return a


def expm_hypercube(dim: int, theta: float) -> np.ndarray:
    """exp(1j*theta*A) for the dim-dimensional hypercube adjacency A.

    Vertex v's bit b is the b-th tensor factor counted from the least
    significant bit, so the Kronecker product runs from bit dim-1 to bit 0.
    """
    f = np.array([[np.cos(theta), 1j * np.sin(theta)],
                  [1j * np.sin(theta), np.cos(theta)]])
    out = np.array([[1.0 + 0j]])
    for _ in range(dim):
        out = np.kron(out, f)
    return out
