"""Independent FP64 references for exp(A) on CTQW inputs.

* ``eig_propagator``: for real-symmetric H, U(t) = V diag(exp(-i t lam)) V^H
  from LAPACK ``eigh`` (numpy).  This is the oracle named for the walk:
  U = expm(-i t H).  Pinned in tests/test_oracle.py against closed forms
  (cycle, complete graph, hypercube) and scipy.linalg.expm.
* ``eig_state``: psi(t) = V (exp(-i t lam) * (V^H psi0)) without forming U.
* ``taylor_expm``: restatement of the reference test's own oracle,
  ``matrix_exp_taylor`` (tests/test_fastexpm_accuracy.py:10-56): s from
  ceil(log2 ||A||_1) (:29-34), Taylor terms until the Frobenius norm of a
  term drops below tol (:40-49), then s squarings (:52-54).
"""
from __future__ import annotations

import numpy as np


def eigh_cached(h: np.ndarray):
    """(lam, V) of a real-symmetric matrix in FP64."""
    h64 = np.asarray(h, dtype=np.float64)
    lam, vec = np.linalg.eigh(h64)
    return lam, vec


def eig_propagator(h: np.ndarray, t: float, eig=None) -> np.ndarray:
    """U = expm(-i t H) for real-symmetric H (complex128)."""
    lam, vec = eig if eig is not None else eigh_cached(h)
    ph = np.exp(-1j * t * lam)
    return (vec * ph[None, :]) @ vec.T.astype(np.complex128)


def eig_state(h: np.ndarray, t: float, psi0: np.ndarray, eig=None) -> np.ndarray:
    """psi(t) = expm(-i t H) psi0 for real-symmetric H (complex128)."""
    lam, vec = eig if eig is not None else eigh_cached(h)
    c = vec.T @ np.asarray(psi0, dtype=np.complex128)
    return vec @ (np.exp(-1j * t * lam) * c)


def taylor_expm(a: np.ndarray, m: int = 100, tol: float = 1e-10) -> np.ndarray:
    """The reference test's Taylor scaling-and-squaring oracle, restated
    (tests/test_fastexpm_accuracy.py:10-56), evaluated in complex128."""
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[0]
    norm_a = np.linalg.norm(a, ord=1)
    if norm_a == 0:
        return np.eye(n, dtype=a.dtype)
    s = max(0, int(np.ceil(np.log2(norm_a))))
    a_scaled = a / (2 ** s)
    out = np.eye(n, dtype=a.dtype)
    term = np.eye(n, dtype=a.dtype)
    for k in range(1, m + 1):
        term = term @ a_scaled / k
        out += term
        if np.linalg.norm(term, ord="fro") < tol:
            break
    for _ in range(s):
        out = out @ out
    return out
