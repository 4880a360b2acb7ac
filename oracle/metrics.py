"""Accuracy metrics used by the parity tests (FP64 throughout)."""
from __future__ import annotations

import numpy as np


def infidelity(psi_ref: np.ndarray, psi: np.ndarray) -> float:
    """1 - |<psi_ref|psi>|^2 / (||psi_ref||^2 ||psi||^2)."""
    a = np.asarray(psi_ref, dtype=np.complex128).ravel()
    b = np.asarray(psi, dtype=np.complex128).ravel()
    num = abs(np.vdot(a, b)) ** 2
    den = np.vdot(a, a).real * np.vdot(b, b).real
    if den == 0:
        return 0.0 if num == 0 else 1.0
    return float(max(0.0, 1.0 - num / den))


def raw_infidelity(psi_ref: np.ndarray, psi: np.ndarray) -> float:
    """1 - |<psi_ref|psi>|^2 without normalisation (norm errors count)."""
    a = np.asarray(psi_ref, dtype=np.complex128).ravel()
    b = np.asarray(psi, dtype=np.complex128).ravel()
    return float(1.0 - abs(np.vdot(a, b)) ** 2)


def rel_fro(ref: np.ndarray, got: np.ndarray) -> float:
    """||got - ref||_F / ||ref||_F (FP64)."""
    r = np.asarray(ref, dtype=np.complex128)
    g = np.asarray(got, dtype=np.complex128)
    den = np.linalg.norm(r)
    num = np.linalg.norm(g - r)
    return float(num / den) if den > 0 else float(num)


def unitarity_defect(u: np.ndarray) -> float:
    """||U^H U - I||_F / sqrt(n)."""
    u = np.asarray(u, dtype=np.complex128)
    n = u.shape[0]
    return float(np.linalg.norm(u.conj().T @ u - np.eye(n)) / np.sqrt(max(n, 1)))
