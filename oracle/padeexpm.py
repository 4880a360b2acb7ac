"""FP64 restatement of the scaling-and-squaring Pade matrix exponential.

The reference engine ``MatrixExpCalculator<T>::run`` has no source in the
reference checkout (its bodies come from the absent ``lib/libexpmcore.a``,
CMakeLists.txt:23-25).  Its declared buffers -- ``d_A, d_A2, d_A4, d_A6,
d_u1, d_u2, d_v1, d_v2, d_mu, d_nrmA`` (include/MatrixExpCalculator.h:52-61)
-- match Higham's Pade-13 evaluation with a trace shift ``mu`` and a single
1-norm.  This module restates that published algorithm:

* N. J. Higham, "The scaling and squaring method for the matrix exponential
  revisited", SIMAX 26(4), 2005 -- Algorithm 2.3: degrees m in
  {3,5,7,9,13} with thresholds theta_m, s = ceil(log2(||A||_1/theta_13)),
  Pade coefficients b_j, U/V assembly, (V-U) X = (V+U), s squarings.
* Trace reduction (Higham, Functions of Matrices, 2008, sec. 10.7.3):
  exp(A) = e^mu exp(A - mu I) with mu = tr(A)/n, applied when it does not
  increase the 1-norm.

* Spectral scaling of normal inputs (``spectral=True``): for A = iG with G
  real symmetric the backward error of the Pade approximant at A/2^s is
  that of its eigenvalues, so Higham's condition may be read in the 2-norm,
  rho(A - mu I)/2^s <= theta_13, with any upper bound of the spectral radius.
  The bounds are rho(B) <= ||B^k||_1^(1/k) (Gelfand; k = 1, 2, 4, 6, the
  powers the Pade-13 evaluation needs anyway; cf. the d_k = ||A^k||^(1/k)
  of Al-Mohy & Higham, SIMAX 31(3), 2009) and, B being symmetric,
  rho(B)^2k <= trace(B^2k) = sum of lambda^2k for 2k = 4, ..., 12, and
  2k = 14, 16 from B^8 for plans that take the Chebyshev route
  (``oracle.chebcs``, which needs B^8 anyway; ``cheb=True``, the engine's
  default).
  When every row sum of G equals the same c, the eigenvector 1 is
  deflated first: B = G - mu I + alpha J
  with alpha = (mu - c)/n has B 1 = 0 and commutes with J = 1 1^T, and
  exp(A) = e^{i mu} (exp(iB) + (e^{i(c - mu)} - 1) J / n).

Stages are returned individually so that GPU tests can compare the engine's
mu, norm, m and s exactly and its U, V, X within rounding.

Pinning: tests/test_oracle.py checks ``pade_expm`` against
scipy.linalg.expm (Al-Mohy-Higham 2009, a different degree/scaling rule),
mpmath.expm at N <= 12 (arbitrary precision), closed forms (complete graph,
cycle, hypercube, diagonal, zero, 1x1) and the FP64 eigendecomposition
oracle; a wrong Pade coefficient, a dropped term, a wrong sign of mu or an
off-by-one in s each fail at least one of those.
"""
from __future__ import annotations

import math

import numpy as np

# theta_m for double precision (Higham 2005, Table 2.3).
THETA = {
    3: 1.495585217958292e-2,
    5: 2.539398330063230e-1,
    7: 9.504178996162932e-1,
    9: 2.097847961257068e0,
    13: 5.371920351148152e0,
}

# Pade [m/m] coefficients b_0..b_m of p_m(x) = sum b_j x^j (Higham 2005, (2.2)),
# scaled so that b_m = 1: b_j = (2m-j)! / (j! (m-j)!), exact integers.
def pade_coefficients(m: int) -> tuple:
    return tuple(float(math.factorial(2 * m - j)
                       // (math.factorial(j) * math.factorial(m - j)))
                 for j in range(m + 1))


PADE_B = {m: pade_coefficients(m) for m in (3, 5, 7, 9, 13)}


def onenorm(a: np.ndarray) -> float:
    """Exact 1-norm: max column sum of |a_ij| (FP64 accumulation)."""
    if a.size == 0:
        return 0.0
    return float(np.abs(a.astype(np.complex128)).sum(axis=0).max())


def trace_shift(a: np.ndarray):
    """Return (mu, shifted) with mu = tr(a)/n, or (0, a) if the shift does
    not reduce the 1-norm.  Matches the engine's rule exactly: the shift is
    kept iff ||a - mu I||_1 <= ||a||_1."""
    n = a.shape[0]
    mu = complex(np.trace(a.astype(np.complex128))) / n
    shifted = a.astype(np.complex128) - mu * np.eye(n)
    if onenorm(shifted) <= onenorm(a):
        return mu, shifted
    return 0.0j, a.astype(np.complex128)


def choose_degree(norm: float):
    """(m, s) of Higham 2005 Algorithm 2.3 for a matrix of 1-norm ``norm``."""
    for m in (3, 5, 7, 9):
        if norm <= THETA[m]:
            return m, 0
    if not math.isfinite(norm):
        raise ValueError("non-finite 1-norm")
    s = max(0, int(math.ceil(math.log2(norm / THETA[13]))))
    # guard against log2 rounding placing norm/2^s just above theta_13
    while norm / (2.0 ** s) > THETA[13]:
        s += 1
    return 13, s


def scaling_for(rho: float) -> int:
    """Squarings of degree 13 for a bound ``rho`` of the 2-norm (the rule of
    choose_degree with rho in place of the 1-norm)."""
    if not rho > THETA[13]:
        return 0
    s = max(0, int(math.ceil(math.log2(rho / THETA[13]))))
    while rho / (2.0 ** s) > THETA[13]:
        s += 1
    return s


def spectral_bound(b: np.ndarray, with8: bool = False) -> float:
    """Upper bound of the spectral radius of a real symmetric b: the least of
    ||b^k||_1^(1/k), k in (1, 2, 4, 6) (Gelfand), and trace(b^2k)^(1/2k),
    2k in (4, 6, 8, 10, 12), and (14, 16) with ``with8`` (a sum of the 2k-th
    powers of the eigenvalues, formed as elementwise products of b^2, b^4,
    b^6, b^8), widened by 1e-9 for rounding."""
    b2 = b @ b
    b4 = b2 @ b2
    b6 = b4 @ b2
    vals = [onenorm(b), onenorm(b2) ** 0.5, onenorm(b4) ** 0.25, onenorm(b6) ** (1.0 / 6.0)]
    pairs = [((b2, b2), 4), ((b2, b4), 6), ((b4, b4), 8), ((b4, b6), 10), ((b6, b6), 12)]
    if with8:
        b8 = b4 @ b4
        pairs += [((b6, b8), 14), ((b8, b8), 16)]
    for (x, y), k in pairs:
        tr = float(np.sum(x * y))
        if tr >= 0.0:
            vals.append(tr ** (1.0 / k))
    return min(vals) * (1.0 + 1e-9)


def structured_part(a: np.ndarray):
    """G if a = iG with G real and exactly symmetric (the CTQW form), else None."""
    a = np.asarray(a)
    if not np.iscomplexobj(a) or np.any(a.real != 0):
        return None
    g = np.ascontiguousarray(a.imag, dtype=np.float64)
    return g if np.array_equal(g, g.T) else None


def equal_row_sums(g: np.ndarray):
    """(True, c) if every row of g has the same exact sum (c = that sum
    rounded).  Each sum is kept as the pair (fl(sum), fl(sum - fl(sum))) of
    correctly rounded math.fsum results, about 106 bits, like the device's
    double-double row sums."""
    pairs = []
    for row in np.asarray(g, dtype=np.float64):
        hi = math.fsum(row)
        pairs.append((hi, math.fsum(list(row) + [-hi])))
    if not pairs:
        return True, 0.0
    return all(p == pairs[0] for p in pairs), pairs[0][0] + pairs[0][1]


def pade_uv(a: np.ndarray, m: int):
    """U (odd part) and V (even part) of the [m/m] Pade approximant at a."""
    n = a.shape[0]
    b = PADE_B[m]
    ident = np.eye(n, dtype=a.dtype)
    a2 = a @ a
    if m == 13:
        a4 = a2 @ a2
        a6 = a4 @ a2
        u1 = a6 @ (b[13] * a6 + b[11] * a4 + b[9] * a2)
        u = a @ (u1 + b[7] * a6 + b[5] * a4 + b[3] * a2 + b[1] * ident)
        v1 = a6 @ (b[12] * a6 + b[10] * a4 + b[8] * a2)
        v = v1 + b[6] * a6 + b[4] * a4 + b[2] * a2 + b[0] * ident
        return u, v
    powers = [ident, a2]
    while len(powers) < (m + 1) // 2:
        powers.append(powers[-1] @ a2)
    odd = sum(b[2 * k + 1] * powers[k] for k in range(len(powers)))
    even = sum(b[2 * k] * powers[k] for k in range(len(powers)))
    return a @ odd, even


def pade_plan(a: np.ndarray, spectral: bool = False, deflate: bool = True, cheb: bool = True) -> dict:
    """The decisions of ``pade_expm`` without the exponential: mu, the
    1-norm, m, s (and, for spectral plans, s_onenorm, rho, deflated, c,
    alpha and the unscaled B).  ``cheb``: the spectral bound includes the
    traces of B^14 and B^16 (the engine's Chebyshev plans, its default)."""
    a = np.asarray(a)
    n = a.shape[0]
    mu, ash = trace_shift(a)
    norm = onenorm(ash)
    m, s = choose_degree(norm)
    plan = {"mu": mu, "norm": norm, "m": m, "s": s, "ash": ash, "spectral": False, "deflated": False,
            "s_onenorm": s}
    g = structured_part(a) if (spectral and m == 13 and 0 < s < 120) else None
    if g is not None:
        equal, c = equal_row_sums(g) if deflate else (False, 0.0)
        alpha = (mu.imag - c) / n if equal else 0.0
        b = (g - mu.imag * np.eye(n)) + alpha
        rho = spectral_bound(b, with8=cheb)
        plan.update({"s": scaling_for(rho), "spectral": True, "deflated": bool(equal), "rho": rho,
                     "c": c if equal else 0.0, "alpha": alpha, "B": b})
    return plan


def pade_expm(a: np.ndarray, stages: bool = False, spectral: bool = False, deflate: bool = True,
              cheb: bool = True):
    """exp(a) in FP64 by trace shift + Higham 2005 scaling and squaring.

    With ``spectral=True`` (the engine's default for single runs), inputs of
    the form a = iG (G real symmetric) that need degree 13 with s > 0 take
    s from the spectral-radius bound, after deflating the eigenvector 1
    when all row sums of G are equal (``deflate``); ``cheb`` selects the
    bound of the engine's Chebyshev plans (see ``pade_plan``), while the
    value stays the Pade approximant's.

    With ``stages=True`` returns a dict with mu, norm, m, s, U, V, X (the
    Pade value before squaring) and E (the result); spectral runs add
    s_onenorm, rho, deflated, c, alpha and B (the unscaled shifted matrix).
    """
    a = np.asarray(a)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("expected a square matrix")
    n = a.shape[0]
    if n == 0:
        out = np.zeros((0, 0), dtype=np.complex128)
        return {"E": out, "m": 0, "s": 0, "mu": 0j, "norm": 0.0} if stages else out
    plan = pade_plan(a, spectral, deflate, cheb)
    mu, norm, m, s = plan["mu"], plan["norm"], plan["m"], plan["s"]
    info = {k: plan[k] for k in ("spectral", "deflated", "s_onenorm", "rho", "c", "alpha", "B") if k in plan}
    if plan["spectral"]:
        scaled = 1j * (plan["B"] / (2.0 ** s))
    else:
        scaled = plan["ash"] / (2.0 ** s)
    u, v = pade_uv(scaled, m)
    x = np.linalg.solve(v - u, v + u)
    e = x
    for _ in range(s):
        e = e @ e
    if info["deflated"]:
        e = e + (np.exp(1j * (info["c"] - mu.imag)) - 1.0) / n
    e = np.exp(mu) * e
    if stages:
        out = {"mu": mu, "norm": norm, "m": m, "s": s, "U": u, "V": v, "X": x, "E": e}
        out.update(info)
        return out
    return e


def work_flops(n: int, s: int, m: int = 13) -> float:
    """Algorithmic real flops of one expm in the 4M complex convention:
    products (2 for m=3, 3 for 5, 4 for 7, 5 for 9, 6 for 13) and s
    squarings at 8 n^3 each, plus LU (8/3 n^3) and the n-RHS solve (8 n^3)."""
    products = {3: 2, 5: 3, 7: 4, 9: 5, 13: 6}[m]
    return (8.0 * (products + s) + 32.0 / 3.0) * float(n) ** 3
