"""FP64 restatement of the Chebyshev route of spectral plans (test infrastructure).

Only the tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline
may import this module, and only as the checker.

The reference computes exp(A) with the degree-13 Pade approximant and a
solve (``MatrixExpCalculator<T>::run``, include/MatrixExpCalculator.h:37-61;
restated in ``oracle.padeexpm``).  For the spectral plans of normal inputs
A = iG (G real symmetric), the engine replaces that approximant and its solve
by a polynomial in X = B / 2^s (B = G - mu I + alpha J as in
``padeexpm.pade_plan``, spectral radius at most R, theta = R / 2^s), which
needs products only.  This route is not the reference's algorithm; its
parity target is the reference's output, ``padeexpm.pade_expm``, to the same
tolerances.

* Polynomial: the Chebyshev (Jacobi-Anger) expansion of e^{ix} on
  [-theta, theta] (Abramowitz & Stegun 9.1.44-45),
      cos x = J0(theta) + 2 sum_{k>=1} (-1)^k J2k(theta) T2k(x/theta),
      sin x = 2 sum_{k>=0} (-1)^k J2k+1(theta) T2k+1(x/theta),
  each truncated at the least degree (dc even, ds odd) whose tail bound
  2 sum (theta/2)^m / m! (|Jm(x)| <= (x/2)^m / m!, A&S 9.1.62) is at most
  2^-54, so that for X normal with spectrum in [-theta, theta] the 2-norm
  error is at most 2^-53, like the backward-error bound behind theta_13
  (Higham 2005).  Written in w = x^2 / theta^2: T2k(x/theta) = Tk(2w - 1),
  T2k+1(x/theta) = (x/theta) Wk(w) with W0 = 1, W1 = 4w - 3,
  Wk+1 = 2 (2w - 1) Wk - Wk-1.
* Evaluation: cos X = sum_j cw[j] Y^j and sin X = (B/R) sum_j sw[j] Y^j,
  Y = B^2 / R^2, each by Horner's rule in Y^4 over blocks of four
  coefficients (Paterson & Stockmeyer, SIAM J. Comput. 2(1), 1973) with
  B^2, B^4, B^6, B^8 from the spectral bound; a top block c I makes the
  first Horner step a linear combination.

Pinning: tests/test_oracle.py checks the degrees against exact rational
arithmetic (the tail bound holds at the degree and fails two below), the
coefficients against an independent float64 route (scipy.special.jv and
numpy's Chebyshev-to-power conversion) and the polynomial against e^{ix} on
a grid in 50-digit arithmetic (error at most 2^-53 plus rounding, and larger
than 2^-53 / 16 two degrees lower), ``cheb_cs`` against the eigendecomposition
e^{iX} = V diag(e^{i lambda}) V^T, cos/sin of a diagonal X and X = 0, and the
product count against ``cheb_products``; the engine's host coefficients
(``fexpm_test_cheb_series``) against these.
"""
from __future__ import annotations

import math

import numpy as np

TAIL = 2.0 ** -54


def _tail_bound(theta: float, m0: int) -> float:
    """2 sum_{m >= m0, m - m0 even} (theta/2)^m / m!."""
    h = theta / 2.0
    t = 1.0
    for i in range(1, m0 + 1):
        t *= h / i
    s = 0.0
    for m in range(m0, m0 + 80, 2):
        s += t
        t *= h * h / ((m + 1) * (m + 2))
    return 2.0 * s


def cheb_tail(s: int, fp32_result: bool) -> float:
    """Tail bound for a result squared s times: 2^-54 for FP64 results,
    max(2^-54, 2^-(24 + s)) for FP32 results, so that the truncation error of
    each eigenvalue's factor times the 2^s of the squarings stays at the FP32
    unit roundoff 2^-24 (the engine's rule for complex64 / float32)."""
    return max(TAIL, 2.0 ** -(24 + max(int(s), 0))) if fp32_result else TAIL


def cheb_degrees(theta: float, tail: float = TAIL):
    """(dc, ds): least even / odd degrees with tail bounds at most `tail`
    (2^-54 by default)."""
    dc, ds = 0, 1
    while dc < 78 and _tail_bound(theta, dc + 2) > tail:
        dc += 2
    while ds < 77 and _tail_bound(theta, ds + 2) > tail:
        ds += 2
    return dc, ds


def cheb_coeffs(theta: float, dc: int | None = None, ds: int | None = None):
    """(cw, sw) as float lists: cos x = sum cw[j] w^j, sin x = (x/theta)
    sum sw[j] w^j, w = x^2/theta^2, from 50-digit Bessel values and an exact
    basis change (mpmath); degrees from cheb_degrees unless given."""
    import mpmath as mp

    if dc is None or ds is None:
        dc, ds = cheb_degrees(theta)
    with mp.workdps(50):
        th = mp.mpf(theta)

        def series(nterms, b, t0, t1):
            acc = [mp.mpf(0)] * nterms
            prev, cur = list(t0), list(t1)
            for k in range(nterms):
                t = prev if k == 0 else cur
                for j, v in enumerate(t):
                    if j < nterms:
                        acc[j] += b(k) * v
                if k >= 1:
                    nxt = [mp.mpf(0)] * (len(cur) + 1)
                    for j in range(len(nxt)):
                        c_j = cur[j] if j < len(cur) else 0
                        c_jm = cur[j - 1] if 0 < j <= len(cur) else 0
                        p_j = prev[j] if j < len(prev) else 0
                        nxt[j] = 2 * (2 * c_jm - c_j) - p_j
                    prev, cur = cur, nxt
            return [float(v) for v in acc]

        cw = series(dc // 2 + 1,
                    lambda k: mp.besselj(0, th) if k == 0 else 2 * (-1) ** k * mp.besselj(2 * k, th),
                    [mp.mpf(1)], [mp.mpf(-1), mp.mpf(2)])
        sw = series((ds - 1) // 2 + 1, lambda k: 2 * (-1) ** k * mp.besselj(2 * k + 1, th),
                    [mp.mpf(1)], [mp.mpf(-3), mp.mpf(4)])
    return cw, sw


def horner_products(nc: int) -> int:
    """Products of Horner's rule in Y^4 over nc coefficients in Y (blocks of
    four; a top block of one coefficient costs none)."""
    nb = -(-nc // 4)
    return max(0, nb - 1 - (1 if nb > 1 and nc % 4 == 1 else 0))


def cheb_products(dc: int, ds: int) -> int:
    """Products after the powers B^2 .. B^8: both Horner evaluations and
    B times the sin polynomial."""
    return horner_products(dc // 2 + 1) + horner_products((ds - 1) // 2 + 1) + 1


def cheb_cs(b: np.ndarray, r: float, s: int):
    """(cos X, sin X, products) for X = b / 2^s, b real symmetric with
    spectral radius at most r, evaluated as the engine does (module
    docstring), in float64 matrix products."""
    b = np.asarray(b, dtype=np.float64)
    n = b.shape[0]
    theta = r * 2.0 ** -s
    dc, ds = cheb_degrees(theta)
    cw, sw = cheb_coeffs(theta, dc, ds)
    ident = np.eye(n)
    b2 = b @ b
    b4 = b2 @ b2
    b6 = b4 @ b2
    b8 = b4 @ b4
    pw = [ident, b2, b4, b6]
    rinv = 1.0 / r if r > 0 else 0.0
    r2 = rinv * rinv
    r8 = r2 ** 4
    products = 0

    def series(a):
        nonlocal products
        nc = len(a)

        def block(i):
            return sum(a[4 * i + j] * r2 ** j * pw[j] for j in range(4) if 4 * i + j < nc)

        nb = -(-nc // 4)
        i = nb - 1
        if nb > 1 and nc % 4 == 1:
            i -= 1
            t = a[nc - 1] * r8 * b8 + block(i)
        else:
            t = block(i)
        for i in range(i - 1, -1, -1):
            t = r8 * (b8 @ t) + block(i)
            products += 1
        return t

    c = series(cw)
    sp = series(sw)
    sn = rinv * (b @ sp)
    products += 1
    return c, sn, products
