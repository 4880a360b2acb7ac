"""GPU parity of the fastexpm engine against the FP64 oracle.

Every case prints its name before touching the GPU so that a failure can be
attributed.  Inputs are seeded.  Tolerances (relative Frobenius error of U
against the oracle) are stated per precision:

* complex128 / float64 (FP64 everywhere): 1e-11
* complex64 / float32 (FP32 squarings, FP64 Pade products and solve):
  2e-6 * 2^(s/2), the random-walk growth of FP32 rounding over s squarings
  of a unitary-scaled matrix, floored at 5e-7: 5-10x the errors measured on
  MI355X (about 5e-8 at s = 0, 5e-7 at s = 2, 1e-6 at s = 4, 2e-6 at s = 6,
  4.5e-6 at s = 7-9).

The degree m and scaling s chosen on the device must equal the oracle's
(single runs: the spectral scaling rule of ``pade_expm(spectral=True)``;
batches: the 1-norm rule).
"""
import math
import sys

import numpy as np
import pytest

from fastctqw_amd.graphs import erdos_renyi, laplacian, random_regular
from oracle.closed import complete_graph, cycle_graph, expm_complete, expm_cycle
from oracle.eig import eig_propagator, taylor_expm
from oracle.metrics import infidelity, rel_fro, unitarity_defect
from oracle.padeexpm import THETA, pade_expm, pade_plan, scaling_for

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fx():
    from fastctqw_amd import fastexpm

    return fastexpm


def _engine(fx, dtype, n):
    return {
        np.complex64: fx.ExpMatComplex64,
        np.complex128: fx.ExpMatComplex128,
        np.float32: fx.ExpMatFloat32,
        np.float64: fx.ExpMatFloat64,
    }[dtype](n)


def _tol(dtype, s):
    if dtype in (np.complex128, np.float64):
        return 1e-11
    return max(5e-7, 2e-6 * 2.0 ** (s / 2.0))


def _same_s(stats, st):
    """The device's s equals the oracle's; with the spectral rule a bound
    within rounding (1e-9) of theta_13 times a power of two may round either way."""
    if stats["s"] == st["s"]:
        return True
    if st.get("spectral"):
        x = math.log2(st["rho"] / THETA[13])
        return abs(x - round(x)) < 1e-9 and abs(stats["s"] - st["s"]) == 1
    return False


def _check(fx, a, dtype, name, ref=None, tol=None):
    n = a.shape[0]
    _say(f"{name} n={n} dtype={np.dtype(dtype).name}")
    a = np.ascontiguousarray(a.astype(dtype))
    st = pade_expm(a.astype(np.complex128), stages=True, spectral=True)
    with _engine(fx, dtype, n) as e:
        u = e.run(a)
        stats = e.stats()
    assert u.shape == (n, n) and u.dtype == np.dtype(dtype)
    assert stats["m"] == st["m"], (stats, st["m"], st["s"])
    assert _same_s(stats, st), (stats, st["m"], st["s"])
    assert stats["spectral"] == int(st["spectral"]) and stats["deflated"] == int(st["deflated"]), stats
    target = st["E"] if ref is None else ref
    err = rel_fro(target, u)
    tol = _tol(dtype, st["s"]) if tol is None else tol
    print(f"  m={stats['m']} s={stats['s']} structured={stats['structured']} relF={err:.3e} tol={tol:.1e}",
          flush=True)
    assert err <= tol, err
    return u, stats


# ---------------------------------------------------------------------------
# CTQW form A = -i t H, H real symmetric (structured path)

@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 127, 128, 129, 300])
def test_ctqw_laplacian_er(fx, dt, n):
    L = laplacian(erdos_renyi(n, 0.3, seed=100 + n, weighted=True))
    t = 2.5
    ref = eig_propagator(L, t)
    u, stats = _check(fx, -1j * t * L, dt, "ER-laplacian", ref=ref)
    assert stats["structured"]


def _spectral_case(case):
    n = 400
    if case == "er_unweighted":  # Laplacian, equal row sums (0): deflated
        return 1j * 4.5 * -laplacian(erdos_renyi(n, 0.5, seed=2)), True
    if case == "er_weighted":  # weighted Laplacian: row sums not exactly equal
        return -1j * 20.0 * laplacian(erdos_renyi(n, 0.5, seed=3, weighted=True)), False
    # adjacency of a 6-regular graph, reference sign A_exp = +i t A: row sums 6 t
    return 1j * 4.5 * random_regular(n, 6, seed=5).astype(np.float64), True


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("case", ["er_unweighted", "er_weighted", "regular"])
def test_spectral_scaling(fx, dt, case):
    """s from the spectral-radius bound (option OPT_SPECTRAL, default) with
    the eigenvector 1 deflated when every row sum is equal: the device's m,
    s, deflation and bound against the oracle's same rule; the Pade value
    (no squarings: output e^mu (X + delta J)) against the FP64 stage
    restatement; U against the eigen oracle; the 1-norm rule (option off)
    against the same oracle with its own, larger s.  The approximant before
    the squarings is the Chebyshev polynomial of degree
    max(cheb_degrees(rho / 2^s)) (option OPT_CHEBYSHEV, default); with the
    option off the Pade approximant and its solve give U to the same
    tolerance."""
    from oracle.chebcs import cheb_degrees, cheb_tail

    core = fx._fastexpm_core
    a, defl = _spectral_case(case)
    n = a.shape[0]
    a = np.ascontiguousarray(a.astype(dt))
    a128 = a.astype(np.complex128)
    st = pade_expm(a128, stages=True, spectral=True)
    assert st["spectral"] and st["deflated"] == defl
    ref = eig_propagator(-a128.imag, 1.0)  # expm(i G) for A = i G
    _say(f"spectral scaling {case} n={n} {np.dtype(dt).name}")
    tol = _tol(dt, st["s"])
    with _engine(fx, dt, n) as e:
        u = e.run(a)
        stats = e.stats()
        assert stats["spectral"] == 1 and stats["deflated"] == int(defl), stats
        assert stats["m"] == 13 and _same_s(stats, st), (stats, st["s"], st["rho"])
        assert stats["s_onenorm"] == st["s_onenorm"] and stats["s"] <= stats["s_onenorm"]
        assert abs(stats["rho_bound"] - st["rho"]) <= 1e-9 * st["rho"], (stats["rho_bound"], st["rho"])
        assert stats["chebyshev"] == max(cheb_degrees(stats["rho_bound"] * 2.0 ** -stats["s"],
                                                      cheb_tail(stats["s"], dt == np.complex64))), stats
        err = rel_fro(ref, u)
        print(f"  s={stats['s']} (1-norm rule {stats['s_onenorm']}) rho={stats['rho_bound']:.2f} "
              f"Chebyshev degree {stats['chebyshev']} relF={err:.3e} tol={tol:.1e}", flush=True)
        assert err <= tol
        e._backend.set_option(core.OPT_CHEBYSHEV, 0)
        up = e.run(a)
        assert e.stats()["chebyshev"] == 0
        errp = rel_fro(ref, up)
        print(f"  Pade and solve relF={errp:.3e}", flush=True)
        assert errp <= tol
        e._backend.set_option(core.OPT_CHEBYSHEV, 1)
        # the Pade value before the squarings
        e._backend.set_option(core.OPT_MAX_SQUARINGS, 0)
        x0 = e.run(a)
        xr = st["X"] + ((np.exp(1j * (st["c"] - st["mu"].imag)) - 1.0) / n if defl else 0.0)
        errx = rel_fro(np.exp(st["mu"]) * xr, x0)
        print(f"  Pade value relF={errx:.3e}", flush=True)
        assert errx <= (1e-6 if dt == np.complex64 else 1e-12)
        e._backend.set_option(core.OPT_MAX_SQUARINGS, -1)
        # the 1-norm rule: more squarings, same exponential
        e._backend.set_option(core.OPT_SPECTRAL, 0)
        u1 = e.run(a)
        s1 = e.stats()
        assert s1["spectral"] == 0 and s1["s"] == st["s_onenorm"]
        err1 = rel_fro(ref, u1)
        print(f"  1-norm rule s={s1['s']} relF={err1:.3e}", flush=True)
        assert err1 <= _tol(dt, s1["s"])


@pytest.mark.parametrize("dt,n", [(np.complex64, 300), (np.complex64, 1100), (np.complex128, 300)])
def test_chebyshev_route(fx, dt, n):
    """Spectral plans with the Chebyshev polynomial in place of Pade and
    solve over ER(n, 0.5) Laplacians (s > 0) and complete-graph Laplacians
    (rho(B) = t: s = 0 and small theta): the degree against cheb_degrees,
    the value before the squarings against the oracle's Horner evaluation
    (to the products' rounding: FP64-accurate residue products for
    complex64) and U against the eigendecomposition oracle and against the
    Pade route."""
    from oracle.chebcs import cheb_cs, cheb_degrees, cheb_tail

    core = fx._fastexpm_core
    graphs = {"er": laplacian(erdos_renyi(n, 0.5, seed=21 + n)), "complete": laplacian(complete_graph(n))}
    seen_s0 = False
    with _engine(fx, dt, n) as e:
        for g, t in (("er", 0.05), ("er", 0.15), ("er", 4.5), ("complete", 0.2), ("complete", 0.3),
                     ("complete", 1.0)):
            _say(f"Chebyshev route {np.dtype(dt).name} n={n} {g} t={t}")
            a = np.ascontiguousarray((-1j * t * graphs[g]).astype(dt))
            a128 = a.astype(np.complex128)
            st = pade_expm(a128, stages=True, spectral=True)
            if not st["spectral"]:
                continue
            ref = eig_propagator(-a128.imag, 1.0)
            u = e.run(a)
            stats = e.stats()
            theta = stats["rho_bound"] * 2.0 ** -stats["s"]
            assert stats["spectral"] == 1, stats
            assert stats["chebyshev"] == max(cheb_degrees(theta, cheb_tail(stats["s"], dt == np.complex64))), stats
            assert _same_s(stats, st), (stats["s"], st["s"])
            seen_s0 |= stats["s"] == 0
            err = rel_fro(ref, u)
            e._backend.set_option(core.OPT_MAX_SQUARINGS, 0)
            x0 = e.run(a)
            e._backend.set_option(core.OPT_MAX_SQUARINGS, -1)
            c, sn, _ = cheb_cs(st["B"], st["rho"], st["s"])
            xr = c + 1j * sn + ((np.exp(1j * (st["c"] - st["mu"].imag)) - 1.0) / n if st["deflated"] else 0.0)
            errx = rel_fro(np.exp(st["mu"]) * xr, x0)
            e._backend.set_option(core.OPT_CHEBYSHEV, 0)
            up = e.run(a)
            e._backend.set_option(core.OPT_CHEBYSHEV, 1)
            diff = rel_fro(up, u)
            # the 1-norm plan (option OPT_SPECTRAL = 0): the polynomial at theta = ||A - mu I||_1 / 2^s
            e._backend.set_option(core.OPT_SPECTRAL, 0)
            u1 = e.run(a)
            s1 = e.stats()
            e._backend.set_option(core.OPT_SPECTRAL, 1)
            assert s1["spectral"] == 0, s1
            assert s1["chebyshev"] == max(cheb_degrees(s1["norm"] * 2.0 ** -s1["s"],
                                                       cheb_tail(s1["s"], dt == np.complex64))), s1
            err1 = rel_fro(ref, u1)
            print(f"  s={stats['s']} theta={theta:.3f} degree={stats['chebyshev']} relF={err:.3e} "
                  f"value relF={errx:.3e} vs Pade {diff:.3e}; 1-norm plan s={s1['s']} degree={s1['chebyshev']} "
                  f"relF={err1:.3e}", flush=True)
            tol = _tol(dt, stats["s"])
            assert err <= tol and diff <= tol and err1 <= _tol(dt, s1["s"])
            assert errx <= (1e-6 if dt == np.complex64 else 1e-12)
    assert seen_s0


@pytest.mark.parametrize("weighted", [False, True])
def test_c3_precision(fx, weighted):
    """The precision study at test scale: the dense ER(n, 0.5) Laplacian
    (unweighted, and weighted with w ~ U(0, 1]), t in {1, 4.5, 20}, complex64
    with FP32 (native), int8 residue (int8crt), fp16x2 and bf16x3 squarings.  Per case: m and s as
    the oracle's rule, U against the FP64 eigendecomposition of the same
    rounded input, and the infidelity of U e_0 and of U psi for a random
    state (device apply with FP64 refinement) at most 1e-9."""
    from fastctqw_amd.graphs import c2_laplacian, random_state

    n = 1200
    L = c2_laplacian(n, 0.5, 2, weighted=weighted)
    psi_r = random_state(n, seed=55)
    e0 = np.zeros(n, dtype=np.complex128)
    e0[0] = 1.0
    with fx.ExpMatComplex64(n) as e:
        for t in (1.0, 4.5, 20.0):
            a = np.ascontiguousarray((-1j * t * L).astype(np.complex64))
            a128 = a.astype(np.complex128)
            plan = pade_plan(a128, spectral=True)
            g = a128.imag
            lam, V = np.linalg.eigh(g)
            ref = (V * np.exp(1j * lam)[None, :]) @ V.T
            for prec in ("native", "int8crt", "fp16x2", "bf16x3"):
                _say(f"C3 n={n} weighted={weighted} t={t} {prec}")
                e.set_precision(prec)
                u = e.run(a)
                st = e.stats()
                assert st["m"] == plan["m"] and _same_s(st, plan), (st, plan["s"])
                assert st["deflated"] == int(plan["deflated"]) == int(not weighted)
                psi = e.apply(np.vstack([e0, psi_r]))
                inf0 = infidelity(ref[:, 0], psi[0])
                infr = infidelity(ref @ psi_r, psi[1])
                err = rel_fro(ref, u)
                print(f"  s={st['s']} (1-norm {st['s_onenorm']}) relF={err:.3e} infidelity e0 {inf0:.2e} "
                      f"random {infr:.2e}", flush=True)
                assert err <= _tol(np.complex64, st["s"])
                assert inf0 <= 1e-9 and infr <= 1e-9
        e.set_precision("native")


def test_c2_full_size(fx):
    """N = 10^4 ER(0.5) Laplacian (seed 2), t = 4.5, complex64 through
    run(numpy) (staged 0.8 GB copies, 128 x 128 squaring tiles, int8-slice
    Pade products at k = 10^4): the 1-norm rule's s = 13 from the host
    1-norm, the deflated spectral bound above a power-iteration estimate of
    rho(B) and s from it; U 1 = 1, U^T = U, unitarity of 64 columns; U e_0
    and U psi against the FP64 (complex128) engine; the group law
    U(2.25)^2 = U(4.5) on states."""
    from fastctqw_amd.graphs import c2_laplacian, random_state
    from oracle.padeexpm import choose_degree, onenorm, trace_shift

    n, t = 10000, 4.5
    L = c2_laplacian(n, 0.5, 2)
    a = np.ascontiguousarray((-1j * t * L).astype(np.complex64))
    mu, ash = trace_shift(a)
    m1, s1 = choose_degree(onenorm(ash))
    del ash
    assert (m1, s1) == (13, 13)
    # a lower estimate of rho(B), B = G - mu I + alpha J (G = -t L, row sums 0)
    g_mu = float(mu.imag)
    alpha = g_mu / n
    rng = np.random.default_rng(7)
    v = rng.standard_normal(n)
    for _ in range(40):
        w = -t * (L @ v) - g_mu * v + alpha * v.sum()
        v = w / np.linalg.norm(w)
    w = -t * (L @ v) - g_mu * v + alpha * v.sum()
    rho_low = abs(v @ w)
    psi_r = random_state(n, seed=55)
    e0 = np.zeros(n, dtype=np.complex128)
    e0[0] = 1.0
    ones = np.ones(n, dtype=np.complex128)
    _say(f"C2 full size n={n} complex64")
    with fx.ExpMatComplex64(n) as e:
        u = e.run(a)
        st = e.stats()
        print(f"  s={st['s']} (1-norm {st['s_onenorm']}) rho bound {st['rho_bound']:.1f} >= estimate "
              f"{rho_low:.1f}, ms {st['ms_total']:.1f}", flush=True)
        assert st["m"] == 13 and st["s_onenorm"] == 13 and st["spectral"] == 1 and st["deflated"] == 1
        assert rho_low <= st["rho_bound"] <= 3.0 * rho_low
        assert st["s"] == scaling_for(st["rho_bound"]) and st["s"] >= scaling_for(rho_low)
        assert np.array_equal(u, u.T)
        u1 = u.astype(np.complex128) @ ones
        assert np.linalg.norm(u1 - ones) / np.linalg.norm(ones) < 2e-5
        cols = u[:, :64].astype(np.complex128)
        assert np.linalg.norm(cols.conj().T @ cols - np.eye(64)) / 8.0 < 1e-5
        del u, cols
        psi64 = e.apply(np.vstack([e0, psi_r]))
        e.upload((-1j * (t / 2) * L).astype(np.complex64))
        e.run_resident()
        half2 = e.apply(e.apply(psi_r))
    a128 = a.astype(np.complex128)
    del a
    _say(f"C2 full size n={n} complex128")
    with fx.ExpMatComplex128(n) as e:
        e.upload(a128)
        e.run_resident()
        assert e.stats()["s"] == st["s"]
        psi128 = e.apply(np.vstack([e0, psi_r]))
    for k, name in enumerate(("e0", "random")):
        inf = infidelity(psi128[k], psi64[k])
        print(f"  {name}: infidelity vs complex128 {inf:.2e}", flush=True)
        assert inf <= 1e-9
    inf_g = infidelity(psi128[1], half2)
    print(f"  group law U(2.25)^2 psi: infidelity {inf_g:.2e}", flush=True)
    assert inf_g <= 1e-9


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("case", ["er_unweighted", "er_weighted"])
def test_two_product_squarings(fx, dt, case):
    """Leading squarings in the two-product form (option OPT_TWO_PRODUCT,
    default 0, never the last): the count in the stats, and U against the
    eigen oracle for 0 (3M throughout), 2 and 4 of them; two stay within
    1.5x of the 3M error (plus rounding)."""
    core = fx._fastexpm_core
    n = 600
    L = laplacian(erdos_renyi(n, 0.5, seed=31, weighted=(case == "er_weighted")))
    a = np.ascontiguousarray((-1j * 20.0 * L).astype(dt))
    ref = eig_propagator(-a.astype(np.complex128).imag, 1.0)
    _say(f"two-product squarings {case} n={n} {np.dtype(dt).name}")
    errs = {}
    with _engine(fx, dt, n) as e:
        for k in (0, 2, 4):
            e._backend.set_option(core.OPT_TWO_PRODUCT, k)
            u = e.run(a)
            st = e.stats()
            assert st["s_two_product"] == min(k, st["s"] - 1), st
            errs[k] = rel_fro(ref, u)
            print(f"  two-product {st['s_two_product']} of s={st['s']}: relF={errs[k]:.3e}", flush=True)
            assert errs[k] <= _tol(dt, st["s"])
        e._backend.set_option(core.OPT_TWO_PRODUCT, 0)
        e.run(a)
        assert e.stats()["s_two_product"] == 0  # the default
    floor = 1e-7 if dt == np.complex64 else 1e-14
    assert errs[2] <= 1.5 * errs[0] + floor


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("n", [5, 63, 200, 520])
def test_structured_cholesky_vs_lu(fx, dt, n):
    """The structured solve (real Cholesky of S = V^2 + Uq^2, X = Z Z^T)
    against the pivoted complex LU of the same Pade system (option
    OPT_SPD_SOLVE = 0) and the eigen oracle (Pade route: OPT_CHEBYSHEV = 0)."""
    L = laplacian(erdos_renyi(n, 0.35, seed=300 + n, weighted=True))
    t = 6.0
    a = np.ascontiguousarray((-1j * t * L).astype(dt))
    ref = eig_propagator(L, t)
    out = {}
    for spd in (1, 0):
        _say(f"structured solve spd={spd} n={n} dtype={np.dtype(dt).name}")
        with _engine(fx, dt, n) as e:
            e._backend.set_option(fx._fastexpm_core.OPT_CHEBYSHEV, 0)
            e._backend.set_option(fx._fastexpm_core.OPT_SPD_SOLVE, spd)
            out[spd] = e.run(a)
            st = e.stats()
        assert st["structured"]
        err = rel_fro(ref, out[spd])
        print(f"  spd={spd} s={st['s']} relF={err:.3e}", flush=True)
        assert err <= _tol(dt, st["s"])
    diff = rel_fro(out[0], out[1])
    assert diff <= _tol(dt, st["s"]), diff


@pytest.mark.parametrize("prec", ["fp16x2", "bf16x3"])
@pytest.mark.parametrize("n", [64, 300, 700])
def test_split_precision_ctqw(fx, prec, n):
    """Structured squarings in split 16-bit precision against the eigen
    oracle (same tolerance as FP32) and against the native FP32 engine."""
    L = laplacian(erdos_renyi(n, 0.3, seed=500 + n, weighted=True))
    t = 5.0
    a = np.ascontiguousarray((-1j * t * L).astype(np.complex64))
    ref = eig_propagator(L, t)
    _say(f"split precision {prec} n={n}")
    with fx.ExpMatComplex64(n) as e:
        u_native = e.run(a)
        e.set_precision(prec)
        u = e.run(a)
        st = e.stats()
    assert st["precision"] == prec and st["structured"]
    err, err_native = rel_fro(ref, u), rel_fro(ref, u_native)
    print(f"  s={st['s']} relF={err:.3e} (native {err_native:.3e})", flush=True)
    assert err <= _tol(np.complex64, st["s"])
    rng = np.random.default_rng(n)
    psi0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    psi0 /= np.linalg.norm(psi0)
    assert infidelity(ref @ psi0, u.astype(np.complex128) @ psi0) < 1e-9


@pytest.mark.parametrize("n", [2, 17, 100, 255, 256, 257, 700, 1500, 3000])
def test_int8crt_squarings(fx, n):
    """Structured squarings from exact int8 residue products (each row of Y
    rounded to beta bits of its largest element, products modulo the default
    8 coprime moduli, Chinese-remainder reconstruction): the same m and s as the FP32
    squarings, U against the eigen oracle within the complex64 tolerance and
    no less accurate than the FP32 squarings (relF at most 1.5x theirs), and
    the infidelity of a random state at most 1e-9.  Orders around the 256-row
    tile and the 64-value k padding."""
    L = laplacian(erdos_renyi(n, 0.5, seed=700 + n, weighted=True))
    t = 6.0
    a = np.ascontiguousarray((-1j * t * L).astype(np.complex64))
    lam, V = np.linalg.eigh(np.ascontiguousarray(a.imag, dtype=np.float64))
    ref = (V * np.exp(1j * lam)[None, :]) @ V.T
    _say(f"int8crt squarings n={n}")
    with fx.ExpMatComplex64(n) as e:
        e.set_precision("native")
        u_native = e.run(a)
        st_native = e.stats()
        e.set_precision("int8crt")
        u = e.run(a)
        st = e.stats()
    assert st["structured"] and (st["m"], st["s"]) == (st_native["m"], st_native["s"])
    assert st["precision"] == "int8crt"
    err, err_native = rel_fro(ref, u), rel_fro(ref, u_native)
    print(f"  s={st['s']} relF={err:.3e} (native {err_native:.3e})", flush=True)
    assert err <= _tol(np.complex64, st["s"])
    assert err <= 1.5 * err_native + 1e-7
    rng = np.random.default_rng(n)
    psi0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    psi0 /= np.linalg.norm(psi0)
    assert infidelity(ref @ psi0, u.astype(np.complex128) @ psi0) < 1e-9


def _crt_beta(k_moduli, n):
    """Bits per row of the residue squarings: the largest beta <= 27 with
    n 2^(2 beta + 1) < M / 2, M the product of the first k_moduli of the
    nine moduli (exact integer arithmetic)."""
    mods = (255, 253, 251, 247, 241, 239, 233, 229, 227)
    M = 1
    for m in mods[:k_moduli]:
        M *= m
    beta = 27
    while beta > 0 and 2 * n * 2 ** (2 * beta + 1) >= M:
        beta -= 1
    return beta


@pytest.mark.parametrize("n,weighted,t", [(1500, True, 6.0), (3000, True, 6.0), (3000, False, 20.0)])
def test_crt_moduli(fx, n, weighted, t):
    """Residue squarings with 7, 8 (default) and 9 moduli (option
    OPT_CRT_MODULI): the run reports K and the beta that keeps every product
    below M / 2; U against the eigen oracle with the accuracy each count was
    measured at (ratios to the FP32 MFMA squarings' relF: 9 moduli 0.14-0.21,
    8 moduli 0.32-0.53, 7 moduli 3.5-3.8; infidelity of a random state at most
    7e-13, 8.2e-12 and 4.4e-10 at these orders; bounds about 10x those), the
    error growing as moduli are removed, and counts outside [7, 9] refused."""
    core = fx._fastexpm_core
    L = laplacian(erdos_renyi(n, 0.5, seed=700 + n, weighted=weighted))
    a = np.ascontiguousarray((-1j * t * L).astype(np.complex64))
    lam, V = np.linalg.eigh(np.ascontiguousarray(a.imag, dtype=np.float64))
    ref = (V * np.exp(1j * lam)[None, :]) @ V.T
    rng = np.random.default_rng(n)
    psi0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    psi0 /= np.linalg.norm(psi0)
    _say(f"crt moduli n={n} weighted={weighted} t={t}")
    errs, infs = {}, {}
    with fx.ExpMatComplex64(n) as e:
        e.set_precision("native")
        u_native = e.run(a)
        st_native = e.stats()
        assert st_native["crt_moduli"] == 0 and st_native["crt_beta"] == 0
        e.set_precision("auto")
        u = e.run(a)
        assert e.stats()["crt_moduli"] == 8  # the default
        for bad in (6, 10):
            with pytest.raises(ValueError):
                e._backend.set_option(core.OPT_CRT_MODULI, bad)
        for K in (9, 8, 7):
            e._backend.set_option(core.OPT_CRT_MODULI, K)
            u = e.run(a)
            st = e.stats()
            assert st["precision"] == "int8crt" and (st["m"], st["s"]) == (st_native["m"], st_native["s"])
            assert st["crt_moduli"] == K and st["crt_beta"] == _crt_beta(K, n)
            errs[K] = rel_fro(ref, u)
            infs[K] = infidelity(ref @ psi0, u.astype(np.complex128) @ psi0)
    err_native = rel_fro(ref, u_native)
    print(f"  native relF={err_native:.3e} " +
          " ".join(f"K={K}: relF={errs[K]:.3e} inf={infs[K]:.3e}" for K in (9, 8, 7)), flush=True)
    assert errs[9] < errs[8] < errs[7]
    assert errs[9] <= 0.5 * err_native and infs[9] < 1e-11
    assert errs[8] <= err_native and infs[8] < 1e-10
    assert errs[7] <= 8.0 * err_native and infs[7] < 5e-9


@pytest.mark.parametrize("weighted", [False, True])
def test_crt64_moduli(fx, weighted):
    """Moduli of the FP64-accurate residue products of the Chebyshev stage
    (option OPT_CRT64_MODULI): automatic 11 for spectral plans and 13 for
    1-norm plans (reported as fp64_moduli), explicit counts 8..13 accepted
    and others refused; with 13, 11 and 10 moduli U agrees with the eigen
    oracle to the same accuracy (relF within 1.2x of the 13-moduli run's,
    infidelity of a random state below 1e-10), with 8 moduli (25 bits per
    row at this order) less so."""
    core = fx._fastexpm_core
    n, t = 2000, 4.5
    L = laplacian(erdos_renyi(n, 0.5, seed=31, weighted=weighted))
    a = np.ascontiguousarray((-1j * t * L).astype(np.complex64))
    lam, V = np.linalg.eigh(np.ascontiguousarray(a.imag, dtype=np.float64))
    ref = (V * np.exp(1j * lam)[None, :]) @ V.T
    rng = np.random.default_rng(5)
    psi0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    psi0 /= np.linalg.norm(psi0)
    _say(f"crt64 moduli n={n} weighted={weighted}")
    errs, infs = {}, {}
    with fx.ExpMatComplex64(n) as e:
        u = e.run(a)
        st = e.stats()
        assert st["spectral"] and st["chebyshev"] and st["fp64_slices"] == 13 and st["fp64_moduli"] == 11
        for bad in (7, 14):
            with pytest.raises(ValueError):
                e._backend.set_option(core.OPT_CRT64_MODULI, bad)
        for K in (13, 11, 10, 8):
            e._backend.set_option(core.OPT_CRT64_MODULI, K)
            u = e.run(a)
            assert e.stats()["fp64_moduli"] == K
            errs[K] = rel_fro(ref, u)
            infs[K] = infidelity(ref @ psi0, u.astype(np.complex128) @ psi0)
        e._backend.set_option(core.OPT_CRT64_MODULI, 0)
        e._backend.set_option(core.OPT_SPECTRAL, 0)
        e.run(a)
        st1 = e.stats()
        assert not st1["spectral"] and st1["fp64_moduli"] == 13
    print("  " + " ".join(f"K={K}: relF={errs[K]:.3e} inf={infs[K]:.3e}" for K in (13, 11, 10, 8)), flush=True)
    for K in (11, 10):
        assert errs[K] <= 1.2 * errs[13] and infs[K] < 1e-10
    assert infs[13] < 1e-10
    assert errs[8] > errs[13]


@pytest.mark.parametrize("n", [1500])
def test_crt_block_maps_agree(fx, n):
    """Both block maps of the residue launches (testing option OPT_CRT_MAP)
    give the same U bit for bit (squarings and FP64-accurate products)."""
    core = fx._fastexpm_core
    _say(f"residue block maps n={n}")
    L = laplacian(erdos_renyi(n, 0.5, seed=77))
    a = np.ascontiguousarray((-1j * 4.5 * L).astype(np.complex64))
    with fx.ExpMatComplex64(n) as e:
        u0 = e.run(a)
        assert e.stats()["precision"] == "int8crt"
        e._backend.set_option(core.OPT_CRT_MAP, 1)
        u1 = e.run(a)
        e._backend.set_option(core.OPT_CRT_MAP, 0)
        with pytest.raises(ValueError):
            e._backend.set_option(core.OPT_CRT_MAP, 2)
    assert np.array_equal(u0, u1)


def test_precision_auto_default(fx):
    """Complex64 engines default to the automatic squaring form: residue
    products for 768 <= n <= 16384 on structured input, FP32 MFMA below (and
    for general input, whose squarings are not the structured ones)."""
    for n, want in ((300, "native"), (767, "native"), (768, "int8crt"), (1100, "int8crt")):
        _say(f"precision auto n={n}")
        L = laplacian(erdos_renyi(n, 0.4, seed=900 + n))
        with fx.ExpMatComplex64(n) as e:
            e.run((-1j * 3.0 * L).astype(np.complex64))
            assert e.stats()["precision"] == want, (n, e.stats()["precision"])
            e.set_precision("native")
            e.run((-1j * 3.0 * L).astype(np.complex64))
            assert e.stats()["precision"] == "native"
            e.set_precision("auto")
            rng = np.random.default_rng(n)
            g = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64) / n
            e.run(g)
            assert not e.stats()["structured"] and e.stats()["precision"] == "native"
    with fx.ExpMatComplex128(800) as e:
        e.run((-1j * 3.0 * laplacian(erdos_renyi(800, 0.4, seed=3))).astype(np.complex128))
        assert e.stats()["precision"] == "native"


@pytest.mark.parametrize("slices,scope", [(13, 1), (13, 3), (6, 1), (5, 1), (6, 3), (5, 3)])
@pytest.mark.parametrize("n", [64, 300, 700, 1100])
def test_fp64_slice_products_ctqw(fx, slices, scope, n):
    """Complex64 engine with the symmetric FP64 products on the int8 MFMA:
    residue products modulo 13 moduli (the default) or int8 slices (the Pade
    products by default, scope 3 also S and X = Z Z^T of the solve): U
    against the eigen oracle within the complex64 tolerance, and close to
    the U of the FP64-MFMA products.  Pade route (OPT_CHEBYSHEV = 0), so that
    scope 3 reaches the solve."""
    from fastctqw_amd.fastexpm import _fastexpm_core as core

    _say(f"fp64 slices {slices} scope {scope} n={n}")
    L = laplacian(erdos_renyi(n, 0.5, seed=40 + n, weighted=True))
    a = (-1j * 4.5 * L).astype(np.complex64)
    lam, V = np.linalg.eigh(np.ascontiguousarray(a.imag, dtype=np.float64))
    ref = (V * np.exp(1j * lam)[None, :]) @ V.T
    with fx.ExpMatComplex64(n) as e:
        e._backend.set_option(core.OPT_CHEBYSHEV, 0)
        e.run(a)
        assert e.stats()["fp64_slices"] == 13 and e.stats()["fp64_slices_scope"] == 1  # the default
        e._backend.set_option(core.OPT_FP64_SLICES, 0)
        u64 = e.run(a)
        assert e.stats()["fp64_slices"] == 0
        e._backend.set_option(core.OPT_FP64_SLICES, slices)
        e._backend.set_option(core.OPT_FP64_SLICES_SCOPE, scope)
        u8 = e.run(a)
        st = e.stats()
        assert st["fp64_slices"] == slices and st["fp64_slices_scope"] == scope
    err = rel_fro(ref, u8)
    print(f"  s={st['s']} relF={err:.3e} vs fp64 path {rel_fro(u64, u8):.3e}", flush=True)
    assert err <= _tol(np.complex64, st["s"])
    assert rel_fro(u64, u8) <= _tol(np.complex64, st["s"])
    psi_ref = ref[:, 0]
    assert infidelity(psi_ref, u8[:, 0].astype(np.complex128)) < 1e-9


@pytest.mark.parametrize("n", [4200])
def test_squaring_tail_split(fx, n):
    """FP32 squarings on 128 x 128 tiles: the last partial wave of tiles as
    64 x 64 sub-tiles gives the same U bit for bit as one launch of 128 x 128
    tiles, and as the 64 x 64 tiling throughout (the automatic choice at this
    order); U matches the eigen oracle; invalid tile edges are rejected."""
    from fastctqw_amd.fastexpm import _fastexpm_core as core

    _say(f"squaring tail split n={n}")
    L = laplacian(erdos_renyi(n, 0.02, seed=60 + n, weighted=True))
    a = (-1j * 4.5 * L).astype(np.complex64)
    with fx.ExpMatComplex64(n) as e:
        e._backend.set_option(core.OPT_SQUARE_TILE, 128)
        u_split = e.run(a)
        st = e.stats()
        e._backend.set_option(core.OPT_TAIL_SPLIT, 0)
        u_one = e.run(a)
        e._backend.set_option(core.OPT_TAIL_SPLIT, 1)
        e._backend.set_option(core.OPT_SQUARE_TILE, 64)
        u_64 = e.run(a)
        e._backend.set_option(core.OPT_SQUARE_TILE, 0)
        u_auto = e.run(a)
        with pytest.raises(ValueError):
            e._backend.set_option(core.OPT_SQUARE_TILE, 32)
    assert st["s"] > 0
    assert np.array_equal(u_split, u_one)
    assert np.array_equal(u_split, u_64) and np.array_equal(u_split, u_auto)
    lam, V = np.linalg.eigh(np.ascontiguousarray(a.imag, dtype=np.float64))
    psi = (V * np.exp(1j * lam)[None, :]) @ (V.T[:, :1])
    err = float(np.linalg.norm(psi[:, 0] - u_split[:, 0]) / np.linalg.norm(psi))
    print(f"  s={st['s']} column-0 relative error {err:.2e}", flush=True)
    assert err <= _tol(np.complex64, st["s"])


@pytest.mark.parametrize("dt,n", [(np.complex64, 700), (np.complex128, 330)])
def test_cholesky_z_overlap(fx, dt, n):
    """Z updates of the Cholesky solve on a second stream give the same U bit
    for bit as one stream (Pade route: OPT_CHEBYSHEV = 0)."""
    from fastctqw_amd.fastexpm import _fastexpm_core as core

    _say(f"z overlap {np.dtype(dt).name} n={n}")
    L = laplacian(erdos_renyi(n, 0.3, seed=70 + n, weighted=True))
    a = (-1j * 6.0 * L).astype(dt)
    with _engine(fx, dt, n) as e:
        e._backend.set_option(core.OPT_CHEBYSHEV, 0)
        u_two = e.run(a)
        e._backend.set_option(core.OPT_Z_OVERLAP, 0)
        u_one = e.run(a)
        e._backend.set_option(core.OPT_Z_OVERLAP, 1)
        u_two2 = e.run(a)
    assert np.array_equal(u_two, u_one) and np.array_equal(u_two, u_two2)


@pytest.mark.parametrize("dt,n", [(np.complex64, 700), (np.complex128, 330)])
def test_cholesky_outer_width(fx, dt, n):
    """Outer block widths of the Cholesky solve (64 .. 1024, wider than the
    order included) change only the FP64 summation order: every width agrees
    with the default to FP64 rounding carried through the squarings, and
    with the eigendecomposition oracle; invalid widths are rejected (Pade
    route: OPT_CHEBYSHEV = 0)."""
    from fastctqw_amd.fastexpm import _fastexpm_core as core

    _say(f"cholesky outer width {np.dtype(dt).name} n={n}")
    L = laplacian(erdos_renyi(n, 0.3, seed=90 + n, weighted=True))
    a = (-1j * 6.0 * L).astype(dt)
    lam, V = np.linalg.eigh(np.ascontiguousarray(a.imag, dtype=np.float64))
    psi = (V * np.exp(1j * lam)[None, :]) @ (V.T[:, :1])
    with _engine(fx, dt, n) as e:
        e._backend.set_option(core.OPT_CHEBYSHEV, 0)
        u_def = e.run(a)
        s = e.stats()["s"]
        for w in (64, 256, 1024, 512):
            e._backend.set_option(core.OPT_CHOL_OUTER, w)
            u = e.run(a)
            assert rel_fro(u, u_def) <= (2e-6 if dt == np.complex64 else 1e-12), w
            err = float(np.linalg.norm(psi[:, 0] - u[:, 0]) / np.linalg.norm(psi))
            assert err <= _tol(dt, s), (w, err)
        for bad in (0, 96, 8192):
            with pytest.raises(ValueError):
                e._backend.set_option(core.OPT_CHOL_OUTER, bad)


@pytest.mark.parametrize("dt,n", [(np.complex64, 700), (np.complex128, 650)])
def test_cholesky_left_looking_z(fx, dt, n):
    """Left-looking Z updates (one long-K product per outer block): the same
    values with and without the second stream, agreement with the
    right-looking form to FP64 rounding carried through the squarings, and
    with the eigendecomposition oracle; small outer blocks give several
    (Pade route: OPT_CHEBYSHEV = 0)."""
    from fastctqw_amd.fastexpm import _fastexpm_core as core

    _say(f"left-looking Z {np.dtype(dt).name} n={n}")
    L = laplacian(erdos_renyi(n, 0.3, seed=110 + n, weighted=True))
    a = (-1j * 6.0 * L).astype(dt)
    lam, V = np.linalg.eigh(np.ascontiguousarray(a.imag, dtype=np.float64))
    psi = (V * np.exp(1j * lam)[None, :]) @ (V.T[:, :1])
    with _engine(fx, dt, n) as e:
        e._backend.set_option(core.OPT_CHEBYSHEV, 0)
        e._backend.set_option(core.OPT_CHOL_OUTER, 128)
        u_right = e.run(a)
        s = e.stats()["s"]
        e._backend.set_option(core.OPT_Z_LEFT, 1)
        u_left = e.run(a)
        e._backend.set_option(core.OPT_Z_OVERLAP, 0)
        u_left1 = e.run(a)
        e._backend.set_option(core.OPT_Z_OVERLAP, 1)
        e._backend.set_option(core.OPT_Z_LEFT, 0)
        e._backend.set_option(core.OPT_CHOL_OUTER, 512)
    assert np.array_equal(u_left, u_left1)
    assert rel_fro(u_left, u_right) <= (2e-6 if dt == np.complex64 else 1e-12)
    err = float(np.linalg.norm(psi[:, 0] - u_left[:, 0]) / np.linalg.norm(psi))
    assert err <= _tol(dt, s), err


def test_fp64_slices_option_errors(fx):
    from fastctqw_amd.fastexpm import _fastexpm_core as core

    with fx.ExpMatComplex128(8) as e:
        with pytest.raises(ValueError):
            e._backend.set_option(core.OPT_FP64_SLICES, 5)
    with fx.ExpMatComplex64(8) as e:
        with pytest.raises(ValueError):
            e._backend.set_option(core.OPT_FP64_SLICES, 9)


def test_split_precision_options(fx):
    _say("split precision options")
    with fx.ExpMatComplex128(8) as e:
        with pytest.raises(Exception):
            e.set_precision("fp16x2")
    with fx.ExpMatComplex64(8) as e:
        with pytest.raises(ValueError):
            e.set_precision("fp8")
        e.set_precision("bf16x3")
        e.set_precision("native")
        e.run(np.zeros((8, 8), np.complex64))
        assert e.stats()["precision"] == "native"


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("t", [0.001, 0.05, 0.3, 1.0, 4.5, 20.0])
def test_ctqw_degrees(fx, dt, t):
    """Small t exercises Pade degrees 3..9 without scaling."""
    a = random_regular(96, 4, seed=7).astype(np.float64)
    ref = eig_propagator(-a, t)
    _check(fx, 1j * t * a, dt, f"R4-adjacency t={t}", ref=ref)


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_c1_four_regular_walk(fx, dt):
    """4-regular graph, N=1000, t=4.5: U and psi against the eigen oracle."""
    n, t = 1000, 4.5
    a = random_regular(n, 4, seed=1)
    H = a  # configuration "H = adjacency"; A_exp = -i t H
    ref = eig_propagator(H, t)
    u, stats = _check(fx, -1j * t * H, dt, "C1 4-regular N=1000", ref=ref)
    assert stats["structured"]
    # uniform superposition is an eigenvector: psi(t) = e^{-i 4 t} psi0
    psi0 = np.full(n, 1 / np.sqrt(n))
    psi = u.astype(np.complex128) @ psi0
    assert infidelity(np.exp(-4j * t) * psi0, psi) < 1e-10
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    phi /= np.linalg.norm(phi)
    inf = infidelity(ref @ phi, u.astype(np.complex128) @ phi)
    print(f"  random-state infidelity {inf:.3e}", flush=True)
    assert inf < 1e-9


def test_closed_forms(fx):
    for n in (5, 64, 200):
        a = 1j * 0.7 * complete_graph(n)
        _check(fx, a, np.complex128, "complete", ref=expm_complete(n, 0.7))
        _check(fx, 1j * 1.3 * cycle_graph(n), np.complex128, "cycle", ref=expm_cycle(n, 1.3))


# ---------------------------------------------------------------------------
# Reference test inputs (tests/test_fastexpm_accuracy.py:83-159):
# A = 1j * adjacency for complete, ER(0.05) and 4-regular graphs, checked
# against the reference's own Taylor oracle restated.

@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("n", [16, 64, 256])
@pytest.mark.parametrize("family", ["complete", "er", "r4"])
def test_reference_accuracy_inputs(fx, dt, n, family):
    """Against the exact closed form (complete graph) or the eigen oracle,
    and against the reference's Taylor oracle, whose own FP64 error grows
    with its ceil(log2 ||A||_1) squarings (1e-9 allowed for FP64 inputs)."""
    if family == "complete":
        adj = complete_graph(n)
        exact = expm_complete(n, 1.0)
    elif family == "er":
        adj = erdos_renyi(n, 0.05, seed=n)
        exact = eig_propagator(-adj, 1.0)
    else:
        adj = random_regular(n, 4, seed=n)
        exact = eig_propagator(-adj, 1.0)
    a = 1j * adj
    u, stats = _check(fx, a, dt, f"ref-{family}", ref=exact)
    assert stats["structured"]
    tay = taylor_expm(a, m=100)
    assert rel_fro(tay, u) <= max(1e-9, _tol(dt, stats["s"]))


# ---------------------------------------------------------------------------
# General complex A (not purely imaginary / not symmetric)

@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("n", [1, 5, 64, 129, 250])
def test_general_complex(fx, dt, n):
    rng = np.random.default_rng(n)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * (3.0 / np.sqrt(n))
    _, stats = _check(fx, a, dt, "general-complex")
    assert not stats["structured"]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n", [4, 100])
def test_real_dts(fx, dt, n):
    rng = np.random.default_rng(11 + n)
    a = rng.standard_normal((n, n)) * (2.0 / np.sqrt(n))
    _check(fx, a, dt, "real")


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("family", ["complete", "er", "r4"])
@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024])
def test_reference_perf_inputs(fx, dt, family, n):
    """The reference's performance inputs A = Adj + 1j * rand(n, n) at its
    default sizes (tests/test_fastexpm_perf.py:10-59, conftest.py:10-16):
    general complex matrices on the LU route.  Where the exact result
    overflows the working precision (the complete graph: |e^A| ~ e^(n-1),
    from n = 128 in complex64 and n = 1024 in complex128) every entry comes
    out non-finite, as from the reference's products, and the call returns
    normally.  Elsewhere U matches the FP64 oracle: measured relF up to
    4e-7 (complex64) and 2.4e-14 (complex128), tolerances 5-10x that."""
    if dt == np.complex128 and n not in (16, 128, 1024):
        pytest.skip("complex128 at three of the sizes")
    from fastctqw_amd.graphs import reference_perf_input

    a = reference_perf_input(family, n, seed=n, dtype=dt)
    overflow = family == "complete" and (n - 1) - math.log(n) > (88.0 if dt == np.complex64 else 709.0)
    if overflow:
        _say(f"perf input {family} n={n} dtype={np.dtype(dt).name} (overflows)")
        with _engine(fx, dt, n) as e:
            u = e.run(a)
            stats = e.stats()
        assert not stats["structured"] and not stats["nonfinite"]
        assert not np.isfinite(u).any()
        return
    _, stats = _check(fx, a, dt, f"perf input {family}", tol=2e-6 if dt == np.complex64 else 2e-13)
    assert not stats["structured"]


def test_reference_perf_inputs_complex(fx):
    """Perf-test inputs (tests/test_fastexpm_perf.py:14-18): Adj + 1j*rand."""
    rng = np.random.default_rng(0)
    n = 48
    a = complete_graph(n) + 1j * rng.random((n, n))
    _check(fx, a, np.complex128, "perf-complete-general")


# ---------------------------------------------------------------------------
# Invariants at a size the oracle does not reach cheaply

def test_group_law_and_unitarity(fx):
    n = 700
    L = laplacian(erdos_renyi(n, 0.5, seed=3))
    _say("group law n=700")
    with fx.ExpMatComplex128(n) as e:
        u1 = e.run(-1j * 1.0 * L)
        u2 = e.run(-1j * 2.0 * L)
        u3 = e.run(-1j * 3.0 * L)
    assert rel_fro(u3, u1 @ u2) < 1e-11
    assert unitarity_defect(u3) < 1e-12
    ones = np.ones(n)
    assert np.allclose(u3 @ ones, ones, atol=1e-10)  # Laplacian: U 1 = 1


# ---------------------------------------------------------------------------
# Edge cases and API behaviour

def test_zero_matrix_and_nonfinite(fx):
    _say("zero / nonfinite")
    with fx.ExpMatComplex64(10) as e:
        u = e.run(np.zeros((10, 10), dtype=np.complex64))
        assert np.array_equal(u, np.eye(10, dtype=np.complex64))
        bad = np.zeros((10, 10), dtype=np.complex64)
        bad[3, 4] = np.nan
        u = e.run(bad)
        assert np.isnan(u).all()
        assert e.stats()["nonfinite"]


def test_fp32_pade_path_and_device_bytes(fx):
    """Complex64 engines run the Pade products and the solve on their FP64
    companion by default; FEXPM_OPT_PADE_FP64 = 0 runs them in the engine's
    own FP32 planes, which are allocated at the first such run.  Both routes
    match the oracle, and device_bytes follows the allocations."""
    from fastctqw_amd.fastexpm import _fastexpm_core as core

    n = 300
    _say(f"fp32 pade route n={n}")
    g = (-1j * 1.3 * laplacian(erdos_renyi(n, 0.3, seed=31))).astype(np.complex64)
    rng = np.random.default_rng(32)
    h = ((rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * (2.0 / np.sqrt(n))).astype(np.complex64)
    with fx.ExpMatComplex64(n) as e:
        b0 = e.stats()["device_bytes"]
        assert b0 > 0
        e.run(g)
        b1 = e.stats()["device_bytes"]
        assert b1 > b0  # the FP64 companion
        e._backend.set_option(core.OPT_PADE_FP64, 0)
        for name, a in (("structured", g), ("general", h)):
            st = pade_expm(a.astype(np.complex128), stages=True, spectral=True)
            u = e.run(a)
            stats = e.stats()
            assert stats["pade_fp64"] == 0
            err = rel_fro(st["E"], u)
            tol = 4.0 * _tol(np.complex64, stats["s"])
            print(f"  {name}: m={stats['m']} s={stats['s']} relF={err:.3e} tol={tol:.1e}", flush=True)
            assert err <= tol, (name, err)
        b2 = e.stats()["device_bytes"]
        assert b2 > b1  # the FP32 Pade and solve buffers
        e.free()
        assert e.stats()["device_bytes"] == 0


def test_free_idempotent_and_errors(fx):
    _say("free / errors")
    e = fx.ExpMatComplex64(8)
    with pytest.raises(TypeError):
        e.run([[0.0]])
    with pytest.raises(ValueError):
        e.run(np.zeros((7, 7), dtype=np.complex64))
    e.free()
    e.free()
    with pytest.raises(RuntimeError):
        e.run(np.zeros((8, 8), dtype=np.complex64))


def test_input_not_modified_and_reuse(fx):
    _say("reuse")
    n = 40
    a = (-1j * 0.9 * laplacian(erdos_renyi(n, 0.4, seed=9))).astype(np.complex64)
    keep = a.copy()
    with fx.ExpMatComplex64(n) as e:
        u1 = e.run(a)
        u2 = e.run(a)
    assert np.array_equal(a, keep)
    assert np.array_equal(u1, u2)


def test_device_resident_and_apply(fx):
    _say("resident + apply")
    n = 333
    L = laplacian(erdos_renyi(n, 0.2, seed=4, weighted=True))
    t = 3.0
    a = (-1j * t * L).astype(np.complex64)
    with fx.ExpMatComplex64(n) as e:
        e.upload(a)
        e.run_resident()
        u = e.download()
        psi0 = np.zeros((2, n), dtype=np.complex128)
        psi0[0, 0] = 1.0
        psi0[1] = 1 / np.sqrt(n)
        psi = e.apply(psi0)
    ref = eig_propagator(L, t)
    assert rel_fro(ref, u) < _tol(np.complex64, 10)
    assert infidelity(ref[:, 0], psi[0]) < 1e-9
    assert np.allclose(psi[0], u[:, 0].astype(np.complex128), atol=1e-6)
    # one-step FP64 refinement (unitary U): every psi keeps the norm of its psi0
    for k in range(2):
        assert abs(np.linalg.norm(psi[k]) - np.linalg.norm(psi0[k])) < 1e-12
    # unnormalised infidelity 1 - |<ref|psi>|^2 equals the normalised one to rounding
    raw = 1.0 - abs(np.vdot(ref[:, 0], psi[0])) ** 2
    assert abs(raw - infidelity(ref[:, 0], psi[0])) < 1e-12


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_propagate_states(fx, dt):
    """Many times x many states (the C5 form): psi_k = U(dt)^k psi0 against
    the eigen oracle of the same rounded input, every k and every state."""
    from fastctqw_amd.graphs import random_state

    n, h, nsteps, ns = 300, 0.5, 6, 5
    _say(f"propagate n={n} steps={nsteps} states={ns} {np.dtype(dt).name}")
    L = laplacian(erdos_renyi(n, 0.3, seed=5, weighted=True))
    a = (-1j * h * L).astype(dt)
    G = np.ascontiguousarray(a.imag, dtype=np.float64)
    lam, V = np.linalg.eigh(G)
    psi0 = np.stack([random_state(n, seed=55, index=k) for k in range(ns)])
    with _engine(fx, dt, n) as e:
        e.run(a)
        out = e.propagate(psi0, nsteps)
        assert e.stats()["ms_propagate"] > 0
    assert out.shape == (nsteps, ns, n)
    tol = 1e-9 if dt == np.complex64 else 1e-13  # FP64 floor of 1 - |<a|b>|^2 is ~1e-16
    worst = 0.0
    for k in range(1, nsteps + 1):
        ref = (V * np.exp(1j * k * lam)[None, :]) @ (V.T @ psi0.T)
        for sidx in range(ns):
            worst = max(worst, infidelity(ref[:, sidx], out[k - 1, sidx]))
            assert abs(np.linalg.norm(out[k - 1, sidx]) - 1.0) < 1e-12
    print(f"  worst infidelity {worst:.2e}", flush=True)
    assert worst < tol


def test_batch_uneven_walks(fx):
    """Many walks of uneven order, graph family and time in one batch (the
    C4 form, small orders): every U and psi against the eigen oracle of the
    same rounded input; lanes reuse one pipeline across orders."""
    from fastctqw_amd.graphs import c4_walks

    walks = c4_walks(count=10, seed=4100, nmin=16, nmax=300)
    _say(f"batch of {len(walks)} walks, n in {[w['n'] for w in walks]}")
    b = fx.ExpMatBatch(nmax=300, dtype=np.complex64, lanes=3)
    mats = [w["A"].astype(np.complex64) for w in walks]
    psi0 = [np.full(w["n"], 1 / np.sqrt(w["n"]), dtype=complex) for w in walks]
    us, psis = b.run(mats, psi0)
    for i, (w, u, psi) in enumerate(zip(walks, us, psis)):
        G = np.ascontiguousarray(mats[i].imag, dtype=np.float64)
        lam, V = np.linalg.eigh(G)
        ref = (V * np.exp(1j * lam)[None, :]) @ V.T
        st = pade_expm(mats[i].astype(np.complex128), stages=True)
        assert b.stats(i)["s"] == st["s"] and b.stats(i)["m"] == st["m"]
        err = rel_fro(ref, u)
        assert err <= _tol(np.complex64, st["s"]), (i, w["kind"], err)
        assert infidelity(ref @ psi0[i], psi) < 1e-9
    # the same walks with fp16x2 split-precision squarings
    b.set_precision("fp16x2")
    us_h, _ = b.run(mats, psi0)
    for i, u in enumerate(us_h):
        G = np.ascontiguousarray(mats[i].imag, dtype=np.float64)
        lam, V = np.linalg.eigh(G)
        ref = (V * np.exp(1j * lam)[None, :]) @ V.T
        assert rel_fro(ref, u) <= _tol(np.complex64, b.stats(i)["s"]), i
        assert b.stats(i)["precision"] in ("fp16x2", "native")  # s = 0 walks have no squaring
    b.set_precision("native")
    # an order above the capacity is rejected
    with pytest.raises(RuntimeError):
        b.upload([np.zeros((301, 301), np.complex64)])


def test_batch_large_upload(fx):
    """A batch whose packed matrices exceed 64 MiB (three walks of n = 1700,
    69 MB of complex64): the upload goes through the pinned staging buffers;
    every U and psi against the eigen oracle."""
    n = 1700
    _say(f"batch of 3 walks n={n} (staged upload)")
    mats = []
    for q in range(3):
        L = laplacian(erdos_renyi(n, 0.03 + 0.02 * q, seed=80 + q, weighted=True))
        mats.append((-1j * (1.0 + q) * L).astype(np.complex64))
    b = fx.ExpMatBatch(nmax=n, dtype=np.complex64)
    psi0 = [np.full(n, 1 / np.sqrt(n), dtype=complex) for _ in mats]
    us, psis = b.run(mats, psi0)
    for i, (u, psi) in enumerate(zip(us, psis)):
        ref = _eig_ref(mats[i])
        assert rel_fro(ref, u) <= _tol(np.complex64, b.stats(i)["s"]), i
        assert infidelity(ref @ psi0[i], psi) < 1e-9


def test_batch_upload_recovery(fx):
    """A rejected upload (an order above nmax) leaves the previous batch as
    it was; a failed allocation (injected on the host after the old matrix
    buffers are released) leaves an empty batch, and the next upload
    allocates again and runs."""
    core = fx._fastexpm_core
    mats = [(-1j * 2.0 * laplacian(erdos_renyi(n, 0.3, seed=90 + n, weighted=True))).astype(np.complex64)
            for n in (40, 70)]
    _say("batch upload recovery")
    b = fx.ExpMatBatch(nmax=128, dtype=np.complex64, lanes=2)
    us = b.run(mats)
    with pytest.raises(RuntimeError):
        b.upload([np.zeros((129, 129), np.complex64)])
    assert b.count() == 2
    u0 = b.download(0)
    assert u0.shape == (40, 40) and np.array_equal(u0, us[0])
    b.run_resident()
    assert rel_fro(us[1], b.download(1)) <= 1e-6
    big = [(-1j * laplacian(erdos_renyi(128, 0.3, seed=99 + q, weighted=True))).astype(np.complex64)
           for q in range(3)]
    b._backend.set_option(core.OPT_FAIL_ALLOC, 1)
    with pytest.raises((RuntimeError, MemoryError)):
        b.upload(big)
    assert b.count() == 0
    with pytest.raises(ValueError):
        b.download(0)
    us2 = b.run(mats)
    for u, v, a in zip(us, us2, mats):
        assert u.shape == v.shape and rel_fro(u, v) <= 1e-6
        assert rel_fro(_eig_ref(a), v) <= _tol(np.complex64, 8)
    us3 = b.run(big)
    for u, a in zip(us3, big):
        assert rel_fro(_eig_ref(a), u) <= _tol(np.complex64, 8)


def _eig_ref(a):
    G = np.ascontiguousarray(np.asarray(a).imag, dtype=np.float64)
    lam, V = np.linalg.eigh(G)
    return (V * np.exp(1j * lam)[None, :]) @ V.T


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_batch_chebyshev_route(fx, dt):
    """Grouped batches take the Chebyshev polynomial of each walk at theta =
    ||A - mu I||_1 / 2^s (the 1-norm plan's bound) in place of the Pade
    products and the solve (default); the per-walk degree against
    oracle.chebcs; U against the eigen oracle and against the Pade route
    (option OPT_CHEBYSHEV = 0) of the same batch."""
    from fastctqw_amd.graphs import c4_walks
    from oracle.chebcs import cheb_degrees, cheb_tail

    core = fx._fastexpm_core
    walks = c4_walks(count=10, seed=4300, nmin=16, nmax=400)
    ts = [1e-3, 0.05, 0.4, 1.5, 3.0, 6.0, 9.0, 14.0, 20.0, 0.8]
    mats = [np.ascontiguousarray((-1j * t * w["H"]).astype(dt)) for w, t in zip(walks, ts)]
    _say(f"batch Chebyshev route {np.dtype(dt).name}: n = {[m.shape[0] for m in mats]}")
    b = fx.ExpMatBatch(nmax=400, dtype=dt, lanes=1)
    us = b.run(mats)
    tol_fn = (lambda s: 1e-11) if dt == np.complex128 else (lambda s: _tol(np.complex64, s))
    for i, (a, u) in enumerate(zip(mats, us)):
        st = b.stats(i)
        ref_st = pade_expm(a.astype(np.complex128), stages=True)
        assert st["m"] == ref_st["m"] and st["s"] == ref_st["s"], (i, st)
        assert st["chebyshev"] == max(cheb_degrees(st["norm"] * 2.0 ** -st["s"],
                                                   cheb_tail(st["s"], dt == np.complex64))), (i, st)
        err = rel_fro(_eig_ref(a), u)
        print(f"  walk {i} n={a.shape[0]} s={st['s']} degree={st['chebyshev']} relF={err:.2e}", flush=True)
        assert err <= tol_fn(st["s"]), (i, err)
    b._backend.set_option(core.OPT_CHEBYSHEV, 0)
    up = b.run(mats)
    for i in range(len(mats)):
        assert b.stats(i)["chebyshev"] == 0
        assert rel_fro(up[i], us[i]) <= 2 * tol_fn(b.stats(i)["s"]), i
    b._backend.set_option(core.OPT_CHEBYSHEV, 1)


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_batch_grouped(fx, dt):
    """Grouped launches over a batch of uneven walks: every Pade degree
    (t from 1e-4 to 20), orders 1 .. 300 around the 64-wide padding, a
    general complex matrix and a non-finite one (both leave the grouped
    program for a lane).  Each U against the eigen oracle of the same
    rounded input (general: the Pade oracle), psi against U psi0, the
    grouped results against the lane results, bitwise replay of the
    captured graph, re-planning after new times, eager launches equal to the
    graph, and default tiles without the autotuner."""
    from fastctqw_amd.graphs import c4_walks

    rng = np.random.default_rng(77)
    walks = c4_walks(count=8, seed=4200, nmin=16, nmax=300)
    mats, kinds = [], []
    for w, t in zip(walks, [1e-4, 0.01, 0.2, 1.0, 3.0, 7.5, 12.0, 20.0]):
        mats.append((-1j * t * w["H"]).astype(dt))
        kinds.append("ctqw")
    for n, t in [(1, 2.0), (63, 4.5), (64, 4.5), (65, 4.5), (129, 0.7)]:
        L = laplacian(erdos_renyi(n, 0.4, seed=900 + n, weighted=True))
        mats.append((-1j * t * L).astype(dt))
        kinds.append("ctqw")
    g = (rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40))) * 0.3
    mats.append(g.astype(dt))
    kinds.append("general")
    bad = (-1j * laplacian(erdos_renyi(20, 0.5, seed=5))).astype(dt)
    bad[3, 4] = np.nan
    mats.append(bad)
    kinds.append("nonfinite")
    ns = [m.shape[0] for m in mats]
    _say(f"grouped batch {np.dtype(dt).name}: n = {ns}")
    psi0 = [np.exp(1j * np.arange(n) * 0.37) / np.sqrt(n) for n in ns]
    b = fx.ExpMatBatch(nmax=300, dtype=dt, lanes=2)
    us, psis = b.run(mats, psi0)
    rep = b.report()
    print(f"  report {rep}", flush=True)
    assert rep["grouped"] == 1 and rep["tuned"] == 1 and rep["graph"] == 1 and rep["graph_built"] == 1
    assert rep["grouped_walks"] == len(mats) - 2 and rep["fallback_walks"] == 2
    tol_fn = (lambda s: 1e-11) if dt == np.complex128 else (lambda s: _tol(np.complex64, s))
    for i, (a, u, psi, kind) in enumerate(zip(mats, us, psis, kinds)):
        st = b.stats(i)
        if kind == "nonfinite":
            assert st["nonfinite"] == 1 and np.isnan(u).all()
            continue
        ref_st = pade_expm(a.astype(np.complex128), stages=True)
        assert st["m"] == ref_st["m"] and st["s"] == ref_st["s"], (i, st, ref_st["m"], ref_st["s"])
        ref = ref_st["E"] if kind == "general" else _eig_ref(a)
        err = rel_fro(ref, u)
        assert err <= tol_fn(ref_st["s"]), (i, kind, ns[i], err)
        assert st["structured"] == (1 if kind == "ctqw" else 0)
        psi_ref = u.astype(np.complex128) @ psi0[i]
        assert np.linalg.norm(psi - psi_ref) <= 1e-5 * max(1.0, np.linalg.norm(psi_ref)), i
        if kind == "ctqw":  # one-step FP64 refinement: psi keeps the norm of psi0
            assert abs(np.linalg.norm(psi) - np.linalg.norm(psi0[i])) < 1e-12, i
    # replay: same plans -> the captured graph again, bitwise the same U
    us2, psis2 = b.run(mats, psi0)
    rep2 = b.report()
    assert rep2["graph"] == 1 and rep2["graph_built"] == 0 and rep2["tuned"] == 0, rep2
    for i in range(len(mats)):
        if kinds[i] != "nonfinite":
            assert np.array_equal(us[i], us2[i]), i
    # lanes on the same inputs agree within the precision's tolerance
    b.set_grouped(False)
    us_l, psis_l = b.run(mats, psi0)
    assert b.report()["grouped"] == 0
    for i in range(len(mats)):
        if kinds[i] == "nonfinite":
            continue
        s_i = b.stats(i)["s"]
        assert rel_fro(us_l[i], us[i]) <= 2 * tol_fn(s_i), i
        if kinds[i] == "ctqw":
            assert abs(np.linalg.norm(psis_l[i]) - np.linalg.norm(psi0[i])) < 1e-12, i
    # new times (same orders): new plans, a new graph, the layout keeps its tuning
    b.set_grouped(True)
    mats3 = [m * (0.5 if k == "ctqw" else 1.0) for m, k in zip(mats, kinds)]
    us3, _ = b.run(mats3, psi0)
    rep3 = b.report()
    assert rep3["graph_built"] == 1 and rep3["tuned"] == 0, rep3
    for i, (a, u, kind) in enumerate(zip(mats3, us3, kinds)):
        if kind != "ctqw":
            continue
        assert rel_fro(_eig_ref(a), u) <= tol_fn(b.stats(i)["s"]), i
    # eager launches of the same program give the graph's results bitwise
    b.set_grouped(True, autotune=True, graph=False)
    us4, _ = b.run(mats3, psi0)
    assert b.report()["graph"] == 0
    for i in range(len(mats)):
        if kinds[i] == "ctqw":
            assert np.array_equal(us3[i], us4[i]), i
    # a fresh batch without the autotuner: default tiles, same accuracy
    b2 = fx.ExpMatBatch(nmax=300, dtype=dt, lanes=1)
    b2.set_grouped(True, autotune=False, graph=True)
    us5 = b2.run(mats[:8])
    assert b2.report()["tuned"] == 0 and b2.report()["candidate"] == [0, 0, 0]
    for i in range(8):
        assert rel_fro(_eig_ref(mats[i]), us5[i]) <= tol_fn(b2.stats(i)["s"]), i
    # every autotuner candidate (tile shape x dispatch) on its own
    from fastctqw_amd.fastexpm import _fastexpm_core as core

    ctqw = [i for i, k in enumerate(kinds) if k == "ctqw"]
    for c in range(6):
        b2._backend.set_option(core.OPT_GROUP_CANDIDATE, c)
        us_c = b2.run([mats[i] for i in ctqw])
        assert b2.report()["candidate"] == [c, c, c]
        for q, i in enumerate(ctqw):
            assert rel_fro(_eig_ref(mats[i]), us_c[q]) <= tol_fn(b2.stats(q)["s"]), (c, i)
    b2._backend.set_option(core.OPT_GROUP_CANDIDATE, -1)
    if dt == np.complex64:
        # Pade products from int8 slices (the complex64 default) against FP64 MFMA products
        assert b2.stats(0)["fp64_slices"] == 5
        b2._backend.set_option(core.OPT_FP64_SLICES, 0)
        us_f = b2.run([mats[i] for i in ctqw])
        assert b2.stats(0)["fp64_slices"] == 0
        for q, i in enumerate(ctqw):
            assert rel_fro(_eig_ref(mats[i]), us_f[q]) <= tol_fn(b2.stats(q)["s"]), ("fp64 pade", i)
        b2._backend.set_option(core.OPT_FP64_SLICES, 5)
        with pytest.raises(RuntimeError):
            b2._backend.set_option(core.OPT_FP64_SLICES, 6)
        # grouped split-precision squaring rounds
        for prec in ("fp16x2", "bf16x3"):
            b2.set_precision(prec)
            for c in (0, 2):
                b2._backend.set_option(core.OPT_GROUP_CANDIDATE, c)
                us_h = b2.run([mats[i] for i in ctqw])
                assert b2.report()["grouped"] == 1
                for q, i in enumerate(ctqw):
                    st_q = b2.stats(q)
                    assert st_q["precision"] == prec
                    assert rel_fro(_eig_ref(mats[i]), us_h[q]) <= tol_fn(st_q["s"]), (prec, c, i)
        b2._backend.set_option(core.OPT_GROUP_CANDIDATE, -1)
        b2.set_precision("native")


def test_ctqw_drop_in(fx):
    """README usage (README.md:61-72) with device='cuda'."""
    _say("Ctqw drop-in")
    from FastCTQW.Ctqw import Ctqw
    from FastCTQW.State import State

    n = 120
    adj = random_regular(n, 4, seed=3).astype(np.float32)
    psi0 = State.getSingleNodeState(n, 0)
    w = Ctqw.from_numpy_array(adj, psi0, device="cuda")
    w.runWalk(time=4.5)
    psi = w.getFinalState().getState()
    # reference convention: H = -gamma*A, A_exp = -1j*H*t = +1j*t*A
    ref = eig_propagator(-adj.astype(np.float64), 4.5)[:, 0]
    assert infidelity(ref, psi) < 1e-9
    u = w.getOperator().getOperator()
    assert u.shape == (n, n)
    w_cpu = Ctqw.from_numpy_array(adj, psi0, device="cpu")
    w_cpu.runWalk(time=4.5)
    assert infidelity(w_cpu.getFinalState().getState(), psi) < 1e-9


@pytest.mark.parametrize("adt", [np.float32, np.float64])
@pytest.mark.parametrize("laplacian", [False, True])
def test_operator_device_hamiltonian(fx, adt, laplacian):
    """H = -gamma*A or -gamma*(D - A) formed on the device (fexpm_set_graph)
    and A_exp = -i t H per walk (fexpm_run_graph): against the eigen oracle
    of the FP64 H, against the host route (setHamiltonian with the same H),
    and the lazily built host H against the reference formula."""
    from fastctqw_amd.Operator import Operator

    n, gamma, t = 150, 0.7, 3.0
    adj = erdos_renyi(n, 0.3, seed=77, weighted=True).astype(adt)
    _say(f"device hamiltonian {np.dtype(adt).name} laplacian={laplacian}")
    op = Operator(adj, gamma=gamma, laplacian=laplacian, device="cuda")
    assert op._graph_on_device and op._h is None
    a64 = adj.astype(np.float64)
    h = -gamma * ((np.diag(a64.sum(axis=1)) - a64) if laplacian else a64)
    ref = eig_propagator(h, t)
    op.buildExpmOperator(t)
    u_dev = op.getOperator()
    st = op.lastRunStats()
    assert st["structured"]
    tol = _tol(np.complex64 if adt == np.float32 else np.complex128, st["s"])
    assert rel_fro(ref, u_dev) <= tol
    # same t twice gives the same U; the host H is built only when read
    op.buildExpmOperator(t)
    assert np.array_equal(op.getOperator(), u_dev)
    hh = op._hamiltonian
    assert hh.dtype == op._dtype and np.allclose(hh, h, rtol=1e-6, atol=1e-6)
    # host route with the same H
    op.setHamiltonian(hh)
    assert not op._graph_on_device
    op.buildExpmOperator(t)
    assert rel_fro(ref, op.getOperator()) <= tol
    # reset to another graph keeps the device route
    adj2 = erdos_renyi(n, 0.2, seed=78, weighted=True).astype(adt)
    op.reset_numpy_array(adj2, gamma=gamma, laplacian=laplacian)
    assert op._graph_on_device
    op.buildExpmOperator(t)
    a2 = adj2.astype(np.float64)
    h2 = -gamma * ((np.diag(a2.sum(axis=1)) - a2) if laplacian else a2)
    assert rel_fro(eig_propagator(h2, t), op.getOperator()) <= tol
    op.free()


@pytest.mark.parametrize("weighted", [False, True])
def test_device_er_generator(fx, weighted):
    """The seeded G(n, p) generated on the device (fexpm_set_graph_er) is
    graphs.erdos_renyi bit for bit (H = A with gamma = -1); with the
    Laplacian, H and U equal those of the dense host adjacency sent through
    fexpm_set_graph, bit for bit."""
    n, p, seed = 1000, 0.3, 7
    _say(f"device ER generator n={n} weighted={weighted}")
    a = erdos_renyi(n, p, seed, weighted=weighted)
    with fx.ExpMatComplex64(n) as e:
        e.set_graph_er(p, seed, weighted=weighted, gamma=-1.0, laplacian=False)
        assert np.array_equal(e._backend.debug_hamiltonian(), a)
        e.set_graph_er(p, seed, weighted=weighted, gamma=1.0, laplacian=True)
        h_dev = e._backend.debug_hamiltonian()
        e.run_graph(2.0)
        u_dev = e.download()
        e.set_graph(a, gamma=1.0, laplacian=True)
        assert np.array_equal(e._backend.debug_hamiltonian(), h_dev)
        e.run_graph(2.0)
        assert np.array_equal(e.download(), u_dev)
    assert np.allclose(h_dev, -laplacian(a), rtol=0, atol=1e-12)


def test_edge_list_graph(fx):
    """An edge list (weighted, one self-loop, one repeated pair) sent
    through fexpm_set_graph_edges gives the H and, bitwise, the U of its
    dense adjacency sent through fexpm_set_graph; U against the eigen
    oracle; vertex indices outside [0, n) are rejected."""
    from fastctqw_amd.graphs import dense_from_edges

    n = 300
    rng = np.random.default_rng(21)
    m = 2500
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    src[0], dst[0] = 5, 5
    src[1], dst[1] = src[2], dst[2] = 7, 9
    w = rng.random(m) + 0.1
    a = dense_from_edges(n, src, dst, w)
    _say(f"edge list graph n={n} m={m}")
    for dt in (np.complex64, np.complex128):
        with _engine(fx, dt, n) as e:
            e.set_graph_edges(src, dst, w, gamma=0.8, laplacian=True)
            h_e = e._backend.debug_hamiltonian()
            e.run_graph(1.5)
            u_e = e.download()
            e.set_graph(a, gamma=0.8, laplacian=True)
            assert np.array_equal(e._backend.debug_hamiltonian(), h_e)
            e.run_graph(1.5)
            assert np.array_equal(e.download(), u_e)
            st = e.stats()
            with pytest.raises(ValueError):
                e.set_graph_edges([0], [n], None)
        assert rel_fro(eig_propagator(h_e, 1.5), u_e) <= _tol(dt, st["s"])


def test_readme_from_networkx_graph(fx):
    """The README call (README.md:61-72): Ctqw.from_networkx_graph(G, psi0,
    device="cuda", dtype=np.complex64) on a random 4-regular graph, runWalk
    at t = 4.5, the final state against the eigen oracle of the reference's
    H = -A (uniform and random initial states); the complex64 adjacency with
    zero imaginary part takes the device graph route."""
    nx = pytest.importorskip("networkx")
    from FastCTQW.Ctqw import Ctqw
    from FastCTQW.State import State
    from fastctqw_amd.graphs import random_state

    for n in (10, 400):
        _say(f"README from_networkx_graph n={n}")
        G = nx.random_regular_graph(d=4, n=n, seed=n)
        adj = nx.to_numpy_array(G)
        ref_u = eig_propagator(-adj, 4.5)
        for init in (State.getUniformSuperposition(n=n), State(random_state(n, seed=3).astype(np.complex64))):
            walker = Ctqw.from_networkx_graph(G, init, device="cuda", dtype=np.complex64)
            assert walker.getOperator()._graph_on_device
            walker.runWalk(time=4.5)
            psi = walker.getFinalState().getState()
            assert psi.dtype == np.complex64
            inf = infidelity(ref_u @ init.getState().astype(np.complex128), psi)
            print(f"  infidelity {inf:.2e}", flush=True)
            assert inf <= 1e-9
            walker.getOperator().free()


def test_from_adjacency_list_and_er_on_device(fx):
    """Ctqw.from_adjacency_list and Ctqw.from_erdos_renyi with
    device="cuda" against the same walks on the CPU route (scipy)."""
    nx = pytest.importorskip("networkx")
    from fastctqw_amd.Ctqw import Ctqw
    from fastctqw_amd.State import State

    G = nx.random_regular_graph(5, 120, seed=8)
    psi0 = State.getUniformSuperposition(120)
    _say("from_adjacency_list / from_erdos_renyi on the device")
    for lap in (False, True):
        wd = Ctqw.from_adjacency_list(list(G.edges), psi0, laplacian=lap, device="cuda")
        wc = Ctqw.from_adjacency_list(list(G.edges), psi0, laplacian=lap, device="cpu")
        wd.runWalk(time=3.0)
        wc.runWalk(time=3.0)
        assert infidelity(wc.getFinalState().getState(), wd.getFinalState().getState()) <= 1e-9
        assert np.allclose(wd.getOperator()._hamiltonian, wc.getOperator()._hamiltonian)
        wd.getOperator().free()
    psi1 = State.getSingleNodeState(200, 3)
    wd = Ctqw.from_erdos_renyi(200, 0.4, 12, psi1, weighted=True, laplacian=True, device="cuda")
    wc = Ctqw.from_erdos_renyi(200, 0.4, 12, psi1, weighted=True, laplacian=True, device="cpu")
    wd.runWalk(time=2.0)
    wc.runWalk(time=2.0)
    assert infidelity(wc.getFinalState().getState(), wd.getFinalState().getState()) <= 1e-9
    wd.getOperator().free()


def test_device_graph_errors(fx):
    _say("device graph errors")
    with fx.ExpMatComplex64(8) as e:
        with pytest.raises(Exception):
            e.run_graph(1.0)  # no graph yet
        with pytest.raises(ValueError):
            e.set_graph(np.zeros((7, 7), np.float32))
        with pytest.raises(TypeError):
            e.set_graph(np.zeros((8, 8), np.complex64))
        e.set_graph(np.zeros((8, 8), np.float32))
        e.run_graph(2.0)
        assert np.allclose(e.download(), np.eye(8))
    with fx.ExpMatFloat32(8) as e:
        with pytest.raises(Exception):
            e.set_graph(np.zeros((8, 8), np.float32))
