"""Component parity on the GPU through the testing C ABI
(include/fastexpm_testing.h): one MFMA GEMM launch, the FP64 LU solve, and
the engine's internal Pade stages, each against an FP64 numpy reference.

Tolerances: FP32 MFMA products 2e-6 relative (Frobenius) for K <= 512, FP64
products 1e-13, LU solve 1e-12 (well-conditioned systems).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from fastctqw_amd.graphs import erdos_renyi, laplacian
from oracle.metrics import rel_fro
from oracle.padeexpm import PADE_B, pade_expm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fastctqw_amd", "fastexpm", "libfastexpm_hip.so")

_D = ctypes.POINTER(ctypes.c_double)


def _say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.fexpm_last_error.restype = ctypes.c_char_p
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def _gemm(lib, wtype, mode, sym, a, b):
    m, k = a.shape
    n = b.shape[0]
    ar = np.ascontiguousarray(a.real, dtype=np.float64)
    br = np.ascontiguousarray(b.real, dtype=np.float64)
    ai = np.ascontiguousarray(a.imag, dtype=np.float64) if mode else None
    bi = np.ascontiguousarray(b.imag, dtype=np.float64) if mode else None
    cr = np.zeros((m, n))
    ci = np.zeros((m, n)) if mode else None
    rc = lib.fexpm_test_gemm(wtype, mode, sym, m, n, k, _p(ar), _p(ai), _p(br), _p(bi), _p(cr), _p(ci))
    assert rc == 0, lib.fexpm_last_error(None)
    return cr if mode == 0 else cr + 1j * ci


GEMM_SHAPES = [(16, 16, 16), (64, 64, 32), (128, 128, 64), (200, 130, 48), (257, 300, 80), (17, 5, 7)]


@pytest.mark.parametrize("wtype", [0, 1])
@pytest.mark.parametrize("shape", GEMM_SHAPES)
def test_gemm_real(lib, wtype, shape):
    m, n, k = shape
    _say(f"gemm real w={wtype} {shape}")
    rng = np.random.default_rng(m * 7 + n)
    a, b = rng.standard_normal((m, k)), rng.standard_normal((n, k))
    c = _gemm(lib, wtype, 0, 0, a, b)
    tol = 2e-6 if wtype == 0 else 1e-13
    assert rel_fro(a @ b.T, c) < tol


@pytest.mark.parametrize("wtype", [0, 1])
@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("shape", GEMM_SHAPES)
def test_gemm_complex(lib, wtype, mode, shape):
    m, n, k = shape
    _say(f"gemm complex w={wtype} mode={mode} {shape}")
    rng = np.random.default_rng(m * 3 + n + mode)
    a = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    b = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    c = _gemm(lib, wtype, mode, 0, a, b)
    tol = 4e-6 if wtype == 0 else 1e-13
    assert rel_fro(a @ b.T, c) < tol


@pytest.mark.parametrize("wtype", [0, 1])
@pytest.mark.parametrize("n", [16, 100, 128, 129, 300])
def test_gemm_symmetric(lib, wtype, n):
    _say(f"gemm sym w={wtype} n={n}")
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n))
    c = _gemm(lib, wtype, 0, 1, x, x)
    tol = 2e-6 if wtype == 0 else 1e-13
    assert rel_fro(x @ x.T, c) < tol
    assert np.array_equal(c, c.T)
    z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    z = z + z.T
    c = _gemm(lib, wtype, 2, 1, z, z)
    assert rel_fro(z @ z, c) < 2 * tol
    assert np.array_equal(c, c.T)


def _lower_mirror(p):
    return np.tril(p) + np.tril(p, -1).T


@pytest.mark.parametrize("wtype", [0, 1])
@pytest.mark.parametrize("n", [16, 100, 129, 300])
def test_gemm_two_product_squaring(lib, wtype, n):
    """Mode 3 (the two-product squaring form): for any symmetric Xr, Xi the
    lower triangle of (Xr + Xi)(Xr - Xi)^T and of 2 Xr Xi^T, mirrored; for
    commuting parts (Xr = cos G, Xi = sin G) that is X^2 = exp(2iG)."""
    _say(f"gemm two-product w={wtype} n={n}")
    rng = np.random.default_rng(n + 11)
    xr = rng.standard_normal((n, n))
    xi = rng.standard_normal((n, n))
    xr, xi = xr + xr.T, xi + xi.T
    c = _gemm(lib, wtype, 3, 1, xr + 1j * xi, xr + 1j * xi)
    ref = _lower_mirror((xr + xi) @ (xr - xi).T) + 1j * _lower_mirror(2 * xr @ xi.T)
    tol = 4e-6 if wtype == 0 else 1e-13
    assert rel_fro(ref, c) < tol
    assert np.array_equal(c, c.T)
    g = rng.standard_normal((n, n))
    lam, v = np.linalg.eigh((g + g.T) / np.sqrt(n))
    x = (v * np.exp(1j * lam)) @ v.T
    c = _gemm(lib, wtype, 3, 1, x, x)
    assert rel_fro((v * np.exp(2j * lam)) @ v.T, c) < tol


@pytest.mark.parametrize("wtype", [2, 3])
@pytest.mark.parametrize("shape", [(16, 16), (100, 100), (128, 128), (129, 129), (300, 300), (200, 77)])
def test_gemm_split_squaring(lib, wtype, shape):
    """Split-precision X X^T (wtype 2: fp16 hi/lo, 3 products; 3: three bf16
    parts, 6 products; FP32 accumulation) for complex X with |x| <= 1, the
    operand range of the squaring, against FP64 numpy; the FP32 3M kernel on
    the same input is printed for comparison."""
    m, k = shape
    _say(f"split squaring w={wtype} {shape}")
    rng = np.random.default_rng(m * 5 + k + wtype)
    x = (rng.uniform(-1, 1, (m, k)) + 1j * rng.uniform(-1, 1, (m, k))) / np.sqrt(2)
    c = _gemm(lib, wtype, 2, 1, x, x)
    ref = x @ x.T
    err = rel_fro(ref, c)
    err32 = rel_fro(ref, _gemm(lib, 0, 2, 0, x, x))
    print(f"  relF={err:.2e} (fp32 3M: {err32:.2e})", flush=True)
    assert err < 4e-6
    assert np.array_equal(c, c.T)


@pytest.mark.parametrize("slices", [4, 5, 6])
@pytest.mark.parametrize("shape,sym", [((300, 200, 250), False), ((129, 129, 1000), True), ((64, 64, 64), True),
                                       ((1, 70, 33), False)])
def test_gemm_int8_slices(lib, slices, shape, sym):
    """FP64 product from exact int8 row slices: every element within the
    slicing bound (S + 2) 2^-7S k 2^(e_i + f_j), with 2^e_i < 2 max_k |a_ik|;
    rows of very different scale keep their own precision."""
    m, n, k = shape
    _say(f"int8-slice product S={slices} m={m} n={n} k={k} sym={sym}")
    rng = np.random.default_rng(m + 7 * n + k + slices)
    a = rng.standard_normal((m, k)) * np.exp2(rng.integers(-30, 30, size=(m, 1)))
    a[0, :] = 0.0  # a zero row
    b = a.copy() if sym else rng.standard_normal((n, k)) * np.exp2(rng.integers(-30, 30, size=(n, 1)))
    c = np.zeros((m, n))
    rc = lib.fexpm_test_ozaki_gemm(slices, int(sym), m, n, k, _p(np.ascontiguousarray(a)),
                                   _p(np.ascontiguousarray(b)), _p(c))
    assert rc == 0, lib.fexpm_last_error(None)
    ref = a @ b.T
    bound = (slices + 2) * 2.0 ** (-7 * slices) * k * 4.0 * np.outer(np.abs(a).max(1), np.abs(b).max(1))
    err = np.abs(c - ref)
    worst = float((err / np.maximum(bound, 1e-300)).max())
    print(f"  worst error / bound = {worst:.3e}", flush=True)
    assert np.all(err <= bound), worst
    if sym:
        assert np.array_equal(c, c.T)


CRT64_MODULI = (255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199)


@pytest.mark.parametrize("shape", [(1, 1), (64, 40), (257, 64), (300, 513), (640, 10016), (300, 20000)])
@pytest.mark.parametrize("same", [True, False])
def test_gemm_int8_residues(lib, shape, same):
    """FP64-accurate product from exact int8 residues (the engine's default
    for the Pade and solve products of complex64 engines): each row rounded
    to beta = min(45, floor((log2 M - 1 - log2 kpad) / 2)) bits of its largest
    element, 13 exact int8 products modulo coprime moduli, Chinese-remainder
    reconstruction.  The lower triangle of A B^T within the rounding bound
    k 2^(e_i + f_j - beta) (2^e_i < 2 max_k |a_ik|), the output mirrored;
    rows of very different scale and a zero row."""
    import math

    m, k = shape
    _say(f"int8 residue product m={m} k={k} same={same}")
    rng = np.random.default_rng(m + 3 * k + int(same))
    a = rng.standard_normal((m, k)) * np.exp2(rng.integers(-30, 30, size=(m, 1)))
    a[0, :] = 0.0  # a zero row
    b = a.copy() if same else rng.standard_normal((m, k)) * np.exp2(rng.integers(-30, 30, size=(m, 1)))
    a = np.ascontiguousarray(a)
    b = a if same else np.ascontiguousarray(b)
    c = np.zeros((m, m))
    rc = lib.fexpm_test_ozaki_gemm(13, 1, m, m, k, _p(a), _p(b), _p(c))
    assert rc == 0, lib.fexpm_last_error(None)
    kpad = (k + 63) // 64 * 64
    log2m = math.log2(math.prod(CRT64_MODULI))
    beta = min(45, math.floor((log2m - 1.0 - math.log2(kpad)) / 2.0))
    ref = a @ b.T
    bound = k * 4.0 * np.outer(np.abs(a).max(1), np.abs(b).max(1)) * 2.0 ** -beta
    low = np.tril(np.ones((m, m), dtype=bool))
    err = np.abs(c - ref)[low]
    worst = float((err / np.maximum(bound[low], 1e-300)).max())
    print(f"  beta={beta} worst error / bound = {worst:.3e}", flush=True)
    assert np.all(err <= bound[low]), worst
    assert np.array_equal(c, c.T)
    # the bound is tight to a few bits, not loose: a product without the
    # rounding (integers well inside beta bits) is exact
    ai = rng.integers(-1000, 1000, size=(m, k)).astype(np.float64)
    ci = np.zeros((m, m))
    assert lib.fexpm_test_ozaki_gemm(13, 1, m, m, k, _p(ai), _p(ai), _p(ci)) == 0
    if m > 1:
        # integer entries: |a| < 1024 = 2^10 <= 2^beta, so every value is kept exactly
        assert np.array_equal(ci, ai @ ai.T)
    assert lib.fexpm_test_ozaki_gemm(13, 0, m, m, k, _p(a), _p(b), _p(c)) != 0  # residue form: sym only


def test_square_issued_plan(lib):
    """Issued MFMA flops of one FP32 structured squaring: 64 x 64 lower tiles
    up to n = 6144; above it 128 x 128 tiles whose last partial wave may run
    as 64 x 64 sub-tiles, so the count lies between the 64-tile and the full
    128-tile figures and is a whole number of 64 x 64 sub-tile products."""
    lib.fexpm_test_square_issued.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]

    def issued(n):
        v = ctypes.c_double(0.0)
        assert lib.fexpm_test_square_issued(n, ctypes.byref(v)) == 0
        return v.value

    def tiles(n, t):
        T = (n + t - 1) // t
        return T * (T + 1) // 2

    for n in (100, 4096, 6144):
        k = (n + 15) // 16 * 16
        assert issued(n) == 6.0 * tiles(n, 64) * 64 * 64 * k
    for n in (6145, 8192, 10000, 13000):
        k = (n + 15) // 16 * 16
        v = issued(n)
        assert 6.0 * tiles(n, 64) * 64 * 64 * k <= v <= 6.0 * tiles(n, 128) * 128 * 128 * k
        assert (v / (6.0 * 64 * 64 * k)).is_integer()
    assert lib.fexpm_test_square_issued(0, ctypes.byref(ctypes.c_double())) != 0


@pytest.mark.parametrize("n", [1500, 3000])
def test_run_host_buffers_in_place(lib, n):
    """fexpm_run on host buffers above the size at which the engine faults
    the output's pages in on a host thread while the device works (16 MiB)
    and, at n = 3000 (72 MB), above the size at which host copies go through
    pinned staging buffers (64 MiB): a fresh output buffer, the input buffer
    reused as the output (plain copies then: the input is still being read),
    and fexpm_upload / fexpm_run_resident / fexpm_download give the same U
    bit for bit; at n = 1500 U matches the eigen oracle."""
    _say(f"fexpm_run host buffers n={n}")
    L = laplacian(erdos_renyi(n, 0.05, seed=71, weighted=True))
    a = np.ascontiguousarray((-1j * 2.0 * L).astype(np.complex64))
    e = ctypes.c_void_p()
    vp = ctypes.c_void_p
    assert lib.fexpm_create(ctypes.c_int64(n), 2, ctypes.byref(e)) == 0
    try:
        u = np.empty((n, n), np.complex64)
        rc = lib.fexpm_run(e, a.ctypes.data_as(vp), u.ctypes.data_as(vp))
        assert rc == 0, lib.fexpm_last_error(e)
        buf = a.copy()
        p = buf.ctypes.data_as(vp)
        rc = lib.fexpm_run(e, p, p)
        assert rc == 0, lib.fexpm_last_error(e)
        u2 = np.empty((n, n), np.complex64)
        assert lib.fexpm_upload(e, a.ctypes.data_as(vp)) == 0, lib.fexpm_last_error(e)
        assert lib.fexpm_run_resident(e) == 0, lib.fexpm_last_error(e)
        assert lib.fexpm_download(e, u2.ctypes.data_as(vp)) == 0, lib.fexpm_last_error(e)
        # the pageable path (taken when pinned memory cannot be allocated)
        assert lib.fexpm_set_option(e, 108, ctypes.c_int64(0)) == 0  # FEXPM_OPT_HOST_STAGING
        u3 = np.empty((n, n), np.complex64)
        rc = lib.fexpm_run(e, a.ctypes.data_as(vp), u3.ctypes.data_as(vp))
        assert rc == 0, lib.fexpm_last_error(e)
    finally:
        lib.fexpm_destroy(e)
    assert np.array_equal(u, buf)
    assert np.array_equal(u, u2)
    assert np.array_equal(u, u3)
    if n <= 1500:
        lam, V = np.linalg.eigh(np.ascontiguousarray(a.imag, dtype=np.float64))
        ref = (V * np.exp(1j * lam)[None, :]) @ V.T
        assert rel_fro(ref, u.astype(np.complex128)) < 1e-4


def test_gemm_split_range_error(lib):
    """fp16 parts hold |x| * 2^14 <= 65504: a larger operand is reported."""
    _say("split squaring range error")
    x = np.full((32, 32), 8.0 + 0j)
    m = x.shape[0]
    ar, ai = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    cr, ci = np.zeros((m, m)), np.zeros((m, m))
    rc = lib.fexpm_test_gemm(2, 2, 1, m, m, m, _p(ar), _p(ai), _p(ar), _p(ai), _p(cr), _p(ci))
    assert rc != 0 and b"range" in lib.fexpm_last_error(None)


def _solve(lib, m, b):
    n = m.shape[0]
    mr, mi = np.ascontiguousarray(m.real), np.ascontiguousarray(m.imag)
    br, bi = np.ascontiguousarray(b.real), np.ascontiguousarray(b.imag)
    xr, xi = np.zeros((n, n)), np.zeros((n, n))
    info = ctypes.c_int(0)
    rc = lib.fexpm_test_lu_solve(n, _p(mr), _p(mi), _p(br), _p(bi), _p(xr), _p(xi), ctypes.byref(info))
    assert rc == 0, lib.fexpm_last_error(None)
    return xr + 1j * xi, info.value


@pytest.mark.parametrize("n", [1, 2, 5, 63, 64, 65, 100, 200, 333, 640, 777])
def test_lu_solve(lib, n):
    _say(f"lu n={n}")
    rng = np.random.default_rng(n)
    m = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 2 * np.sqrt(n) * np.eye(n)
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    x, info = _solve(lib, m, b)
    assert info == 0
    assert rel_fro(np.linalg.solve(m, b), x) < 1e-12


def test_lu_needs_pivoting(lib):
    _say("lu pivoting")
    n = 150
    rng = np.random.default_rng(1)
    perm = rng.permutation(n)
    m = np.zeros((n, n), complex)
    m[np.arange(n), perm] = 1.0 + 0.5j  # zero diagonal: fails without row interchanges
    m += 1e-3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    b = rng.standard_normal((n, n)) + 0j
    x, info = _solve(lib, m, b)
    assert info == 0
    assert rel_fro(np.linalg.solve(m, b), x) < 1e-11


def test_lu_singular_flag(lib):
    _say("lu singular")
    n = 20
    m = np.zeros((n, n), complex)
    m[:, :10] = np.eye(n)[:, :10]
    _, info = _solve(lib, m, np.eye(n, dtype=complex))
    assert info != 0


# ---------------------------------------------------------------------------
# Engine stages (structured path) against the FP64 stage restatement

def _planes(lib_fx, e, which, n):
    re, im = np.zeros((n, n)), np.zeros((n, n))
    lib = ctypes.CDLL(LIB)
    return re, im


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_structured_stages(lib, dt):
    from fastctqw_amd.fastexpm import ExpMatComplex64, ExpMatComplex128

    n, t = 200, 3.0
    _say(f"stages n={n} {np.dtype(dt).name}")
    L = laplacian(erdos_renyi(n, 0.5, seed=2))
    a = (-1j * t * L).astype(dt)
    st = pade_expm(a.astype(np.complex128), stages=True)
    e = (ExpMatComplex64 if dt == np.complex64 else ExpMatComplex128)(n)
    e._backend.set_option(100, 0)  # no squarings: output = e^mu X
    e._backend.set_option(12, 0)  # 1-norm scaling rule (the stage restatement above)
    u0 = e.run(a)
    s = e.stats()
    assert s["structured"] and s["m"] == st["m"] and s["s"] == st["s"]
    tol = 1e-5 if dt == np.complex64 else 1e-12
    # FP64 C (imag part of the shifted, scaled A)
    c_ref = np.imag(a.astype(np.complex128) - st["mu"] * np.eye(n)) / 2 ** st["s"]
    assert abs(s["mu"] - st["mu"]) <= 1e-12 * abs(st["mu"])
    x_ref = np.exp(st["mu"]) * st["X"]
    err = rel_fro(x_ref, u0)
    print(f"  X relF={err:.3e}", flush=True)
    assert err < tol
    del c_ref
    e.free()
