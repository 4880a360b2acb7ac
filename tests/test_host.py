"""Host-side tests (no GPU): generators, the C ABI library's exports, the
Python wrappers' validation and the CPU walk path of Ctqw/Operator/State."""
import ctypes
import os
import re
import warnings

import numpy as np
import pytest

from fastctqw_amd.graphs import (erdos_renyi, laplacian, random_regular, random_state, splitmix64,
                                 uniform01)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fastctqw_amd", "fastexpm", "libfastexpm_hip.so")
HEADER = os.path.join(ROOT, "include", "fastexpm.h")


# --------------------------------------------------------------------------- generators

def test_splitmix64_known_values():
    # splitmix64 finaliser of 0 and 1 (first outputs of the reference
    # generator seeded with 0 are the finaliser of 0x9E3779B97F4A7C15*k)
    assert int(splitmix64(np.uint64(0))) == 0xE220A8397B1DCDAF
    assert int(splitmix64(np.uint64(0x9E3779B97F4A7C15))) == 0x6E789E6AA1B965F4


def test_uniform_range_and_determinism():
    u = uniform01(3, 1, np.arange(100000, dtype=np.uint64))
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    assert np.array_equal(u, uniform01(3, 1, np.arange(100000, dtype=np.uint64)))
    assert not np.array_equal(u, uniform01(4, 1, np.arange(100000, dtype=np.uint64)))


def test_erdos_renyi_properties():
    a = erdos_renyi(400, 0.3, seed=2)
    assert np.array_equal(a, a.T) and np.all(np.diag(a) == 0)
    assert set(np.unique(a)) <= {0.0, 1.0}
    dens = a.sum() / (400 * 399)
    assert abs(dens - 0.3) < 0.01
    assert np.array_equal(a, erdos_renyi(400, 0.3, seed=2))
    w = erdos_renyi(200, 0.5, seed=3, weighted=True)
    assert np.array_equal(w, w.T) and w.max() <= 1.0 and w[w > 0].min() > 0
    assert np.array_equal(w > 0, erdos_renyi(200, 0.5, seed=3) > 0)


@pytest.mark.parametrize("n,d", [(10, 3), (64, 4), (500, 8), (1000, 4)])
def test_random_regular(n, d):
    a = random_regular(n, d, seed=n + d)
    assert np.array_equal(a, a.T) and np.all(np.diag(a) == 0)
    assert np.all(a.sum(axis=1) == d)
    assert np.array_equal(a, random_regular(n, d, seed=n + d))


def test_random_regular_rejects_impossible():
    with pytest.raises(ValueError):
        random_regular(5, 3, seed=0)


def test_laplacian_and_state():
    a = erdos_renyi(50, 0.4, seed=1, weighted=True)
    L = laplacian(a)
    assert np.allclose(L @ np.ones(50), 0)
    psi = random_state(50, seed=55)
    assert abs(np.linalg.norm(psi) - 1) < 1e-14
    assert not np.allclose(psi, random_state(50, seed=55, index=1))


# --------------------------------------------------------------------------- C ABI

PROBES = os.path.join(os.path.dirname(LIB), "libfastexpm_probes.so")


def _declared_symbols(*headers):
    text = "".join(open(os.path.join(ROOT, "include", h)).read() for h in headers)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|double|const char\*)\s+(fexpm_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    """Each header's declarations are exported by its library: fastexpm.h
    and fastexpm_testing.h by the engine, fastexpm_probes.h by the probes
    library (and the engine exports none of the probes)."""
    if not os.path.exists(LIB):
        pytest.skip("HIP library not built")
    syms = _declared_symbols("fastexpm.h", "fastexpm_testing.h")
    assert len(syms) >= 14
    assert "fexpm_batch_get_report" in syms and "fexpm_test_group_program" in syms
    lib = ctypes.CDLL(LIB)
    for s in syms:
        assert hasattr(lib, s), s
    probes = _declared_symbols("fastexpm_probes.h")
    assert "fexpm_probe_gemm_time" in probes and "fexpm_probe_copy" in probes
    for s in probes:
        assert not hasattr(lib, s), s
    if not os.path.exists(PROBES):
        pytest.skip("probes library not built")
    plib = ctypes.CDLL(PROBES)
    for s in probes:
        assert hasattr(plib, s), s
    # argument errors are reported without a device call
    plib.fexpm_probe_last_error.restype = ctypes.c_char_p
    ms = ctypes.c_double(0.0)
    assert plib.fexpm_probe_copy(ctypes.c_int64(17), 1, ctypes.byref(ms)) == -1
    assert plib.fexpm_probe_gemm_time(0, 2, 1, 0, 64, 32, 64, 1, ctypes.byref(ms)) == -1


def test_library_no_device_and_argument_errors():
    if not os.path.exists(LIB):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(LIB)
    lib.fexpm_status_string.restype = ctypes.c_char_p
    assert lib.fexpm_status_string(0) == b"ok"
    assert lib.fexpm_status_string(-4) == b"singular Pade denominator"
    h = ctypes.c_void_p()
    # invalid dtype / size are rejected before any device call
    assert lib.fexpm_create(ctypes.c_int64(4), 9, ctypes.byref(h)) == -1
    assert lib.fexpm_create(ctypes.c_int64(-1), 2, ctypes.byref(h)) == -1
    assert lib.fexpm_create(ctypes.c_int64(4), 2, None) == -1
    assert lib.fexpm_free(None) == -1
    lib.fexpm_get_n.restype = ctypes.c_int64
    assert lib.fexpm_get_n(None) == -1


def _grouped_order(T, g=8):
    """The 8 x 8 grouped lower-triangular enumeration, written out directly."""
    out = []
    for gi in range((T + g - 1) // g):
        for gj in range(gi + 1):
            for ti in range(gi * g, min(T, gi * g + g)):
                for tj in range(gj * g, min(gj * g + g, ti + 1)):
                    out.append((ti, tj))
    return out


@pytest.mark.parametrize("T", [1, 2, 7, 8, 9, 16, 17, 79, 158, 159])
def test_grouped_tile_order(T):
    """The symmetric GEMM launches visit every lower tile exactly once, in
    the grouped order (evaluated on the host through the testing ABI)."""
    if not os.path.exists(LIB):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(LIB)
    nb = T * (T + 1) // 2
    buf = (ctypes.c_int * nb)()
    assert lib.fexpm_test_grouped_tile_order(T, buf) == 0
    got = [(v >> 16, v & 0xFFFF) for v in buf]
    assert got == _grouped_order(T)
    assert sorted(got) == sorted((i, j) for i in range(T) for j in range(i + 1))
    assert lib.fexpm_test_grouped_tile_order(0, buf) == -1


@pytest.mark.parametrize("nbm,nbn", [(1, 1), (1, 5), (5, 1), (8, 3), (9, 4), (17, 154), (316, 8)])
def test_grouped_rect_order(nbm, nbn):
    """Rectangular GEMM launches visit every tile once, 8 tile rows at a
    time, each group column by column."""
    if not os.path.exists(LIB):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(LIB)
    buf = (ctypes.c_int * (nbm * nbn))()
    assert lib.fexpm_test_grouped_rect_order(nbm, nbn, buf) == 0
    got = [(v >> 16, v & 0xFFFF) for v in buf]
    want = [(i, j) for g0 in range(0, nbm, 8) for j in range(nbn) for i in range(g0, min(nbm, g0 + 8))]
    assert got == want
    assert sorted(got) == [(i, j) for i in range(nbm) for j in range(nbn)]


# --------------------------------------------------------------------------- wrappers

class _FakeBackend:
    def __init__(self, n):
        self.n = n
        self.freed = 0

    def run(self, a):
        return np.eye(self.n, dtype=a.dtype).ravel()

    def free(self):
        self.freed += 1
        raise RuntimeError("boom")


def _fake_wrapper(monkeypatch):
    pytest.importorskip("fastctqw_amd.fastexpm._fastexpm_core")
    from fastctqw_amd.fastexpm import _wrapper

    class W(_wrapper._ExpMatBase):
        _core_cls = _FakeBackend

    return W


def test_wrapper_validation_and_exit(monkeypatch):
    W = _fake_wrapper(monkeypatch)
    w = W(3)
    with pytest.raises(TypeError):
        w.run([[1.0]])
    with pytest.raises(ValueError):
        w.run(np.zeros((2, 2)))
    assert w.run(np.zeros((3, 3))).shape == (3, 3)  # reshape of a flat return (_wrapper.py:96)
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        with pytest.raises(KeyError):
            with W(2):
                raise KeyError("original")
    assert any("Failed to free" in str(r.message) for r in rec)


# --------------------------------------------------------------------------- CPU walk path

def test_ctqw_cpu_path_matches_oracle():
    from fastctqw_amd.Ctqw import Ctqw
    from fastctqw_amd.State import State
    from oracle.eig import eig_propagator
    from oracle.metrics import infidelity

    n = 30
    adj = random_regular(n, 4, seed=9).astype(np.float64)
    w = Ctqw.from_numpy_array(adj, State.getUniformSuperposition(n), gamma=0.5, laplacian=True)
    w.runWalk(time=2.0)
    psi = w.getFinalState().getState()
    Lg = laplacian(adj)
    ref = eig_propagator(-0.5 * Lg, 2.0) @ np.full(n, 1 / np.sqrt(n))
    assert infidelity(ref, psi) < 1e-12
    assert w.getOperator().getOperator().dtype == np.complex128


def test_operator_errors_and_dtype_map():
    from fastctqw_amd.Operator import Operator

    with pytest.raises(ValueError):
        Operator(np.array([[0, 1], [0, 0]], dtype=np.float32))
    with pytest.raises(TypeError):
        Operator(np.zeros((3, 3), dtype=np.int32))
    op = Operator(np.zeros((3, 3), dtype=np.float32))
    assert op._dtype == np.complex64
    op = Operator(np.zeros((3, 3), dtype=np.float64))
    assert op._dtype == np.complex128


def test_state_semantics():
    from fastctqw_amd.State import State

    assert State.getUniformSuperposition(4).getState().dtype == np.complex64
    s = State.getLocalSuperposition(5, [1, 3]).getState()
    assert np.allclose(np.abs(s) ** 2, [0, 0.5, 0, 0.5, 0])
    with pytest.raises(ValueError):
        State(np.zeros((2, 2)))
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        State(np.array([1.0, 1.0]))
    assert rec


def test_directed_graph_rejected():
    nx = pytest.importorskip("networkx")
    from fastctqw_amd.Ctqw import Ctqw
    from fastctqw_amd.State import State

    with pytest.raises(TypeError):
        Ctqw.from_networkx_graph(nx.DiGraph([(0, 1)]), State.getSingleNodeState(2, 0))


def test_product_does_not_import_oracle():
    import subprocess
    import sys

    code = ("import sys; import fastctqw_amd, fastctqw_amd.Ctqw, fastctqw_amd.graphs;"
            "import FastCTQW; print(any(m == 'oracle' or m.startswith('oracle.') for m in sys.modules))")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT)
    assert out.stdout.strip() == "False", out.stderr


def test_edge_list_forms_match_networkx():
    """Adjacency-list inputs (networkx graph with weights and a self-loop,
    dict of lists, dict of dicts, pairs and triples) give the dense matrix of
    nx.to_numpy_array; repeated pairs in a sequence add up."""
    nx = pytest.importorskip("networkx")
    from fastctqw_amd.graphs import dense_from_edges, edge_list

    G = nx.random_regular_graph(4, 20, seed=2)
    for k, (u, v) in enumerate(G.edges):
        G[u][v]["weight"] = 0.25 + 0.01 * k
    G.add_edge(5, 5, weight=3.0)
    n, s, d, w = edge_list(G)
    assert np.array_equal(dense_from_edges(n, s, d, w), nx.to_numpy_array(G))
    dl = nx.to_dict_of_lists(G)
    n, s, d, w = edge_list(dl)
    assert np.array_equal(dense_from_edges(n, s, d, w), nx.to_numpy_array(nx.from_dict_of_lists(dl)))
    dd = {0: {1: {"weight": 2.0}}, 1: {2: {"weight": 0.5}}}
    n, s, d, w = edge_list(dd)
    assert np.array_equal(dense_from_edges(n, s, d, w), nx.to_numpy_array(nx.from_dict_of_dicts(dd)))
    n, s, d, w = edge_list([(0, 1), (1, 2, 0.5), (0, 1)], num_nodes=4)
    a = dense_from_edges(n, s, d, w)
    assert n == 4 and a[0, 1] == a[1, 0] == 2.0 and a[1, 2] == a[2, 1] == 0.5 and a.sum() == 5.0
    with pytest.raises(ValueError):
        edge_list([(0, 3)], num_nodes=3)
    with pytest.raises(TypeError):
        edge_list(nx.DiGraph([(0, 1)]))


def test_adjacency_list_and_er_cpu_routes():
    """Operator/Ctqw.from_adjacency_list and from_erdos_renyi on the CPU give
    the walks of the dense adjacency (scipy path)."""
    nx = pytest.importorskip("networkx")
    from fastctqw_amd.Ctqw import Ctqw
    from fastctqw_amd.Operator import Operator
    from fastctqw_amd.State import State

    G = nx.random_regular_graph(3, 16, seed=4)
    A = nx.to_numpy_array(G, dtype=np.float32)
    psi0 = State.getSingleNodeState(16, 0)
    w1 = Ctqw.from_adjacency_list(list(G.edges), psi0, laplacian=True)
    w2 = Ctqw.from_numpy_array(A, psi0, laplacian=True)
    w1.runWalk(time=2.0)
    w2.runWalk(time=2.0)
    assert np.array_equal(w1.getFinalState().getState(), w2.getFinalState().getState())
    op = Operator.from_adjacency_list(G, gamma=0.5, dtype=np.float64)
    assert op._dtype == np.complex128
    op.buildExpmOperator(1.5)
    ref = Operator.from_numpy_array(nx.to_numpy_array(G), gamma=0.5)
    ref.buildExpmOperator(1.5)
    assert np.array_equal(op.getOperator(), ref.getOperator())
    op.reset_adjList([(0, 1), (1, 2)], gamma=1.0, time=None, laplacian=False)
    assert op._n == 3
    op.buildExpmOperator(1.0)
    assert op.getOperator().shape == (3, 3)
    er = Operator.from_erdos_renyi(30, 0.4, 9, weighted=True, laplacian=True)
    dense = Operator.from_numpy_array(erdos_renyi(30, 0.4, 9, weighted=True).astype(np.float32), laplacian=True)
    assert np.array_equal(er._hamiltonian, dense._hamiltonian)


@pytest.mark.parametrize("theta", [0.0, 1e-3, 0.3, 1.0, 2.7, 4.08, 5.371920351148152, 12.0])
def test_cheb_series_host(theta):
    """The engine's host routine for the Chebyshev polynomial of spectral
    plans (Bessel power series and basis change in long double) against the
    oracle's 50-digit coefficients: the same degrees, coefficients to a few
    units of the float rounding of the largest term, and an out-of-range
    theta rejected."""
    pytest.importorskip("mpmath")
    from oracle.chebcs import cheb_coeffs, cheb_degrees

    if not os.path.exists(LIB):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(LIB)
    lib.fexpm_test_cheb_series.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double)]
    dc, ds = ctypes.c_int(), ctypes.c_int()
    cw, sw = (ctypes.c_double * 40)(), (ctypes.c_double * 40)()
    assert lib.fexpm_test_cheb_series(theta, ctypes.byref(dc), ctypes.byref(ds), cw, sw) == 0
    assert (dc.value, ds.value) == cheb_degrees(theta)
    rc, rs = cheb_coeffs(theta)
    scale = max(1.0, max(abs(v) for v in rc + rs))
    assert max(abs(cw[j] - v) for j, v in enumerate(rc)) <= 8 * 2.0 ** -53 * scale
    assert max(abs(sw[j] - v) for j, v in enumerate(rs)) <= 8 * 2.0 ** -53 * scale
    assert all(cw[j] == 0.0 for j in range(len(rc), 40)) and all(sw[j] == 0.0 for j in range(len(rs), 40))
    assert lib.fexpm_test_cheb_series(17.0, ctypes.byref(dc), ctypes.byref(ds), cw, sw) != 0
