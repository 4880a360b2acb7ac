"""The no-allocation entry point fexpm_run_device_noalloc (its work buffers
live in a static device array of the library) against the engine entry
point fexpm_run_device on the same device arrays: two calls in a row, the
second with a smaller order, for complex64 and complex128 CTQW inputs and a
general complex64 input; each result also against the FP64 oracle.  Device
arrays come from the HIP runtime the library itself links (ctypes), so no
other runtime shares the process."""
import ctypes
import os
import sys

import numpy as np
import pytest

from fastctqw_amd.graphs import erdos_renyi, laplacian
from oracle.eig import eig_propagator
from oracle.metrics import rel_fro
from oracle.padeexpm import pade_expm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fastctqw_amd", "fastexpm", "libfastexpm_hip.so")
NMAX = 256
C64, C128 = 2, 3


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def libs():
    lib = ctypes.CDLL(LIB)
    hip = ctypes.CDLL("libamdhip64.so.7")
    lib.fexpm_last_error.restype = ctypes.c_char_p
    lib.fexpm_last_error.argtypes = [ctypes.c_void_p]
    lib.fexpm_create.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.fexpm_run_device.argtypes = [ctypes.c_void_p] * 4
    lib.fexpm_destroy.argtypes = [ctypes.c_void_p]
    lib.fexpm_run_device_noalloc.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib, hip


class _Dev:
    """A device array of nbytes from the library's own HIP runtime."""

    def __init__(self, hip, nbytes):
        self.hip, self.nbytes, self.p = hip, nbytes, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device

    def get(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _engine_run(lib, hip, a, dt_code):
    n = a.shape[0]
    e = ctypes.c_void_p()
    assert lib.fexpm_create(n, dt_code, ctypes.byref(e)) == 0
    da, du = _Dev(hip, a.nbytes), _Dev(hip, a.nbytes)
    try:
        da.put(a)
        rc = lib.fexpm_run_device(e, da.p, du.p, None)
        assert rc == 0, lib.fexpm_last_error(e)
        return du.get(a.dtype, a.shape)
    finally:
        da.free()
        du.free()
        lib.fexpm_destroy(e)


def _noalloc_run(lib, hip, a, dt_code):
    da, du = _Dev(hip, a.nbytes), _Dev(hip, a.nbytes)
    try:
        da.put(a)
        rc = lib.fexpm_run_device_noalloc(a.shape[0], dt_code, da.p, du.p, None)
        assert rc == 0, lib.fexpm_last_error(None)
        return du.get(a.dtype, a.shape)
    finally:
        da.free()
        du.free()


def _ctqw(n, seed, t):
    L = laplacian(erdos_renyi(n, 0.5, seed=seed))
    return -1j * t * L, eig_propagator(L, t)


@pytest.mark.parametrize("dt,code,tol_eng,tol_ref", [(np.complex64, C64, 1e-6, 4e-5),
                                                      (np.complex128, C128, 1e-13, 1e-11)])
def test_noalloc_matches_engine(libs, dt, code, tol_eng, tol_ref):
    lib, hip = libs
    for n, seed, t in ((220, 11, 4.5), (100, 12, 20.0)):  # the second call is smaller
        _say(f"no-alloc vs engine n={n} {np.dtype(dt).name}")
        a, ref = _ctqw(n, seed, t)
        a = a.astype(dt)
        u_eng = _engine_run(lib, hip, a, code)
        u_na = _noalloc_run(lib, hip, a, code)
        d, e = rel_fro(u_eng, u_na), rel_fro(ref, u_na)
        print(f"  vs engine {d:.2e}, vs eigen oracle {e:.2e}", flush=True)
        assert d <= tol_eng and e <= tol_ref


def test_noalloc_general_input_and_limits(libs):
    lib, hip = libs
    rng = np.random.default_rng(3)
    for n in (150, 60):
        _say(f"no-alloc general complex64 n={n}")
        a = ((rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)).astype(np.complex64)
        u_eng = _engine_run(lib, hip, a, C64)
        u_na = _noalloc_run(lib, hip, a, C64)
        ref = pade_expm(a.astype(np.complex128))
        assert rel_fro(u_eng, u_na) <= 1e-6 and rel_fro(ref, u_na) <= 4e-5
    _say("no-alloc limits")
    assert lib.fexpm_run_device_noalloc(NMAX + 1, C64, None, None, None) == -1
    assert b"256" in lib.fexpm_last_error(None)
    assert lib.fexpm_run_device_noalloc(0, C64, None, None, None) == 0
    assert lib.fexpm_run_device_noalloc(8, 9, None, None, None) == -1
