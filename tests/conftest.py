"""Test configuration.

Markers: ``gpu`` for tests that need an MI355X (run with ``-m gpu``).
Options mirror the reference's tests/conftest.py:5-49: ``--matrix-sizes``
and ``--dtype`` parametrise tests that take ``N`` / ``dtype``.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


def pytest_addoption(parser):
    parser.addoption("--matrix-sizes", action="store", default="16,32,64,128,256,512,1024", type=str,
                     help="comma-separated matrix sizes for N-parametrised tests")
    parser.addoption("--dtype", action="store", default="complex64", type=str,
                     help="complex64 or complex128 for dtype-parametrised tests")


def pytest_generate_tests(metafunc):
    dtype_map = {"complex64": np.complex64, "complex128": np.complex128}
    if "N" in metafunc.fixturenames:
        sizes = [int(s) for s in metafunc.config.getoption("--matrix-sizes").split(",") if s.strip()]
        metafunc.parametrize("N", sizes)
    if "dtype" in metafunc.fixturenames:
        d = metafunc.config.getoption("--dtype")
        if d in dtype_map:
            metafunc.parametrize("dtype", [dtype_map[d]])
