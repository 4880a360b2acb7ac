"""Build the gfx950 engine and its pybind11 module in-tree.

Outputs (next to each other so that the module finds the library through
its $ORIGIN run path):

    fastctqw_amd/fastexpm/libfastexpm_hip.so      HIP kernels + C ABI
    fastctqw_amd/fastexpm/_fastexpm_core*.so      pybind11 layer (host only)

Run ``python -m fastctqw_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "fastexpm")
LIB = os.path.join(OUT, "libfastexpm_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are left to no code path of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(verbose: bool = True) -> str:
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(CSRC, "fexpm_engine.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


def build_pybind() -> str:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(OUT, "_fastexpm_core" + suffix)
    inc_py = sysconfig.get_paths()["include"]
    cmd = [
        os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-shared", "-fPIC",
        "-fvisibility=hidden",
        f"-I{pybind11.get_include()}", f"-I{inc_py}",
        os.path.join(CSRC, "fexpm_pybind.cpp"),
        f"-L{OUT}", "-lfastexpm_hip", "-Wl,-rpath,$ORIGIN",
        "-o", out + ".tmp",
    ]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build() -> None:
    build_hip()
    build_pybind()


if __name__ == "__main__":
    build()
    sys.exit(0)
