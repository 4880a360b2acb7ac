"""Build the gfx950 engine and its pybind11 module in-tree.

Outputs (next to each other so that the module finds the library through
its $ORIGIN run path):

    fastctqw_amd/fastexpm/libfastexpm_hip.so      HIP kernels + C ABI
    fastctqw_amd/fastexpm/_fastexpm_core*.so      pybind11 layer (host only)
    fastctqw_amd/fastexpm/libfastexpm_probes.so   measurement probes (kernel
                                                  timing, MFMA rates, copy);
                                                  not loaded by the engine

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
PROBES = os.path.join(OUT, "libfastexpm_probes.so")

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


def _build_one(src_name: str, out: str) -> str:
    os.makedirs(OUT, exist_ok=True)
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, src_name)])
    os.replace(tmp, out)
    return out


def build_hip(verbose: bool = True) -> str:
    return _build_one("fexpm_engine.hip", LIB)


def build_probes() -> str:
    return _build_one("fexpm_probes.hip", PROBES)


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
    build_probes()


if __name__ == "__main__":
    build()
    sys.exit(0)
