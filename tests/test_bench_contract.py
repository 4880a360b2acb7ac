"""bench.py contract checks that need no GPU: one GPU only, so any --gpus
other than 1 ends with a JSON error line and exit status 2 before the GPU or
the FP64 reference is touched."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("extra", [[], ["--config", "c4"], ["--config", "c5"]])
def test_more_than_one_gpu_is_refused(extra):
    cp = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", *extra],
                        capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert cp.returncode == 2, cp.stderr
    line = cp.stdout.strip().splitlines()[-1]
    assert "one GPU" in json.loads(line)["error"]


def test_roofline_fraction_is_at_most_the_issued_fraction():
    """The squaring roofline counts the structured algorithm's 3 n^3 flops
    per 3M squaring and 2 n^3 per two-product squaring; the kernels issue
    whole lower 128 x 128 (or 64 x 64) tiles over the padded k, never fewer,
    so frac <= issued_frac, and a phase time at the issued peak rate gives
    frac <= 1."""
    sys.path.insert(0, ROOT)
    import bench

    for n in (1000, 4097, 6144, 8192, 10000, 13001):
        for tile in (64, 128):
            T = (n + tile - 1) // tile
            kpad = (n + 15) // 16 * 16
            issued = 3 * 2.0 * T * (T + 1) / 2 * tile * tile * kpad
            for s, two in ((8, 0), (8, 2), (3, 2), (1, 0)):
                total = issued * ((s - two) + two * 2.0 / 3.0)
                ms = total / (bench.FP32_PEAK_TFLOPS * 1e12) * 1e3
                r = bench.squaring_roofline(n, ms, s, two, issued, bench.FP32_PEAK_TFLOPS)
                assert r["frac"] <= r["issued_frac"] <= 1.0 + 1e-12, (n, tile, s, two, r)
                assert r["achieved_4m_convention"] > r["achieved"]


def test_crt_roofline_fraction():
    """The residue squaring roofline counts 27 n^3 int8 operations per launch
    (27 products, lower half of each); the kernel issues whole lower
    256 x 256 tiles over k padded to 64, never fewer, so frac <= issued_frac,
    and a launch at the issued peak rate gives frac <= 1.  The FP32-equivalent
    rate counts the structured FP32 algorithm's 3 n^3 flops per squaring over
    the whole phase."""
    sys.path.insert(0, ROOT)
    import bench

    for n in (768, 1000, 4097, 8192, 10000, 16384):
        rows = (n + 255) // 256 * 256
        kpad = (n + 63) // 64 * 64
        T = rows // 256
        issued = 27 * 2.0 * T * (T + 1) / 2 * 256 * 256 * kpad
        for s in (1, 8):
            ms = issued / (bench.INT8_PEAK_TOPS * 1e12) * 1e3 * s
            r = bench.crt_roofline(n, ms, s, 2.0 * ms, s)
            assert r["unit"] == "TOP/s" and r["peak"] == bench.INT8_PEAK_TOPS
            assert r["frac"] <= r["issued_frac"] <= 1.0 + 1e-12, (n, s, r)
            assert r["algorithmic_ops_per_launch"] == 27.0 * n ** 3
            fe = r["fp32_equivalent"]
            assert abs(fe["tflops"] - 3.0 * n ** 3 * s / (2.0 * ms * 1e-3) / 1e12) < 1e-9 * fe["tflops"]
