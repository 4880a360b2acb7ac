#!/usr/bin/env python3
"""Benchmark: seconds per N = 10^4 complex CTQW matrix exponential on one MI355X.

Workload (default): Laplacian L = D - A of the seeded Erdos-Renyi graph
G(N = 10000, p = 0.5) (seed 2), A_exp = -i * 4.5 * L in complex64, the full
U = expm(A_exp) materialised in device memory.  One step = one expm with the
input already resident in HBM (upload happens before the timed region).

Prints ONE JSON line (see the keys below).  ``roofline`` describes the
dominant kernel, the squaring GEMM: its average launch duration comes from
HIP events recorded by the engine on the stream the squarings run on
(phase time / s).  ``cpu_baseline`` times scipy.linalg.expm (the reference's
CPU path, FastCTQW/Operator.py:76) on the same N = 10^4 input on the host's
share of cores.

Other workloads (--config): c3 precision study, c4 256 uneven walks, c5
times x states, general (dense general complex input on the LU route and the
reference's performance inputs beside scipy).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
FP16_PEAK_TFLOPS = 2500.0  # dense fp16 / bf16 MFMA
INT8_PEAK_TOPS = 5000.0  # dense int8 MFMA
TRAFFIC_PROFILE = "profiles/squaring_er10000_c64_counters.md"


def _profiled_traffic(key):
    """HBM bytes per launch of a squaring kernel at N = 10^4, as measured by
    the profiler counter runs summarised in TRAFFIC_PROFILE (None if absent)."""
    if key is None:
        return None
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == key:
                    return float(parts[2])
    except OSError:
        pass
    return None


MFMA_PROFILE = "profiles/mfma_sustained_rates.md"


def _sustained_peak(key):
    """Sustained MFMA TFLOP/s of the device measured by the in-register
    probe (bench.py --mfma-rate), as recorded in MFMA_PROFILE (None if absent)."""
    try:
        with open(os.path.join(ROOT, MFMA_PROFILE)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "sustained_tflops" and parts[1] == key:
                    return float(parts[2])
    except OSError:
        pass
    return None


def squaring_roofline(n, ms_square, s, two, issued_3m, peak, split_parts=0):
    """Roofline fields of the s structured squarings X -> X^2 of order n
    (complex symmetric X = Xr + i Xi) that took ms_square in all.  The
    algorithm's own flops are those of three real symmetric products per
    squaring (3M rule, half of each n x n x n product by symmetry), 3 n^3,
    and 2 n^3 for the ``two`` leading squarings in the two-product form,
    times the part products of a split-precision form (split_parts).
    ``issued_3m`` is what one 3M squaring's MFMAs execute (whole lower tiles,
    padded k; the two-product form issues 2/3 of it), never less than the
    algorithmic count, so frac <= issued_frac.  The 4M convention (8 n^3 per
    complex product) is kept under its own key."""
    s = max(int(s), 1)
    two = min(max(int(two), 0), s)
    parts = split_parts if split_parts else 1
    alg = (3.0 * (s - two) + 2.0 * two) * float(n) ** 3 * parts
    issued = issued_3m * ((s - two) + two * 2.0 / 3.0)
    t = ms_square * 1e-3
    return {
        "achieved": alg / t / 1e12, "peak": peak, "unit": "TFLOP/s", "frac": alg / t / 1e12 / peak,
        "algorithmic_flops_per_launch": alg / s, "issued_flops_per_launch": issued / s,
        "issued_tflops": issued / t / 1e12, "issued_frac": issued / t / 1e12 / peak,
        "achieved_4m_convention": 8.0 * float(n) ** 3 * s / t / 1e12,
        "two_product_launches": two,
    }


def crt_roofline(n, kernel_ms, launches, ms_square, s, moduli=9):
    """Roofline fields of the int8 residue squarings: the dominant kernel is
    the batched residue product (3 K int8 products per squaring: yr', yi' and
    s' = yr' + yi' against themselves, modulo K moduli, lower 256 x 256
    tiles).  Its algorithmic work per launch is 3 K n^3 int8 operations (the
    lower half of each n x n x n product: n^2/2 outputs, n multiply-adds,
    2 operations each); ``issued`` counts the whole lower tiles and the padded
    k.  kernel_ms is the launches' own device time (HIP events around each
    launch on the squaring stream).  The squaring phase is also expressed as
    the FP32-equivalent rate of the structured algorithm it replaces (3 n^3
    real flops per squaring), against the FP32 MFMA peak."""
    launches = max(int(launches), 1)
    rows = (n + 255) // 256 * 256
    kpad = (n + 63) // 64 * 64
    T = rows // 256
    nprob = 3.0 * moduli
    alg = nprob * float(n) ** 3
    issued = nprob * 2.0 * (T * (T + 1) // 2) * 256.0 * 256.0 * kpad
    t_launch = kernel_ms / launches * 1e-3
    t_phase = ms_square * 1e-3
    return {
        "achieved": alg / t_launch / 1e12, "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
        "frac": alg / t_launch / 1e12 / INT8_PEAK_TOPS,
        "algorithmic_ops_per_launch": alg, "issued_ops_per_launch": issued,
        "issued_tops": issued / t_launch / 1e12, "issued_frac": issued / t_launch / 1e12 / INT8_PEAK_TOPS,
        "fp32_equivalent": {
            "flops_per_squaring": 3.0 * float(n) ** 3,
            "tflops": 3.0 * float(n) ** 3 * max(int(s), 1) / t_phase / 1e12,
            "frac_of_fp32_peak": 3.0 * float(n) ** 3 * max(int(s), 1) / t_phase / 1e12 / FP32_PEAK_TFLOPS,
            "note": "whole squaring phase (row exponents, residues, products, reconstruction) as the rate of "
                    "the FP32 structured algorithm it replaces, whose results it reproduces more accurately",
        },
    }


def work_flops(n, s, products=6):
    # 4M complex convention: 8 n^3 per complex product/squaring, LU 8/3 n^3, solve 8 n^3
    return (8.0 * (products + s) + 32.0 / 3.0) * float(n) ** 3


def _cpu_share():
    """Host threads for the CPU baseline: the cores this process may run on,
    at most 16 (the host CPU share of one GPU on the measurement boxes)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline(n, t, p, seed, repeats=1, n_est=2000):
    """The reference's CPU path (FastCTQW/Operator.py:76 scipy.linalg.expm,
    then FastCTQW/Ctqw.py:94 U @ psi0) timed on the same N = 10^4 input as
    the GPU run, complex64, psi0 = e_0, OpenBLAS on the host share of cores;
    median of ``repeats`` runs.  The earlier work-model estimate from an
    N = 2000 sample is kept beside it as ``estimate``."""
    import scipy.linalg as sl
    from threadpoolctl import threadpool_limits

    from fastctqw_amd.graphs import c2_laplacian
    from oracle.padeexpm import choose_degree, onenorm, trace_shift

    cores = _cpu_share()
    L = c2_laplacian(n, p, seed, dtype=np.float64)
    a = (-1j * t * L).astype(np.complex64)
    del L
    psi0 = np.zeros(n, dtype=np.complex64)
    psi0[0] = 1.0
    times = []
    with threadpool_limits(limits=cores):
        for _ in range(max(1, repeats)):
            t0 = time.perf_counter()
            u = sl.expm(a)
            _ = u @ psi0
            times.append(time.perf_counter() - t0)
            del u
        del a
        # work-model estimate from a small sample (the figure reported before)
        Ls = c2_laplacian(n_est, p, seed, dtype=np.float64)
        a_s = (-1j * t * Ls).astype(np.complex64)
        _, ash = trace_shift(a_s.astype(np.complex128))
        _, s_s = choose_degree(onenorm(ash))
        sl.expm(a_s)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            sl.expm(a_s)
            ts.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {
        "value": med,
        "unit": "s per N=10k complex expm",
        "cores": cores,
        "kind": "port",
        "host": _host_info(cores),
        "sample": (f"scipy.linalg.expm + U@psi0 on the same complex64 ER({n},{p}) seed {seed} Laplacian, "
                   f"t={t}, psi0=e_0: {len(times)} run(s) at N={n}, median {med:.2f} s"),
        "runs_s": times,
        "estimate": {"n_sample": n_est, "s_sample": s_s, "median_s": float(np.median(ts)),
                     "note": "the small sample behind the earlier work-model extrapolation; the value "
                             "above is measured at the full N"},
    }


def _eigh_reference(n, p, seed, t, out_path, dtype="complex64"):
    """Child process: FP64 eigendecomposition of the matrix the engine
    actually receives.  The engine's input is A = -i*t*L rounded to
    ``dtype``; A = i*G with G = Im(A) real symmetric (exact in FP64), so
    expm(A) = V diag(e^{i lam}) V^T with (lam, V) = eigh(G), computed with
    torch.linalg.eigh on the GPU in FP64.  Writes psi_ref for e_0 and for a
    seeded random state.  Runs in its own process so that torch's HIP
    runtime never shares a process with the engine."""
    import torch

    from fastctqw_amd.graphs import c2_laplacian, random_state

    a = (-1j * t * c2_laplacian(n, p, seed, dtype=np.float64)).astype(dtype)
    G = torch.from_numpy(np.ascontiguousarray(a.imag, dtype=np.float64)).cuda()
    del a
    t0 = time.perf_counter()
    lam, V = torch.linalg.eigh(G)
    torch.cuda.synchronize()
    t_eig = time.perf_counter() - t0
    res = torch.linalg.norm(G @ V - V * lam[None, :]) / torch.linalg.norm(G)
    ph = torch.exp(1j * lam.to(torch.complex128))
    Vc = V.to(torch.complex128)
    phi = torch.from_numpy(random_state(n, seed=55)).cuda()
    e0 = torch.zeros(n, dtype=torch.complex128, device="cuda")
    e0[0] = 1.0
    refs = [Vc @ (ph * (Vc.T @ v)) for v in (e0, phi)]
    np.savez(out_path, psi_e0=refs[0].cpu().numpy(), psi_rand=refs[1].cpu().numpy(),
             eig_residual=float(res), eig_seconds=t_eig)


def _c5_inputs(n, p, seed, dt, nstates, dtype="complex64"):
    from fastctqw_amd.graphs import erdos_renyi, laplacian, random_state

    L = laplacian(erdos_renyi(n, p, seed, weighted=True))
    a = (-1j * dt * L).astype(dtype)
    psi0 = np.stack([random_state(n, seed=55, index=k) for k in range(nstates)])
    return a, psi0


def _c5_reference(n, p, seed, dt, nstates, nsteps, out_path):
    """Child process: FP64 eigendecomposition of G = Im(A) for the same
    rounded input and psi_ref(t_k) = V e^{i k lam} V^T psi0 for every k and
    state (torch on the GPU)."""
    import torch

    a, psi0 = _c5_inputs(n, p, seed, dt, nstates)
    G = torch.from_numpy(np.ascontiguousarray(a.imag, dtype=np.float64)).cuda()
    lam, V = torch.linalg.eigh(G)
    Vc = V.to(torch.complex128)
    C = Vc.T @ torch.from_numpy(psi0.T.copy()).cuda()
    refs = np.empty((nsteps, nstates, n), dtype=np.complex128)
    for k in range(1, nsteps + 1):
        refs[k - 1] = (Vc @ (torch.exp(1j * k * lam.to(torch.complex128))[:, None] * C)).T.cpu().numpy()
    np.save(out_path, refs)


def run_c5(args):
    """Many times x many states: one weighted ER(8192, 0.3) Laplacian, U(dt)
    with dt = 0.5, psi(t_k) = U^k psi0 for k = 1..32 and 128 random states."""
    import subprocess
    import tempfile

    n, p, seed, dt, nsteps, nstates = args.n, 0.3, 5, 0.5, 32, 128
    ref, ref_error = None, None
    if not args.no_accuracy:
        path = os.path.join(tempfile.gettempdir(), f"fexpm_c5_{os.getpid()}.npy")
        cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--c5-ref", path, "--n", str(n)],
                            timeout=900)
        if cp.returncode == 0 and os.path.exists(path):
            ref = np.load(path)
            os.remove(path)
        else:
            ref_error = f"reference failed (exit code {cp.returncode})"
    from fastctqw_amd.fastexpm import ExpMatComplex64

    a, psi0 = _c5_inputs(n, p, seed, dt, nstates)
    eng = ExpMatComplex64(n)
    _engine_options(eng, args)
    eng.upload(a)
    for _ in range(args.warmup):
        eng.run_resident()
        eng.propagate(psi0, nsteps)
    dev_ms, wall = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        eng.run_resident()
        out = eng.propagate(psi0, nsteps)
        wall.append(time.perf_counter() - t0)
        st = eng.stats()
        dev_ms.append(st["ms_total"] + st["ms_propagate"])
    walks = nsteps * nstates
    dev = float(np.median(dev_ms))
    res = {
        "metric": "batched walks/s (32 times x 128 states, N=8192)",
        "value": walks / (dev / 1000.0), "unit": "walks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dev, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (fp32 mfma squarings and propagation, pade products from 6 exact int8 slices, "
                                            "fp64 solve/refinement)",
        "data": "synthetic: seeded weighted ER(8192,0.3) Laplacian, 128 seeded random states",
        "config": {"workload": "C5: U(0.5) once, psi(t_k)=U^k psi0, k=1..32, 128 states", "n": n,
                   "m": st["m"], "s": st["s"], "expm_ms": st["ms_total"], "propagate_ms": st["ms_propagate"],
                   "wall_ms_incl_pcie": 1000.0 * float(np.median(wall))},
    }
    if ref is not None:
        inf = [_infidelity(ref[k, j], out[k, j]) for k in range(nsteps) for j in range(nstates)]
        res["accuracy"] = {"worst_infidelity": float(max(inf)), "median_infidelity": float(np.median(inf)),
                           "oracle": "FP64 eigendecomposition of G = Im(A) (same rounded input)"}
    elif ref_error:
        res["accuracy"] = {"oracle_error": ref_error}
    if not args.no_split:
        # the same walks with the expm squarings in fp16x2
        eng.set_precision("fp16x2")
        eng.run_resident()
        eng.propagate(psi0, nsteps)
        hm = []
        for _ in range(args.steps):
            eng.run_resident()
            out_h = eng.propagate(psi0, nsteps)
            sth = eng.stats()
            hm.append(sth["ms_total"] + sth["ms_propagate"])
        sp = {"precision": "fp16x2", "walks_per_s": walks / (float(np.median(hm)) / 1000.0),
              "ms_per_step": float(np.median(hm))}
        if ref is not None:
            sp["worst_infidelity"] = float(max(_infidelity(ref[k, j], out_h[k, j])
                                               for k in range(nsteps) for j in range(nstates)))
        res["split_precision"] = sp
        eng.set_precision("native")
    eng.free()
    print(json.dumps(res), flush=True)


GENERAL_SIZES = (16, 32, 64, 128, 256, 512, 1024)


def _gaussian_general(n, seed=7):
    """Dense general complex64 matrix (G_r + i G_i) * 3 / sqrt(n) with standard
    normal G: spectrum in a disk of radius about 3, 1-norm about 3.8 sqrt(n)."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n), dtype=np.float32) + 1j * rng.standard_normal((n, n), dtype=np.float32)
    return (g * np.float32(3.0 / np.sqrt(n))).astype(np.complex64)


def run_general(args):
    """General complex inputs on the LU route.

    Timed metric: one expm of a dense general complex64 matrix of order N
    (default 10^4, input resident).  Accuracy by the group law
    e^{-A} (e^{A} x) = x on 4 random vectors (FP64 applies on the device).
    Beside it, the reference's performance inputs A = Adj + 1j * rand
    (complete graph, G(n, 0.05), 4-regular; tests/test_fastexpm_perf.py) at
    its default sizes 16..1024, timed as its benchmark does (host array in,
    host U out) next to scipy.linalg.expm on the same array, with the
    relative error against scipy's complex128 expm of that array where the
    result fits complex64."""
    import scipy.linalg

    from fastctqw_amd.fastexpm import ExpMatComplex64
    from fastctqw_amd.graphs import reference_perf_input

    n = args.n
    a = _gaussian_general(n)
    eng = ExpMatComplex64(n)
    eng.upload(a)
    for _ in range(args.warmup):
        eng.run_resident()
    ms, phases = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        eng.run_resident()
        ms.append(1000.0 * (time.perf_counter() - t0))
        phases.append(eng.stats())
    st = phases[-1]
    rng = np.random.default_rng(11)
    x = rng.standard_normal((4, n)) + 1j * rng.standard_normal((4, n))
    y = eng.apply(x)
    eng.upload((-a).astype(np.complex64))
    eng.run_resident()
    back = eng.apply(y)
    eng.free()
    del a
    group_law = float(np.linalg.norm(back - x) / np.linalg.norm(x))
    ms_step = float(np.median(ms))
    s = st["s"]
    npad = (n + 127) // 128 * 128
    sq_ms = float(np.mean([p["ms_square"] for p in phases])) / max(s, 1)
    issued = 3 * 2.0 * npad ** 3  # 3M: three real products of npad^3 (all tiles)
    res = {
        "metric": f"s per N={n} general complex expm (LU route)",
        "value": ms_step / 1000.0, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp64 pade products and pivoted LU solve, fp32 mfma squarings)",
        "data": "synthetic: seeded dense (G_r + i G_i) * 3/sqrt(N), G standard normal",
        "config": {"workload": f"expm of a dense general complex64 matrix, N={n}", "n": n,
                   "m": st["m"], "s": s, "structured": st["structured"]},
        "phases_ms": {k: float(np.median([p[k] for p in phases]))
                      for k in ("ms_analyze", "ms_pade", "ms_solve", "ms_square", "ms_output", "ms_total")},
        "squaring": {"kernel": "gemm_nt_kernel (3M complex, all tiles, FP32 MFMA)", "avg_launch_ms": sq_ms,
                     "issued_tflops": issued / (sq_ms * 1e-3) / 1e12,
                     "issued_frac_of_fp32_peak": issued / (sq_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS},
        "accuracy": {"group_law_residual": group_law,
                     "check": "||e^{-A} (e^{A} x) - x|| / ||x|| over 4 random x (FP64 applies)"},
    }
    if args.no_perf_inputs:
        print(json.dumps(res), flush=True)
        return
    from threadpoolctl import threadpool_limits

    rows = []
    cores = _cpu_share()
    for fam in ("complete", "er", "r4"):
        for m in GENERAL_SIZES:
            am = reference_perf_input(fam, m, seed=m)
            with np.errstate(all="ignore"), threadpool_limits(limits=cores):
                t0 = time.perf_counter()
                ref = scipy.linalg.expm(am.astype(np.complex128))
                ref_first = time.perf_counter() - t0
            reps = max(3, min(50, int(0.5 / max(ref_first, 1e-4))))
            # scipy on one host thread and on the host share: the faster
            # (BLAS threads cost more than they gain at the small orders)
            cpu_by = {}
            for th in (1, cores):
                with np.errstate(all="ignore"), threadpool_limits(limits=th):
                    ts = []
                    for _ in range(reps):
                        t0 = time.perf_counter()
                        scipy.linalg.expm(am)
                        ts.append(1000.0 * (time.perf_counter() - t0))
                cpu_by[th] = float(np.median(ts))
            best_th = min(cpu_by, key=cpu_by.get)
            cpu = [cpu_by[best_th]]
            with ExpMatComplex64(m) as e:
                u = e.run(am)
                gpu = []
                for _ in range(max(5, reps)):
                    t0 = time.perf_counter()
                    u = e.run(am)
                    gpu.append(1000.0 * (time.perf_counter() - t0))
                se = e.stats()
            fits = bool(np.isfinite(ref).all() and np.abs(ref).max() < np.finfo(np.float32).max)
            row = {"family": fam, "n": m, "m": se["m"], "s": se["s"], "fastexpm_ms": float(np.median(gpu)),
                   "scipy_complex64_ms": cpu[0], "scipy_threads": best_th,
                   "speedup": cpu[0] / float(np.median(gpu))}
            if fits:
                row["rel_err_vs_scipy_c128"] = float(np.linalg.norm(u - ref) / np.linalg.norm(ref))
            else:
                row["result_overflows_complex64"] = True
                row["output_nonfinite_fraction"] = float((~np.isfinite(u)).mean())
            rows.append(row)
    res["reference_perf_inputs"] = {
        "note": "A = Adj + 1j*rand(n,n), complex64 (the reference's performance inputs); times are host "
                "array in, host U out (median), as the reference's benchmark calls run(A); scipy: the faster "
                f"of 1 and {cores} host threads; the complete graph overflows complex64 from n=128",
        "rows": rows}
    print(json.dumps(res), flush=True)


C3_TIMES = (1.0, 4.5, 20.0)
C3_PRECISIONS = ("native", "int8crt", "fp16x2", "bf16x3")
# C3 cases: the squaring forms, the residue form with each moduli count
C3_CASES = ("native", "int8crt:9", "int8crt:8", "int8crt:7", "fp16x2", "bf16x3")
C3_PROBES = 8


def _c3_graph(n, weighted):
    from fastctqw_amd.graphs import c2_laplacian

    return c2_laplacian(n, 0.5, 2, weighted=weighted, dtype=np.float64)


def _c3_probes(n):
    """Gaussian probe columns Z (n x C3_PROBES): E||E Z||_F^2 = C3_PROBES ||E||_F^2."""
    rng = np.random.Generator(np.random.PCG64(33))
    return rng.standard_normal((n, C3_PROBES)) + 1j * rng.standard_normal((n, C3_PROBES))


def _c3_reference(n, out_path):
    """Child process: per graph (unweighted, weighted) and t, FP64 eigh of
    G = Im(A) for the complex64 input A = -i t L, psi_ref for e_0 and a
    random state, and U_ref Z for the probe columns (torch on the GPU)."""
    import torch

    from fastctqw_amd.graphs import random_state

    res = {}
    Z = torch.from_numpy(_c3_probes(n)).cuda()
    psi_r = torch.from_numpy(random_state(n, seed=55)).cuda()
    for weighted in (False, True):
        L = _c3_graph(n, weighted)
        for t in C3_TIMES:
            a_im = (-1j * t * L).astype(np.complex64).imag
            G = torch.from_numpy(np.ascontiguousarray(a_im, dtype=np.float64)).cuda()
            lam, V = torch.linalg.eigh(G)
            del G
            Vc = V.to(torch.complex128)
            del V
            ph = torch.exp(1j * lam.to(torch.complex128))
            key = f"{int(weighted)}_{t}"
            res[f"e0_{key}"] = (Vc @ (ph * Vc[0, :])).cpu().numpy()
            res[f"rand_{key}"] = (Vc @ (ph * (Vc.T @ psi_r))).cpu().numpy()
            res[f"UZ_{key}"] = (Vc @ (ph[:, None] * (Vc.T @ Z))).cpu().numpy()
            del Vc
            torch.cuda.empty_cache()
    np.savez(out_path, **res)


def run_c3(args):
    """Precision study at N=10^4: unweighted and weighted ER(N,0.5)
    Laplacians, t in {1, 4.5, 20}, complex64; squarings in FP32 MFMA
    (native), from exact int8 residue products (int8crt, the default at this
    order) and as fp16x2 / bf16x3 split products with FP32 accumulation
    (Pade products, solve and norm FP64-accurate in every mode).  Per case:
    device ms per expm, infidelity of U e_0 and of U psi_rand, and
    ||U - U_ref||_F / ||U_ref||_F estimated with Gaussian probes."""
    import subprocess
    import tempfile

    from fastctqw_amd.fastexpm import ExpMatComplex64, _fastexpm_core
    from fastctqw_amd.graphs import random_state

    n = args.n
    ref, ref_error = None, None
    if not args.no_accuracy:
        path = os.path.join(tempfile.gettempdir(), f"fexpm_c3_{os.getpid()}.npz")
        print("[bench] FP64 eigh references (child process) ...", file=sys.stderr, flush=True)
        cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--c3-ref", path, "--n", str(n)],
                            timeout=1200)
        if cp.returncode == 0 and os.path.exists(path):
            ref = dict(np.load(path))
            os.remove(path)
        else:
            ref_error = f"reference failed (exit code {cp.returncode})"
    Z = _c3_probes(n)
    psi_r = random_state(n, seed=55)
    e0 = np.zeros(n, dtype=np.complex128)
    e0[0] = 1.0
    eng = ExpMatComplex64(n)
    _engine_options(eng, args)
    c3_cases = tuple(args.c3_cases.split(",")) if args.c3_cases else C3_CASES
    c3_graphs = (False, True) if args.c3_graphs == "both" else ((args.c3_graphs == "weighted"),)
    cases = []
    for weighted in c3_graphs:
        L = _c3_graph(n, weighted)
        for t in C3_TIMES:
            a = (-1j * t * L).astype(np.complex64)
            eng.upload(a)
            del a
            key = f"{int(weighted)}_{t}"
            for case_name in c3_cases:
                prec, _, kmod = case_name.partition(":")
                eng.set_precision(prec)
                if kmod:
                    eng._backend.set_option(_fastexpm_core.OPT_CRT_MODULI, int(kmod))
                for _ in range(args.warmup):
                    eng.run_resident()
                ms = []
                for _ in range(args.steps):
                    eng.run_resident()
                    ms.append(eng.stats()["ms_total"])
                st = eng.stats()
                last_st = st
                case = {"graph": "weighted" if weighted else "unweighted", "t": t, "precision": prec,
                        "s": st["s"], "ms": float(np.median(ms)), "ms_square": st["ms_square"]}
                if prec == "int8crt":
                    case["moduli"], case["beta"] = st["crt_moduli"], st["crt_beta"]
                if ref is not None:
                    out = eng.apply(np.vstack([e0, psi_r, Z.T]))
                    uz = out[2:].T
                    case["infidelity_e0"] = _infidelity(ref[f"e0_{key}"], out[0])
                    case["infidelity_random_state"] = _infidelity(ref[f"rand_{key}"], out[1])
                    case["U_relF_probe"] = float(np.linalg.norm(uz - ref[f"UZ_{key}"]) /
                                                 np.linalg.norm(ref[f"UZ_{key}"]))
                cases.append(case)
                print(f"[bench] {case}", file=sys.stderr, flush=True)
        del L
    eng.free()
    head = [c for c in cases if c["graph"] == "unweighted" and c["t"] == 4.5] or cases
    native = next((c for c in head if c["precision"] == "native"), None)
    best = min(head, key=lambda c: c["ms"])
    res = {
        "metric": "s per N=10k complex expm, squaring precision study (fastest mode at t=4.5, unweighted)",
        "value": best["ms"] / 1000.0, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": f"c64 ({_square_phrase(best['precision'], last_st)}, {_slice_phrase(last_st)}",
        "data": "synthetic: seeded ER(10000,0.5) Laplacian, unweighted and weighted U(0,1]",
        "config": {"workload": "C3: expm(-i t L), t in {1, 4.5, 20}, precision native/int8crt (7, 8, 9 moduli)/"
                               "fp16x2/bf16x3",
                   "n": n, "native_ms_t4.5": native["ms"] if native else None,
                   "speedup_vs_native": native["ms"] / best["ms"] if native else None},
        "cases": cases,
    }
    if ref_error:
        res["accuracy"] = {"oracle_error": ref_error}
    else:
        res["accuracy"] = {"oracle": "FP64 eigendecomposition of G = Im(A) per (graph, t), same rounded input; "
                                     f"U_relF_probe from {C3_PROBES} Gaussian probe columns"}
    print(json.dumps(res), flush=True)


C4_COUNT, C4_SEED = 256, 4000
C4_CHECK_U = (0, 1, 2, 3, 64, 65, 128, 129)  # walks whose full U is compared with U_ref


def _c4_states(walks_n, first=0):
    """Uniform superposition and a seeded random state for walks first, first+1, ..."""
    from fastctqw_amd.graphs import random_state

    uni = [np.full(n, 1.0 / np.sqrt(n), dtype=np.complex128) for n in walks_n]
    rnd = [random_state(n, seed=44, index=first + i) for i, n in enumerate(walks_n)]
    return uni, rnd


def _c4_reference(out_path, count):
    """Child process: for each C4 walk, FP64 eigendecomposition of
    G = Im(A) of the complex64 input (torch on the GPU), psi_ref for the
    uniform and a random state, and U_ref for a few walks."""
    import torch

    from fastctqw_amd.graphs import iter_c4_walks

    res = {}
    ns = []
    for i, w in enumerate(iter_c4_walks(count, C4_SEED)):
        a = w["A"].astype(np.complex64)
        ns.append(w["n"])
        G = torch.from_numpy(np.ascontiguousarray(a.imag, dtype=np.float64)).cuda()
        lam, V = torch.linalg.eigh(G)
        Vc = V.to(torch.complex128)
        ph = torch.exp(1j * lam.to(torch.complex128))
        uni, rnd = _c4_states([w["n"]], first=i)
        for tag, v in (("uni", uni[0]), ("rnd", rnd[0])):
            x = torch.from_numpy(v).cuda()
            res[f"{tag}{i}"] = (Vc @ (ph * (Vc.T @ x))).cpu().numpy()
        if i in C4_CHECK_U:
            res[f"U{i}"] = ((Vc * ph[None, :]) @ Vc.T).cpu().numpy()
    res["n"] = np.asarray(ns)
    np.savez(out_path, **res)


def c4_cpu_baseline(mats, s_list, budget_s=20.0):
    """scipy.linalg.expm + U@psi0 (FastCTQW/Operator.py:76, Ctqw.py:94) on
    the first walks of the batch until the budget is spent; the batch time
    is extrapolated by the work model summed over all walks."""
    import scipy.linalg as sl
    from threadpoolctl import threadpool_limits

    cores = min(16, os.cpu_count() or 1)
    done, spent = 0, 0.0
    with threadpool_limits(limits=cores):
        sl.expm(mats[0])
        for a in mats:
            t0 = time.perf_counter()
            u = sl.expm(a)
            _ = u @ np.full(a.shape[0], 1.0 / np.sqrt(a.shape[0]), dtype=a.dtype)
            spent += time.perf_counter() - t0
            done += 1
            if spent > budget_s:
                break
    w_all = sum(work_flops(a.shape[0], s) for a, s in zip(mats, s_list))
    w_done = sum(work_flops(a.shape[0], s) for a, s in zip(mats[:done], s_list[:done]))
    batch_s = spent * w_all / w_done
    return {
        "value": len(mats) / batch_s, "unit": "walks/s", "cores": cores, "kind": "port", "host": _host_info(cores),
        "sample": (f"scipy.linalg.expm + U@psi0 (complex64) on the first {done} of {len(mats)} walks: "
                   f"{spent:.2f} s, extrapolated by sum W(N_i, s_i) to {batch_s:.1f} s per batch"),
    }


def run_c4(args):
    """256 independent walks of uneven order, graph, H and t: U_i and
    psi_i = U_i psi0 (uniform superposition) for every walk."""
    import subprocess
    import tempfile

    from fastctqw_amd.graphs import iter_c4_walks

    count = args.c4_count
    # one hardware queue per lane (read when HIP initialises in this process):
    # lanes that share a queue execute their kernels one after another
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(4, args.lanes)))
    ref_path = os.path.join(tempfile.gettempdir(), f"fexpm_c4_{os.getpid()}.npz")
    child = None
    if not args.no_accuracy:
        # reference runs on the GPU in a child while this process generates
        # the inputs on the host (this process touches the GPU afterwards)
        child = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--c4-ref", ref_path,
                                  "--c4-count", str(count)])
    t_gen = time.perf_counter()
    mats, kinds = [], []
    for w in iter_c4_walks(count, C4_SEED):
        mats.append(w["A"].astype(np.complex64))
        kinds.append((w["kind"], w["laplacian"], w["t"]))
    ns = [a.shape[0] for a in mats]
    print(f"[bench] {count} walks generated in {time.perf_counter() - t_gen:.1f} s, "
          f"sum N^3 = {sum(float(n) ** 3 for n in ns):.3e}", file=sys.stderr, flush=True)
    ref, ref_error = None, None
    if child is not None:
        try:
            rc = child.wait(timeout=900)
        except subprocess.TimeoutExpired:
            child.kill()
            rc = "timeout"
        if rc == 0 and os.path.exists(ref_path):
            ref = dict(np.load(ref_path))
            os.remove(ref_path)
        else:
            ref_error = f"reference failed ({rc})"
        print(f"[bench] reference done: {ref_error or 'ok'}", file=sys.stderr, flush=True)

    from fastctqw_amd.fastexpm import ExpMatBatch

    uni, rnd = _c4_states(ns)
    b = ExpMatBatch(nmax=max(ns), dtype=np.complex64, lanes=args.lanes)
    b.set_grouped(not args.c4_lanes)
    b.upload(mats)
    for _ in range(args.warmup):
        b.run_resident(uni)
        print(f"[bench] warmup {b.last_ms():.1f} ms {b.report()}", file=sys.stderr, flush=True)
    ms, dev_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        psi = b.run_resident(uni)
        ms.append(b.last_ms())
        dev_ms.append(b.report()["device_ms"])
    wall = time.perf_counter() - t0
    report = b.report()
    st = [b.stats(i) for i in range(count)]
    dev = float(np.median(ms))
    s_list = [x["s"] for x in st]
    flops = sum(work_flops(n, s) for n, s in zip(ns, s_list))
    res = {
        "metric": f"batched walks/s ({count} uneven walks, N in [64, 2048])",
        "value": count / (dev / 1000.0), "unit": "walks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dev, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c64 (fp32 mfma squarings; " +
                  ("chebyshev polynomial of each walk at its 1-norm bound from fp64-accurate int8-slice products"
                   if all(x.get("chebyshev", 0) for x in st) else "fp64 pade/solve") + ", fp64 psi)"),
        "data": "synthetic: seeded d-regular / ER graphs, weights U(0,1], adjacency or Laplacian, t in [0.1,20]",
        "config": {"workload": f"C4: {count} walks, U_i=expm(-i t_i H_i), psi_i=U_i psi0 (uniform)",
                   "form": "lanes" if args.c4_lanes else "grouped launches, hipGraph, autotuned tiles",
                   "lanes": args.lanes, "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                   "sum_n3": float(sum(float(n) ** 3 for n in ns)),
                   "s_histogram": {str(k): s_list.count(k) for k in sorted(set(s_list))},
                   "structured_walks": int(sum(x["structured"] for x in st)),
                   "model_tflops": flops / (dev * 1e-3) / 1e12,
                   "wall_ms_per_step_incl_psi_copies": 1000.0 * wall / args.steps},
        "batch_report": report,
        "grouped_device_ms": float(np.median(dev_ms)) if report.get("grouped") else None,
    }
    # the dominant class of grouped launches: the FP32 squaring rounds, three
    # real symmetric half products (3 n^3 flops) per squaring and walk, over
    # the class's GEMM device time of the chosen autotuner candidate
    cms = report.get("class_ms") or []
    if report.get("grouped") and cms and cms[0] > 0:
        sq_alg = sum(3.0 * float(n) ** 3 * x["s"] for n, x in zip(ns, st) if x["structured"])
        res["squaring_class"] = {"algorithmic_flops": sq_alg, "class_ms": cms[0],
                                 "tflops": sq_alg / (cms[0] * 1e-3) / 1e12,
                                 "frac_of_fp32_peak": sq_alg / (cms[0] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS}
    # the same batch with fp16x2 squarings, and walk by walk on one lane (the unbatched form)
    variants = {}
    if not args.no_variants:
        if not args.c4_lanes:
            # the previous form: concurrent lanes (streams + host threads)
            b.set_grouped(False)
            b.run_resident(uni)
            lm = []
            for _ in range(args.steps):
                b.run_resident(uni)
                lm.append(b.last_ms())
            variants["lanes"] = {"walks_per_s": count / (float(np.median(lm)) / 1000.0),
                                 "ms_per_step": float(np.median(lm)), "lanes": args.lanes}
            b.set_grouped(True)
        for prec in ("fp16x2", "bf16x3"):
            # split-precision squaring rounds (grouped unless --c4-lanes); the first run tunes
            b.set_precision(prec)
            b.run_resident(uni)
            vm = []
            for _ in range(args.steps):
                psi_h = b.run_resident(uni)
                vm.append(b.last_ms())
            variants[prec] = {"walks_per_s": count / (float(np.median(vm)) / 1000.0),
                              "ms_per_step": float(np.median(vm)), "report": b.report()}
            if ref is not None:
                variants[prec]["worst_infidelity_uniform"] = float(
                    max(_infidelity(ref[f"uni{i}"], psi_h[i]) for i in range(count)))
        b.set_precision("native")
        b1 = ExpMatBatch(nmax=max(ns), dtype=np.complex64, lanes=1)
        b1.set_grouped(False)
        b1.upload(mats)
        b1.run_resident(uni)
        t1 = []
        for _ in range(max(1, args.steps - 1)):
            b1.run_resident(uni)
            t1.append(b1.last_ms())
        variants["one_lane"] = {"walks_per_s": count / (float(np.median(t1)) / 1000.0),
                                "ms_per_step": float(np.median(t1)),
                                "note": "same walks one after another on one stream (no batching)"}
        del b1
    res["variants"] = variants
    if ref is not None:
        psi_r = b.run_resident(rnd)
        inf_u = [_infidelity(ref[f"uni{i}"], psi[i]) for i in range(count)]
        inf_r = [_infidelity(ref[f"rnd{i}"], psi_r[i]) for i in range(count)]
        relf = {}
        for i in C4_CHECK_U:
            if i < count:
                u = b.download(i).astype(np.complex128)
                relf[str(i)] = float(np.linalg.norm(u - ref[f"U{i}"]) / np.linalg.norm(ref[f"U{i}"]))
        worst = int(np.argmax(inf_r))
        res["accuracy"] = {
            "oracle": "FP64 eigendecomposition of G = Im(A) per walk (same rounded complex64 input)",
            "worst_infidelity_uniform": float(max(inf_u)),
            "worst_infidelity_random_state": float(max(inf_r)),
            "median_infidelity_random_state": float(np.median(inf_r)),
            "worst_walk": {"i": worst, "n": ns[worst], "kind": kinds[worst][0], "laplacian": kinds[worst][1],
                           "t": kinds[worst][2], "s": s_list[worst]},
            "U_relF": relf,
        }
    elif ref_error:
        res["accuracy"] = {"oracle_error": ref_error}
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        res["cpu_baseline"] = c4_cpu_baseline(mats, s_list)
    print(json.dumps(res), flush=True)


def _c128_run(a64, args, ref):
    """The same expm in complex128 (FP64 MFMA throughout).  The C2 input
    -4.5i L has entries 4.5 * integer, exact in complex64, so the complex64
    eigen reference is the reference of this input too."""
    from fastctqw_amd.fastexpm import ExpMatComplex128
    from fastctqw_amd.graphs import random_state

    n = a64.shape[0]
    a = a64.astype(np.complex128)
    exact = bool(np.array_equal(a.astype(np.complex64), a64))
    eng = ExpMatComplex128(n)
    eng.upload(a)
    del a
    for _ in range(args.warmup):
        eng.run_resident()
    ms = []
    for _ in range(args.steps):
        eng.run_resident()
        ms.append(eng.stats()["ms_total"])
    st = eng.stats()
    res = {"ms_per_step": float(np.median(ms)), "s": st["s"],
           "phases_ms": {k: st[k] for k in ("ms_pade", "ms_solve", "ms_square", "ms_total")},
           "input_equals_complex64": exact}
    if ref is not None and exact:
        psi = eng.apply(np.vstack([np.eye(1, n, dtype=np.complex128), random_state(n, seed=55)[None, :]]))
        res["infidelity_e0"] = _infidelity(ref["psi_e0"], psi[0])
        res["infidelity_random_state"] = _infidelity(ref["psi_rand"], psi[1])
        res["rel_err_e0"] = float(np.linalg.norm(psi[0] - ref["psi_e0"]) / np.linalg.norm(ref["psi_e0"]))
        res["rel_err_random_state"] = float(np.linalg.norm(psi[1] - ref["psi_rand"]) / np.linalg.norm(ref["psi_rand"]))
    eng.free()
    return res


def _api_timing(eng, a, args):
    """Host-inclusive times (not the headline value): ExpMatComplex64.run on
    a host array (PCIe both ways), and Ctqw.runWalk end to end (adjacency
    upload, H and A_exp on the device, expm, device U psi0) on the same
    graph: adjacency A, Laplacian, gamma = -1 so that H = L, psi0 = e_0."""
    from fastctqw_amd.Ctqw import Ctqw
    from fastctqw_amd.State import State
    from fastctqw_amd.graphs import erdos_renyi

    t0 = time.perf_counter()
    eng.run(a)
    api_ms = 1000.0 * (time.perf_counter() - t0)
    adj = erdos_renyi(args.n, args.p, args.seed, dtype=np.float32)
    walk = Ctqw.from_numpy_array(adj, State.getSingleNodeState(args.n, 0), gamma=-1.0, laplacian=True,
                                 device="cuda")
    del adj
    walk.runWalk(args.t)  # warm-up (engine creation)
    t0 = time.perf_counter()
    walk.runWalk(args.t)
    rw_ms = 1000.0 * (time.perf_counter() - t0)
    walk.getOperator().free()
    return {"run_numpy_ms": api_ms, "ctqw_runwalk_ms": rw_ms,
            "note": "run_numpy: host array in, host U out (PCIe both ways); runWalk: float32 adjacency upload, "
                    "H and A_exp formed on the device, expm, U psi0 on the device"}


def _probes():
    """The measurement probes library (built beside the engine by
    fastctqw_amd.build; not used by the expm path)."""
    import ctypes

    path = os.path.join(ROOT, "fastctqw_amd", "fastexpm", "libfastexpm_probes.so")
    if not os.path.exists(path):
        raise RuntimeError("libfastexpm_probes.so is not built: run python -m fastctqw_amd.build")
    lib = ctypes.CDLL(path)
    lib.fexpm_probe_last_error.restype = ctypes.c_char_p
    return lib


def _copy_calibration(mib):
    """A plain device copy of known size (probes library); profiler counters
    of this launch calibrate FETCH_SIZE / WRITE_SIZE."""
    import ctypes

    lib = _probes()
    ms = ctypes.c_double(0.0)
    nbytes = int(mib) << 20
    rc = lib.fexpm_probe_copy(ctypes.c_int64(nbytes), 3, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(lib.fexpm_probe_last_error().decode())
    print(json.dumps({"copy_bytes": nbytes, "launches": 4, "ms_per_launch": ms.value,
                      "GB_s": 2.0 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)


def _mfma_rates():
    """Sustained MFMA rates of the device (in-register MFMA loop, probes
    library), beside the datasheet peaks the roofline uses."""
    import ctypes

    lib = _probes()
    out = {}
    for kind, name, peak, iters in ((0, "fp32_16x16x4", FP32_PEAK_TFLOPS, 20000),
                                    (1, "fp64_16x16x4", FP64_PEAK_TFLOPS, 20000),
                                    (2, "fp16_32x32x16", FP16_PEAK_TFLOPS, 8000),
                                    (3, "bf16_32x32x16", FP16_PEAK_TFLOPS, 8000),
                                    (4, "int8_32x32x32", INT8_PEAK_TOPS, 4000),
                                    (5, "fp32_32x32x2", FP32_PEAK_TFLOPS, 20000)):
        tf, ms = ctypes.c_double(0.0), ctypes.c_double(0.0)
        rc = lib.fexpm_probe_mfma_rate(kind, iters, ctypes.byref(tf), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(lib.fexpm_probe_last_error().decode())
        out[name] = {"tflops": tf.value, "datasheet_tflops": peak, "fraction_of_datasheet": tf.value / peak,
                     "kernel_ms": ms.value}
        # the same MFMA stream at 1, 2 and 4 waves per SIMD
        by_waves = {}
        for w in (1, 2, 4):
            rc = lib.fexpm_probe_mfma_rate_waves(kind, w, iters, ctypes.byref(tf), ctypes.byref(ms))
            if rc != 0:
                raise RuntimeError(lib.fexpm_probe_last_error().decode())
            by_waves[w] = tf.value
        out[name]["tflops_by_waves_per_simd"] = by_waves
    # the GEMM kernels at the N = 10^4 shapes, issued MFMA flops per launch time
    gem = {}
    n, kp = 10000, 10112
    lib.fexpm_probe_gemm_time.argtypes = [ctypes.c_int] * 8 + [ctypes.POINTER(ctypes.c_double)]
    for name, w, mode, cfg, tile, mult in (("fp32_c3m_sym_128x128_8waves", 0, 2, 0, 128, 3),
                                           ("fp32_c3m_sym_64x64_4waves", 0, 2, 4, 64, 3),
                                           ("fp32_c3m_sym_128x128_16waves", 0, 2, 12, 128, 3),
                                           ("fp32_c3m_sym_128x128_8waves_pad8", 0, 2, 5, 128, 3),
                                           ("fp32_c3m_sym_128x128_8waves_32x64", 0, 2, 6, 128, 3),
                                           ("fp32_c3m_sym_128x128_8waves_pad12", 0, 2, 7, 128, 3),
                                           ("fp32_c3m_sym_128x128_8waves_32x64_pad8", 0, 2, 8, 128, 3),
                                           ("fp32_c3m_sym_128x128_8waves_mfma16x16x4", 0, 2, 9, 128, 3),
                                           ("fp32_c3m_sym_128x128_8waves_pad8_mfma32", 0, 2, 10, 128, 3),
                                           ("fp32_c3m_sym_128x128_8waves_32x64_mfma32", 0, 2, 11, 128, 3),
                                           ("fp64_real_sym_64x64_8waves", 1, 0, 1, 64, 1),
                                           ("fp64_real_sym_64x64_4waves", 1, 0, 0, 64, 1),
                                           ("fp64_c3m_sym_64x64_8waves", 1, 2, 1, 64, 3),
                                           ("fp64_c3m_sym_64x64_4waves", 1, 2, 0, 64, 3),
                                           ("fp64_real_sym_64x64_8waves_pad4", 1, 0, 4, 64, 1),
                                           ("fp64_c3m_sym_64x64_8waves_pad4", 1, 2, 4, 64, 3),
                                           ("fp64_real_sym_64x64_8waves_w4", 1, 0, 5, 64, 1),
                                           ("fp64_c3m_sym_64x64_8waves_w4", 1, 2, 5, 64, 3),
                                           ("int8_slices6_sym_128x128_8waves", 4, 0, 6, 128, 21),
                                           # split squarings: 4M x the part products (3 for fp16x2, 6 bf16x3)
                                           ("fp16x2_c4m_sym_128x128_16waves_2kblocks", 2, 2, 84, 128, 12),
                                           ("fp16x2_c4m_sym_128x128_16waves_1kblock", 2, 2, 83, 128, 12),
                                           ("fp16x2_c4m_sym_128x128_8waves_1kblock", 2, 2, 80, 128, 12),
                                           ("bf16x3_c4m_sym_128x128_8waves", 3, 2, 80, 128, 24)):
        ms = ctypes.c_double(0.0)
        rc = lib.fexpm_probe_gemm_time(w, mode, 1, cfg, n, n, kp, 3, ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(lib.fexpm_probe_last_error().decode())
        T = (n + tile - 1) // tile
        issued = mult * 2.0 * T * (T + 1) / 2 * tile * tile * kp
        gem[name] = {"ms": ms.value, "issued_tflops": issued / (ms.value * 1e-3) / 1e12}
    print(json.dumps({"mfma_sustained": out, "gemm_n10000": gem}), flush=True)


def _memory_report(n, st):
    """Device memory the engine holds after the run (fexpm_stats.device_bytes)
    and the largest order whose N^2-scaled footprint fits the device's memory
    (HIP's total, 3% left to the runtime)."""
    import ctypes

    total = None
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        free_b, total_b = ctypes.c_size_t(), ctypes.c_size_t()
        if hip.hipMemGetInfo(ctypes.byref(free_b), ctypes.byref(total_b)) == 0:
            total = int(total_b.value)
    except OSError:
        pass
    held = int(st.get("device_bytes", 0))
    out = {"device_bytes": held, "device_total_bytes": total}
    if total and held:
        out["largest_n_estimate"] = int(n * (0.97 * total / held) ** 0.5)
    return out


def _issued_structured_squaring(n, kpad, fp32):
    """MFMA flops one structured 3M squaring issues: lower tiles of 64 x 64
    for FP64; for FP32 the engine's own launch plan (64 x 64 tiles up to
    N = 6144, above it 128 x 128 tiles and, when the last wave of workgroups
    is less than half full, 64 x 64 sub-tiles for it)."""
    if not fp32:
        T = (n + 63) // 64
        return 3 * 2.0 * (T * (T + 1) // 2) * 64 * 64 * kpad
    import ctypes

    lib = ctypes.CDLL(os.path.join(ROOT, "fastctqw_amd", "fastexpm", "libfastexpm_hip.so"))
    lib.fexpm_test_square_issued.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    fl = ctypes.c_double(0.0)
    rc = lib.fexpm_test_square_issued(n, ctypes.byref(fl))
    if rc != 0:
        raise RuntimeError(f"fexpm_test_square_issued: {rc}")
    return fl.value


def _host_info(cores):
    import platform

    import scipy
    from threadpoolctl import threadpool_info

    model = platform.processor()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = [f"{d.get('internal_api')} {d.get('version')}" for d in threadpool_info() if d.get("user_api") == "blas"]
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "threads_used": cores, "numpy": np.__version__,
            "scipy": scipy.__version__, "blas": blas}


def _crt64_beta(moduli, k):
    """Bits per operand row of the FP64-accurate residue products with the
    first `moduli` of the thirteen moduli at product length k (padded to 64):
    min(45, floor((log2 M - 1 - log2 kpad) / 2))."""
    import math
    mods = (255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199)
    kpad = (k + 63) // 64 * 64
    return min(45, int(math.floor((math.log2(math.prod(mods[:moduli])) - 1.0 - math.log2(kpad)) * 0.5)))


def _slice_phrase(st):
    """How the FP64 stages of a complex64 run computed their symmetric products."""
    k, scope = st.get("fp64_slices", 0), st.get("fp64_slices_scope", 0)
    deg = st.get("chebyshev", 0)
    if k == 13:
        km = int(st.get("fp64_moduli", 0) or 13)
        residue = (f"exact int8 residue products ({km} moduli, rows rounded to {_crt64_beta(km, int(st['n']))} "
                   "bits, chinese-remainder reconstruction)")
    if deg:
        how = residue if k == 13 else (f"{k} exact int8 slices" if k else "fp64 mfma")
        return (f"degree-{deg} chebyshev polynomial of exp(iX) in place of pade and solve, fp64-accurate "
                f"products from {how})")
    if not k:
        return "fp64 pade and solve)"
    where = {1: "pade products", 2: "solve products", 3: "pade and solve products"}.get(scope, "products")
    rest = {1: ", fp64 cholesky solve", 2: ", fp64 pade", 3: ", fp64 cholesky factorisation"}.get(scope, "")
    if k == 13:
        how = residue
    else:
        how = f"{k} exact int8 slices (int32 accumulation)"
    return f"fp64-accurate {where} from {how}{rest})"


def _engine_options(eng, args, cdt=np.complex64):
    """The residue options of the command line on a complex64 engine."""
    if cdt != np.complex64:
        return
    from fastctqw_amd.fastexpm import _fastexpm_core
    if getattr(args, "crt_moduli", None) is not None:
        eng._backend.set_option(_fastexpm_core.OPT_CRT_MODULI, args.crt_moduli)
    if getattr(args, "crt64_moduli", None) is not None:
        eng._backend.set_option(_fastexpm_core.OPT_CRT64_MODULI, args.crt64_moduli)


def _square_phrase(prec, st=None):
    """How the squarings of a complex64 run computed their products."""
    K = int((st or {}).get("crt_moduli", 0) or 9)
    beta = int((st or {}).get("crt_beta", 0) or 27)
    return {"native": "fp32 mfma squarings",
            "int8crt": f"squarings from exact int8 residue products ({K} moduli, rows rounded to {beta} bits, "
                       "exact int32 accumulation, chinese-remainder reconstruction)",
            "fp16x2": "fp16x2 split squarings (fp32 accumulation)",
            "bf16x3": "bf16x3 split squarings (fp32 accumulation)"}.get(prec, prec + " squarings")


def _infidelity(a, b):
    num = abs(np.vdot(a, b)) ** 2
    return float(max(0.0, 1.0 - num / (np.vdot(a, a).real * np.vdot(b, b).real)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--p", type=float, default=0.5)
    ap.add_argument("--t", type=float, default=4.5)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--dtype", choices=["complex64", "complex128"], default="complex64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=2000,
                    help="order of the small scipy sample behind the cpu_baseline estimate")
    ap.add_argument("--cpu-repeats", type=int, default=1,
                    help="scipy.linalg.expm runs at the full N for cpu_baseline.value (median)")
    ap.add_argument("--no-batched", action="store_true",
                    help="c2: skip the C4 (256 uneven walks), C5 (times x states) and general-input child runs")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP64 eigendecomposition reference (torch.linalg.eigh, child process)")
    ap.add_argument("--eigh-ref", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--c5-ref", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--c3-ref", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--c4-ref", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--c4-count", type=int, default=C4_COUNT)
    ap.add_argument("--lanes", type=int, default=16,
                    help="C4: concurrent lanes (streams + host threads, at most 16); sets GPU_MAX_HW_QUEUES")
    ap.add_argument("--no-variants", action="store_true", help="c4: skip the lanes, fp16x2 and one-lane runs")
    ap.add_argument("--c4-lanes", action="store_true", help="c4: time the concurrent-lanes form instead of "
                                                             "the grouped launches")
    ap.add_argument("--config", choices=["c2", "c3", "c4", "c5", "general"], default="c2",
                    help="c2: N=10^4 single expm (default); c3: precision study at N=10^4; "
                         "c4: 256 uneven walks; c5: 32 times x 128 states, N=8192; general: dense general "
                         "complex64 expm at N (LU route) and the reference's performance inputs at 16..1024")
    ap.add_argument("--precision", choices=list(C3_PRECISIONS) + ["auto"], default="auto",
                    help="c2: squaring precision of the complex64 engine")
    ap.add_argument("--no-split", action="store_true",
                    help="c2: skip the fp16x2 split-precision comparison run")
    ap.add_argument("--no-c128", action="store_true", help="c2: skip the complex128 run")
    ap.add_argument("--no-perf-inputs", action="store_true",
                    help="general: skip the reference's performance inputs (16..1024) beside scipy")
    ap.add_argument("--crt-moduli", type=int, default=None, choices=[7, 8, 9],
                    help="residue squarings: number of moduli (engine default when omitted)")
    ap.add_argument("--crt64-moduli", type=int, default=None, choices=[0, 8, 9, 10, 11, 12, 13],
                    help="FP64-accurate residue products: number of moduli (0 = automatic; engine default "
                         "when omitted)")
    ap.add_argument("--c3-cases", default=None, help="c3: comma-separated cases (default: all of C3_CASES)")
    ap.add_argument("--c3-graphs", choices=["both", "unweighted", "weighted"], default="both")
    ap.add_argument("--fp64-slices", type=int, default=None,
                    help="c2: FP64 Pade products from this many exact int8 slices (4..6) on the int8 MFMA; "
                         "0 = FP64 MFMA (default: the engine's, 6)")
    ap.add_argument("--fp64-slices-scope", type=int, default=0, choices=[0, 1, 2, 3],
                    help=argparse.SUPPRESS)
    ap.add_argument("--kblock", type=int, default=None,
                    help="c2: accumulation block of the squarings in 16-k tiles (0 = one block)")
    ap.add_argument("--no-api-timing", action="store_true",
                    help="c2: skip the host-inclusive run() and Ctqw.runWalk timings")
    ap.add_argument("--mfma-rate", action="store_true",
                    help="only measure the sustained MFMA rates (in-register loop) and the GEMM kernels "
                         "at the N=10^4 shapes, then exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        _copy_calibration(args.copy_calibration)
        return
    if args.mfma_rate:
        _mfma_rates()
        return
    if args.c3_ref:
        _c3_reference(args.n, args.c3_ref)
        return
    if args.config == "c3":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}), flush=True)
            sys.exit(2)
        run_c3(args)
        return

    if args.c4_ref:
        _c4_reference(args.c4_ref, args.c4_count)
        return
    if args.config == "c4":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}), flush=True)
            sys.exit(2)
        run_c4(args)
        return

    if args.c5_ref:
        _c5_reference(args.n, 0.3, 5, 0.5, 128, 32, args.c5_ref)
        return
    if args.config == "c5":
        if args.n == 10000:
            args.n = 8192
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}), flush=True)
            sys.exit(2)
        run_c5(args)
        return

    if args.config == "general":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}), flush=True)
            sys.exit(2)
        run_general(args)
        return

    if args.eigh_ref:
        _eigh_reference(args.n, args.p, args.seed, args.t, args.eigh_ref, args.dtype)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}), flush=True)
        sys.exit(2)

    # FP64 reference first, in a separate process, before this process
    # touches the GPU (torch and the engine never share a process)
    ref, ref_error = None, None
    if not args.no_accuracy:
        import subprocess
        import tempfile

        ref_path = os.path.join(tempfile.gettempdir(), f"fexpm_ref_{os.getpid()}.npz")
        print("[bench] FP64 eigh reference (child process) ...", file=sys.stderr, flush=True)
        try:
            cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--eigh-ref", ref_path,
                                 "--n", str(args.n), "--p", str(args.p), "--seed", str(args.seed),
                                 "--t", str(args.t), "--dtype", args.dtype], timeout=900)
            if cp.returncode == 0 and os.path.exists(ref_path):
                ref = dict(np.load(ref_path))
                os.remove(ref_path)
            else:
                ref_error = f"eigh reference failed (exit code {cp.returncode})"
        except subprocess.TimeoutExpired:
            ref_error = "eigh reference timed out"
        print(f"[bench] reference done: {ref_error or 'ok'}", file=sys.stderr, flush=True)

    from fastctqw_amd.fastexpm import ExpMatComplex64, ExpMatComplex128
    from fastctqw_amd.graphs import c2_laplacian

    t_gen = time.perf_counter()
    L = c2_laplacian(args.n, args.p, args.seed, dtype=np.float64)
    cdt = np.complex64 if args.dtype == "complex64" else np.complex128
    a = (-1j * args.t * L).astype(cdt)
    del L
    print(f"[bench] input ready in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)

    eng = (ExpMatComplex64 if cdt == np.complex64 else ExpMatComplex128)(args.n)
    if cdt == np.complex64:
        eng.set_precision(args.precision)
    if args.kblock is not None:
        from fastctqw_amd.fastexpm import _fastexpm_core
        eng._backend.set_option(_fastexpm_core.OPT_KBLOCK, args.kblock)
    _engine_options(eng, args, cdt)
    if args.fp64_slices is not None and cdt == np.complex64:
        from fastctqw_amd.fastexpm import _fastexpm_core
        eng._backend.set_option(_fastexpm_core.OPT_FP64_SLICES, args.fp64_slices)
        eng._backend.set_option(_fastexpm_core.OPT_FP64_SLICES_SCOPE, args.fp64_slices_scope)
    eng.upload(a)
    for _ in range(args.warmup):
        eng.run_resident()
        print(f"[bench] warmup {eng.stats()['ms_total']:.1f} ms", file=sys.stderr, flush=True)
    phase = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.run_resident()
        phase.append(eng.stats())
    wall = time.perf_counter() - t0
    st = phase[-1]
    ms_step = 1000.0 * wall / args.steps
    n, s = args.n, st["s"]

    # dominant kernel: the squaring (s launches per step); the phase time
    # comes from HIP events on the stream the squarings run on
    sq_ms = float(np.mean([p["ms_square"] for p in phase])) / max(s, 1)
    T = (n + 127) // 128
    np_pad = T * 128
    kpad = (n + 15) // 16 * 16
    used = st.get("precision", "native")  # the squaring form the run used (the default resolves by n)
    kernel_ms = float(np.mean([p.get("ms_square_kernel", 0.0) for p in phase]))
    kernel_launches = st.get("square_kernel_launches", 0)
    split = cdt == np.complex64 and used in ("fp16x2", "bf16x3") and st["structured"]
    crt = cdt == np.complex64 and used == "int8crt" and st["structured"]
    if crt:
        K = int(st.get("crt_moduli", 9) or 9)
        roof = crt_roofline(n, kernel_ms, kernel_launches, float(np.mean([p["ms_square"] for p in phase])), s, K)
        roof["moduli"] = K
        roof["beta"] = int(st.get("crt_beta", 0))
        kname = (f"crt_gemm_kernel ({3 * K} batched int8 products per squaring modulo {K} moduli on "
                 "v_mfma_i32_32x32x32_i8: lower 256x256 tiles, 4 waves of 128x128, direct-to-LDS operand "
                 "copies in a 4-stage ring interleaved with the MFMAs; byte residues out)")
        traffic_key = "int8crt_squaring_gemm"
        peak = INT8_PEAK_TOPS
        issued_tf = roof["issued_tops"]
    elif split:
        npair = 3 if used == "fp16x2" else 6
        peak = FP16_PEAK_TFLOPS
        issued = 4 * npair * 2.0 * (T * (T + 1) // 2) * 128 * 128 * kpad
        kname = f"sq_split_kernel<{used}> (4M complex, 16-bit MFMA, FP32 accumulation)"
        traffic_key = f"{used}_squaring"
    else:
        peak = FP32_PEAK_TFLOPS if cdt == np.complex64 else FP64_PEAK_TFLOPS
        issued = (_issued_structured_squaring(n, kpad, cdt == np.complex64) if st["structured"]
                  else 3 * 2.0 * np_pad ** 3)
        kname = ("gemm_nt_kernel (3M symmetric squaring on v_mfma_f32_32x32x2, four waves per SIMD: up to "
                 "N=6144 lower 64x64 tiles of 4 waves, above it 128x128 tiles of 16 waves plus a 64x64 launch "
                 "for a last partial wave of tiles; 8x8 tile groups; each wave 32x32)")
        traffic_key = "fp32_squaring" if cdt == np.complex64 else None
    if not crt:
        # the squaring GEMM launches' own time (HIP events), else the phase's
        sq_time = kernel_ms if kernel_launches == s and kernel_ms > 0 else float(np.mean([p["ms_square"]
                                                                                          for p in phase]))
        roof = squaring_roofline(n, sq_time, s, st.get("s_two_product", 0), issued, peak,
                                 split_parts=(3 if used == "fp16x2" else 6) if split else 0)
        issued_tf = roof["issued_tflops"]
    # the profiled counters are those of the default residue squaring (8 moduli)
    traffic = _profiled_traffic(traffic_key) if n == 10000 and (not crt or K == 8) else None
    sustained = (_sustained_peak("int8" if crt else "fp16" if split and used == "fp16x2" else
                                 "bf16" if split else "fp32") if cdt == np.complex64 else None)

    # accuracy: invariants and, unless disabled, infidelity against an FP64
    # eigendecomposition of the same H computed in a child process
    from fastctqw_amd.graphs import random_state

    u = eng.download()
    ones = np.ones(n, dtype=np.complex128)
    u1 = u.astype(np.complex128) @ ones
    lap_defect = float(np.linalg.norm(u1 - ones) / np.linalg.norm(ones))
    cols = u[:, :64].astype(np.complex128)
    unit_defect = float(np.linalg.norm(cols.conj().T @ cols - np.eye(64)) / 8.0)
    psi_e0 = u[:, 0].astype(np.complex128)
    psi_rand = eng.apply(random_state(n, seed=55))  # FP64 accumulation on the device
    del u
    acc = {"laplacian_U1_defect": lap_defect, "unitarity_defect_64cols": unit_defect}
    if ref is not None:
        acc.update({
            "oracle": "FP64 eigendecomposition of G = Im(A) for the same rounded input A = iG "
                      "(torch.linalg.eigh on the GPU), psi = V e^{i lam} V^T psi0",
            "eig_residual": float(ref["eig_residual"]), "eig_seconds": float(ref["eig_seconds"]),
            "infidelity_e0": _infidelity(ref["psi_e0"], psi_e0),
            "infidelity_random_state": _infidelity(ref["psi_rand"], psi_rand),
            "rel_err_e0": float(np.linalg.norm(psi_e0 - ref["psi_e0"]) / np.linalg.norm(ref["psi_e0"])),
            "raw_infidelity_e0": float(1.0 - abs(np.vdot(ref["psi_e0"], psi_e0)) ** 2),
            # |‖psi‖ - 1|: psi_e0 is column 0 of the complex64 U as stored; the random
            # state comes from the device apply, refined to the norm of psi0 in FP64
            "norm_deviation_e0": float(abs(np.linalg.norm(psi_e0) - 1.0)),
            "norm_deviation_random_state": float(abs(np.linalg.norm(psi_rand) - 1.0)),
        })
    elif ref_error:
        acc["oracle_error"] = ref_error

    out = {
        "metric": "s per N=10k complex expm",
        "value": ms_step / 1000.0,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (f"c64 ({_square_phrase(used, st)}, {_slice_phrase(st)}" if cdt == np.complex64 else "c128 (fp64)"),
        "data": "synthetic: seeded ER(N,p) Laplacian, A=-i*t*L",
        "config": {"workload": f"expm(-i*{args.t}*L), L Laplacian of ER({n},{args.p}) seed {args.seed}, "
                               f"{args.dtype}, full U resident", "n": n, "t": args.t,
                   "m": st["m"], "s": s, "structured": st["structured"], "chebyshev_degree": st.get("chebyshev", 0),
                   "pade_int8_slices": st.get("fp64_slices", 0), "int8_slice_scope": st.get("fp64_slices_scope", 0)},
        "phases_ms": {k: st[k] for k in ("ms_analyze", "ms_pade", "ms_solve", "ms_square", "ms_output",
                                          "ms_total")},
        "roofline": {"kernel": kname, "bound": "mfma", **roof,
                     "traffic": traffic, "traffic_source": TRAFFIC_PROFILE if traffic else None,
                     "avg_launch_ms": (kernel_ms / kernel_launches if kernel_launches else sq_ms),
                     "squaring_ms": sq_ms, "launches_per_step": s,
                     "sustained_peak": sustained, "sustained_source": MFMA_PROFILE if sustained else None,
                     "issued_frac_of_sustained": issued_tf / sustained if sustained else None,
                     "note": ("achieved = algorithmic int8 operations of one launch of the residue products "
                              "(3 K N^3, K moduli) / its device time (HIP events around each launch on the squaring "
                              "stream); issued = the whole lower tiles and padded k; fp32_equivalent = the "
                              "squaring phase as the rate of the FP32 structured algorithm (3 N^3 per "
                              "squaring); traffic = HBM bytes per launch from FETCH_SIZE / WRITE_SIZE scaled by "
                              "a copy of known size" if crt else
                              "achieved = the structured algorithm's flops of the s squarings (3 N^3 per 3M "
                              "squaring, 2 N^3 per two-product squaring: real symmetric half products) / "
                              "their device time from HIP events around the launches on the squaring stream; "
                              "issued = MFMA flops the kernels execute (whole lower tiles; equal to the "
                              "profiler's SQ_INSTS_VALU_MFMA_MOPS x 512); achieved_4m_convention = 8 N^3 per "
                              "squaring, the general complex convention; traffic = HBM bytes per 3M launch "
                              "from FETCH_SIZE / WRITE_SIZE scaled by a copy of known size")},
        "accuracy": acc,
        "memory": _memory_report(n, st),
    }
    if cdt == np.complex64 and args.precision == "auto" and not args.no_split:
        # the same expm with other squaring forms, same input and reference:
        # FP32 MFMA (the plain form the default is measured against) and
        # fp16x2 split precision
        def variant(prec):
            eng.set_precision(prec)
            for _ in range(args.warmup):
                eng.run_resident()
            ms, vst = [], []
            for _ in range(args.steps):
                eng.run_resident()
                vst.append(eng.stats())
                ms.append(vst[-1]["ms_total"])
            v = {"precision": vst[-1]["precision"], "ms_per_step": float(np.median(ms)),
                 "ms_square": float(np.median([q["ms_square"] for q in vst])),
                 "ms_per_squaring_launch": float(np.median([q["ms_square_kernel"] for q in vst])) /
                 max(vst[-1]["square_kernel_launches"], 1),
                 "default_speedup": float(np.median(ms)) / ms_step}
            if ref is not None:
                psi = eng.apply(np.vstack([np.eye(1, n, dtype=np.complex128),
                                           random_state(n, seed=55)[None, :]]))
                v["infidelity_e0"] = _infidelity(ref["psi_e0"], psi[0])
                v["infidelity_random_state"] = _infidelity(ref["psi_rand"], psi[1])
            return v, vst[-1]

        if used != "native":
            fp32, fst = variant("native")
            kl = max(fst["square_kernel_launches"], 1)
            fp32["roofline"] = squaring_roofline(n, fst["ms_square_kernel"], fst["s"], 0,
                                                 _issued_structured_squaring(n, kpad, True), FP32_PEAK_TFLOPS)
            fp32["roofline"]["avg_launch_ms"] = fst["ms_square_kernel"] / kl
            out["fp32_squarings"] = fp32
        if used == "int8crt":
            # the residue squarings with the other moduli counts (rows rounded
            # to fewer bits, fewer int8 products), same input and reference
            from fastctqw_amd.fastexpm import _fastexpm_core
            K0 = int(st.get("crt_moduli", 9))
            sweep = {}
            for K in (7, 8, 9):
                if K == K0:
                    continue
                eng._backend.set_option(_fastexpm_core.OPT_CRT_MODULI, K)
                v, vs = variant("auto")
                v["beta"] = vs.get("crt_beta", 0)
                sweep[str(K)] = v
            eng._backend.set_option(_fastexpm_core.OPT_CRT_MODULI, K0)
            out["crt_moduli_sweep"] = sweep
        sp, _ = variant("fp16x2")
        sp["speedup_vs_fp32"] = (out["fp32_squarings"]["ms_per_step"] if "fp32_squarings" in out
                                 else ms_step) / sp["ms_per_step"]
        out["split_precision"] = sp
        eng.set_precision("auto")
    if not args.no_api_timing:
        out["api_timing"] = _api_timing(eng, a, args)
    if cdt == np.complex64 and not args.no_c128:
        out["complex128"] = _c128_run(a, args, ref)
    eng.free()
    del a
    if not args.no_batched:
        # the batched workloads in child processes (each its own GPU context)
        for cfg, extra in (("c4", ["--no-variants", "--no-cpu-baseline"]), ("c5", ["--no-split"]),
                           ("general", [])):
            print(f"[bench] {cfg} (child process) ...", file=sys.stderr, flush=True)
            out[cfg] = _child_summary(cfg, extra, args)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, args.t, args.p, args.seed, repeats=args.cpu_repeats,
                                           n_est=args.cpu_sample_n)
    print(json.dumps(out), flush=True)


def _child_summary(cfg, extra, args):
    """Runs ``bench.py --config cfg`` in a child process and keeps the
    figures of its JSON line that the default line reports."""
    import subprocess

    cmd = [sys.executable, os.path.abspath(__file__), "--config", cfg, "--steps", str(max(3, args.steps)),
           "--warmup", "1", *extra]
    try:
        cp = subprocess.run(cmd, capture_output=True, text=True, timeout=420)
    except subprocess.TimeoutExpired:
        return {"error": "timed out"}
    sys.stderr.write(cp.stderr[-2000:])
    try:
        r = json.loads(cp.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": f"exit code {cp.returncode}"}
    keep = {"metric": r.get("metric"), "value": r.get("value"), "unit": r.get("unit"),
            "ms_per_step": r.get("ms_per_step"), "steps": r.get("steps"), "config": r.get("config"),
            "accuracy": r.get("accuracy")}
    if cfg == "c4":
        keep["batch_report"] = r.get("batch_report")
        keep["squaring_class"] = r.get("squaring_class")
    if cfg == "general":
        keep["phases_ms"] = r.get("phases_ms")
        keep["reference_perf_inputs"] = r.get("reference_perf_inputs")
    return keep


if __name__ == "__main__":
    main()
