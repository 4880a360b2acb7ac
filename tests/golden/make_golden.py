"""Write tests/golden/expm_golden.npz: small matrices and exp(A) computed by
mpmath at 40 significant digits (independent of every Pade/Taylor code path
in this repository).  Rerun with ``python tests/golden/make_golden.py``.

Inputs (seeded): the reference test families at tiny N
(tests/test_fastexpm_accuracy.py:83-159: 1j * complete / ER / 4-regular
adjacency), a CTQW Laplacian walk, and a general complex matrix like the
reference perf inputs (tests/test_fastexpm_perf.py:14-18: Adj + 1j*rand).
"""
import os
import sys

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from fastctqw_amd.graphs import erdos_renyi, laplacian, random_regular  # noqa: E402


def mp_expm(a):
    mpmath.mp.dps = 40
    return np.array(mpmath.expm(mpmath.matrix(a.tolist())).tolist(), dtype=np.complex128)


def main():
    rng = np.random.default_rng(2024)
    cases = {
        "complete8": 1j * (np.ones((8, 8)) - np.eye(8)),
        "er10": 1j * erdos_renyi(10, 0.3, seed=10),
        "r4_10": 1j * random_regular(10, 4, seed=10),
        "lap12_t4p5": -1j * 4.5 * laplacian(erdos_renyi(12, 0.5, seed=2, weighted=True)),
        "general6": (np.ones((6, 6)) - np.eye(6)) + 1j * rng.random((6, 6)),
        "scalar": np.array([[0.25 - 3.0j]]),
    }
    out = {}
    for k, a in cases.items():
        a = np.asarray(a, dtype=np.complex128)
        out[k + "_A"] = a
        out[k + "_U"] = mp_expm(a)
    np.savez(os.path.join(HERE, "expm_golden.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
