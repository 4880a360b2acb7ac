"""CPU oracle for the fastexpm hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the thing measured or shipped.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``fastctqw_amd``) never
imports it and fails loudly when the HIP library is missing.

Contents
--------
padeexpm   FP64 numpy restatement of the scaling-and-squaring Pade expm as the
           HIP engine runs it (trace shift, exact 1-norm, degree/scaling
           choice of Higham 2005, U/V assembly, LU solve, squarings), stage
           by stage.  The reference's own implementation of
           ``MatrixExpCalculator<T>::run`` (include/MatrixExpCalculator.h:40,68,96)
           is absent from the reference checkout (CMakeLists.txt:24 links a
           missing lib/libexpmcore.a); the stage structure follows the buffer
           set it declares (include/MatrixExpCalculator.h:52-61).
eig        FP64 eigendecomposition oracle U = V diag(exp(-i t lam)) V^H for
           real-symmetric H, and the reference test's Taylor oracle
           (tests/test_fastexpm_accuracy.py:10-56).
closed     Closed forms: complete graph, cycle, hypercube, diagonal, zero.
metrics    Infidelity and relative Frobenius error.

How each function is pinned (beyond restating formulas) is listed in the
module docstrings and exercised by tests/test_oracle.py: scipy.linalg.expm,
mpmath at tiny N, closed forms, group-law / unitarity invariants and
committed golden vectors.
"""
