"""The grouped batch program, executed on the CPU.

The engine's builder of grouped batch programs (every stage of the structured
pipeline as one launch over all walks) is called through the testing ABI on a
virtual address layout, no GPU involved.  This interpreter executes the listed
stages in FP64 numpy with the kernels' semantics (row-major NT products with
the epilogue, lower-triangle products mirrored, in-place column panels, the
diagonal-split squaring, the 64x64 Cholesky diagonal block and its inverse)
and the results are compared with the Pade oracle and the eigen oracle.  It
checks the program's addressing, aliasing, stage order and per-walk plans:
a wrong offset, a stage missing a walk, a plane reused too early or a wrong
coefficient shows up as a large error.  The kernels themselves are checked
against the same oracles on the GPU (tests/test_gpu_expm.py).
"""
import ctypes
import os

import numpy as np
import pytest

from fastctqw_amd.graphs import erdos_renyi, laplacian, random_regular
from oracle.metrics import rel_fro
from oracle.padeexpm import pade_expm

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fastctqw_amd", "fastexpm",
                   "libfastexpm_hip.so")
COMPLEX64, COMPLEX128 = 2, 3
PREC = {"native": 0, "fp16x2": 1, "bf16x3": 2}


def _program(dtype, precision, n, m, s, mu, with_psi, slices=-1, norm=None):
    """Listing of the batch program; with `norm` (1-norms of the shifted
    inputs) the Chebyshev route, else the Pade route."""
    lib = ctypes.CDLL(LIB)
    cnt = len(n)
    n_ = np.asarray(n, dtype=np.int64)
    m_ = np.asarray(m, dtype=np.int32)
    s_ = np.asarray(s, dtype=np.int32)
    sh = np.asarray([1 if x != 0 else 0 for x in mu], dtype=np.int32)
    mr = np.asarray([complex(x).real for x in mu], dtype=np.float64)
    mi = np.asarray([complex(x).imag for x in mu], dtype=np.float64)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    nrm = None if norm is None else np.asarray(norm, dtype=np.float64)
    ln = ctypes.c_int64(0)
    args = (ctypes.c_int(dtype), ctypes.c_int(precision), ctypes.c_int(slices), ctypes.c_int64(cnt), ptr(n_), ptr(m_),
            ptr(s_), ptr(sh),
            ptr(mr), ptr(mi), None if nrm is None else ptr(nrm), ctypes.c_int(1 if with_psi else 0))
    lib.fexpm_test_group_program(*args, None, ctypes.c_int64(0), ctypes.byref(ln))
    buf = ctypes.create_string_buffer(ln.value + 1)
    rc = lib.fexpm_test_group_program(*args, buf, ctypes.c_int64(ln.value + 1), ctypes.byref(ln))
    assert rc == 0, rc
    return buf.value.decode()


class Interp:
    """FP64 execution of a program listing over numpy "regions"."""

    def __init__(self, text, mats, psi0):
        lines = text.splitlines()
        hdr = dict(kv.split("=") for kv in lines[0].split()[1:])
        self.wsize, self.insize = int(hdr["wsize"]), int(hdr["insize"])
        self.reg = {
            1: (np.zeros(int(hdr["f64"]) // 8), 8),
            2: (np.zeros(int(hdr["w"]) // self.wsize), self.wsize),
            3: (np.zeros(int(hdr["small"]) // 8), 8),
            5: (np.concatenate([a.ravel() for a in mats]).astype(np.complex128), self.insize),
            6: (np.full(sum(a.size for a in mats), np.nan, dtype=np.complex128), self.insize),
            7: (np.concatenate(psi0).astype(np.complex128) if psi0 else np.zeros(1, complex), 16),
            8: (np.full(sum(a.shape[0] for a in mats), np.nan, dtype=np.complex128), 16),
        }
        self.lines = lines[1:]
        self.kinds = []
        self.oz = {}  # int8 slice buffer -> FP64 values of the operand it holds

    def _elem(self, tok):
        r, off = (int(x) for x in tok.split(":"))
        arr, es = self.reg[r]
        assert off % es == 0, tok
        return arr, off // es

    def get(self, tok, ld, rows, cols):
        arr, e = self._elem(tok)
        idx = e + np.arange(rows)[:, None] * ld + np.arange(cols)[None, :]
        assert idx.size == 0 or idx.max() < arr.size, ("read beyond region", tok, ld, rows, cols)
        return arr[idx].copy()

    def put(self, tok, ld, val):
        arr, e = self._elem(tok)
        rows, cols = val.shape
        idx = e + np.arange(rows)[:, None] * ld + np.arange(cols)[None, :]
        assert idx.max() < arr.size, ("write beyond region", tok)
        arr[idx] = val

    def pairs(self, tok, n):
        arr, e = self._elem(tok)
        return arr[e:e + 2 * n:2] + 1j * arr[e + 1:e + 2 * n:2]

    def set_pairs(self, tok, vals):
        arr, e = self._elem(tok)
        arr[e:e + 2 * len(vals):2] = vals.real
        arr[e + 1:e + 2 * len(vals):2] = vals.imag

    @staticmethod
    def _sym(full):
        low = np.tril(full)
        return low + np.tril(full, -1).T

    def _split_square(self, Yc, d):
        """Diagonal-split squaring: X = diag(d) + Y, X^2 = diag(d_out) + Y'."""
        P = Yc @ Yc.T
        out = P + (d[:, None] + d[None, :]) * Yc
        d_out = d * d + np.diag(P)
        np.fill_diagonal(out, 0.0)
        return out, d_out

    def run(self):
        for line in self.lines:
            f = line.split()
            if f[0] == "STAGE":
                self.kinds.append(f[1])
                continue
            kv = dict(x.split("=", 1) for x in f[1:])
            getattr(self, "op_" + f[0].lower())(kv)

    def op_prepare(self, kv):
        n, ld = int(kv["n"]), int(kv["ld"])
        arr, e = self._elem(kv["a"])
        A = arr[e:e + n * n].reshape(n, n)
        g = (A.imag - float(kv["mu_im"]) * np.eye(n)) * float(kv["scale"])
        self.put(kv["g"], ld, g)

    def op_lincomb(self, kv):
        n, ld = int(kv["n"]), int(kv["ld"])
        acc = float(kv["diag"]) * np.eye(n)
        for t in range(int(kv["nin"])):
            acc = acc + float(kv[f"c{t}"]) * self.get(kv[f"in{t}"], ld, n, n)
        self.put(kv["out"], ld, acc)

    def op_pack(self, kv):
        n, np_, ld = int(kv["n"]), int(kv["np"]), int(kv["ld"])
        zb = np.zeros((2 * np_, np_))
        zb[:n, :n] = self.get(kv["v"], ld, n, n)
        zb[np_:np_ + n, :n] = self.get(kv["u"], ld, n, n)
        self.put(kv["zb"], ld, zb)

    def op_gemm(self, kv):
        m, n, k = int(kv["m"]), int(kv["n"]), int(kv["k"])
        lda, ldb, ldc = int(kv["lda"]), int(kv["ldb"]), int(kv["ldc"])
        mode, sym = int(kv["mode"]), int(kv["sym"]) == 1
        Ar = self.get(kv["ar"], lda, m, k)
        Br = self.get(kv["br"], ldb, n, k)
        if mode == 0:
            Pr, Pi = Ar @ Br.T, None
        else:
            Ai = self.get(kv["ai"], lda, m, k)
            Bi = self.get(kv["bi"], ldb, n, k)
            if kv["din"] != "0:0":
                assert m == n and kv["ar"] == kv["br"] and kv["ai"] == kv["bi"]
                d = self.pairs(kv["din"], n)
                Y = (Ar + 1j * Ai)[:, :n]
                out, d_out = self._split_square(Y, d)
                self.set_pairs(kv["dout"], d_out)
                Pr, Pi = out.real, out.imag
            else:
                Pr = Ar @ Br.T - Ai @ Bi.T
                Pi = Ar @ Bi.T + Ai @ Br.T
        alpha, beta = float(kv["alpha"]), float(kv["beta"])
        Cr = alpha * Pr
        if beta != 0.0:
            Cr = Cr + beta * self.get(kv["cr"], ldc, m, n)
        for t in range(int(kv["nadd"])):
            Cr = Cr + float(kv[f"c{t}"]) * self.get(kv[f"add{t}"], int(kv["ld_add"]), m, n)
        Cr = Cr + float(kv["diag"]) * np.eye(m, n)
        if sym:
            Cr = self._sym(Cr)
        self.put(kv["cr"], ldc, Cr)
        if Pi is not None:
            Ci = alpha * Pi
            if beta != 0.0:
                Ci = Ci + beta * self.get(kv["ci"], ldc, m, n)
            if sym:
                Ci = self._sym(Ci)
            self.put(kv["ci"], ldc, Ci)

    @staticmethod
    def _shift(tok, elems, esize):
        r, off = (int(x) for x in tok.split(":"))
        return f"{r}:{off + elems * esize}"

    def op_chol(self, kv):
        ld, k0, n = int(kv["ld"]), int(kv["k0"]), int(kv["n"])
        blk_tok = self._shift(kv["s"], k0 * ld + k0, 8)  # diagonal block (k0, k0)
        blk = self.get(blk_tok, ld, 64, 64)
        valid = np.zeros((64, 64), bool)
        nv = max(0, min(64, n - k0))
        valid[:nv, :nv] = True
        L = np.linalg.cholesky(self._sym(np.where(valid, blk, np.eye(64))))  # padding factored as identity
        keep = valid & (np.tril(np.ones((64, 64))) > 0)  # the factor is written inside n, lower part
        self.put(blk_tok, ld, np.where(keep, L, blk))
        self.put(kv["inv"], 64, np.tril(np.linalg.inv(L)))

    def op_solution(self, kv):
        n, ldx, ldw = int(kv["n"]), int(kv["ldx"]), int(kv["ldw"])
        X = self.get(kv["xr"], ldx, n, n) + 1j * self.get(kv["xi"], ldx, n, n)
        self.set_pairs(kv["d"], np.diag(X).copy())
        np.fill_diagonal(X, 0.0)
        self.put(kv["wr"], ldw, X.real)
        self.put(kv["wi"], ldw, X.imag)

    def op_ozslice(self, kv):
        """Record the operand a slice buffer holds (its FP64 values now);
        the slicing itself is checked on the GPU against its error bound."""
        nr, nk, rows, kpad = int(kv["nr"]), int(kv["nk"]), int(kv["rows"]), int(kv["kpad"])
        k1, ld = int(kv["k1"]), int(kv["ld"])
        x = np.zeros((rows, kpad))
        x[:nr, :min(k1, nk)] = float(kv["s1"]) * self.get(kv["x1"], ld, nr, min(k1, nk))
        if nk > k1:
            x[:nr, k1:nk] = float(kv["s2"]) * self.get(kv["x2"], ld, nr, nk - k1)
        self.oz[kv["out"]] = x

    def op_ozgemm(self, kv):
        m, n, kp, ldc = int(kv["m"]), int(kv["n"]), int(kv["kpad"]), int(kv["ldc"])
        A, B = self.oz[kv["a"]][:m, :kp], self.oz[kv["b"]][:n, :kp]
        assert self.oz[kv["a"]].shape[0] == int(kv["rows_a"]) and self.oz[kv["b"]].shape[0] == int(kv["rows_b"])
        C = float(kv["alpha"]) * (A @ B.T)
        if float(kv["beta"]) != 0.0:
            C = C + float(kv["beta"]) * self.get(kv["cr"], ldc, m, n)
        for t in range(int(kv["nadd"])):
            C = C + float(kv[f"c{t}"]) * self.get(kv[f"add{t}"], int(kv["ld_add"]), m, n)
        C = C + float(kv["diag"]) * np.eye(m, n)
        if int(kv["sym"]) == 1:
            C = self._sym(C)
        self.put(kv["cr"], ldc, C)

    def op_split(self, kv):
        pass  # the 16-bit parts are a device representation of Y; SQ reads Y itself

    def op_sq(self, kv):
        m, ldy, ldc = int(kv["m"]), int(kv["ldy"]), int(kv["ldc"])
        Y = self.get(kv["yr"], ldy, m, m) + 1j * self.get(kv["yi"], ldy, m, m)
        out, d_out = self._split_square(Y, self.pairs(kv["din"], m))
        self.set_pairs(kv["dout"], d_out)
        self.put(kv["cr"], ldc, out.real)
        self.put(kv["ci"], ldc, out.imag)

    def _u(self, kv):
        n, ld = int(kv["n"]), int(kv["ld"])
        Y = self.get(kv["yr"], ld, n, n) + 1j * self.get(kv["yi"], ld, n, n)
        return complex(float(kv["phr"]), float(kv["phi"])) * (Y + np.diag(self.pairs(kv["d"], n)))

    def op_output(self, kv):
        n = int(kv["n"])
        arr, e = self._elem(kv["out"])
        arr[e:e + n * n] = self._u(kv).ravel()

    def op_apply(self, kv):
        n = int(kv["n"])
        p0, e0 = self._elem(kv["psi0"])
        arr, e = self._elem(kv["out"])
        arr[e:e + n] = self._u(kv) @ p0[e0:e0 + n]

    def op_renorm(self, kv):
        n = int(kv["n"])
        p0, e0 = self._elem(kv["psi0"])
        arr, e = self._elem(kv["out"])
        a, b = np.linalg.norm(p0[e0:e0 + n]), np.linalg.norm(arr[e:e + n])
        if a > 0 and b > 0:
            arr[e:e + n] *= a / b

    def result(self, i_off, n):
        arr = self.reg[6][0]
        return arr[i_off:i_off + n * n].reshape(n, n)


def _walks(seed):
    rng = np.random.default_rng(seed)
    mats = []
    for n, t in [(1, 2.0), (5, 1e-4), (63, 0.003), (64, 0.05), (65, 0.056), (70, 1.3), (130, 4.5), (200, 0.9)]:
        adj = erdos_renyi(n, 0.4, seed=int(rng.integers(1 << 30)), weighted=True) if n > 8 else \
            random_regular(n, 2, seed=3, weighted=True) if n > 2 else np.zeros((n, n))
        h = laplacian(adj) if n % 2 else adj
        mats.append(-1j * t * h)
    return mats


@pytest.mark.parametrize("cheb", [False, True])
@pytest.mark.parametrize("dt,precision,slices", [(COMPLEX128, "native", -1), (COMPLEX64, "native", -1),
                                                 (COMPLEX64, "native", 0), (COMPLEX64, "fp16x2", -1),
                                                 (COMPLEX64, "bf16x3", 0)])
def test_group_program_matches_oracle(dt, precision, slices, cheb):
    """Pade route and Chebyshev route (the batch default): every walk's U
    and psi against the Pade oracle and U against the eigen oracle."""
    mats = _walks(11)
    plans = []
    for a in mats:
        st = pade_expm(a, stages=True)
        m, s = (st["m"], st["s"]) if st["m"] else (3, 0)
        plans.append((m, s, st["mu"], st["norm"]))
    ms = sorted({p[0] for p in plans})
    assert ms == [3, 5, 7, 9, 13], ms  # every Pade degree is exercised
    psi0 = [np.exp(0.3j * np.arange(a.shape[0])) / np.sqrt(a.shape[0]) for a in mats]
    text = _program(dt, PREC[precision], [a.shape[0] for a in mats], [p[0] for p in plans],
                    [p[1] for p in plans], [p[2] for p in plans], True, slices,
                    [p[3] for p in plans] if cheb else None)
    assert ("OZGEMM" in text) == (dt == COMPLEX64 and slices != 0)
    assert ("CHOL" in text) == (not cheb)
    it = Interp(text, mats, psi0)
    it.run()
    off = 0
    voff = 0
    for i, a in enumerate(mats):
        n = a.shape[0]
        u = it.result(off, n)
        ref = pade_expm(a)
        # complex64 programs truncate the Chebyshev series at 2^-(24 + s)
        # (each eigenvalue's factor within 2^-24 after the squarings)
        tol = 2.0 ** -23 if (cheb and dt == COMPLEX64) else 1e-12
        assert rel_fro(ref, u) < tol, (i, n, plans[i], rel_fro(ref, u))
        lam, V = np.linalg.eigh(a.imag)
        assert rel_fro((V * np.exp(1j * lam)) @ V.T, u) < max(tol, 1e-11), i
        psi = it.reg[8][0][voff:voff + n]
        assert np.linalg.norm(psi - ref @ psi0[i]) < tol * max(1.0, np.linalg.norm(psi)), i
        off += n * n
        voff += n


def test_group_program_structure():
    """Launch counts of the grouped program: one launch per stage over all
    walks, one squaring round per r < max s, three launches per 64-wide
    Cholesky panel of the largest walk."""
    mats = _walks(12)
    plans = [pade_expm(a, stages=True) for a in mats]
    m = [p["m"] or 3 for p in plans]
    s = [p["s"] for p in plans]
    text = _program(COMPLEX64, 0, [a.shape[0] for a in mats], m, s, [p["mu"] for p in plans], False)
    stages = [ln for ln in text.splitlines() if ln.startswith("STAGE")]
    kinds = [int(ln.split()[1].split("=")[1]) for ln in stages]
    counts = [int(ln.split()[2].split("=")[1]) for ln in stages]
    np_max = max((a.shape[0] + 63) // 64 * 64 for a in mats)
    assert kinds.count(9) == np_max // 64                        # CHOL stages
    squaring = [c for k, c in zip(kinds, counts) if k == 10][-max(s):]
    assert squaring == [sum(1 for x in s if x > r) for r in range(max(s))]
    assert counts[kinds.index(3)] == len(mats)                   # one PREPARE problem per walk
    assert counts[kinds.index(7)] == len(mats)                   # one OUTPUT problem per walk
    assert counts[kinds.index(8)] == 0                           # no APPLY without states
    # the split form: a split stage before each squaring round
    text_h = _program(COMPLEX64, 1, [a.shape[0] for a in mats], m, s, [p["mu"] for p in plans], False)
    kinds_h = [int(ln.split()[1].split("=")[1]) for ln in text_h.splitlines() if ln.startswith("STAGE")]
    assert kinds_h.count(11) == max(s) and kinds_h.count(13) == max(s)


def test_group_program_rejects_bad_plans():
    lib = ctypes.CDLL(LIB)
    n = np.array([4], np.int64)
    bad_m = np.array([4], np.int32)
    z = np.zeros(1, np.int32)
    d = np.zeros(1)
    ln = ctypes.c_int64(0)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    assert lib.fexpm_test_group_program(2, 0, -1, ctypes.c_int64(1), p(n), p(bad_m), p(z), p(z), p(d), p(d), None, 0,
                                        None, ctypes.c_int64(0), ctypes.byref(ln)) == -1
    good_m = np.array([13], np.int32)
    # split precision is for complex64 only
    assert lib.fexpm_test_group_program(3, 1, -1, ctypes.c_int64(1), p(n), p(good_m), p(z), p(z), p(d), p(d), None, 0,
                                        None, ctypes.c_int64(0), ctypes.byref(ln)) == -1
    # int8-slice products: 0 or 6, complex64 only
    assert lib.fexpm_test_group_program(3, 0, 6, ctypes.c_int64(1), p(n), p(good_m), p(z), p(z), p(d), p(d), None, 0,
                                        None, ctypes.c_int64(0), ctypes.byref(ln)) == -1
    assert lib.fexpm_test_group_program(2, 0, 5, ctypes.c_int64(1), p(n), p(good_m), p(z), p(z), p(d), p(d), None, 0,
                                        None, ctypes.c_int64(0), ctypes.byref(ln)) == -1


def test_group_program_detects_planted_mistakes():
    """The interpreter is sensitive to the mistakes it is meant to catch: a
    listing with one squaring problem dropped, or the leading Pade
    coefficient b_1 of one walk changed by one part in 10^9, no longer
    matches the oracle."""
    mats = _walks(13)[5:]
    plans = [pade_expm(a, stages=True) for a in mats]
    args = ([a.shape[0] for a in mats], [p["m"] for p in plans], [p["s"] for p in plans], [p["mu"] for p in plans])
    text = _program(COMPLEX128, 0, *args, False)
    lines = text.splitlines()

    def worst(listing):
        it = Interp(listing, mats, None)
        it.run()
        off, err = 0, 0.0
        for a in mats:
            n = a.shape[0]
            err = max(err, rel_fro(pade_expm(a), it.result(off, n)))
            off += n * n
        return err

    assert worst(text) < 1e-12
    sq = [i for i, ln in enumerate(lines) if ln.startswith("GEMM") and "din=3:" in ln]
    dropped = "\n".join(ln for i, ln in enumerate(lines) if i != sq[-1])
    assert worst(dropped) > 1e-3
    # b_1 of the first walk's odd polynomial (the diagonal term of its Wb product)
    lc = next(i for i, ln in enumerate(lines) if ln.startswith("GEMM") and "nadd=3" in ln)
    f = lines[lc].split()
    k = next(j for j, x in enumerate(f) if x.startswith("diag="))
    f[k] = "diag=" + repr(float(f[k][5:]) * (1 + 1e-9))
    changed = "\n".join(" ".join(f) if i == lc else ln for i, ln in enumerate(lines))
    assert worst(changed) > 1e-11


def test_group_program_chebyshev_structure_and_mistakes():
    """Chebyshev route: no Cholesky or pack stages; each walk's degree and
    product count follow oracle.chebcs at theta = norm / 2^s; a dropped
    Horner product or a coefficient changed by one part in 10^9 is caught."""
    from oracle.chebcs import cheb_degrees, horner_products

    mats = _walks(14)[3:]
    plans = [pade_expm(a, stages=True) for a in mats]
    m = [p["m"] or 3 for p in plans]
    s = [p["s"] for p in plans]
    norm = [p["norm"] for p in plans]
    args = ([a.shape[0] for a in mats], m, s, [p["mu"] for p in plans])
    text = _program(COMPLEX128, 0, *args, False, -1, norm)
    kinds = [int(ln.split()[1].split("=")[1]) for ln in text.splitlines() if ln.startswith("STAGE")]
    assert 9 not in kinds and 5 not in kinds  # no CHOL, no PACK
    gemm = [ln for ln in text.splitlines() if ln.startswith("GEMM") and "mode=0" in ln]
    want = 0
    for i, a in enumerate(mats):
        dc, ds = cheb_degrees(norm[i] * 2.0 ** -s[i])
        nc, ns = dc // 2 + 1, (ds - 1) // 2 + 1
        powers = 0 if max(nc, ns) == 1 else (1 if max(nc, ns) == 2 else (2 if max(nc, ns) == 3 else 3))
        if nc > 4 or ns > 4:
            powers = 4
        want += powers + horner_products(nc) + horner_products(ns) + 1
    assert len(gemm) == want, (len(gemm), want)
    lines = text.splitlines()

    def worst(listing):
        it = Interp(listing, mats, None)
        it.run()
        off, err = 0, 0.0
        for a in mats:
            n = a.shape[0]
            err = max(err, rel_fro(pade_expm(a), it.result(off, n)))
            off += n * n
        return err

    assert worst(text) < 1e-12
    horner = [i for i, ln in enumerate(lines) if ln.startswith("GEMM") and "nadd=3" in ln]
    assert horner
    dropped = "\n".join(ln for i, ln in enumerate(lines) if i != horner[0])
    assert worst(dropped) > 1e-6
    # the constant term of a last Horner step (block 0: cw[0] or sw[0], of order 1)
    f = lines[horner[-1]].split()
    k = next(j for j, x in enumerate(f) if x.startswith("diag="))
    f[k] = "diag=" + repr(float(f[k][5:]) * (1 + 1e-9))
    changed = "\n".join(" ".join(f) if i == horner[-1] else ln for i, ln in enumerate(lines))
    assert worst(changed) > 1e-11
