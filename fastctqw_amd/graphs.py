"""Seeded, reproducible graph and state generators for CTQW workloads.

The reference densifies networkx graphs (``Ctqw.from_networkx_graph``,
FastCTQW/Ctqw.py:46-57) and builds ``H = -gamma*A`` or ``-gamma*(D - A)``
(FastCTQW/Operator.py:49-53).  Its own tests draw graphs from unseeded
networkx generators (tests/test_fastexpm_accuracy.py:110,137), so nothing
they produce is reproducible.  This module gives every workload a seed:

* Erdos-Renyi edges, edge weights and random states come from a
  counter-based hash (splitmix64 of (seed, stream, i, j)), so an N = 10^4
  graph is reproducible from three integers without shipping 800 MB, and the
  same bits can be produced by a device kernel.
* Random d-regular graphs use the Steger-Wormald pairing procedure driven by
  numpy's PCG64 stream (bit-stable across numpy releases).
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_MIX1 = np.uint64(0xBF58476D1CE4E5B9)
_MIX2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    """Vectorised splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _MIX1
        z = (z ^ (z >> np.uint64(27))) * _MIX2
        return z ^ (z >> np.uint64(31))


def _stream_key(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        k = splitmix64(np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
        return splitmix64(k ^ np.uint64(stream & 0xFFFFFFFFFFFFFFFF))


def uniform01(seed: int, stream: int, counters) -> np.ndarray:
    """U[0,1) doubles from 53 hash bits of (seed, stream, counter)."""
    key = _stream_key(seed, stream)
    with np.errstate(over="ignore"):
        h = splitmix64(np.asarray(counters, dtype=np.uint64) ^ key)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# Hash streams: one per quantity so that edges and weights are independent.
STREAM_EDGE = 1
STREAM_WEIGHT = 2
STREAM_STATE_RE = 3
STREAM_STATE_IM = 4


def _pair_counter(i, j, n):
    return i.astype(np.uint64) * np.uint64(n) + j.astype(np.uint64)


def erdos_renyi(n: int, p: float, seed: int, weighted: bool = False,
                dtype=np.float64) -> np.ndarray:
    """Dense symmetric adjacency of G(n, p); edge (i<j) iff u(seed,i,j) < p.

    Weighted graphs carry w_ij = 1 - u'(seed,i,j) in (0, 1].  Rows are
    produced in blocks so that n = 10^4 stays within a few hundred MB.
    """
    a = np.zeros((n, n), dtype=dtype)
    blk = max(1, min(n, (1 << 24) // max(n, 1)))
    cols = np.arange(n, dtype=np.uint64)
    for r0 in range(0, n, blk):
        r1 = min(n, r0 + blk)
        rows = np.arange(r0, r1, dtype=np.uint64)[:, None]
        ctr = rows * np.uint64(n) + cols[None, :]
        mask = uniform01(seed, STREAM_EDGE, ctr) < p
        mask &= cols[None, :] > rows  # strict upper triangle
        if weighted:
            w = 1.0 - uniform01(seed, STREAM_WEIGHT, ctr)
            a[r0:r1] = np.where(mask, w, 0.0)
        else:
            a[r0:r1] = mask
    # mirror the strict upper triangle into the lower one
    a += a.T
    return a


def random_regular(n: int, d: int, seed: int, weighted: bool = False,
                   dtype=np.float64, max_restarts: int = 1000) -> np.ndarray:
    """Dense adjacency of a random simple d-regular graph (Steger-Wormald).

    Points (vertex, stub) are paired one random pair at a time; a pair is
    accepted when it joins two distinct vertices not yet adjacent.  When no
    suitable pair is left the procedure restarts with the same stream.
    """
    if (n * d) % 2 != 0 or d >= n or d < 0:
        raise ValueError(f"no simple {d}-regular graph on {n} vertices")
    rng = np.random.Generator(np.random.PCG64(seed))
    for _ in range(max_restarts):
        edges = _sw_attempt(n, d, rng)
        if edges is not None:
            break
    else:  # pragma: no cover - astronomically unlikely for d <= 8
        raise RuntimeError("random_regular: too many restarts")
    a = np.zeros((n, n), dtype=dtype)
    if not edges:
        return a
    e = np.asarray(edges, dtype=np.int64)
    if weighted:
        lo = np.minimum(e[:, 0], e[:, 1]).astype(np.uint64)
        hi = np.maximum(e[:, 0], e[:, 1]).astype(np.uint64)
        w = 1.0 - uniform01(seed, STREAM_WEIGHT, lo * np.uint64(n) + hi)
    else:
        w = 1.0
    a[e[:, 0], e[:, 1]] = w
    a[e[:, 1], e[:, 0]] = w
    return a


def _sw_attempt(n, d, rng):
    points = np.repeat(np.arange(n), d)
    adj = [set() for _ in range(n)]
    edges = []
    pts = list(points)
    while pts:
        # try a bounded number of random pairs before declaring a dead end
        ok = False
        m = len(pts)
        for _ in range(64):
            i, j = rng.integers(0, m, size=2)
            if i == j:
                continue
            u, v = pts[i], pts[j]
            if u != v and v not in adj[u]:
                ok = True
                break
        if not ok:
            # exhaustive search for any suitable pair among the remaining points
            found = None
            for i in range(m):
                for j in range(i + 1, m):
                    u, v = pts[i], pts[j]
                    if u != v and v not in adj[u]:
                        found = (i, j)
                        break
                if found:
                    break
            if found is None:
                return None
            i, j = found
            u, v = pts[i], pts[j]
        adj[u].add(v)
        adj[v].add(u)
        edges.append((u, v))
        for k in sorted((i, j), reverse=True):
            pts[k] = pts[-1]
            pts.pop()
    return edges


def edge_list(adj_list, num_nodes=None):
    """(n, src, dst, w) of an undirected graph given as an adjacency list.

    Accepted forms: a networkx graph (vertices indexed in ``G.nodes`` order,
    as ``nx.to_numpy_array`` does; weights from the 'weight' attribute,
    default 1); a dict of lists ``{i: [j, ...]}`` or of dicts ``{i: {j: w}}``
    (networkx's dict-of-lists / dict-of-dicts forms, each edge taken once);
    or a sequence of pairs ``(i, j)`` / triples ``(i, j, w)`` of integer
    vertices (a pair listed twice adds its weights).  n is ``num_nodes``, or
    one more than the largest vertex index.
    """
    try:
        import networkx as nx
    except ImportError:  # pragma: no cover - networkx is a dependency of the package
        nx = None
    if isinstance(adj_list, dict) and nx is not None:
        vals = list(adj_list.values())
        if vals and isinstance(vals[0], dict):
            adj_list = nx.from_dict_of_dicts(adj_list)
        else:
            adj_list = nx.from_dict_of_lists(adj_list)
    if nx is not None and isinstance(adj_list, nx.Graph):
        if adj_list.is_directed():
            raise TypeError("graph is a directed graph, but Ctqw recive a undirected graph")
        index = {v: k for k, v in enumerate(adj_list.nodes)}
        ed = list(adj_list.edges(data="weight", default=1.0))
        src = np.fromiter((index[u] for u, _, _ in ed), dtype=np.int64, count=len(ed))
        dst = np.fromiter((index[v] for _, v, _ in ed), dtype=np.int64, count=len(ed))
        w = np.fromiter((float(x) for _, _, x in ed), dtype=np.float64, count=len(ed))
        n = adj_list.number_of_nodes() if num_nodes is None else int(num_nodes)
        return n, src, dst, w
    rows = [tuple(e) for e in adj_list]
    if any(len(r) not in (2, 3) for r in rows):
        raise ValueError("adjacency list entries must be (i, j) or (i, j, w)")
    src = np.asarray([r[0] for r in rows], dtype=np.int64)
    dst = np.asarray([r[1] for r in rows], dtype=np.int64)
    w = np.asarray([float(r[2]) if len(r) == 3 else 1.0 for r in rows], dtype=np.float64)
    n = int(max(src.max(initial=-1), dst.max(initial=-1)) + 1) if num_nodes is None else int(num_nodes)
    if src.size and (min(src.min(), dst.min()) < 0 or max(src.max(), dst.max()) >= n):
        raise ValueError("vertex index outside [0, n)")
    return n, src, dst, w


def dense_from_edges(n: int, src, dst, w=None, dtype=np.float64) -> np.ndarray:
    """Dense symmetric adjacency of an undirected edge list: a_ij = a_ji = w,
    a self-loop on the diagonal once, repeated pairs adding up (the matrix
    nx.to_numpy_array gives for the graph)."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    w = np.ones(src.shape, dtype=np.float64) if w is None else np.asarray(w, dtype=np.float64)
    a = np.zeros((n, n), dtype=np.float64)
    np.add.at(a, (src, dst), w)
    off = src != dst
    np.add.at(a, (dst[off], src[off]), w[off])
    return a.astype(dtype, copy=False)


def laplacian(a: np.ndarray) -> np.ndarray:
    """L = D - A with D = diag(row sums) (FastCTQW/Operator.py:51)."""
    return np.diag(a.sum(axis=1)) - a


def random_state(n: int, seed: int, index: int = 0) -> np.ndarray:
    """Normalised complex128 state with Re, Im ~ N(0,1) (Box-Muller on hash bits)."""
    base = np.uint64(index) * np.uint64(2 * n) if index else np.uint64(0)
    k = np.arange(2 * n, dtype=np.uint64) + base
    u1 = uniform01(seed, STREAM_STATE_RE, k)
    u2 = uniform01(seed, STREAM_STATE_IM, k)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    g = r * np.cos(2.0 * np.pi * u2)
    psi = g[:n] + 1j * g[n:]
    return psi / np.linalg.norm(psi)


def c2_laplacian(n: int = 10000, p: float = 0.5, seed: int = 2,
                 weighted: bool = False, dtype=np.float64) -> np.ndarray:
    """Laplacian of the seeded G(n, p) graph used for the N = 10^4 walk."""
    return laplacian(erdos_renyi(n, p, seed, weighted=weighted, dtype=dtype))


def reference_perf_input(family: str, n: int, seed: int = 0, dtype=np.complex64) -> np.ndarray:
    """General complex input of the reference's performance tests:
    A = Adj + 1j * rand(n, n) with Adj the adjacency of the complete graph,
    of G(n, 0.05) or of a random 4-regular graph and rand uniform in [0, 1)
    (tests/test_fastexpm_perf.py:10-59).  Not skew-Hermitian: for the
    complete graph ||e^A|| grows like e^(n-1), so the complex64 result
    overflows from n of about 90 on.  Seeded here (numpy PCG64 for rand);
    the reference draws from unseeded networkx generators.  The reference
    skips ER below n = 2 and 4-regular graphs for odd n or n < 5; this
    raises ValueError there."""
    if family == "complete":
        adj = np.ones((n, n)) - np.eye(n)
    elif family == "er":
        if n < 2:
            raise ValueError("ER input needs n >= 2")
        adj = erdos_renyi(n, 0.05, seed)
    elif family == "r4":
        if n % 2 or n < 5:
            raise ValueError("4-regular input needs even n >= 6")
        adj = random_regular(n, 4, seed)
    else:
        raise ValueError(f"unknown family {family!r}")
    rng = np.random.default_rng(seed)
    rdt = np.float32 if np.dtype(dtype) == np.complex64 else np.float64
    a = adj.astype(rdt) + 1j * rng.random((n, n)).astype(rdt)
    return a.astype(dtype)


def c4_walks(count: int = 256, seed: int = 4000, nmin: int = 64, nmax: int = 2048):
    """Seeded batch of uneven CTQW instances (the many-walks workload).

    Instance i draws from PCG64(seed + i): N ~ U{nmin..nmax}; even i: a random
    d-regular graph (d ~ U{3..8}, N bumped by one when N*d is odd), odd i:
    G(N, p) with p ~ U[0.05, 0.5]; edge weights uniform(0, 1]; H = adjacency
    or Laplacian at random; t ~ U[0.1, 20].  Returns a list of dicts with
    keys n, kind, laplacian, t, H (float64) and A = -1j*t*H (complex128).
    """
    return list(iter_c4_walks(count, seed, nmin, nmax))


def iter_c4_walks(count: int = 256, seed: int = 4000, nmin: int = 64, nmax: int = 2048):
    """Generator form of :func:`c4_walks` (one instance in memory at a time)."""
    for i in range(count):
        rng = np.random.Generator(np.random.PCG64(seed + i))
        n = int(rng.integers(nmin, nmax + 1))
        lap = bool(rng.integers(0, 2))
        t = float(rng.uniform(0.1, 20.0))
        if i % 2 == 0:
            d = int(rng.integers(3, 9))
            if (n * d) % 2:
                n += 1 if n < nmax else -1
            adj = random_regular(n, d, seed=seed + i, weighted=True)
            kind = f"regular(d={d})"
        else:
            p = float(rng.uniform(0.05, 0.5))
            adj = erdos_renyi(n, p, seed=seed + i, weighted=True)
            kind = f"er(p={p:.3f})"
        h = laplacian(adj) if lap else adj
        yield {"n": n, "kind": kind, "laplacian": lap, "t": t, "H": h, "A": -1j * t * h}
