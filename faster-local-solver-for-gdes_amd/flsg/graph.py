"""Graphs: symmetric CSR, slack CSR, synthetic generators and the reference's
harness conventions (seed draw, alpha round-trip).

CSR convention (the reference's, ``demo.py:39-44``): ``indptr`` (n+1),
``indices`` int32 with sorted rows, ``degree`` = row lengths.  Graphs built here
are undirected, without self-loops, with int32 degrees -- so ``eps*alpha*degree``
is an exact FP64 product, as in the reference's solvers.

Slack CSR (``data_process_GNN.py:92-119``): ``indptr`` is sized for the final
graph, the live adjacency of u is ``indices[indptr[u] : indptr[u]+degree[u]]``
and unused slots hold the sentinel ``n``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _build

__all__ = [
    "CSRGraph", "SlackGraph", "chung_lu", "powerlaw_graph_small", "powerlaw_graph_large",
    "citation_dynamic_graph", "demo_seeds", "uniform_seeds", "mixed_seeds",
    "alpha_roundtrip", "save_npz", "load_npz",
]


@dataclass
class CSRGraph:
    n: int
    indptr: np.ndarray      # int64 [n+1]
    indices: np.ndarray     # int32 [m]
    degree: np.ndarray      # int32 [n]
    name: str = "graph"
    meta: dict = field(default_factory=dict)

    @property
    def m(self) -> int:
        return int(self.indptr[-1])

    @property
    def d_max(self) -> int:
        return int(self.degree.max()) if self.n else 0

    @staticmethod
    def from_edges(n: int, u: np.ndarray, v: np.ndarray, name: str = "graph") -> "CSRGraph":
        """Symmetric, de-duplicated, self-loop-free CSR from an undirected edge list."""
        u = np.asarray(u, dtype=np.int64)
        v = np.asarray(v, dtype=np.int64)
        keep = u != v
        u, v = u[keep], v[keep]
        src = np.concatenate([u, v])
        dst = np.concatenate([v, u])
        key = np.unique(src * np.int64(n) + dst)
        src = key // n
        dst = key - src * n
        counts = np.bincount(src, minlength=n).astype(np.int64)
        indptr = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(counts, out=indptr[1:])
        return CSRGraph(n=n, indptr=indptr, indices=dst.astype(np.int32),
                        degree=counts.astype(np.int32), name=name)

    def to_scipy(self):
        import scipy.sparse as sp
        data = np.ones(self.m, dtype=np.float64)
        return sp.csr_matrix((data, self.indices, self.indptr), shape=(self.n, self.n))


@dataclass
class SlackGraph:
    """Slack CSR for in-place edge insertion (``data_process_GNN.py:92-119``)."""
    n: int
    indptr: np.ndarray      # int64 [n+1], capacity per row
    indices: np.ndarray     # int32 [cap], sentinel n in unused slots
    degree: np.ndarray      # int32 [n], live row length
    us: np.ndarray          # int32 [snapshots, batch]
    vs: np.ndarray          # int32 [snapshots, batch]
    name: str = "dynamic"

    def live_csr(self) -> CSRGraph:
        rows = [self.indices[self.indptr[u]:self.indptr[u] + self.degree[u]] for u in range(self.n)]
        indptr = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(self.degree.astype(np.int64), out=indptr[1:])
        idx = np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32)
        return CSRGraph(self.n, indptr, idx, self.degree.astype(np.int32).copy(), self.name)

    def copy(self) -> "SlackGraph":
        return SlackGraph(self.n, self.indptr.copy(), self.indices.copy(), self.degree.copy(),
                          self.us.copy(), self.vs.copy(), self.name)


def _lib():
    return _build.load_graphgen()


def chung_lu(w: np.ndarray, seed: int, permute: bool = True, name: str = "chung-lu") -> CSRGraph:
    """Chung-Lu graph with expected degrees ``w`` (edge {u,v} with probability
    min(1, w_u w_v / sum w)); node ids are randomly permuted so that degree does
    not follow id order, as in a real relabelled graph."""
    w = np.asarray(w, dtype=np.float64)
    n = w.size
    order = np.argsort(-w, kind="stable")
    ws = np.ascontiguousarray(w[order])
    cap = int(ws.sum() * 0.6 + 8 * np.sqrt(ws.sum()) + 1024)
    ou = np.empty(cap, dtype=np.int32)
    ov = np.empty(cap, dtype=np.int32)
    cnt = ctypes.c_int64(0)
    lib = _lib()
    rc = lib.flsg_gen_chung_lu(ctypes.c_int64(n), ws.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(seed),
                               ou.ctypes.data_as(ctypes.c_void_p), ov.ctypes.data_as(ctypes.c_void_p),
                               ctypes.c_int64(cap), ctypes.byref(cnt))
    if rc != 0:
        raise RuntimeError("chung_lu: edge buffer too small")
    c = cnt.value
    u = order[ou[:c]]
    v = order[ov[:c]]
    if permute:
        perm = np.random.default_rng(seed + 1).permutation(n)
        u, v = perm[u], perm[v]
    g = CSRGraph.from_edges(n, u, v, name=name)
    return g


def powerlaw_graph_small(n: int = 10_000, gamma: float = 2.1, wmin: float = 1.0, wmax: float = 500.0,
                         seed: int = 0) -> CSRGraph:
    """Chung-Lu graph whose expected degrees are i.i.d. draws from a power law
    with exponent ``gamma`` truncated to [wmin, wmax] (inverse-CDF sampling)."""
    rng = np.random.default_rng(seed)
    a = 1.0 - gamma
    lo, hi = wmin ** a, wmax ** a
    w = (lo + rng.random(n) * (hi - lo)) ** (1.0 / a)
    g = chung_lu(w, seed=seed + 11, name=f"chunglu-n{n}")
    g.meta.update(gamma=gamma, wmin=wmin, wmax=wmax, seed=seed)
    return g


def powerlaw_graph_large(n: int = 2_394_385, nnz_target: float = 9.5e6, gamma: float = 2.1,
                         floor: float = 1.6, i0: float = 1.4, seed: int = 1) -> CSRGraph:
    """Chung-Lu graph with the deterministic power-law weight sequence
    w_i = floor + c (i + i0)^(-1/(gamma-1)): tail exponent ``gamma``, max expected
    degree ~1e5, about 10% isolated nodes and ``nnz_target`` CSR entries."""
    b = 1.0 / (gamma - 1.0)
    base = (np.arange(n, dtype=np.float64) + i0) ** (-b)
    c = (nnz_target - floor * n) / base.sum()
    w = floor + c * base
    g = chung_lu(w, seed=seed, name=f"chunglu-n{n}")
    g.meta.update(gamma=gamma, floor=floor, i0=i0, seed=seed, nnz_target=nnz_target)
    return g


def citation_dynamic_graph(n: int = 169_343, m_directed: int = 1_166_243, delnum: int = 950_000,
                           snapshots: int = 16, seed: int = 17) -> SlackGraph:
    """Dynamic-graph workload built with the procedure of
    ``data_process_GNN.py:43-171`` on a synthetic citation graph:
    symmetrise, keep a random spanning forest plus all but ``delnum`` of the
    remaining edges in a slack CSR sized for the final graph, and split the
    ``delnum`` removed edges into ``snapshots`` equal insert batches."""
    lib = _lib()
    cap = int(m_directed * 1.1) + 1024
    src = np.empty(cap, dtype=np.int32)
    dst = np.empty(cap, dtype=np.int32)
    cnt = ctypes.c_int64(0)
    rc = lib.flsg_gen_pref_attach(ctypes.c_int64(n), ctypes.c_int64(m_directed), ctypes.c_uint64(seed),
                                  src.ctypes.data_as(ctypes.c_void_p), dst.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_int64(cap), ctypes.byref(cnt))
    if rc != 0:
        raise RuntimeError("pref_attach: buffer too small")
    c = cnt.value
    full = CSRGraph.from_edges(n, src[:c], dst[:c])
    # undirected edge list (u < v), shuffled, then Kruskal
    rows = np.repeat(np.arange(n, dtype=np.int64), full.degree)
    cols = full.indices.astype(np.int64)
    up = rows < cols
    eu, ev = rows[up], cols[up]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(eu.size)
    eu, ev = eu[perm].astype(np.int32), ev[perm].astype(np.int32)
    in_tree = np.zeros(eu.size, dtype=np.uint8)
    lib.flsg_spanning_forest(ctypes.c_int64(n), ctypes.c_int64(eu.size), eu.ctypes.data_as(ctypes.c_void_p),
                             ev.ctypes.data_as(ctypes.c_void_p), in_tree.ctypes.data_as(ctypes.c_void_p))
    tree = in_tree.astype(bool)
    rest = np.flatnonzero(~tree)
    delnum = min(delnum, rest.size)
    rest = rest[rng.permutation(rest.size)]
    removed = rest[:delnum]
    kept = np.concatenate([np.flatnonzero(tree), rest[delnum:]])
    # slack CSR: row capacity = final degree; live = tree + kept edges
    indptr = full.indptr.copy()
    indices = np.full(full.m, n, dtype=np.int32)
    degree = np.zeros(n, dtype=np.int32)
    ku = np.concatenate([eu[kept], ev[kept]]).astype(np.int64)
    kv = np.concatenate([ev[kept], eu[kept]]).astype(np.int64)
    order = np.lexsort((kv, ku))
    ku, kv = ku[order], kv[order]
    cnts = np.bincount(ku, minlength=n)
    start = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cnts, out=start[1:])
    pos = indptr[ku] + (np.arange(ku.size) - start[ku])
    indices[pos] = kv.astype(np.int32)
    degree[:] = cnts.astype(np.int32)
    batch = delnum // snapshots
    sel = removed[:batch * snapshots]
    us = eu[sel].reshape(snapshots, batch).astype(np.int32)
    vs = ev[sel].reshape(snapshots, batch).astype(np.int32)
    return SlackGraph(n=n, indptr=indptr, indices=indices, degree=degree, us=us, vs=vs,
                      name=f"citation-n{n}")


def edge_ids(us, vs, n: int) -> tuple[np.ndarray, np.ndarray]:
    """Edge endpoints as int32 arrays, checked to lie in [0, n) before the
    cast (an id of 2^31 or more would otherwise wrap to another node)."""
    a = np.asarray(us).reshape(-1)
    b = np.asarray(vs).reshape(-1)
    if a.shape != b.shape:
        raise ValueError("us and vs must have the same length")
    for x in (a, b):
        if x.size and (x.dtype.kind not in "iu" or int(x.min()) < 0 or int(x.max()) >= n):
            raise ValueError(f"edge endpoints must be integers in [0, {n})")
    return np.ascontiguousarray(a, dtype=np.int32), np.ascontiguousarray(b, dtype=np.int32)


def demo_seeds(n: int, test_num: int = 50, seed: int = 17) -> np.ndarray:
    """``demo.py:19,45-46``: ``np.random.seed(17); np.unique(randint(n, 50))``."""
    rs = np.random.RandomState(seed)
    s = rs.randint(n, size=test_num)
    return np.array(list(np.unique(s).astype(np.int32)))


def uniform_seeds(g: CSRGraph, k: int, seed: int = 0) -> np.ndarray:
    """k seeds drawn uniformly (with replacement) from nodes with degree >= 1."""
    pool = np.flatnonzero(g.degree >= 1)
    rng = np.random.default_rng(seed)
    return pool[rng.integers(0, pool.size, size=k)].astype(np.int64)


def mixed_seeds(g: CSRGraph, k: int = 1024, frac_uniform: float = 0.7, seed: int = 0) -> np.ndarray:
    """Batched workload seeds: ``frac_uniform`` uniform over non-isolated nodes,
    the rest drawn proportional to degree (uneven frontier sizes)."""
    rng = np.random.default_rng(seed)
    ku = int(round(k * frac_uniform))
    pool = np.flatnonzero(g.degree >= 1)
    a = pool[rng.integers(0, pool.size, size=ku)]
    p = g.degree.astype(np.float64)
    p /= p.sum()
    b = rng.choice(g.n, size=k - ku, p=p)
    out = np.concatenate([a, b]).astype(np.int64)
    return out[rng.permutation(k)]


def alpha_roundtrip(alpha: float):
    """``demo.py:31-35``: returns (alpha', omega) as the reference's harness
    computes them; alpha=0.1 gives 0.10000000000000002."""
    a = alpha / (2 - alpha)
    mu = (1. - a) / (1. + a)
    omega = 1. + (mu / (1. + np.sqrt(1. - mu ** 2.))) ** 2.
    return 2 * a / (1 + a), float(omega)


def save_exp_npz(path: str, g: SlackGraph, node_feature: np.ndarray, labels: np.ndarray | None = None,
                 split: dict | None = None) -> None:
    """Write a dynamic-graph workload in the file layout the reference's
    InstantGNN experiment reads (``data_process_GNN.py:86-87``, read at
    ``instantGNN-arxiv-exp.py:174-190``): slack CSR ``indptr``/``indices``
    (free slots hold n) /``degree`` (int64 live lengths), the insert batches
    ``Us``/``Vs`` (int32 [snapshots, batch]), ``node_feature`` (float32),
    ``labels`` (int64 [n, 1]) and ``split_idx`` (a dict of train/valid/test
    index arrays, stored as an object array as the reference stores it)."""
    n = g.n
    if labels is None:
        labels = np.zeros((n, 1), dtype=np.int64)
    if split is None:
        idx = np.arange(n, dtype=np.int64)
        a, b = int(0.54 * n), int(0.72 * n)
        split = {"train": idx[:a], "valid": idx[a:b], "test": idx[b:]}
    np.savez(path, indices=np.asarray(g.indices, dtype=np.int32), indptr=np.asarray(g.indptr, dtype=np.int64),
             degree=np.asarray(g.degree, dtype=np.int64), Us=np.asarray(g.us, dtype=np.int32),
             Vs=np.asarray(g.vs, dtype=np.int32), node_feature=np.asarray(node_feature, dtype=np.float32),
             labels=np.asarray(labels, dtype=np.int64), split_idx=np.array([split], dtype=object))


def load_exp_npz(path: str, name: str | None = None) -> tuple[SlackGraph, np.ndarray]:
    """Read ``save_exp_npz`` / ``data_process_GNN.py`` output: (slack graph with
    its insert batches, node features).  Only the numeric members are read, so
    the pickled ``split_idx`` is never loaded."""
    with np.load(path, allow_pickle=False) as d:
        indptr = d["indptr"].astype(np.int64)
        indices = d["indices"].astype(np.int32)
        degree = d["degree"].astype(np.int32)
        us = d["Us"].astype(np.int32)
        vs = d["Vs"].astype(np.int32)
        feat = d["node_feature"].astype(np.float32)
    n = degree.size
    g = SlackGraph(n=n, indptr=indptr, indices=indices, degree=degree, us=us, vs=vs,
                   name=name or os.path.basename(path))
    return g, feat


def save_npz(path: str, g: CSRGraph) -> None:
    """Write a file ``scipy.sparse.load_npz`` reads (the reference's
    ``demo.py:39`` input format)."""
    import scipy.sparse as sp
    sp.save_npz(path, g.to_scipy().astype(np.float32))


def load_npz(path: str, name: str | None = None) -> CSRGraph:
    """Read a ``scipy.sparse.save_npz`` CSR (``demo.py:39``), with int32
    degrees = row lengths."""
    import scipy.sparse as sp
    a = sp.load_npz(path).tocsr()
    a.sort_indices()
    n = a.shape[0]
    indptr = a.indptr.astype(np.int64)
    return CSRGraph(n=n, indptr=indptr, indices=a.indices.astype(np.int32),
                    degree=np.diff(indptr).astype(np.int32), name=name or os.path.basename(path))
