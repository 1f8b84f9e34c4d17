"""Oracle: CPU restatement of the reference's solvers (test infrastructure).

Only the tests, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
use this module, always as the checker or the timed CPU baseline -- never as
part of the product path.

C restatement (flsg_oracle.c, built -O2 -ffp-contract=off), one function per
reference function:

  ppr_gd          ppr_solver.py:208-328      pinned: dense solve identity,
                                             hand-traced tiny graphs, invariants
  ppr_cheby       ppr_solver.py:582-661      pinned: exact-rational local hand
                                             trace (queues, x, r), global sweeps =
                                             Chebyshev semi-iterates (closed form)
  katz_gd         katz_solver.py:167-248     pinned: dense Jacobi frontier replay
                                             (eps*d thresholds), dense solve
  katz_cheby      katz_solver.py:340-431     pinned: as ppr_cheby, Katz interval
  estimate_n      hk_solver.py:28-48         pinned: Taylor-tail bound property
  hk_taylor       hk_solver.py:51-80         pinned: scipy expm
  hk_push         hk_solver.py:206-295       pinned: level identity vs. the
                                             truncated Taylor series, dense level
                                             replay with closed-form psi, hk-relax
                                             bound vs expm
  ignn_column     InstantGNN.py:12-71        pinned: invariant p + a(I-(1-a)M)^-1 r
  update_graph    InstantGNN.py:137-180      pinned: resulting adjacency sets
  ppr_sor         ppr_solver.py:331-461      pinned: one global sweep = the dense SOR
                                             update in natural order, invariants,
                                             hand-traced path graph
  katz_sor        katz_solver.py:252-337     pinned: the same matrix form with
                                             Q = I - bA, dense Katz solve
  ignn_sor_column InstantGNN.py:74-105       pinned: propagation invariant,
                                             certificate, hand trace
  ignn_sor_coloured (this build's GPU variant, restated)  pinned: the same
                                             invariant and certificate, hand trace
  colour          (this build's GPU colouring, restated) pinned: proper colouring,
                                             first-fit minimality, rounds by hand
  ppr_sor_coloured (this build's GPU SOR, restated)  pinned: one global sweep = the
                                             dense SOR update in colour order,
                                             certificate vs. ppr_sor's solution
  ignn_update     InstantGNN.py:298-329      pinned: invariant on the new graph
  ppr_batch_l1, residual_cert (checkers of sparse results, no reference
                  counterpart)            pinned: numpy dense L1 and dense
                                             residual b - Qx on a small graph
The reference itself cannot run here (numba / cupy are not importable:
an ordinary ImportError, not a denial), so the pins are the mathematics and
hand-worked cases in tests/test_oracle.py and tests/test_oracle_pins.py;
oracle/mutate.py shows that each of a list of plausible restatement slips
fails at least one of them.

numpy helpers (``ppr_dense_solve``, ``ppr_residual`` ...) give the FP64 global
solves and residual recomputations the certificates are checked against.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FLSG_ORACLE_LIB names another build of the same source (oracle/mutate.py
# loads deliberately altered copies through it to show which test fails)
LIB_PATH = os.environ.get("FLSG_ORACLE_LIB") or os.path.join(HERE, "libflsg_oracle.so")

ORA_OK, ORA_EASSERT, ORA_ENOMEM, ORA_EARG, ORA_ETRACE = 0, 1, 2, 3, 4

_lib = None


class _Trace(C.Structure):
    _fields_ = [("off", C.c_void_p), ("nodes", C.c_void_p), ("cap_iters", C.c_int64), ("cap_nodes", C.c_int64),
                ("n_iters", C.c_int64), ("n_nodes", C.c_int64), ("overflow", C.c_int32)]


def build() -> str:
    r = subprocess.run(["make", "-s", "-B", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i64, dbl, i32 = C.c_void_p, C.c_int64, C.c_double, C.c_int
        L.ora_ppr_gd.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, i32, vp, vp, vp, vp, vp]
        L.ora_ppr_cheby.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, i32, vp, vp, vp, vp, vp]
        L.ora_katz_gd.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, i32, vp, vp, vp, vp, vp]
        L.ora_katz_cheby.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, dbl, dbl, i32, vp, vp, vp, vp, vp]
        L.ora_estimate_n.argtypes = [dbl, dbl]
        L.ora_estimate_n.restype = i64
        L.ora_estimate_coefficient.argtypes = [dbl, i64]
        L.ora_estimate_coefficient.restype = dbl
        L.ora_hk_taylor.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, i64, vp, vp]
        L.ora_hk_push.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, dbl, i32, i64, vp, vp, vp, vp, vp, vp]
        L.ora_ignn_column.argtypes = [i64, vp, vp, vp, dbl, dbl, dbl, vp, vp, vp]
        L.ora_update_graph.argtypes = [i64, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp]
        L.ora_ignn_update_column.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp, dbl, dbl, dbl, vp, vp, vp, i64, vp]
        L.ora_ppr_batch.argtypes = [i64, vp, vp, vp, vp, i64, dbl, dbl, i32, i32, vp, vp]
        L.ora_ppr_batch_l1.argtypes = [i64, vp, vp, vp, vp, i64, dbl, dbl, i32, vp, vp, vp, vp, vp, vp, vp]
        L.ora_residual_cert.argtypes = [i64, vp, vp, vp, i32, dbl, dbl, vp, i64, i32, vp, vp, vp, vp, vp, vp]
        L.ora_max_threads.restype = i32
        L.ora_ppr_sor.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, dbl, i32, vp, vp, vp, vp, vp]
        L.ora_colour.argtypes = [i64, vp, vp, vp, vp, vp]
        L.ora_colour.restype = i64
        L.ora_katz_sor.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, dbl, i32, vp, vp, vp, vp, vp]
        L.ora_ignn_sor_push.argtypes = [i64, vp, vp, vp, dbl, vp, dbl, vp, vp, vp, i64, i64, vp, dbl, vp]
        L.ora_ignn_sor_coloured.argtypes = [i64, vp, vp, vp, vp, i64, dbl, dbl, vp, vp, vp, vp, vp, i64, vp, vp,
                                            i64, vp, vp]
        L.ora_sor_coloured.argtypes = [i32, i64, vp, vp, vp, vp, i64, i64, dbl, dbl, dbl, i32, i64, vp, vp, vp, vp,
                                       vp]
        L.ora_near_threshold.argtypes = []
        L.ora_near_threshold.restype = i64
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _csr(g):
    indptr = np.ascontiguousarray(g.indptr, dtype=np.int64)
    indices = np.ascontiguousarray(g.indices, dtype=np.int32)
    degree = np.ascontiguousarray(g.degree, dtype=np.int32)
    return indptr, indices, degree


def _mk_trace(n: int, cap_nodes: int | None, cap_iters: int = 100000):
    cap_nodes = cap_nodes if cap_nodes is not None else max(64 * n, 1 << 20)
    off = np.zeros(cap_iters + 1, dtype=np.int64)
    nodes = np.zeros(cap_nodes, dtype=np.int32)
    tr = _Trace(off=_p(off), nodes=_p(nodes), cap_iters=cap_iters, cap_nodes=cap_nodes)
    return tr, off, nodes


def _unpack_trace(tr, off, nodes):
    if tr.overflow:
        raise RuntimeError("oracle trace overflow")
    return [nodes[off[t]:off[t + 1]].copy() for t in range(tr.n_iters)]


def _run_local(fn, g, s, *args, trace=False, trace_cap=None):
    if trace and trace_cap is None:
        cap = max(32 * g.n, 1 << 20)
        for _ in range(4):
            try:
                return _run_local(fn, g, s, *args, trace=True, trace_cap=cap)
            except OverflowError:
                cap *= 8
        raise RuntimeError("oracle trace overflow")
    indptr, indices, degree = _csr(g)
    n = g.n
    x = np.zeros(n)
    r = np.zeros(n)
    oper = np.zeros(1, dtype=np.int64)
    t = np.zeros(1)
    tr = off = nodes = None
    if trace:
        tr, off, nodes = _mk_trace(n, trace_cap)
    rc = fn(n, _p(indptr), _p(indices), _p(degree), int(s), *args, _p(x), _p(r), _p(oper), _p(t),
            C.byref(tr) if trace else None)
    if rc == ORA_EASSERT:
        raise AssertionError("eps * degree[s_node] <= 1.")
    if rc == ORA_ETRACE:
        raise OverflowError("trace capacity")
    if rc not in (ORA_OK,):
        raise RuntimeError(f"oracle error {rc}")
    # enqueue tests within one ulp of the threshold (local methods; same thread)
    out = {"x": x, "r": r, "oper": int(oper[0]), "time": float(t[0]), "near_thr": int(lib().ora_near_threshold())}
    if trace:
        out["trace"] = _unpack_trace(tr, off, nodes)
    return out


def ppr_gd(g, s, alpha, eps, local=True, trace=False):
    return _run_local(lib().ora_ppr_gd, g, s, float(alpha), float(eps), int(local), trace=trace)


def ppr_cheby(g, s, alpha, eps, local=True, trace=False):
    return _run_local(lib().ora_ppr_cheby, g, s, float(alpha), float(eps), int(local), trace=trace)


def katz_gd(g, s, alpha, eps, local=True, trace=False):
    return _run_local(lib().ora_katz_gd, g, s, float(alpha), float(eps), int(local), trace=trace)


def katz_cheby(g, s, alpha, l_max, mu_min, eps, local=True, trace=False):
    return _run_local(lib().ora_katz_cheby, g, s, float(alpha), float(l_max), float(mu_min), float(eps),
                      int(local), trace=trace)


def estimate_n(t, eps) -> int:
    return int(lib().ora_estimate_n(float(t), float(eps)))


def estimate_coefficient(t, k) -> float:
    return float(lib().ora_estimate_coefficient(float(t), int(k)))


def hk_taylor(g, s, tau, eps, n_terms=0):
    indptr, indices, degree = _csr(g)
    x = np.zeros(g.n)
    est = np.zeros(1, dtype=np.int64)
    rc = lib().ora_hk_taylor(g.n, _p(indptr), _p(indices), _p(degree), int(s), float(tau), float(eps),
                             int(n_terms), _p(x), _p(est))
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    return {"x": x, "est_n": int(est[0])}


def hk_push(g, s, tau, eps, omega=1.0, local=True, n_terms=0, trace=False):
    indptr, indices, degree = _csr(g)
    n = g.n
    x = np.zeros(n)
    r = np.zeros(2 * n)
    oper = np.zeros(1)
    est = np.zeros(1, dtype=np.int64)
    t = np.zeros(1)
    tr = off = nodes = None
    if trace:
        tr, off, nodes = _mk_trace(n, None)
    rc = lib().ora_hk_push(n, _p(indptr), _p(indices), _p(degree), int(s), float(tau), float(eps), float(omega),
                           int(local), int(n_terms), _p(x), _p(r), _p(oper), _p(est), _p(t),
                           C.byref(tr) if trace else None)
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    out = {"x": x, "r": r.reshape(2, n), "oper": float(oper[0]), "est_n": int(est[0]), "time": float(t[0])}
    if trace:
        out["trace"] = _unpack_trace(tr, off, nodes)
    return out


def ignn_column(indptr, indices, degree, alpha, eps, beta, r_col):
    """One column of instantGNNParallel: returns (p float32, r float64, edge visits)."""
    n = degree.size
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    degree = np.ascontiguousarray(degree, dtype=np.int32)
    p = np.zeros(n, dtype=np.float32)
    r = np.ascontiguousarray(r_col, dtype=np.float64).copy()
    oper = np.zeros(1, dtype=np.int64)
    rc = lib().ora_ignn_column(n, _p(indptr), _p(indices), _p(degree), float(alpha), float(eps), float(beta),
                               _p(p), _p(r), _p(oper))
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    return p, r, int(oper[0])


def update_graph(indptr, indices, degree, us, vs):
    """updetaGraph(onlyinsert=True): mutates indices/degree (int32 arrays) in
    place, returns the insert_neighbor dict as (keys, nbr_off, nbr)."""
    n = degree.size
    k = int(np.asarray(us).size)
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    assert indices.dtype == np.int32 and degree.dtype == np.int32
    us = np.ascontiguousarray(us, dtype=np.int32)
    vs = np.ascontiguousarray(vs, dtype=np.int32)
    keys = np.zeros(2 * k + 1, dtype=np.int64)
    nbr_off = np.zeros(2 * k + 2, dtype=np.int64)
    nbr = np.zeros(2 * k + 1, dtype=np.int64)
    nk = np.zeros(1, dtype=np.int64)
    rc = lib().ora_update_graph(n, _p(indptr), _p(indices), _p(degree), _p(us), _p(vs), k, _p(keys), _p(nbr_off),
                                _p(nbr), _p(nk))
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    nkv = int(nk[0])
    return keys[:nkv].copy(), nbr_off[:nkv + 1].copy(), nbr[:int(nbr_off[nkv])].copy()


def ignn_update_column(indptr, indices, degree, old_degree, p, r, x, alpha, eps, beta, keys, nbr_off, nbr):
    """instantGNNUpdateBatch for one column (p, r float64 updated in place)."""
    n = degree.size
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    oper = np.zeros(1, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert p.dtype == np.float64 and r.dtype == np.float64 and p.flags.c_contiguous and r.flags.c_contiguous
    rc = lib().ora_ignn_update_column(n, _p(indptr), _p(indices), _p(degree), _p(old_degree), _p(p), _p(r), _p(x),
                                      float(alpha), float(eps), float(beta), _p(keys), _p(nbr_off), _p(nbr),
                                      int(keys.size), _p(oper))
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    return int(oper[0])


def ppr_batch(g, seeds, alpha, eps, kind=0, nthreads=0):
    """OpenMP fan-out over seeds (local GD kind=0 / local Chebyshev kind=1):
    per-seed (oper, seconds)."""
    indptr, indices, degree = _csr(g)
    seeds = np.ascontiguousarray(seeds, dtype=np.int64)
    opers = np.zeros(seeds.size, dtype=np.int64)
    times = np.zeros(seeds.size)
    rc = lib().ora_ppr_batch(g.n, _p(indptr), _p(indices), _p(degree), _p(seeds), seeds.size, float(alpha),
                             float(eps), int(kind), int(nthreads), _p(opers), _p(times))
    if rc == ORA_EASSERT:
        raise AssertionError("eps * degree[s_node] <= 1.")
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    return opers, times


def _sparse_set(off, cnt, ids, vals):
    return (np.ascontiguousarray(off, dtype=np.int64), np.ascontiguousarray(cnt, dtype=np.int64),
            np.ascontiguousarray(ids, dtype=np.int32), np.ascontiguousarray(vals, dtype=np.float64))


def ppr_batch_l1(g, seeds, alpha, eps, off, cnt, ids, vals, nthreads=0):
    """Checker of batched results: per query q, the FP64 local GD of seeds[q]
    (ppr_gd) and the L1 distance sum_v |x_v - y_v| of the sparse result
    y = (ids, vals)[off[q] : off[q] + cnt[q]] to it; OpenMP over queries.
    Returns (l1, sum |x|, oper) arrays."""
    indptr, indices, degree = _csr(g)
    seeds = np.ascontiguousarray(seeds, dtype=np.int64)
    off, cnt, ids, vals = _sparse_set(off, cnt, ids, vals)
    l1 = np.zeros(seeds.size)
    norm = np.zeros(seeds.size)
    opers = np.zeros(seeds.size, dtype=np.int64)
    rc = lib().ora_ppr_batch_l1(g.n, _p(indptr), _p(indices), _p(degree), _p(seeds), seeds.size, float(alpha),
                                float(eps), int(nthreads), _p(off), _p(cnt), _p(ids), _p(vals), _p(l1), _p(norm),
                                _p(opers))
    if rc == ORA_EASSERT:
        raise AssertionError("eps * degree[s_node] <= 1.")
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    return l1, norm, opers


def residual_cert(g, kind, a, eps, seeds, off, cnt, ids, vals, nthreads=0):
    """Independent FP64 certificate of sparse results: per query the residual
    r = b - Q x recomputed from the graph (PPR kind 0: a = alpha, r = a e_s -
    x + (1-a) A D^-1 x, thr = eps a d; Katz kind 1: a = beta, x the full
    solution, r = e_s - x + a A x, thr = eps d), scattered over the support's
    live adjacency.  Returns (max_v |r_v| / thr_v, sum_v |r_v|) arrays."""
    indptr, indices, degree = _csr(g)
    seeds = np.ascontiguousarray(seeds, dtype=np.int64)
    off, cnt, ids, vals = _sparse_set(off, cnt, ids, vals)
    cert = np.zeros(seeds.size)
    rl1 = np.zeros(seeds.size)
    rc = lib().ora_residual_cert(g.n, _p(indptr), _p(indices), _p(degree), int(kind), float(a), float(eps),
                                 _p(seeds), seeds.size, int(nthreads), _p(off), _p(cnt), _p(ids), _p(vals),
                                 _p(cert), _p(rl1))
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    return cert, rl1


def ppr_sor(g, s, alpha, eps, omega, local=True, _fn="ora_ppr_sor"):
    """The reference's ppr_sor (local FIFO or natural-order sweeps)."""
    indptr, indices, degree = _csr(g)
    x = np.zeros(g.n)
    r = np.zeros(g.n)
    oper = np.zeros(1)
    iters = np.zeros(1, dtype=np.int64)
    t = np.zeros(1)
    rc = getattr(lib(), _fn)(g.n, _p(indptr), _p(indices), _p(degree), int(s), float(alpha), float(eps),
                             float(omega), int(local), _p(x), _p(r), _p(oper), _p(iters), _p(t))
    if rc == ORA_EASSERT:
        raise AssertionError("eps * degree[s_node] <= 1.")
    if rc != ORA_OK:
        raise RuntimeError(f"oracle error {rc}")
    return {"x": x, "r": r, "oper": float(oper[0]), "iters": int(iters[0]), "time": float(t[0])}


def katz_sor(g, s, beta, eps, omega, local=True):
    """The reference's katz_sor (katz_solver.py:252-337)."""
    return ppr_sor(g, s, beta, eps, omega, local=local, _fn="ora_katz_sor")


def colour(g):
    """Jones-Plassmann colouring restated from the GPU variant:
    (colour[n] int32, number of colours, rounds)."""
    indptr, indices, degree = _csr(g)
    col = np.zeros(g.n, dtype=np.int32)
    rounds = np.zeros(1, dtype=np.int64)
    k = lib().ora_colour(g.n, _p(indptr), _p(indices), _p(degree), _p(col), _p(rounds))
    if k < 0:
        raise MemoryError("oracle colouring")
    return col, int(k), int(rounds[0])


def ppr_sor_coloured(g, s, alpha, eps, omega, col, ncol, local=True, max_iters=100000, kind=0):
    """Sequential restatement of the GPU coloured SOR push (kind 0 PPR,
    1 Katz with alpha = beta)."""
    indptr, indices, degree = _csr(g)
    col = np.ascontiguousarray(col, dtype=np.int32)
    x = np.zeros(g.n)
    r = np.zeros(g.n)
    out = np.zeros(3, dtype=np.int64)
    rc = lib().ora_sor_coloured(int(kind), g.n, _p(indptr), _p(indices), _p(degree), _p(col), int(ncol), int(s),
                                float(alpha), float(eps), float(omega), int(local), int(max_iters),
                                _p(x), _p(r), _p(out[0:]), _p(out[1:]), _p(out[2:]))
    if rc == ORA_EASSERT:
        raise AssertionError("eps * degree[s_node] <= 1.")
    if rc != ORA_OK:
        raise RuntimeError(f"oracle error {rc}")
    return {"x": x, "r": r, "oper": int(out[0]), "iters": int(out[1]), "nodes": int(out[2])}


def katz_sor_coloured(g, s, beta, eps, omega, col, ncol, local=True, max_iters=100000):
    return ppr_sor_coloured(g, s, beta, eps, omega, col, ncol, local=local, max_iters=max_iters, kind=1)


def _slack_arrays(indptr, indices, degree):
    return (np.ascontiguousarray(indptr, dtype=np.int64), np.ascontiguousarray(indices, dtype=np.int32),
            np.ascontiguousarray(degree, dtype=np.int32))


def ignn_eps_vec(degree, eps, beta):
    """eps_vec = np.power(degree, 1 - beta) * eps (InstantGNN.py:57, 118)."""
    return np.power(np.asarray(degree, dtype=np.float64), 1 - beta) * eps


def ignn_sor_column(indptr, indices, degree, alpha, eps, omega, beta, p, r, nodes=None):
    """instantGNNSorPush on one column (p, r FP64, updated in place), the
    queue filled as instantGNNSorParallel does (every node with |r| >=
    eps_vec, in node order, InstantGNN.py:118-123) or from ``nodes``.
    Returns the sum of degrees pushed."""
    ip, ix, dg = _slack_arrays(indptr, indices, degree)
    n = dg.size
    ev = ignn_eps_vec(dg, eps, beta)
    qn = n + 2
    queue = np.zeros(qn, dtype=np.int64)
    q_mark = np.zeros(qn, dtype=np.uint8)
    rear = 0
    cand = np.arange(n) if nodes is None else np.asarray(nodes, dtype=np.int64)
    for i in cand:
        if np.abs(r[i]) >= ev[i] and not q_mark[i]:
            rear = (rear + 1) % qn
            queue[rear] = i
            q_mark[i] = 1
    oper = np.zeros(1, dtype=np.int64)
    rc = lib().ora_ignn_sor_push(qn, _p(ip), _p(ix), _p(dg), float(alpha), _p(ev), float(omega), _p(p), _p(r),
                                 _p(queue), 0, int(rear), _p(q_mark), float(beta), _p(oper))
    if rc != ORA_OK:
        raise RuntimeError(f"oracle error {rc}")
    return int(oper[0])


def self_loops(indptr, indices, degree):
    ip, ix, dg = _slack_arrays(indptr, indices, degree)
    rows = np.repeat(np.arange(dg.size), dg)
    cols = np.concatenate([ix[ip[u]:ip[u] + dg[u]] for u in range(dg.size)]) if dg.size else np.zeros(0, int)
    return np.bincount(rows[rows == cols], minlength=dg.size).astype(np.int32)


def ignn_sor_coloured(indptr, indices, degree, col, ncol, alpha, eps, omega, beta, p, r, nodes,
                      max_iters=100000):
    """The GPU's coloured InstantGNN SOR on one column, restated (p, r
    updated in place); nodes = the starting frontier (distinct)."""
    ip, ix, dg = _slack_arrays(indptr, indices, degree)
    d = dg.astype(np.float64)
    ev = ignn_eps_vec(dg, eps, beta)
    dp1 = np.power(d, 1 - beta)
    dpb = np.power(d, beta)
    loops = self_loops(ip, ix, dg)
    F0 = np.ascontiguousarray(nodes, dtype=np.int32)
    out = np.zeros(2, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    rc = lib().ora_ignn_sor_coloured(dg.size, _p(ip), _p(ix), _p(dg), _p(col), int(ncol), float(alpha),
                                     float(omega), _p(ev), _p(dp1), _p(dpb), _p(loops), _p(F0), F0.size, _p(p),
                                     _p(r), int(max_iters), _p(out[0:]), _p(out[1:]))
    if rc != ORA_OK:
        raise RuntimeError(f"oracle error {rc}")
    return {"oper": int(out[0]), "iters": int(out[1])}


def max_threads() -> int:
    return int(lib().ora_max_threads())


# --------------------------------------------------------------------- numpy helpers
def adjacency(g):
    import scipy.sparse as sp
    data = np.ones(int(g.indptr[-1]), dtype=np.float64)
    return sp.csr_matrix((data, np.asarray(g.indices), np.asarray(g.indptr)), shape=(g.n, g.n))


def ppr_operator(g, alpha):
    """Q = I - (1-alpha) A D^{-1} (isolated columns left as identity)."""
    import scipy.sparse as sp
    A = adjacency(g)
    d = np.asarray(g.degree, dtype=np.float64)
    dinv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    return sp.identity(g.n, format="csr") - (1.0 - alpha) * (A @ sp.diags(dinv))


def ppr_residual(g, s, alpha, x):
    """r = alpha e_s - Q x in FP64 (the certificate residual)."""
    Q = ppr_operator(g, alpha)
    b = np.zeros(g.n)
    b[s] = alpha
    return b - Q @ x


def ppr_dense_solve(g, s, alpha):
    """Exact PPR vector by a dense FP64 solve (small graphs only)."""
    Q = ppr_operator(g, alpha).toarray()
    b = np.zeros(g.n)
    b[s] = alpha
    return np.linalg.solve(Q, b)


def ppr_global_solve(g, s, alpha, tol=1e-15, max_iter=100000):
    """FP64 global solve by power iteration x <- alpha e_s + (1-alpha) A D^-1 x
    until ||r||_1 <= tol (residual recomputed in FP64)."""
    import scipy.sparse as sp
    A = adjacency(g)
    d = np.asarray(g.degree, dtype=np.float64)
    dinv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    P = (1.0 - alpha) * (A @ sp.diags(dinv))
    x = np.zeros(g.n)
    x[s] = alpha
    b = x.copy()
    for _ in range(max_iter):
        xn = b + P @ x
        r = xn - x
        x = xn
        if np.abs(r).sum() <= tol:
            break
    return x
