"""The drop-in modules called as the reference's harness and drivers call
them (positional arguments in the reference's order, the reference's return
tuples), checked against the oracle:
  * PPR local GD: operation count equal, x within relative L1 1e-12; global GD
    and local Chebyshev: x bit-identical (the kernels' own parity bars);
  * Katz Chebyshev through solve_a_graph_cheby with (alpha, l_max, mu_min, eps)
    in the reference's order: per-seed operation counts equal;
  * heat kernel: est_n and oper equal, x within relative L1 1e-12;
  * InstantGNN's per-column push and update, driven like instantGNNParallel /
    instantGNNUpdateBatchParallel: every |r_u| below eps_vec_u (or r_u = 0
    where eps_vec_u = 0) and r equal to the FP64 recompute of the invariant
    X - P/alpha + ((1-alpha)/alpha) M P (relative 1e-9);
  * the reference's AssertionError for eps * d_s > 1 (ppr_solver.py:213).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

import InstantGNN  # noqa: E402
import hk_solver  # noqa: E402
import katz_solver  # noqa: E402
import ppr_solver  # noqa: E402
from flsg import graph  # noqa: E402
from flsg.ignn import ignn_residual  # noqa: E402
from oracle import oracle as O  # noqa: E402

ALPHA = 0.1


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def _csr(g):
    return g.n, g.indptr, g.indices, g.degree


def test_ppr_dropins_match_oracle(small_graph):
    g = small_graph
    n, ip, ind, deg = _csr(g)
    for s in graph.uniform_seeds(g, 3, seed=31):
        s = int(s)
        xt, rt, oper, _ = ppr_solver.ppr_gd(n, ip, ind, deg, s, ALPHA, 1e-6, True, None)
        o = O.ppr_gd(g, s, ALPHA, 1e-6, local=True)
        assert oper == o["oper"] and _rel_l1(xt, o["x"]) <= 1e-12
        xt, rt, oper, _ = ppr_solver.ppr_gd(n, ip, ind, deg, s, ALPHA, 1e-6, False, None)
        o = O.ppr_gd(g, s, ALPHA, 1e-6, local=False)
        assert oper == o["oper"] and np.array_equal(xt, o["x"]) and np.array_equal(rt, o["r"])
        xt, rt, oper, _ = ppr_solver.ppr_cheby(n, ip, ind, deg, s, ALPHA, 1e-6, True)
        o = O.ppr_cheby(g, s, ALPHA, 1e-6, local=True)
        assert oper == o["oper"] and np.array_equal(xt, o["x"]) and np.array_equal(rt, o["r"])


def test_ppr_dropin_asserts_like_reference(small_graph):
    n, ip, ind, deg = _csr(small_graph)
    s = int(np.argmax(deg))
    with pytest.raises(AssertionError):
        ppr_solver.ppr_gd(n, ip, ind, deg, s, ALPHA, 2.0 / float(deg[s]), True, None)


def test_ppr_drivers_return_the_reference_tuples(small_graph):
    g = small_graph
    n, ip, ind, deg = _csr(g)
    seeds = graph.demo_seeds(n, 6)
    seeds = seeds[deg[seeds] > 0]
    out = ppr_solver.solve_a_graph_cheby(n, ip, ind, deg, ALPHA, 1e-6, 1.0, "g", seeds)
    assert len(out) == 5
    out13 = ppr_solver.solve_a_graph_all(n, ip, ind, deg, ALPHA, 1e-6, 1.0, "g", seeds)
    assert len(out13) == 13
    for i, s in enumerate(seeds):
        assert out[1][i] == O.ppr_cheby(g, int(s), ALPHA, 1e-6, local=True)["oper"]
        assert out[3][i] == O.ppr_cheby(g, int(s), ALPHA, 1e-6, local=False)["oper"]
        assert out13[9][i] == O.ppr_gd(g, int(s), ALPHA, 1e-6, local=True)["oper"]
        assert out13[11][i] == O.ppr_gd(g, int(s), ALPHA, 1e-6, local=False)["oper"]


def test_katz_cheby_driver_argument_order(small_graph):
    g = small_graph
    n, ip, ind, deg = _csr(g)
    beta = 1.0 / (int(deg.max()) + 1)
    l_max, mu_min = katz_solver.eigen_bounds(n, ip, ind, beta)
    seeds = graph.uniform_seeds(g, 3, seed=32)
    s_nodes, lo, lt, go, gt = katz_solver.solve_a_graph_cheby(n, ip, ind, deg, beta, l_max, mu_min, 1e-8, "g",
                                                              seeds)
    for i, s in enumerate(seeds):
        o = O.katz_cheby(g, int(s), beta, l_max, mu_min, 1e-8, local=True)
        assert lo[i] == o["oper"]
        xt, rt, oper, _ = katz_solver.katz_cheby(n, ip, ind, deg, int(s), beta, l_max, mu_min, 1e-8, True)
        assert oper == o["oper"] and np.array_equal(xt, o["x"])


def test_hk_dropin_matches_oracle(small_graph):
    g = small_graph
    n, ip, ind, deg = _csr(g)
    for s in graph.uniform_seeds(g, 2, seed=33):
        est_n, xt, rt, oper, beta_t, run_time, algo_time = hk_solver.hk_sor_queue(n, ip, ind, deg, int(s), 5.0,
                                                                                   1e-6, 1.0, True)
        o = O.hk_push(g, int(s), 5.0, 1e-6, local=True)
        assert est_n == o["est_n"] and oper == o["oper"]
        assert _rel_l1(xt, o["x"]) <= 1e-12
    # local=False: the level sweep over all nodes, bit-identical; beta_t
    # stays [1.0] as the reference leaves it there
    s = int(graph.uniform_seeds(g, 1, seed=34)[0])
    est_n, xt, rt, oper, beta_t, run_time, algo_time = hk_solver.hk_sor_queue(n, ip, ind, deg, s, 5.0, 1e-6, 1.0,
                                                                               False)
    o = O.hk_push(g, s, 5.0, 1e-6, local=False)
    assert est_n == o["est_n"] and oper == o["oper"] and beta_t == [1.0]
    assert np.array_equal(xt, o["x"]) and not rt.any()


def _certified(ip, ind, deg, x, p, r, eps, beta):
    eps_vec = np.power(deg.astype(np.float64), 1 - beta) * eps
    assert np.all((np.abs(r) < eps_vec) | (r == 0.0))
    Rh = ignn_residual(ip, ind, deg, x[:, None], p[:, None], ALPHA, beta)[:, 0]
    assert np.abs(Rh - r).max() <= 1e-9 * np.abs(x).max()


def test_instantgnn_column_dropins():
    """instantGNNPush seeded as instantGNNParallel seeds it
    (InstantGNN.py:52-62), then one snapshot through updetaGraph and
    instantGNNUpdateBatch as instantGNNUpdateBatchParallel calls them
    (InstantGNN.py:334-351)."""
    sg = graph.citation_dynamic_graph(n=400, m_directed=2400, delnum=300, snapshots=4, seed=5)
    n, ip = sg.n, sg.indptr
    ind, deg = sg.indices.copy(), sg.degree.copy()
    beta, eps = 0.5, 1e-7
    rng = np.random.default_rng(4)
    x = rng.standard_normal(n)
    r, p = x.copy(), np.zeros(n)
    eps_vec = np.power(deg.astype(np.float64), 1 - beta) * eps
    queue = np.zeros(n + 2, dtype=np.int32)
    front, rear = 0, 0
    q_mark = np.zeros(n)
    for i in range(n):
        if np.abs(r[i]) >= eps_vec[i]:
            rear = (rear + 1) % (n + 2)
            queue[rear] = i
            q_mark[i] = True
    InstantGNN.instantGNNPush(n + 2, ip, ind, deg, ALPHA, eps_vec, p, r, queue, front, rear, q_mark, beta)
    assert not q_mark.any()
    _certified(ip, ind, deg, x, p, r, eps, beta)
    old = deg.copy()
    ins, dele = InstantGNN.updetaGraph(ip, ind, deg, sg.us[0], sg.vs[0], onlyinsert=True)
    assert (deg >= old).all() and deg.sum() > old.sum()
    eps_vec = np.power(deg.astype(np.float64), 1 - beta) * eps
    queue = np.zeros(n + 1, dtype=np.int32)
    q_mark = np.zeros(n)
    InstantGNN.instantGNNUpdateBatch(n, ip, ind, deg, old, p, r, x, ALPHA, queue, 0, 0, q_mark, eps_vec, ins, dele,
                                     algo="fwd_insert", omega=1, beta=beta)
    _certified(ip, ind, deg, x, p, r, eps, beta)


def test_dropin_sees_arrays_changed_in_place(small_graph):
    """The resident-graph cache is keyed by content: after the caller rewires
    edges in the same arrays (a degree-preserving double edge swap, rows kept
    sorted), ppr_gd / katz_gd solve the new graph."""
    g = small_graph
    ip, ind, deg = g.indptr.copy(), g.indices.copy(), g.degree.copy()
    s = int(graph.uniform_seeds(g, 1, seed=41)[0])
    xt0, _, oper0, _ = ppr_solver.ppr_gd(g.n, ip, ind, deg, s, ALPHA, 1e-6, True, None)
    nbrs = lambda u: set(ind[ip[u]:ip[u + 1]].tolist())
    # a swap among the seed's neighbourhood, so the PPR vector changes
    a = s
    b = int(ind[ip[a]])
    rng = np.random.default_rng(0)
    for c in rng.permutation(g.n):
        c = int(c)
        if c in (a, b) or deg[c] == 0:
            continue
        d = int(ind[ip[c]])
        if d in (a, b) or d in nbrs(a) or b in nbrs(c):
            continue
        break
    for u, old, new in ((a, b, d), (b, a, c), (c, d, b), (d, c, a)):
        row = ind[ip[u]:ip[u + 1]]
        row[row == old] = new
        row.sort()
    g2 = graph.CSRGraph(n=g.n, indptr=ip, indices=ind, degree=deg)
    xt1, _, oper1, _ = ppr_solver.ppr_gd(g.n, ip, ind, deg, s, ALPHA, 1e-6, True, None)
    o = O.ppr_gd(g2, s, ALPHA, 1e-6, local=True)
    assert oper1 == o["oper"] and _rel_l1(xt1, o["x"]) <= 1e-12
    assert _rel_l1(xt1, xt0) > 1e-6
    # (beta, eps) with no enqueue test within an ulp of its threshold, so the
    # operation counts must agree exactly
    xk, _, ok, _ = katz_solver.katz_gd(g.n, ip, ind, deg, s, 0.0037, 1.3e-8, True)
    ko = O.katz_gd(g2, s, 0.0037, 1.3e-8, local=True)
    assert ko["near_thr"] == 0
    assert ok == ko["oper"] and _rel_l1(xk, ko["x"]) <= 1e-12
