"""Pins of the oracle (the CPU restatement of the reference's solvers).

The reference cannot run here (numba / cupy are not importable) and ships no
golden vectors for its solvers, so each oracle function is pinned by checks a
plausible restatement mistake would fail:
  * hand-traced tiny graphs with exact (dyadic) arithmetic: operation counts,
    frontier order, the >= threshold test, Jacobi (zero-then-push) semantics;
  * exact identities against dense FP64 linear algebra (x* - x = Q^-1 r,
    mass conservation, residual consistency r == b - Q x);
  * closed forms (Chebyshev coefficient limit, estimate_n's geometric
    recurrence, scipy expm for the heat kernel).
"""
import math

import numpy as np
import pytest

from flsg.graph import CSRGraph
from oracle import oracle as O


def path3():
    return CSRGraph.from_edges(3, np.array([0, 1]), np.array([1, 2]))


def random_graph(n=40, p=0.12, seed=0, connected=True):
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(iu[0].size) < p
    u, v = iu[0][keep], iu[1][keep]
    if connected:   # add a Hamiltonian path so no node is isolated
        u = np.concatenate([u, np.arange(n - 1)])
        v = np.concatenate([v, np.arange(1, n)])
    return CSRGraph.from_edges(n, u, v)


# ------------------------------------------------------------------ PPR GD
def test_ppr_gd_hand_traced_path():
    """P3, seed 0, alpha=1/2, eps=1/100 -- traced by hand (all values dyadic):
    frontiers [0],[1],[0,2],[1],[0,2],[1]; oper = 1+2+2+2+2+2+0 = 11."""
    g = path3()
    o = O.ppr_gd(g, 0, 0.5, 0.01, local=True, trace=True)
    assert [list(t) for t in o["trace"]] == [[0], [1], [0, 2], [1], [0, 2], [1]]
    assert o["oper"] == 11
    assert np.array_equal(o["x"], [0.578125, 0.328125, 0.078125])
    assert np.array_equal(o["r"], [0.00390625, 0.0, 0.00390625])


def test_ppr_gd_threshold_is_inclusive():
    """eps = 2^-7 makes r_0 = r_2 = 2^-8 land exactly on eps*alpha*d: the
    reference enqueues on equality (eps_vec[v] <= |r_v|), giving one more
    iteration [0, 2] and oper = 13."""
    g = path3()
    o = O.ppr_gd(g, 0, 0.5, 2.0 ** -7, local=True, trace=True)
    assert [list(t) for t in o["trace"]] == [[0], [1], [0, 2], [1], [0, 2], [1], [0, 2]]
    assert o["oper"] == 13
    assert np.array_equal(o["r"], [0.0, 0.00390625, 0.0])


def test_ppr_gd_assert():
    g = path3()
    with pytest.raises(AssertionError):
        O.ppr_gd(g, 1, 0.1, 0.6, local=True)   # eps * d_s = 1.2 > 1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ppr_gd_dense_identity(seed):
    g = random_graph(seed=seed)
    s, alpha, eps = 3, 0.15, 1e-4
    xs = O.ppr_dense_solve(g, s, alpha)
    Qinv = np.linalg.inv(O.ppr_operator(g, alpha).toarray())
    for local in (True, False):
        o = O.ppr_gd(g, s, alpha, eps, local=local)
        # r is the exact residual of x, so x* - x = Q^-1 r
        assert np.abs(xs - o["x"] - Qinv @ o["r"]).max() < 1e-14
        assert np.abs(O.ppr_residual(g, s, alpha, o["x"]) - o["r"]).max() < 1e-15
        # 1^T Q = alpha 1^T: L1 error = sum(r) / alpha (r >= 0 for GD)
        assert (o["r"] >= 0).all() and (o["x"] >= 0).all()
        assert abs(np.abs(xs - o["x"]).sum() - o["r"].sum() / alpha) < 1e-14
        # termination certificate
        thr = eps * alpha * g.degree
        assert (np.abs(o["r"]) < thr).all()
    og = O.ppr_gd(g, s, alpha, eps, local=False)
    assert og["oper"] % (g.n + g.m) == 0


def test_ppr_gd_local_frontier_is_threshold_set():
    """Monotone pushes: the next frontier is exactly {v : r_v >= eps*alpha*d_v}
    after each sweep, so the frontier count of every iteration can be
    rebuilt from a dense Jacobi replay."""
    g = random_graph(n=60, seed=5)
    s, alpha, eps = 0, 0.2, 1e-3
    o = O.ppr_gd(g, s, alpha, eps, local=True, trace=True)
    A = O.adjacency(g).toarray()
    d = g.degree.astype(float)
    r = np.zeros(g.n)
    r[s] = alpha
    front = np.array([s])
    total = 0
    for t, fr in enumerate(o["trace"]):
        assert set(fr.tolist()) == set(front.tolist()), t
        total += int(d[front].sum())
        push = np.zeros(g.n)
        push[front] = (1 - alpha) * r[front] / d[front]
        r[front] = 0.0
        r = r + A @ push
        front = np.flatnonzero(r >= eps * alpha * d)
    assert front.size == 0
    assert total == o["oper"]


# ------------------------------------------------------------------ PPR Chebyshev
def test_cheby_coefficient_limit():
    """beta_t -> 2 / (1 + sqrt(1 - (1-alpha)^2)) (the optimal Chebyshev /
    SOR weight); the recurrence ppr_solver.py:636-637 restated."""
    alpha = 0.1
    d = 1 - alpha
    for _ in range(200):
        d = 1.0 / (2.0 / (1 - alpha) - d)
    beta = 2 * d / (1 - alpha)
    assert abs(beta - 2 / (1 + math.sqrt(1 - (1 - alpha) ** 2))) < 1e-14
    assert abs(beta - 1.3928644583850192) < 1e-15


@pytest.mark.parametrize("seed", [0, 3])
def test_ppr_cheby_dense(seed):
    g = random_graph(seed=seed)
    s, alpha, eps = 2, 0.15, 1e-6
    xs = O.ppr_dense_solve(g, s, alpha)
    for local in (True, False):
        o = O.ppr_cheby(g, s, alpha, eps, local=local)
        # r stays the true residual b - Q x through the three-term recurrence
        assert np.abs(O.ppr_residual(g, s, alpha, o["x"]) - o["r"]).max() < 1e-14
        Qinv = np.linalg.inv(O.ppr_operator(g, alpha).toarray())
        assert np.abs(xs - o["x"] - Qinv @ o["r"]).max() < 1e-13
        assert np.abs(xs - o["x"]).sum() < 1e-3
    og = O.ppr_cheby(g, s, alpha, eps, local=False)
    assert og["oper"] % (g.n + g.m) == 0
    assert (np.abs(og["r"]) < eps * alpha * g.degree).all()


def test_ppr_cheby_initial_queue_and_uncounted_init():
    """Star with centre 0: the initial push (not counted, ppr_solver.py:597-601)
    sends (1-alpha)*alpha/k to each leaf; the first queue is the leaves in
    adjacency order; the first counted iteration adds their degrees."""
    k = 5
    g = CSRGraph.from_edges(k + 1, np.zeros(k, dtype=np.int64), np.arange(1, k + 1))
    alpha = 0.5
    o = O.ppr_cheby(g, 0, alpha, 1e-3, local=True, trace=True)
    assert list(o["trace"][0]) == [1, 2, 3, 4, 5]
    assert o["oper"] == sum(int(g.degree[t].sum()) for t in o["trace"])


# ------------------------------------------------------------------ Katz
def test_katz_gd_dense():
    g = random_graph(seed=7)
    s = 4
    lam = float(np.abs(np.linalg.eigvalsh(O.adjacency(g).toarray())).max())
    beta = 0.5 / lam
    A = O.adjacency(g).toarray()
    M = np.eye(g.n) - beta * A
    es = np.zeros(g.n)
    es[s] = 1
    xs = np.linalg.solve(M, es) - es
    for local in (True, False):
        o = O.katz_gd(g, s, beta, 1e-9, local=local)
        # invariant (I - beta A)(x + e_s) + r = e_s
        assert np.abs(M @ (o["x"] + es) + o["r"] - es).max() < 1e-14
        assert np.abs(xs - o["x"]).sum() < 1e-5


def test_katz_cheby_dense():
    g = random_graph(seed=8)
    s = 1
    ev = np.linalg.eigvalsh(O.adjacency(g).toarray())
    beta = 0.5 / np.abs(ev).max()
    l_max, mu_min = 1 - beta * ev.min(), 1 - beta * ev.max()
    A = O.adjacency(g).toarray()
    M = np.eye(g.n) - beta * A
    es = np.zeros(g.n)
    es[s] = 1
    xs = np.linalg.solve(M, es)
    for local in (True, False):
        o = O.katz_cheby(g, s, beta, l_max, mu_min, 1e-10, local=local)
        # katz_cheby returns x including e_s (no subtraction, katz_solver.py:431)
        assert np.abs(M @ o["x"] + o["r"] - es).max() < 1e-13
        assert np.abs(xs - o["x"]).sum() < 1e-6


# ------------------------------------------------------------------ heat kernel
def test_estimate_n_closed_form():
    """estimate_n multiplies its bound by a constant factor q built from the
    initial k (hk_solver.py:43), so N = k + ceil(log(eps e^t / c0) / log q)."""
    for t in (1.0, 2.5, 5.0, 10.0):
        for eps in (1e-3, 1e-6, 5e-8):
            k = math.ceil(t - 1)
            c0 = t ** (k + 1) / math.factorial(k + 1) * ((k + 2) / (k + 2 - t))
            target = eps * math.exp(t)
            if c0 < target:
                want = k
            else:
                q = ((k + 2 - t) * t / (k + 2) ** 2) * ((k + 3) / (k + 3 - t))
                want = k + math.ceil(math.log(target / c0) / math.log(q) - 1e-12)
                if c0 * q ** (want - k) >= target:   # guard the rounding of the log
                    want += 1
            assert O.estimate_n(t, eps) == want, (t, eps)
    assert [O.estimate_n(5.0, e) for e in (1e-6, 1e-7, 5e-8)] == [24, 27, 28]


def test_estimate_coefficient():
    for t in (1.0, 3.0, 5.0):
        for k in range(0, 12):
            want = t ** (k + 1) / math.factorial(k + 1)
            assert abs(O.estimate_coefficient(t, k) - want) <= 1e-13 * want


def test_hk_taylor_vs_expm():
    import scipy.linalg as sla
    g = random_graph(n=30, seed=9)
    s, t = 0, 3.0
    A = O.adjacency(g).toarray()
    P = A / g.degree[None, :]
    es = np.zeros(g.n)
    es[s] = 1
    h = sla.expm(t * (P - np.eye(g.n))) @ es
    o = O.hk_taylor(g, s, t, 1e-10)
    N = o["est_n"]
    tail = sum(t ** k / math.factorial(k) for k in range(N + 1, N + 60)) * math.exp(-t)
    assert np.abs(o["x"] - h).sum() <= tail * 1.0001 + 1e-14


def test_hk_push_levels_exact_when_nothing_dropped():
    """With a threshold below every residual, every level pushes everything:
    x = e^{-t} [sum_{k<N} t^k/k! P^k e_s + t^{N-1}/(N-1)! P^N e_s]
    (the last level adds r/d without the t/N factor, hk_solver.py:254-256)."""
    g = random_graph(n=25, seed=10)
    s, t, N = 2, 2.0, 9
    o = O.hk_push(g, s, t, 1e-300, local=True, n_terms=N)
    A = O.adjacency(g).toarray()
    P = A / g.degree[None, :]
    v = np.zeros(g.n)
    v[s] = 1
    acc = np.zeros(g.n)
    vol = 0
    for k in range(N):
        acc += t ** k / math.factorial(k) * v
        vol += int(g.degree[v > 0].sum())   # every node with residual at level k is pushed
        v = P @ v
    acc += t ** (N - 1) / math.factorial(N - 1) * v
    assert np.abs(o["x"] - acc / math.exp(t)).max() < 1e-15
    assert o["est_n"] == N
    assert o["oper"] == vol


# ------------------------------------------------------------------ InstantGNN
def _ignn_operator(indptr, indices, degree, beta):
    n = degree.size
    M = np.zeros((n, n))
    for u in range(n):
        for e in range(indptr[u], indptr[u] + degree[u]):
            v = indices[e]
            M[v, u] += 1.0 / (degree[v] ** beta * degree[u] ** (1 - beta))
    return M


def test_ignn_column_invariant():
    g = random_graph(n=35, seed=11)
    alpha, beta, eps = 0.1, 0.5, 1e-6
    rng = np.random.default_rng(0)
    X = rng.standard_normal(g.n)
    p, r, _ = O.ignn_column(g.indptr, g.indices, g.degree, alpha, eps, beta, X)
    M = _ignn_operator(g.indptr, g.indices, g.degree, beta)
    K = alpha * np.linalg.inv(np.eye(g.n) - (1 - alpha) * M)
    pi = K @ X
    # p + alpha (I - (1-alpha) M)^-1 r = pi (p accumulated in float32)
    assert np.abs(p.astype(np.float64) + K @ r - pi).max() < 1e-5
    assert (np.abs(r) < eps * g.degree ** (1 - beta)).all()


def test_ignn_isolated_node_takes_its_feature():
    g = random_graph(n=20, seed=12)
    # append an isolated node
    indptr = np.concatenate([g.indptr, [g.indptr[-1]]])
    degree = np.concatenate([g.degree, [0]]).astype(np.int32)
    X = np.linspace(-1, 1, g.n + 1)
    p, r, _ = O.ignn_column(indptr, g.indices, degree, 0.1, 1e-6, 0.5, X)
    assert p[-1] == np.float32(X[-1]) and r[-1] == 0.0


def test_ignn_update_keeps_invariant():
    """Insert edges into a slack CSR, apply instantGNNUpdateBatch: the
    invariant p + K' r = K' X holds on the new graph."""
    full = random_graph(n=30, p=0.2, seed=13)
    rng = np.random.default_rng(1)
    rows = np.repeat(np.arange(full.n), full.degree)
    cols = full.indices
    up = np.flatnonzero(rows < cols)
    held = rng.choice(up, size=12, replace=False)
    held_set = set((int(rows[i]), int(cols[i])) for i in held)
    indptr = full.indptr.copy()
    indices = np.full(full.m, full.n, dtype=np.int32)
    degree = np.zeros(full.n, dtype=np.int32)
    for u in range(full.n):
        for e in range(full.indptr[u], full.indptr[u + 1]):
            v = int(full.indices[e])
            if (min(u, v), max(u, v)) in held_set:
                continue
            indices[indptr[u] + degree[u]] = v
            degree[u] += 1
    alpha, beta, eps = 0.1, 0.5, 1e-7
    X = rng.standard_normal(full.n)
    p32, r, _ = O.ignn_column(indptr, indices, degree, alpha, eps, beta, X)
    p = p32.astype(np.float64)
    old_degree = degree.copy()
    us = np.array([a for a, b in held_set], dtype=np.int32)
    vs = np.array([b for a, b in held_set], dtype=np.int32)
    keys, nbr_off, nbr = O.update_graph(indptr, indices, degree, us, vs)
    # the live graph is now the full graph
    for u in range(full.n):
        assert set(indices[indptr[u]:indptr[u] + degree[u]].tolist()) == set(
            full.indices[full.indptr[u]:full.indptr[u + 1]].tolist())
    O.ignn_update_column(indptr, indices, degree, old_degree, p, r, X, alpha, eps, beta, keys, nbr_off, nbr)
    M = _ignn_operator(indptr, indices, degree, beta)
    K = alpha * np.linalg.inv(np.eye(full.n) - (1 - alpha) * M)
    assert np.abs(p + K @ r - K @ X).max() < 1e-5
    assert (np.abs(r) < eps * degree ** (1 - beta)).all()


def test_update_graph_skips_existing_and_orders_keys():
    g = random_graph(n=10, p=0.0, seed=0)   # path 0-1-...-9
    indptr = np.zeros(11, dtype=np.int64)
    cap = 4
    indptr[1:] = np.cumsum(np.full(10, cap))
    indices = np.full(10 * cap, 10, dtype=np.int32)
    degree = np.zeros(10, dtype=np.int32)
    for u in range(10):
        row = g.indices[g.indptr[u]:g.indptr[u + 1]]
        indices[indptr[u]:indptr[u] + row.size] = row
        degree[u] = row.size
    us = np.array([5, 0, 3, 0], dtype=np.int32)
    vs = np.array([7, 1, 9, 1], dtype=np.int32)   # (0,1) exists; (0,1) repeated
    keys, off, nbr = O.update_graph(indptr, indices, degree, us, vs)
    assert keys.tolist() == [5, 7, 3, 9]
    assert [nbr[off[k]:off[k + 1]].tolist() for k in range(4)] == [[7], [5], [9], [3]]
    assert degree[5] == 3 and degree[7] == 3 and degree[0] == 1


def test_near_threshold_count_star():
    """Star with centre 0 and 4 leaves, alpha = 1/2, eps = 1/8: the centre's
    push gives every leaf r = (1-alpha) alpha / 4 = 1/16, exactly its
    threshold eps alpha d = 1/16 (dyadic, no rounding), so the first
    iteration's four enqueue tests sit on the threshold and are counted; the
    leaves still join the frontier (the inclusive test eps_vec <= |r|)."""
    k = 4
    g = CSRGraph.from_edges(k + 1, np.zeros(k, dtype=np.int64), np.arange(1, k + 1))
    o = O.ppr_gd(g, 0, 0.5, 0.125, local=True, trace=True)
    assert o["near_thr"] >= k
    assert sorted(o["trace"][1].tolist()) == [1, 2, 3, 4]
    # a threshold well away from every residual counts nothing
    o2 = O.ppr_gd(g, 0, 0.5, 0.1, local=True)
    assert o2["near_thr"] == 0
