"""Pins of the SOR oracle functions (CPU only).

* ``ppr_sor`` restates the reference's SOR / Gauss-Southwell push
  (ppr_solver.py:331-461).  Pinned by a hand trace (FIFO with the iteration
  flag, the float op count) and, for the global sweeps, by the matrix form of
  SOR: with Q = I + L + U (L strictly lower in the sweep order) one sweep is
  x <- (I + w L)^-1 [((1 - w) I - w U) x + w b], iterated with dense numpy
  until the reference's stopping test holds.
* ``colour`` / ``ppr_sor_coloured`` restate the GPU's coloured variant.
  Pinned by properness and first-fit minimality of the colouring, a hand
  trace, and the same matrix form with L taken in colour order.
* Both end on the certificate |r_u| < eps*alpha*d_u with r == b - Q x.
"""
import numpy as np
import pytest

from flsg.graph import CSRGraph
from oracle import oracle as O

OMEGA_OPT = 1.3928644583850192   # demo.py:31-35 at alpha = 0.1


def path3():
    return CSRGraph.from_edges(3, np.array([0, 1]), np.array([1, 2]))


def random_graph(n=40, p=0.12, seed=0):
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(iu[0].size) < p
    u = np.concatenate([iu[0][keep], np.arange(n - 1)])
    v = np.concatenate([iu[1][keep], np.arange(1, n)])
    return CSRGraph.from_edges(n, u, v)


def rows(g):
    return np.repeat(np.arange(g.n), np.asarray(g.degree, dtype=np.int64))


def sor_matrix_form(g, s, alpha, eps, omega, order_key):
    """Dense SOR on Q x = alpha e_s in the sweep order given by order_key
    (row u is updated before row v when order_key[u] < order_key[v]; equal
    keys are never adjacent).  Returns (x, sweeps)."""
    Q = O.ppr_operator(g, alpha).toarray()
    n = g.n
    b = np.zeros(n)
    b[s] = alpha
    key = np.asarray(order_key)
    before = key[None, :] < key[:, None]          # column v is updated before row u
    L = np.where(before, Q, 0.0)
    np.fill_diagonal(L, 0.0)
    U = Q - np.eye(n) - L
    M = np.eye(n) + omega * L
    N = (1.0 - omega) * np.eye(n) - omega * U
    thr = eps * alpha * np.asarray(g.degree, dtype=np.float64)
    x = np.zeros(n)
    for sweeps in range(1, 100000):
        x = np.linalg.solve(M, N @ x + omega * b)
        r = b - Q @ x
        if not np.any(thr <= np.abs(r)):
            return x, sweeps
    raise RuntimeError("no convergence")


# ------------------------------------------------------------------ reference ppr_sor
def test_ppr_sor_hand_traced_path():
    """P3, seed 0, alpha = 1/2, eps = 1/100, omega = 1 (Gauss-Southwell).
    Thresholds (1/200) d = [.005, .01, .005].  Pops: flag, 0 (r0 = .5 ->
    r1 = .25), flag, 1 (-> r0 = r2 = .0625), flag, 0 (-> r1 = .03125),
    2 (-> r1 = .0625), flag, 1 (-> r0 = r2 = .015625), flag, 0 (-> r1 =
    .0078125 < .01, not queued), 2 (-> r1 = .015625, queued), flag, 1
    (-> r0 = r2 = .00390625 < .005), flag with the queue empty: stop.
    oper = 1+2+1+1+2+1+1+2 = 11 (a float), 6 iterations."""
    o = O.ppr_sor(path3(), 0, 0.5, 0.01, 1.0, local=True)
    assert o["oper"] == 11.0 and o["iters"] == 6
    assert np.array_equal(o["x"], [0.578125, 0.328125, 0.078125])
    assert np.array_equal(o["r"], [0.00390625, 0.0, 0.00390625])


def test_ppr_sor_overrelaxed_step_by_hand():
    """omega = 3/2 on P3 (alpha = 1/2, eps = 1/100): the first push moves
    res = .75 out of r0 = .5, leaving r0 = -.25, which is still active
    (|r0| >= .005) and is re-queued after the push (ppr_solver.py:415-419)."""
    o = O.ppr_sor(path3(), 0, 0.5, 0.01, 1.5, local=True)
    # first pops by hand: 0 -> r0 = -.25, r1 = .375 (queue 1, 0); both active
    assert o["oper"] >= 1 + 2 + 1
    r = O.ppr_residual(path3(), 0, 0.5, o["x"])
    assert np.abs(r - o["r"]).max() < 1e-15
    assert np.all(np.abs(r) < 0.005 * np.array([1, 2, 1]))


@pytest.mark.parametrize("omega", [1.0, OMEGA_OPT])
@pytest.mark.parametrize("seed", [0, 1])
def test_ppr_sor_global_matches_matrix_form(seed, omega):
    g = random_graph(seed=seed)
    s = 3
    o = O.ppr_sor(g, s, 0.1, 1e-4, omega, local=False)
    x, sweeps = sor_matrix_form(g, s, 0.1, 1e-4, omega, np.arange(g.n))
    assert o["iters"] == sweeps
    assert o["oper"] == sweeps * (g.n + int(g.degree.sum()))
    assert np.abs(o["x"] - x).sum() <= 1e-12 * np.abs(x).sum()


def test_jacobi_order_would_fail_the_matrix_form():
    """The matrix-form pin distinguishes sweep orders: natural order and the
    reversed order give different iterates on the same graph."""
    g = random_graph(seed=2)
    x1, k1 = sor_matrix_form(g, 3, 0.1, 1e-4, OMEGA_OPT, np.arange(g.n))
    x2, k2 = sor_matrix_form(g, 3, 0.1, 1e-4, OMEGA_OPT, -np.arange(g.n))
    o = O.ppr_sor(g, 3, 0.1, 1e-4, OMEGA_OPT, local=False)
    assert np.abs(o["x"] - x1).sum() < 1e-12 < np.abs(o["x"] - x2).sum()


@pytest.mark.parametrize("omega", [0.7, 1.0, OMEGA_OPT, 1.8])
@pytest.mark.parametrize("local", [True, False])
def test_ppr_sor_certificate_and_identity(local, omega):
    g = random_graph(n=60, p=0.08, seed=5)
    s = 7
    o = O.ppr_sor(g, s, 0.1, 1e-5, omega, local=local)
    r = O.ppr_residual(g, s, 0.1, o["x"])
    assert np.abs(r - o["r"]).max() < 1e-15
    thr = 1e-6 * g.degree
    assert np.all(np.abs(r) < thr)
    # x* - x = Q^-1 r exactly
    xs = O.ppr_dense_solve(g, s, 0.1)
    Q = O.ppr_operator(g, 0.1).toarray()
    assert np.abs((xs - o["x"]) - np.linalg.solve(Q, r)).max() < 1e-14


def test_ppr_sor_assert():
    with pytest.raises(AssertionError):
        O.ppr_sor(path3(), 1, 0.1, 0.6, 1.0, local=True)


# ------------------------------------------------------------------ colouring
def check_colouring(g, col, k):
    r = rows(g)
    c = np.asarray(col)
    assert c.min() >= 0 and c.max() == k - 1
    loops = r == g.indices
    assert not np.any((c[r] == c[g.indices]) & ~loops), "adjacent nodes share a colour"
    # first fit: a node of colour c has neighbours of every colour below c
    for u in range(g.n):
        nb = set(c[g.indices[g.indptr[u]:g.indptr[u] + g.degree[u]]].tolist()) - {c[u]}
        assert set(range(c[u])) <= nb, u


def test_colour_path_by_hand():
    """P3: node 1 (degree 2) has the highest key and takes colour 0 in round
    1; nodes 0 and 2 take colour 1 in round 2."""
    col, k, rounds = O.colour(path3())
    assert list(col) == [1, 0, 1] and k == 2 and rounds == 2


def test_colour_clique_and_isolated():
    """K5 plus two isolated nodes: five colours, one per round (keys are
    distinct); isolated nodes take colour 0 without a round."""
    u, v = np.triu_indices(5, 1)
    g = CSRGraph.from_edges(7, u, v)
    col, k, rounds = O.colour(g)
    assert sorted(col[:5]) == [0, 1, 2, 3, 4] and list(col[5:]) == [0, 0]
    assert k == 5 and rounds == 5


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_colour_proper_first_fit(seed):
    g = random_graph(n=80, p=0.1, seed=seed)
    col, k, rounds = O.colour(g)
    check_colouring(g, col, k)
    assert rounds >= k


def test_colour_powerlaw_small(small_graph):
    col, k, rounds = O.colour(small_graph)
    check_colouring(small_graph, col, k)


# ------------------------------------------------------------------ coloured SOR
def test_coloured_sor_hand_traced_path():
    """P3 with colours [1, 0, 1], omega = 1, same parameters as the FIFO
    trace: the frontier passes are [0], [1], [0, 2], [1], [0, 2], [1] and the
    values equal the FIFO run's (every push here sees the same residuals)."""
    g = path3()
    col, k, _ = O.colour(g)
    o = O.ppr_sor_coloured(g, 0, 0.5, 0.01, 1.0, col, k, local=True)
    assert o["oper"] == 11 and o["iters"] == 6 and o["nodes"] == 8
    assert np.array_equal(o["x"], [0.578125, 0.328125, 0.078125])
    assert np.array_equal(o["r"], [0.00390625, 0.0, 0.00390625])


@pytest.mark.parametrize("omega", [1.0, OMEGA_OPT])
@pytest.mark.parametrize("seed", [0, 3])
def test_coloured_sor_global_matches_matrix_form(seed, omega):
    g = random_graph(seed=seed)
    col, k, _ = O.colour(g)
    o = O.ppr_sor_coloured(g, 3, 0.1, 1e-4, omega, col, k, local=False)
    x, sweeps = sor_matrix_form(g, 3, 0.1, 1e-4, omega, col)
    assert o["iters"] == sweeps
    assert np.abs(o["x"] - x).sum() <= 1e-12 * np.abs(x).sum()


@pytest.mark.parametrize("omega", [1.0, OMEGA_OPT, 1.8])
@pytest.mark.parametrize("local", [True, False])
def test_coloured_sor_certificate_and_reference_distance(local, omega, small_graph):
    g = small_graph
    col, k, _ = O.colour(g)
    s = 17 if g.degree[17] > 0 else int(np.argmax(g.degree > 0))
    a = O.ppr_sor(g, s, 0.1, 1e-6, omega, local=local)
    b = O.ppr_sor_coloured(g, s, 0.1, 1e-6, omega, col, k, local=local)
    live = g.degree > 0
    thr = 1e-7 * g.degree
    for o in (a, b):
        r = O.ppr_residual(g, s, 0.1, o["x"])
        assert np.abs(r - o["r"]).max() < 1e-15
        assert np.all(np.abs(r[live]) < thr[live])
    # both certified: ||x_a - x_b||_1 <= (||r_a||_1 + ||r_b||_1) / alpha
    bound = (np.abs(a["r"]).sum() + np.abs(b["r"]).sum()) / 0.1
    assert np.abs(a["x"] - b["x"]).sum() <= bound


# ------------------------------------------------------------------ Katz SOR
def katz_matrix_form(g, s, beta, eps, omega, order_key):
    """Dense SOR on (I - beta A) y = e_s in the given sweep order, stopped by
    the reference's test (eps*d_v <= |r_v| for no live v).  Returns (y - e_s,
    sweeps): katz_sor's x convention."""
    A = O.adjacency(g).toarray()
    n = g.n
    Q = np.eye(n) - beta * A
    b = np.zeros(n)
    b[s] = 1.0
    key = np.asarray(order_key)
    L = np.where(key[None, :] < key[:, None], Q, 0.0)
    np.fill_diagonal(L, 0.0)
    U = Q - np.eye(n) - L
    M = np.eye(n) + omega * L
    N = (1.0 - omega) * np.eye(n) - omega * U
    thr = eps * np.asarray(g.degree, dtype=np.float64)
    y = np.zeros(n)
    for sweeps in range(1, 100000):
        y = np.linalg.solve(M, N @ y + omega * b)
        if not np.any(thr <= np.abs(b - Q @ y)):
            y[s] -= 1.0
            return y, sweeps
    raise RuntimeError("no convergence")


def test_katz_sor_hand_traced_path():
    """P3, seed 0, beta = 1/4, eps = 1/16, omega = 1; thresholds eps*d =
    [1/16, 1/8, 1/16].  Pops: flag, 0 (x0 = 1, r1 = 1/4), flag, 1 (x1 = 1/4,
    r0 = r2 = 1/16, both queued on equality), flag, 0 (r1 = 1/64), 2
    (r1 = 1/32), flag with the queue empty.  oper = 1+2+1+1 = 5, 3
    iterations, x_s -= 1 at the end."""
    o = O.katz_sor(path3(), 0, 0.25, 1.0 / 16, 1.0, local=True)
    assert o["oper"] == 5.0 and o["iters"] == 3
    assert np.array_equal(o["x"], [0.0625, 0.25, 0.0625])
    assert np.array_equal(o["r"], [0.0, 0.03125, 0.0])


@pytest.mark.parametrize("omega", [1.0, 1.3])
@pytest.mark.parametrize("seed", [0, 1])
def test_katz_sor_global_matches_matrix_form(seed, omega):
    g = random_graph(seed=seed)
    beta = 1.0 / (g.degree.max() + 1)
    o = O.katz_sor(g, 3, beta, 1e-6, omega, local=False)
    x, sweeps = katz_matrix_form(g, 3, beta, 1e-6, omega, np.arange(g.n))
    assert o["iters"] == sweeps
    assert o["oper"] == sweeps * (g.n + int(g.degree.sum()))
    assert np.abs(o["x"] - x).sum() <= 1e-12 * np.abs(x).sum()


@pytest.mark.parametrize("omega", [1.0, 1.3])
@pytest.mark.parametrize("local", [True, False])
def test_katz_sor_certificate(local, omega):
    g = random_graph(n=60, p=0.08, seed=5)
    beta = 1.0 / (g.degree.max() + 1)
    s = 7
    o = O.katz_sor(g, s, beta, 1e-7, omega, local=local)
    A = O.adjacency(g)
    y = o["x"].copy()
    y[s] += 1.0
    r = -(y - beta * (A @ y))
    r[s] += 1.0
    assert np.abs(r - o["r"]).max() < 1e-15
    assert np.all(np.abs(r) < 1e-7 * g.degree)


@pytest.mark.parametrize("omega", [1.0, 1.3])
def test_coloured_katz_global_matches_matrix_form(omega):
    g = random_graph(seed=4)
    beta = 1.0 / (g.degree.max() + 1)
    col, k, _ = O.colour(g)
    o = O.katz_sor_coloured(g, 3, beta, 1e-6, omega, col, k, local=False)
    x, sweeps = katz_matrix_form(g, 3, beta, 1e-6, omega, col)
    assert o["iters"] == sweeps
    assert np.abs(o["x"] - x).sum() <= 1e-12 * np.abs(x).sum()


def test_coloured_katz_hand_traced_path():
    """Colours [1, 0, 1]: iteration 1 pushes 0 (r1 = 1/4), iteration 2 pushes
    1 (r0 = r2 = 1/16 = their thresholds), iteration 3 pushes 0 and 2 (r1 =
    1/32 < 1/8): the same pushes as the FIFO run."""
    g = path3()
    col, k, _ = O.colour(g)
    o = O.katz_sor_coloured(g, 0, 0.25, 1.0 / 16, 1.0, col, k, local=True)
    assert o["oper"] == 5 and o["iters"] == 3
    assert np.array_equal(o["x"], [0.0625, 0.25, 0.0625])
    assert np.array_equal(o["r"], [0.0, 0.03125, 0.0])


# ------------------------------------------------------------------ InstantGNN SOR (instantGNNSorPush)
def ignn_graph(n=30, p=0.15, seed=21, loops=(0, 3), isolated=True):
    """Slack CSR (capacity 2 spare slots per row) of a random graph with
    self-loops at ``loops``, a self-loop-only node and an isolated node."""
    g = random_graph(n=n, p=p, seed=seed)
    rows = [list(g.indices[g.indptr[u]:g.indptr[u + 1]]) for u in range(g.n)]
    for u in loops:
        rows[u] = sorted(rows[u] + [u])
    rows.append([g.n])          # node n: only a self-loop
    if isolated:
        rows.append([])         # node n+1: isolated
    N = len(rows)
    cap = max(len(r) for r in rows) + 2
    indptr = np.arange(0, cap * (N + 1), cap, dtype=np.int64)
    indices = np.full(cap * N, N, dtype=np.int32)
    degree = np.zeros(N, dtype=np.int32)
    for u, rw in enumerate(rows):
        indices[cap * u:cap * u + len(rw)] = rw
        degree[u] = len(rw)
    return indptr, indices, degree


def ignn_operator(indptr, indices, degree, beta):
    n = degree.size
    M = np.zeros((n, n))
    for u in range(n):
        for v in indices[indptr[u]:indptr[u] + degree[u]]:
            M[v, u] += 1.0 / (degree[v] ** beta * degree[u] ** (1 - beta))
    return M


def check_ignn_column(indptr, indices, degree, X, p, r, alpha, eps, beta):
    n = degree.size
    live = ~((degree == 0) | ((degree == 1) & (O.self_loops(indptr, indices, degree) == 1)))
    M = ignn_operator(indptr, indices, degree, beta)
    K = alpha * np.linalg.inv(np.eye(n) - (1 - alpha) * M)
    # invariant p + K r = K X on the pushed part; absorbing nodes take p = X
    assert np.abs((p + K @ r - K @ X)[live]).max() < 1e-13
    assert np.array_equal(p[~live], X[~live]) and np.all(r[~live] == 0)
    ev = O.ignn_eps_vec(degree, eps, beta)
    assert np.all((np.abs(r) < ev) | ((ev == 0) & (r == 0)))   # isolated nodes: threshold 0, r = 0


@pytest.mark.parametrize("omega", [1.0, 1.3])
def test_ignn_sor_column_invariant(omega):
    ip, ix, dg = ignn_graph()
    alpha, beta, eps = 0.1, 0.5, 1e-6
    X = np.random.default_rng(3).standard_normal(dg.size)
    p, r = np.zeros(dg.size), X.copy()
    O.ignn_sor_column(ip, ix, dg, alpha, eps, omega, beta, p, r)
    check_ignn_column(ip, ix, dg, X, p, r, alpha, eps, beta)


def test_ignn_sor_hand_traced():
    """Node 0 with a self-loop and an edge to node 1 (degrees 2, 1), alpha =
    1/2, beta = 0 (eps_vec = eps d = [1/4, 1/8]), eps = 1/8, omega = 1,
    r = [1, 0].  Pop 0: p0 = 1/2, push (1/2)(1)/2 = 1/4 to 0 itself (r0 =
    1/4, queued on equality) and to 1 (queued).  Pop 0: p0 = 5/8, push 1/16
    -> r0 = 1/16 (not queued), r1 = 5/16.  Pop 1: p1 = 5/32, push 5/32 to 0
    -> r0 = 7/32 < 1/4.  Queue empty."""
    ip = np.array([0, 2, 3], dtype=np.int64)
    ix = np.array([0, 1, 0], dtype=np.int32)
    dg = np.array([2, 1], dtype=np.int32)
    p, r = np.zeros(2), np.array([1.0, 0.0])
    oper = O.ignn_sor_column(ip, ix, dg, 0.5, 0.125, 1.0, 0.0, p, r)
    assert np.array_equal(p, [0.625, 0.15625]) and np.array_equal(r, [0.21875, 0.0]) and oper == 5
    check_ignn_column(ip, ix, dg, np.array([1.0, 0.0]), p, r, 0.5, 0.125, 0.0)


@pytest.mark.parametrize("omega", [1.0, 1.3])
def test_ignn_sor_coloured_invariant(omega):
    from flsg.graph import CSRGraph
    ip, ix, dg = ignn_graph(seed=22)
    alpha, beta, eps = 0.1, 0.5, 1e-6
    X = np.random.default_rng(4).standard_normal(dg.size)
    live_rows = CSRGraph(n=dg.size, indptr=ip, indices=ix, degree=dg)
    col, k, _ = O.colour(live_rows)
    p, r = np.zeros(dg.size), X.copy()
    ev = O.ignn_eps_vec(dg, eps, beta)
    nodes = np.flatnonzero((np.abs(r) >= ev) & (r != 0))
    o = O.ignn_sor_coloured(ip, ix, dg, col, k, alpha, eps, omega, beta, p, r, nodes)
    assert o["iters"] > 0
    check_ignn_column(ip, ix, dg, X, p, r, alpha, eps, beta)
