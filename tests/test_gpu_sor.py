"""GPU parity of the graph-coloured SOR push (ppr_sor, ppr_solver.py:331-461).

Bars (written in each test):
  * colouring: identical to the oracle's restatement, node for node (integer
    work, bit-exact), and proper;
  * against the oracle's sequential restatement of the same coloured sweep
    (same colouring): iteration count and operation count equal, FP64
    estimate within 1e-12 relative L1 (the GPU adds one colour's pushes in
    atomic order, so sums may round differently);
  * certificate: |r_u| < eps*alpha*d_u for every live u, with r recomputed in
    FP64 on the host from the returned x, and the returned r equal to it
    within 1e-15 absolute on the 10,000-node graph and 1e-14 on the
    2.39M-node one (the running residual accumulates a rounding error per
    add; 1e-14 is 1e-6 of the smallest threshold, 1e-8);
  * against the reference-order SOR (the oracle's ppr_sor, FIFO or natural
    order): both certified, so ||x_gpu - x_ref||_1 <= (||r_gpu||_1 +
    ||r_ref||_1) / alpha.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from flsg import engine, graph  # noqa: E402
from flsg.engine import FlsgError  # noqa: E402
from flsg.graph import CSRGraph  # noqa: E402
from oracle import oracle as O  # noqa: E402

ALPHA = 0.1
OMEGA_OPT = 1.3928644583850192
REL_L1 = 1e-12


@pytest.fixture(scope="module")
def dev_small(small_graph):
    return engine.DeviceGraph(small_graph)


def _certified(g, s, x, r_ret, eps, tol=1e-15):
    r = O.ppr_residual(g, s, ALPHA, x)
    live = g.degree > 0
    thr = eps * ALPHA * g.degree
    assert np.abs(r - r_ret).max() < tol
    worst = float((np.abs(r[live]) / thr[live]).max())
    assert worst < 1.0, worst
    return r


def test_colouring_equals_oracle_small(small_graph, dev_small):
    col, k, rounds = engine.colouring(dev_small)
    oc, ok, orounds = O.colour(small_graph)
    assert np.array_equal(col, oc) and k == ok and rounds == orounds


def test_colouring_relabelled_graph(small_graph):
    """A degree-relabelled device graph colours its own ids; mapped back the
    colouring is still proper."""
    dg = engine.DeviceGraph(small_graph, relabel=True)
    col, k, _ = engine.colouring(dg)
    g = small_graph
    rows = np.repeat(np.arange(g.n), g.degree)
    assert not np.any(col[rows] == col[g.indices])
    assert col.max() == k - 1
    dg.close()


@pytest.mark.parametrize("omega", [1.0, OMEGA_OPT])
@pytest.mark.parametrize("local", [True, False])
def test_sor_parity_small(small_graph, dev_small, local, omega):
    g = small_graph
    col, k, _ = engine.colouring(dev_small)
    eps = 1e-6
    for s in graph.uniform_seeds(g, 6, seed=3):
        s = int(s)
        out = engine.sor_solve(dev_small, ALPHA, eps, omega, s, local=local, max_iters=5000)
        x = out.x.cpu().numpy()
        oc = O.ppr_sor_coloured(g, s, ALPHA, eps, omega, col, k, local=local)
        res = out.result
        assert res["iters"] == oc["iters"], (s, res["iters"], oc["iters"])
        assert res["oper"] == oc["oper"], (s, res["oper"], oc["oper"])
        if local:
            assert res["nodes"] == oc["nodes"]
        rel = np.abs(x - oc["x"]).sum() / np.abs(oc["x"]).sum()
        assert rel <= REL_L1, (s, rel)
        r = _certified(g, s, x, out.r.cpu().numpy(), eps)
        ref = O.ppr_sor(g, s, ALPHA, eps, omega, local=local)
        bound = (np.abs(r).sum() + np.abs(ref["r"]).sum()) / ALPHA
        assert np.abs(x - ref["x"]).sum() <= bound


def test_sor_overrelaxation_range(small_graph, dev_small):
    """omega = 0.6 and 1.9 (inside (0, 2)) converge to the certificate."""
    g = small_graph
    s = int(graph.uniform_seeds(g, 1, seed=9)[0])
    for omega in (0.6, 1.9):
        out = engine.sor_solve(dev_small, ALPHA, 1e-6, omega, s, local=True, max_iters=5000)
        _certified(g, s, out.x.cpu().numpy(), out.r.cpu().numpy(), 1e-6)


def test_sor_rejects_bad_arguments(small_graph, dev_small):
    s = int(graph.uniform_seeds(small_graph, 1, seed=0)[0])
    for omega in (0.0, 2.0, -1.0):
        with pytest.raises(FlsgError):
            engine.sor_solve(dev_small, ALPHA, 1e-6, omega, s, max_iters=5000)
    # path 0-1-2-3 with a self-loop on 2 (from_edges would drop it)
    loops = CSRGraph(n=4, indptr=np.array([0, 1, 3, 6, 7], dtype=np.int64),
                     indices=np.array([1, 0, 2, 1, 2, 3, 2], dtype=np.int32),
                     degree=np.array([1, 2, 3, 1], dtype=np.int32))
    dg = engine.DeviceGraph(loops)
    with pytest.raises(FlsgError):
        engine.sor_solve(dg, ALPHA, 1e-3, 1.0, 0, max_iters=5000)
    dg.close()


def test_sor_dropin_tuple(small_graph):
    import ppr_solver
    g = small_graph
    s = int(graph.uniform_seeds(g, 1, seed=4)[0])
    x, r, oper, t = ppr_solver.ppr_sor(g.n, g.indptr, g.indices, g.degree, s, ALPHA, 1e-6, OMEGA_OPT, True, None)
    assert isinstance(oper, float) and oper > 0 and t > 0
    _certified(g, s, x, r, 1e-6)
    out = ppr_solver.solve_a_graph_omega(g.n, g.indptr, g.indices, g.degree, ALPHA, 1e-6, OMEGA_OPT, "synthetic",
                                         np.array([s], dtype=np.int32))
    assert len(out) == 13 and out[1][0] == int(oper) and out[5][0] > 0 and out[3][0] == 0
    with pytest.raises(AssertionError):
        hub = int(np.argmax(g.degree))
        ppr_solver.ppr_sor(g.n, g.indptr, g.indices, g.degree, hub, ALPHA, 1.0, 1.0, True, None)


def test_colouring_follows_edge_inserts():
    """After an edge insert the handle recolours: the new edge's two ends,
    which shared a colour before, differ afterwards."""
    import ctypes as C

    import torch
    n, cap = 6, 4
    adj = {u: [] for u in range(n)}
    for u in range(n - 1):   # path 0-1-2-3-4-5
        adj[u].append(u + 1)
        adj[u + 1].append(u)
    indptr = np.arange(0, cap * (n + 1), cap, dtype=np.int64)
    indices = np.full(cap * n, n, dtype=np.int32)
    degree = np.zeros(n, dtype=np.int32)
    for u in range(n):
        row = sorted(adj[u])
        indices[cap * u:cap * u + len(row)] = row
        degree[u] = len(row)
    none = np.zeros((0, 0), dtype=np.int32)
    sg = graph.SlackGraph(n=n, indptr=indptr, indices=indices, degree=degree, us=none, vs=none)
    dg = engine.DeviceGraph(sg)
    col0, _, _ = engine.colouring(dg)
    pairs = [(a, b) for a in range(n) for b in range(a + 2, n) if col0[a] == col0[b]]
    assert pairs
    a, b = pairs[0]
    us = torch.tensor([a], dtype=torch.int32, device=dg.device)
    vs = torch.tensor([b], dtype=torch.int32, device=dg.device)
    n_ins = C.c_int64(0)
    rc = dg.lib.flsg_graph_insert_edges(dg.handle, us.data_ptr(), vs.data_ptr(), 1, None, None, None,
                                        C.byref(n_ins), None)
    assert rc == 0 and n_ins.value == 1
    col1, _, _ = engine.colouring(dg)
    assert col1[a] != col1[b]
    dg.close()


@pytest.fixture(scope="module")
def dev_large(large_graph):
    return engine.DeviceGraph(large_graph)


def test_colouring_equals_oracle_large(large_graph, dev_large):
    col, k, rounds = engine.colouring(dev_large)
    oc, ok, orounds = O.colour(large_graph)
    assert np.array_equal(col, oc) and k == ok and rounds == orounds
    print(f"colours {k} rounds {rounds}")


@pytest.mark.parametrize("omega", [1.0, OMEGA_OPT])
def test_sor_local_parity_large(large_graph, dev_large, omega):
    """Full size (2.39M nodes, eps = 1e-7): coloured local SOR against its
    sequential restatement and the certificate."""
    g = large_graph
    col, k, _ = engine.colouring(dev_large)
    eps = 1e-7
    for s in graph.uniform_seeds(g, 2, seed=11):
        s = int(s)
        out = engine.sor_solve(dev_large, ALPHA, eps, omega, s, local=True, max_iters=5000)
        x = out.x.cpu().numpy()
        oc = O.ppr_sor_coloured(g, s, ALPHA, eps, omega, col, k, local=True)
        res = out.result
        rel = np.abs(x - oc["x"]).sum() / np.abs(oc["x"]).sum()
        print(f"seed {s} omega {omega}: iters {res['iters']}/{oc['iters']} oper {res['oper']}/{oc['oper']} "
              f"rel L1 {rel:.2e} ms {res['ms']:.2f}")
        _certified(g, s, x, out.r.cpu().numpy(), eps, tol=1e-14)
        if res["oper"] == oc["oper"]:
            assert res["iters"] == oc["iters"] and rel <= REL_L1
        else:
            # a node whose residual lands within rounding of its threshold may
            # be pushed by one and not the other (sums of mixed signs round
            # differently in atomic order): one push of degree d moves at most
            # ~eps*alpha*d of mass, so the estimates stay within 1e-6 relative
            assert abs(res["oper"] - oc["oper"]) <= 1e-3 * oc["oper"] and rel <= 1e-6


def test_sor_global_certified_large(large_graph, dev_large):
    g = large_graph
    s = int(graph.uniform_seeds(g, 1, seed=12)[0])
    out = engine.sor_solve(dev_large, ALPHA, 1e-7, OMEGA_OPT, s, local=False, max_iters=5000)
    _certified(g, s, out.x.cpu().numpy(), out.r.cpu().numpy(), 1e-7, tol=1e-14)
    print(f"global SOR sweeps {out.result['iters']} ms {out.result['ms']:.1f}")


# ------------------------------------------------------------------ Katz (katz_sor)
def _katz_certified(g, s, beta, x, r_ret, eps, tol=1e-15):
    A = O.adjacency(g)
    y = x.copy()
    y[s] += 1.0
    r = -(y - beta * (A @ y))
    r[s] += 1.0
    live = g.degree > 0
    assert np.abs(r - r_ret).max() < tol
    worst = float((np.abs(r[live]) / (eps * g.degree[live])).max())
    assert worst < 1.0, worst


@pytest.mark.parametrize("omega", [1.0, 1.3])
@pytest.mark.parametrize("local", [True, False])
def test_katz_sor_parity_small(small_graph, dev_small, local, omega):
    """Katz (beta = 1/(d_max+1), eps = 1e-7) against the coloured
    restatement (same bars as PPR) and the reference-order katz_sor (both
    certified: ||x - x_ref||_1 <= ||(I - bA)^-1||_1 (||r||_1 + ||r_ref||_1) with
    ||(I - bA)^-1||_1 <= 1 / (1 - b d_max))."""
    from flsg import _abi
    g = small_graph
    beta = 1.0 / (int(g.degree.max()) + 1)
    col, k, _ = engine.colouring(dev_small)
    eps = 1e-7
    for s in graph.uniform_seeds(g, 4, seed=5):
        s = int(s)
        out = engine.sor_solve(dev_small, beta, eps, omega, s, local=local, max_iters=5000, kind=_abi.FLSG_KATZ)
        x = out.x.cpu().numpy()
        oc = O.katz_sor_coloured(g, s, beta, eps, omega, col, k, local=local)
        res = out.result
        assert res["iters"] == oc["iters"] and res["oper"] == oc["oper"], (s, res, oc["iters"], oc["oper"])
        rel = np.abs(x - oc["x"]).sum() / np.abs(oc["x"]).sum()
        assert rel <= REL_L1, (s, rel)
        r = out.r.cpu().numpy()
        _katz_certified(g, s, beta, x, r, eps)
        ref = O.katz_sor(g, s, beta, eps, omega, local=local)
        bound = (np.abs(r).sum() + np.abs(ref["r"]).sum()) / (1.0 - beta * g.degree.max())
        assert np.abs(x - ref["x"]).sum() <= bound


def test_katz_sor_dropin(small_graph):
    import katz_solver
    g = small_graph
    beta = 1.0 / (int(g.degree.max()) + 1)
    s = int(graph.uniform_seeds(g, 1, seed=6)[0])
    x, r, oper, t = katz_solver.katz_sor(g.n, g.indptr, g.indices, g.degree, s, beta, 1e-7, 1.2, True)
    assert isinstance(oper, float) and oper > 0
    _katz_certified(g, s, beta, x, r, 1e-7)
    out = katz_solver.solve_a_graph_all(g.n, g.indptr, g.indices, g.degree, beta, 1e-7, 1.2, "synthetic",
                                        np.array([s], dtype=np.int32))
    assert len(out) == 13 and out[1][0] == int(oper) and min(o[0] for o in out[1::2]) > 0
    om = katz_solver.solve_a_graph_omega(g.n, g.indptr, g.indices, g.degree, beta, 1e-7, 1.2, "synthetic",
                                         np.array([s], dtype=np.int32))
    assert len(om) == 13 and om[1][0] == int(oper) and om[5][0] > 0 and om[3][0] == 0


def test_katz_sor_local_parity_large(large_graph, dev_large):
    from flsg import _abi
    g = large_graph
    beta = 1.0 / (int(g.degree.max()) + 1)
    col, k, _ = engine.colouring(dev_large)
    s = int(graph.uniform_seeds(g, 1, seed=13)[0])
    out = engine.sor_solve(dev_large, beta, 1e-7, 1.0, s, local=True, max_iters=5000, kind=_abi.FLSG_KATZ)
    x = out.x.cpu().numpy()
    oc = O.katz_sor_coloured(g, s, beta, 1e-7, 1.0, col, k, local=True)
    rel = np.abs(x - oc["x"]).sum() / np.abs(oc["x"]).sum()
    print(f"katz seed {s}: iters {out.result['iters']}/{oc['iters']} oper {out.result['oper']}/{oc['oper']} "
          f"rel L1 {rel:.2e} ms {out.result['ms']:.2f}")
    assert out.result["oper"] == oc["oper"] and rel <= REL_L1
    _katz_certified(g, s, beta, x, out.r.cpu().numpy(), 1e-7, tol=1e-14)


# ------------------------------------------------------------------ InstantGNN (instantGNNSorPush)
def _ignn_slack(n=400, p=0.02, seed=31, loops=(0, 5, 17)):
    """Slack CSR (2 spare slots per row) of a random connected graph with
    self-loops, one self-loop-only node and one isolated node."""
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(iu[0].size) < p
    g = CSRGraph.from_edges(n, np.concatenate([iu[0][keep], np.arange(n - 1)]),
                            np.concatenate([iu[1][keep], np.arange(1, n)]))
    rows = [list(g.indices[g.indptr[u]:g.indptr[u + 1]]) for u in range(n)]
    for u in loops:
        rows[u] = sorted(rows[u] + [u])
    rows.append([n])
    rows.append([])
    N = len(rows)
    cap = max(len(r) for r in rows) + 2
    indptr = np.arange(0, cap * (N + 1), cap, dtype=np.int64)
    indices = np.full(cap * N, N, dtype=np.int32)
    degree = np.zeros(N, dtype=np.int32)
    for u, rw in enumerate(rows):
        indices[cap * u:cap * u + len(rw)] = rw
        degree[u] = len(rw)
    return indptr, indices, degree


def _ignn_check(indptr, indices, degree, X, P, R, alpha, eps, beta):
    from flsg.ignn import ignn_residual
    loops = O.self_loops(indptr, indices, degree)
    live = ~((degree == 0) | ((degree == 1) & (loops == 1)))
    Rr = ignn_residual(indptr, indices, degree, X, P, alpha, beta)
    # invariant R == X - P/alpha + ((1-alpha)/alpha) M P on pushing nodes
    assert np.abs((Rr - R)[live]).max() < 1e-12
    ev = O.ignn_eps_vec(degree, eps, beta)
    assert np.all((np.abs(R) < ev[:, None]) | ((ev[:, None] == 0) & (R == 0)))


@pytest.mark.parametrize("omega", [1.0, 1.3])
def test_ignn_sor_column_parity(omega):
    """One column against the coloured restatement (same colouring): equal
    iteration and operation counts, p and r within 1e-12 relative L1; the
    invariant and certificate on the host."""
    import torch
    ip, ix, dg = _ignn_slack()
    N = dg.size
    alpha, eps, beta = 0.1, 1e-7, 0.5
    X = np.random.default_rng(7).standard_normal(N)
    sg = graph.SlackGraph(n=N, indptr=ip, indices=ix, degree=dg, us=np.zeros((0, 0), np.int32),
                          vs=np.zeros((0, 0), np.int32))
    dev = engine.DeviceGraph(sg)
    col, k, _ = engine.colouring(dev)
    p = torch.zeros(N, dtype=torch.float64, device=dev.device)
    r = torch.from_numpy(X.copy()).to(dev.device)
    st = engine.ignn_sor_columns(dev, alpha, eps, beta, omega, p, r, max_iters=5000)
    po, ro = np.zeros(N), X.copy()
    ev = O.ignn_eps_vec(dg, eps, beta)
    nodes = np.flatnonzero((np.abs(X) >= ev) & (X != 0))
    oc = O.ignn_sor_coloured(ip, ix, dg, col, k, alpha, eps, omega, beta, po, ro, nodes)
    assert st["iters"] == oc["iters"] and st["oper"] == oc["oper"], (st, oc)
    pg, rg = p.cpu().numpy(), r.cpu().numpy()
    assert np.abs(pg - po).sum() <= 1e-12 * np.abs(po).sum()
    assert np.abs(rg - ro).sum() <= 1e-12 * max(np.abs(ro).sum(), 1e-300) + 1e-18
    _ignn_check(ip, ix, dg, X[:, None], pg[:, None], rg[:, None], alpha, eps, beta)
    dev.close()


def test_ignn_sor_parallel_dropin():
    """instantGNNSorParallel on 3 columns: P, R overwritten in place, the
    invariant and certificate hold, and P is within the certificates of the
    reference-order SOR (||p - p_ref||_1 <= ||K||_1 (||r||_1 + ||r_ref||_1))."""
    import InstantGNN
    ip, ix, dg = _ignn_slack(n=120, p=0.05, seed=32)
    N = dg.size
    alpha, eps, beta, omega = 0.1, 1e-6, 0.5, 1.2
    X = np.random.default_rng(8).standard_normal((N, 3))
    P = np.full((N, 3), 7.0)   # the reference starts every column from p = 0
    R = X.copy()
    times = InstantGNN.instantGNNSorParallel(N, ip, ix.copy(), dg.copy(), alpha, eps, omega, P, R, beta)
    assert times.shape == (3,) and np.all(times >= 0)
    _ignn_check(ip, ix, dg, X, P, R, alpha, eps, beta)
    M = np.zeros((N, N))
    for u in range(N):
        for v in ix[ip[u]:ip[u] + dg[u]]:
            M[v, u] += 1.0 / (dg[v] ** beta * dg[u] ** (1 - beta))
    K = alpha * np.linalg.inv(np.eye(N) - (1 - alpha) * M)
    kn = np.abs(K).sum(axis=0).max()
    live = ~((dg == 0) | ((dg == 1) & (O.self_loops(ip, ix, dg) == 1)))
    for c in range(3):
        pr, rr = np.zeros(N), X[:, c].copy()
        O.ignn_sor_column(ip, ix, dg, alpha, eps, omega, beta, pr, rr)
        bound = kn * (np.abs(R[:, c]).sum() + np.abs(rr).sum())
        assert np.abs((P[:, c] - pr)[live]).sum() <= bound + 1e-15


def test_ignn_sor_update_insert():
    """instantGNNUpdateBatchParallel with a 'sor' algorithm: after inserting
    edges the invariant and certificate hold on the new graph."""
    import InstantGNN
    ip, ix, dg = _ignn_slack(n=150, p=0.03, seed=33)
    N = dg.size
    alpha, eps, beta = 0.1, 1e-6, 0.5
    X = np.random.default_rng(9).standard_normal((N, 2))
    P, R = np.zeros((N, 2)), X.copy()
    ix, dg = ix.copy(), dg.copy()
    InstantGNN.instantGNNParallel(N, ip, ix, dg, alpha, eps, P, R, beta)
    rng = np.random.default_rng(10)
    us = rng.integers(0, N - 2, size=10).astype(np.int32)
    vs = rng.integers(0, N - 2, size=10).astype(np.int32)
    keep = us != vs
    InstantGNN.instantGNNUpdateBatchParallel(N, ip, ix, dg, alpha, eps, 1.2, P, R, X, us[keep], vs[keep], beta,
                                             "sorbatchinsert")
    _ignn_check(ip, ix, dg, X, P, R, alpha, eps, beta)


def test_sor_iteration_cap_then_recovers(small_graph, dev_small):
    """A solve stopped by its iteration cap raises, and the next solve on the
    same resident graph (its workspace reset at the start of every solve)
    equals a solve that ran before the failure: iteration and operation
    counts equal, estimate within 1e-12 relative L1."""
    g = small_graph
    s = int(graph.uniform_seeds(g, 1, seed=11)[0])
    for local in (True, False):
        before = engine.sor_solve(dev_small, ALPHA, 1e-6, OMEGA_OPT, s, local=local, max_iters=5000)
        assert before.result["iters"] > 2
        with pytest.raises(FlsgError):
            engine.sor_solve(dev_small, ALPHA, 1e-6, OMEGA_OPT, s, local=local, max_iters=2)
        after = engine.sor_solve(dev_small, ALPHA, 1e-6, OMEGA_OPT, s, local=local, max_iters=5000)
        assert after.result["iters"] == before.result["iters"]
        assert after.result["oper"] == before.result["oper"]
        xb, xa = before.x.cpu().numpy(), after.x.cpu().numpy()
        assert np.abs(xa - xb).sum() / np.abs(xb).sum() <= REL_L1
        _certified(g, s, xa, after.r.cpu().numpy(), 1e-6)
