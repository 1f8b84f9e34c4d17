"""GPU parity of the push kernels against the oracle (C restatement of the
reference's numba solvers).

Bars (written in each test):
  * frontier sets per iteration: identical to the oracle's (integer work is
    bit-exact; a node whose residual sits within rounding of its threshold
    could legitimately flip -- such nodes would be counted and reported);
  * operation counts: equal;
  * FP64 estimate: relative L1 difference <= 1e-12;
  * FP32 batched path: every query certified in FP64 (max |r_v|/thr_v < 1,
    recomputed independently on the host), and the L1 distance to the oracle
    bounded by the two certified residuals (||x - x*||_1 <= ||r||_1 / alpha).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from flsg import _abi, engine, graph  # noqa: E402
from oracle import oracle as O  # noqa: E402

ALPHA = 0.1
REL_L1 = 1e-12


@pytest.fixture(scope="module")
def dev_small(small_graph):
    return engine.DeviceGraph(small_graph)


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def _compare_traces(gpu_tr, ora_tr):
    """Number of iterations and per-iteration symmetric-difference sizes."""
    assert len(gpu_tr) == len(ora_tr), f"iteration count {len(gpu_tr)} != oracle {len(ora_tr)}"
    diffs = []
    for t, (a, b) in enumerate(zip(gpu_tr, ora_tr)):
        sa, sb = np.sort(a), np.sort(b)
        assert np.unique(sa).size == sa.size, f"duplicate frontier entries at iteration {t}"
        if sa.size != sb.size or not np.array_equal(sa, sb):
            diffs.append((t, np.setxor1d(sa, sb).size))
    return diffs


def test_ppr_local_gd_parity_small(small_graph, dev_small):
    g = small_graph
    seeds = graph.uniform_seeds(g, 20, seed=0)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-6)
    for s in seeds:
        o = O.ppr_gd(g, s, ALPHA, 1e-6, local=True, trace=True)
        out = engine.solve(dev_small, p, int(s), trace=True)
        diffs = _compare_traces(out.trace, o["trace"])
        assert diffs == [], f"seed {s}: frontier set differences {diffs}"
        assert out.result["oper"] == o["oper"], f"seed {s}"
        x = out.x.cpu().numpy()
        r = out.r.cpu().numpy()
        assert _rel_l1(x, o["x"]) <= REL_L1, f"seed {s}: rel L1 {_rel_l1(x, o['x'])}"
        assert np.abs(r - o["r"]).sum() <= REL_L1 * np.abs(o["r"]).sum() + 1e-300
        # certificate: |r_v| < eps*alpha*d_v for every v, recomputed in FP64
        rr = O.ppr_residual(g, s, ALPHA, x)
        thr = 1e-6 * ALPHA * np.maximum(g.degree, 1)
        assert (np.abs(rr) / thr).max() < 1.0 + 1e-9


def test_local_push_graph_replays(small_graph, dev_small):
    """The local push runs as a captured 8-iteration hipGraph, replayed until
    the frontier empties and kept across solves while the arguments match.
    Seeds A, A, B, A, Katz A, A with eager host work and synchronisations in
    between: every solve matches the oracle (frontier sets, oper, x)."""
    import torch
    g = small_graph
    sa, sb = (int(s) for s in graph.uniform_seeds(g, 2, seed=41))
    beta = 1.0 / (g.d_max + 1)
    runs = [(_abi.FLSG_PPR, sa), (_abi.FLSG_PPR, sa), (_abi.FLSG_PPR, sb), (_abi.FLSG_PPR, sa),
            (_abi.FLSG_KATZ, sa), (_abi.FLSG_PPR, sa)]
    for kind, s in runs:
        if kind == _abi.FLSG_PPR:
            p = engine.make_params(kind, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-6)
            o = O.ppr_gd(g, s, ALPHA, 1e-6, local=True, trace=True)
        else:
            p = engine.make_params(kind, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=beta, eps=1e-7)
            o = O.katz_gd(g, s, beta, 1e-7, local=True, trace=True)
        out = engine.solve(dev_small, p, s, trace=True)
        assert _compare_traces(out.trace, o["trace"]) == []
        assert out.result["oper"] == o["oper"]
        assert _rel_l1(out.x.cpu().numpy(), o["x"]) <= REL_L1
        torch.ones(1 << 20, device="cuda").sum().item()   # unrelated eager work between solves
        torch.cuda.synchronize()


def test_ppr_local_gd_alpha_roundtrip(small_graph, dev_small):
    """The harness alpha (demo.py:31-35): 0.1 -> 0.10000000000000002."""
    a, _ = graph.alpha_roundtrip(0.1)
    assert a == 0.10000000000000002
    g = small_graph
    s = int(graph.uniform_seeds(g, 1, seed=3)[0])
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=a, eps=1e-6)
    o = O.ppr_gd(g, s, a, 1e-6, local=True, trace=True)
    out = engine.solve(dev_small, p, s, trace=True)
    assert _compare_traces(out.trace, o["trace"]) == []
    assert out.result["oper"] == o["oper"]
    assert _rel_l1(out.x.cpu().numpy(), o["x"]) <= REL_L1


@pytest.mark.parametrize("method", ["gd", "cheb"])
def test_ppr_global_parity_small(small_graph, dev_small, method):
    g = small_graph
    m = _abi.FLSG_GD if method == "gd" else _abi.FLSG_CHEB
    p = engine.make_params(_abi.FLSG_PPR, m, False, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-6)
    fn = O.ppr_gd if method == "gd" else O.ppr_cheby
    for s in graph.uniform_seeds(g, 3, seed=1):
        o = fn(g, s, ALPHA, 1e-6, local=False)
        out = engine.solve(dev_small, p, int(s))
        assert out.result["oper"] == o["oper"], f"seed {s}"
        assert out.result["oper"] % (g.n + g.m) == 0
        x = out.x.cpu().numpy()
        # sequential row-order pull sums: bit-identical to the oracle
        assert np.array_equal(x, o["x"]), f"seed {s}: rel L1 {_rel_l1(x, o['x'])}"
        assert np.array_equal(out.r.cpu().numpy(), o["r"])


@pytest.fixture(scope="module")
def hub_graph():
    """20,000 nodes with five rows over 1024 entries and 193 of 65-1024: every
    row class of the global sweeps' pull (thread, wave, block)."""
    return graph.powerlaw_graph_small(n=20000, wmax=5000.0, seed=5)


@pytest.mark.parametrize("method", ["gd", "cheb"])
def test_ppr_global_parity_hub_rows(hub_graph, method):
    g = hub_graph
    assert int(g.degree.max()) > 1024
    dg = engine.DeviceGraph(g)
    m = _abi.FLSG_GD if method == "gd" else _abi.FLSG_CHEB
    p = engine.make_params(_abi.FLSG_PPR, m, False, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-6)
    fn = O.ppr_gd if method == "gd" else O.ppr_cheby
    hubs = np.argsort(-np.asarray(g.degree))[:2]
    for s in list(graph.uniform_seeds(g, 2, seed=3)) + [int(h) for h in hubs]:
        o = fn(g, s, ALPHA, 1e-6, local=False)
        out = engine.solve(dg, p, int(s))
        assert out.result["oper"] == o["oper"], f"seed {s}"
        assert np.array_equal(out.x.cpu().numpy(), o["x"]), f"seed {s}"
        assert np.array_equal(out.r.cpu().numpy(), o["r"]), f"seed {s}"
    dg.close()


@pytest.mark.parametrize("local", [True, False])
def test_katz_gd_parity_small(small_graph, dev_small, local):
    g = small_graph
    beta = 1.0 / (g.d_max + 1)
    eps = 1e-7
    p = engine.make_params(_abi.FLSG_KATZ, _abi.FLSG_GD, local, _abi.FLSG_FP64, alpha=beta, eps=eps)
    for s in graph.uniform_seeds(g, 5, seed=2):
        o = O.katz_gd(g, s, beta, eps, local=local, trace=local)
        out = engine.solve(dev_small, p, int(s), trace=local)
        if local:
            assert _compare_traces(out.trace, o["trace"]) == []
        assert out.result["oper"] == o["oper"]
        assert _rel_l1(out.x.cpu().numpy(), o["x"]) <= REL_L1


def test_katz_global_cheb_parity_small(small_graph, dev_small):
    g = small_graph
    beta = 1.0 / (g.d_max + 1)
    import scipy.sparse.linalg as sla
    A = O.adjacency(g)
    lam_max = float(sla.eigsh(A, 1, which="LA")[0][0])
    lam_min = float(sla.eigsh(A, 1, which="SA")[0][0])
    l_max, mu_min = 1 - beta * lam_min, 1 - beta * lam_max
    p = engine.make_params(_abi.FLSG_KATZ, _abi.FLSG_CHEB, False, _abi.FLSG_FP64, alpha=beta, eps=1e-7,
                           l_max=l_max, mu_min=mu_min)
    for s in graph.uniform_seeds(g, 2, seed=4):
        o = O.katz_cheby(g, s, beta, l_max, mu_min, 1e-7, local=False)
        out = engine.solve(dev_small, p, int(s))
        assert out.result["oper"] == o["oper"]
        assert np.array_equal(out.x.cpu().numpy(), o["x"])


@pytest.mark.parametrize("n_terms", [0, 40])
def test_hk_push_parity_small(small_graph, dev_small, n_terms):
    g = small_graph
    tau, eps = 5.0, 1e-7
    p = engine.make_params(_abi.FLSG_HK, _abi.FLSG_GD, True, _abi.FLSG_FP64, eps=eps, tau=tau, n_terms=n_terms)
    for s in graph.uniform_seeds(g, 5, seed=5):
        o = O.hk_push(g, s, tau, eps, local=True, n_terms=n_terms, trace=True)
        out = engine.solve(dev_small, p, int(s), trace=True)
        assert out.extra["est_n"] == o["est_n"]
        assert _compare_traces(out.trace, o["trace"]) == []
        assert out.result["oper"] == int(o["oper"])
        assert _rel_l1(out.x.cpu().numpy(), o["x"]) <= REL_L1
        rg = out.r.cpu().numpy()
        assert np.abs(rg - o["r"]).sum() <= REL_L1 * max(np.abs(o["r"]).sum(), 1e-300)


@pytest.mark.parametrize("which,n_terms", [("small", 0), ("small", 40), ("hub", 0), ("hub", 1)])
def test_hk_global_sweep_parity(small_graph, dev_small, hub_graph, which, n_terms):
    """hk_sor_queue(local=False): one sweep per Taylor level, the pull
    summing each sorted row in the reference's scatter order, so x is
    bit-identical to the oracle's; the last level's pull inserts the node's
    own residual at its place in the row (rows of one thread, a wave and a
    block on the hub graph; n_terms 1 makes the first level the last)."""
    g = small_graph if which == "small" else hub_graph
    dg = dev_small if which == "small" else engine.DeviceGraph(g)
    tau, eps = 5.0, 1e-7
    p = engine.make_params(_abi.FLSG_HK, _abi.FLSG_GD, False, _abi.FLSG_FP64, eps=eps, tau=tau, n_terms=n_terms)
    for s in graph.uniform_seeds(g, 3, seed=15):
        o = O.hk_push(g, s, tau, eps, local=False, n_terms=n_terms)
        out = engine.solve(dg, p, int(s))
        assert out.extra["est_n"] == o["est_n"]
        assert out.result["oper"] == int(o["oper"])
        assert np.array_equal(out.x.cpu().numpy(), o["x"]), f"seed {s}: rel L1 {_rel_l1(out.x.cpu().numpy(), o['x'])}"
        assert not out.r.cpu().numpy().any()


def _check_batch(g, seeds, bo, eps, alpha=ALPHA, n_check=None):
    n_check = len(seeds) if n_check is None else n_check
    thr = eps * alpha * np.maximum(g.degree, 1)
    worst = 0.0
    for q in range(n_check):
        res = bo.results[q]
        assert res["status"] == 0
        assert res["cert"] < 1.0
        x = bo.dense(q, g.n)
        rr = O.ppr_residual(g, int(seeds[q]), alpha, x)
        ratio = float((np.abs(rr) / thr).max())
        worst = max(worst, ratio)
        # host FP64 recompute agrees with the device certificate up to rounding
        assert ratio < 1.0 + 1e-9, f"query {q}: host residual ratio {ratio}"
    return worst


def test_batch_fp32_certified_small(small_graph, dev_small):
    g = small_graph
    eps = 1e-6
    seeds = graph.mixed_seeds(g, 64, seed=7)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=eps)
    bo = engine.batch_solve(dev_small, p, seeds)
    _check_batch(g, seeds, bo, eps)
    for q in range(8):
        s = int(seeds[q])
        o = O.ppr_gd(g, s, ALPHA, eps, local=True)
        x = bo.dense(q, g.n)
        rr = O.ppr_residual(g, s, ALPHA, x)
        bound = (np.abs(rr).sum() + np.abs(o["r"]).sum()) / ALPHA
        err = np.abs(x - o["x"]).sum()
        assert err <= bound * (1 + 1e-9) + 1e-15, f"query {q}: L1 {err} > bound {bound}"
        # the FP32 frontier work tracks the FP64 one closely
        assert abs(bo.results[q]["oper"] - o["oper"]) <= 0.05 * o["oper"] + 100


@pytest.mark.parametrize("split,piece", [(64, 64), (4096, 0), (64, 1 << 16)])
def test_batch_degree_bins(small_graph, split, piece):
    """Any degree bins (entries split into pieces from 64 edges up, pieces of
    64 edges up) give certified results, in both precisions."""
    g = small_graph
    dg = engine.DeviceGraph(g)
    try:
        assert dg.lib.flsg_batch_set_degree_bins(dg.handle, split, piece) == 0
        seeds = graph.mixed_seeds(g, 48, seed=31)
        for prec in (_abi.FLSG_FP32, _abi.FLSG_FP64):
            p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, prec, alpha=ALPHA, eps=1e-6)
            bo = engine.batch_solve(dg, p, seeds)
            _check_batch(g, seeds, bo, 1e-6, n_check=12)
            if prec == _abi.FLSG_FP64:
                for q in range(4):
                    o = O.ppr_gd(g, int(seeds[q]), ALPHA, 1e-6, local=True)
                    assert bo.results[q]["oper"] == o["oper"]
        assert dg.lib.flsg_batch_set_degree_bins(dg.handle, 8, 0) != 0     # below one wavefront: rejected
    finally:
        dg.close()


def test_batch_arena_clean_after_launches(small_graph, dev_small):
    """Every launch resets each record it touched (through the touched list or
    by streaming the slot): FP32 and FP64 pushes, PPR and Katz, small and
    large supports, leave the whole arena in its between-queries state."""
    g = small_graph
    seeds = graph.mixed_seeds(g, 96, seed=5)
    modes = set()
    for kind, a, eps in ((_abi.FLSG_PPR, ALPHA, 1e-6), (_abi.FLSG_PPR, ALPHA, 1e-3),
                         (_abi.FLSG_KATZ, 1.0 / (g.d_max + 1), 1e-7)):
        for prec in (_abi.FLSG_FP32, _abi.FLSG_FP64):
            p = engine.make_params(kind, _abi.FLSG_GD, True, prec, alpha=a, eps=eps)
            bo = engine.batch_solve(dev_small, p, seeds)
            assert all(r["status"] == 0 and r["cert"] < 1.0 for r in bo.results)
            assert engine.arena_dirty(dev_small) == 0, (kind, eps, prec)
            modes |= {r["touched"] * 4 > g.n for r in bo.results}
    assert modes == {True, False}   # both reset modes ran


@pytest.mark.parametrize("alpha,eps", [(0.01, 1e-6), (0.005, 1e-7)])
def test_batch_fp32_boundary_regimes(small_graph, dev_small, alpha, eps):
    """FP32 certificate in both regimes of the boundary bound: at alpha 0.01
    (about 295 push iterations) boundary nodes well below their threshold keep
    their FP32 residual plus the rounding-error bound; at alpha 0.005 (about
    1050 iterations, past the 494 up to which the bound is used at degree 64)
    every boundary node is recomputed in FP64.  Certified on the device and
    re-certified by a host FP64 recompute in both."""
    g = small_graph
    seeds = graph.mixed_seeds(g, 24, seed=17)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=alpha, eps=eps)
    bo = engine.batch_solve(dev_small, p, seeds)
    its = bo.results["iters"]
    if alpha == 0.01:
        assert its.max() < 494
    else:
        assert its.min() > 494
    _check_batch(g, seeds, bo, eps, alpha=alpha, n_check=8)
    assert engine.arena_dirty(dev_small) == 0


def test_batch_fp32_graph_preconditions(small_graph):
    """The FP32 certificate's preconditions are decided per graph.  On a graph
    whose rows repeat entries (every edge listed twice: the same PPR vector)
    the push-count cap of its error bounds is off and every query is still
    certified, by the device and by a host FP64 recompute.  A directed graph
    (leaves pointing at a centre that points back at one of them) is refused
    for the FP32 push with FLSG_EARG; the FP64 push runs on it."""
    g = small_graph
    ip = np.asarray(g.indptr, dtype=np.int64)
    rows = [np.asarray(g.indices[ip[u]:ip[u + 1]]) for u in range(g.n)]
    ind2 = np.concatenate([np.concatenate([r, r]) for r in rows]).astype(np.int32)
    deg2 = (2 * np.asarray(g.degree)).astype(np.int32)
    ip2 = np.zeros(g.n + 1, dtype=np.int64)
    ip2[1:] = np.cumsum(deg2)
    g2 = graph.CSRGraph(n=g.n, indptr=ip2, indices=ind2, degree=deg2, name="doubled")
    dg2 = engine.DeviceGraph(g2)
    try:
        eps = 1e-6
        seeds = graph.mixed_seeds(g, 24, seed=23)
        p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=eps)
        bo = engine.batch_solve(dg2, p, seeds)
        _check_batch(g2, seeds, bo, eps)
        assert engine.arena_dirty(dg2) == 0
    finally:
        dg2.close()
    k = 50
    # row 0 = [1] (the centre points at leaf 1), rows 1..k = [0]
    ipd = np.arange(k + 2, dtype=np.int64)
    ind = np.concatenate([[1], np.zeros(k, dtype=np.int64)]).astype(np.int32)
    degd = np.diff(ipd).astype(np.int32)
    gd = graph.CSRGraph(n=k + 1, indptr=ipd, indices=ind, degree=degd, name="directed star")
    dgd = engine.DeviceGraph(gd)
    try:
        s = np.array([1, 2], dtype=np.int64)
        p32 = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=1e-4)
        with pytest.raises(engine.FlsgError) as ei:
            engine.batch_solve(dgd, p32, s)
        assert ei.value.code == _abi.FLSG_EARG
        p64 = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-4)
        bo = engine.batch_solve(dgd, p64, s)
        assert all(r["status"] == 0 for r in bo.results)
    finally:
        dgd.close()


def test_batch_bad_seed_fails_alone(small_graph, dev_small):
    """A query whose seed lies outside the graph fails with FLSG_EARG on its
    own record (the kernel checks before touching a slot); the rest of the
    launch is unaffected."""
    g = small_graph
    seeds = graph.mixed_seeds(g, 16, seed=41).astype(np.int64)
    seeds[3] = g.n + 5
    seeds[9] = -1
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=1e-6)
    with pytest.raises(engine.FlsgError):
        engine.batch_solve(dev_small, p, seeds)
    bo = engine.batch_solve(dev_small, p, seeds, allow_failed=True)
    for q in range(len(seeds)):
        if q in (3, 9):
            assert bo.results[q]["status"] == _abi.FLSG_EARG
        else:
            assert bo.results[q]["status"] == 0 and bo.results[q]["cert"] < 1.0
    good = [q for q in range(len(seeds)) if q not in (3, 9)]
    bo2 = engine.batch_solve(dev_small, p, seeds[good])
    assert all(r["status"] == 0 and r["cert"] < 1.0 for r in bo2.results)


def test_batch_same_alone_and_in_mix(small_graph, dev_small):
    """A query's certified result does not depend on what else is in the batch."""
    g = small_graph
    eps = 1e-6
    seeds = graph.mixed_seeds(g, 32, seed=9)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=eps)
    mix = engine.batch_solve(dev_small, p, seeds)
    for q in (0, 5, 17):
        s = int(seeds[q])
        alone = engine.batch_solve(dev_small, p, seeds[q:q + 1])
        a = alone.dense(0, g.n)
        b = mix.dense(q, g.n)
        # both are certified solutions of the same system: they differ by at
        # most the sum of their residual bounds, ||x - x*||_1 <= ||r||_1 / alpha
        ra = np.abs(O.ppr_residual(g, s, ALPHA, a)).sum()
        rb = np.abs(O.ppr_residual(g, s, ALPHA, b)).sum()
        assert np.abs(a - b).sum() <= (ra + rb) / ALPHA * (1 + 1e-9)
        assert alone.results[0]["supp"] > 0


def test_batch_katz_certified_small(small_graph, dev_small):
    g = small_graph
    beta = 1.0 / (g.d_max + 1)
    eps = 1e-7
    seeds = graph.uniform_seeds(g, 32, seed=11)
    p = engine.make_params(_abi.FLSG_KATZ, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=beta, eps=eps)
    bo = engine.batch_solve(dev_small, p, seeds)
    for q in range(len(seeds)):
        assert bo.results[q]["status"] == 0 and bo.results[q]["cert"] < 1.0
        o = O.katz_gd(g, int(seeds[q]), beta, eps, local=True)
        x = bo.dense(q, g.n)
        x[int(seeds[q])] -= 1.0   # katz_gd returns x - e_s
        assert np.abs(x - o["x"]).sum() <= 1e-5 * max(np.abs(o["x"]).sum(), 1e-12) + 1e-9


@pytest.mark.parametrize("kind", [_abi.FLSG_PPR, _abi.FLSG_KATZ])
def test_batch_fp64_parity_small(small_graph, dev_small, kind):
    """Batched FP64 push: the same frontier semantics as the plain FP64 path,
    so the operation count equals the oracle's and x agrees to rel L1 1e-12
    (only the order of the atomic sums differs)."""
    g = small_graph
    if kind == _abi.FLSG_PPR:
        a, eps = ALPHA, 1e-6
    else:
        a, eps = 1.0 / (g.d_max + 1), 1e-7
    seeds = graph.mixed_seeds(g, 48, seed=13)
    p = engine.make_params(kind, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=a, eps=eps)
    bo = engine.batch_solve(dev_small, p, seeds)
    for q in range(len(seeds)):
        s = int(seeds[q])
        res = bo.results[q]
        assert res["status"] == 0 and res["cert"] < 1.0 and res["n_corr"] == 0
        x = bo.dense(q, g.n)
        if kind == _abi.FLSG_PPR:
            o = O.ppr_gd(g, s, a, eps, local=True)
        else:
            o = O.katz_gd(g, s, a, eps, local=True)
            x[s] -= 1.0   # katz_gd returns x - e_s
        assert res["oper"] == o["oper"], f"query {q}"
        assert _rel_l1(x, o["x"]) <= REL_L1, f"query {q}: {_rel_l1(x, o['x'])}"
    if kind == _abi.FLSG_PPR:
        _check_batch(g, seeds, bo, eps, n_check=8)


@pytest.mark.parametrize("hub_bytes,bin_min,dense_min", [(16384, 1, 0), (16384, 1, 1 << 30), (24576, 2000, 64),
                                                       (-1, 1, 0), (-1, 0, 256)])
def test_batch_binned_push_small(small_graph, hub_bytes, bin_min, dense_min):
    """The binned push (pairs per destination range summed in LDS) keeps the
    FP64 frontier semantics: with a 16 KiB slice the 10,000-node graph has 5
    ranges (FP64) / 3 (FP32), every range summed in LDS and drained by its
    pairs or swept in id order (dense_min 0), or every one applied with
    memory-side atomics (dense_min 2^30), or a mix; with the default 128 KiB
    slice the graph is one range, swept; bin_min 0 pushes every iteration
    directly.  FP64: operation counts equal
    to the oracle's and x within relative L1 1e-12; FP32: certified, and the
    arena is clean afterwards."""
    g = small_graph
    dg = engine.DeviceGraph(g)
    try:
        engine._check(dg.lib.flsg_batch_set_hub_cache(dg.handle, hub_bytes))
        engine._check(dg.lib.flsg_batch_set_binning(dg.handle, bin_min, dense_min))
        engine._check(dg.lib.flsg_batch_set_bin_load(dg.handle, 1))   # bin with any number of busy slots
        seeds = graph.mixed_seeds(g, 40, seed=77)
        for kind, a, eps in ((_abi.FLSG_PPR, ALPHA, 1e-6), (_abi.FLSG_KATZ, 1.0 / (g.d_max + 1), 1e-7)):
            p = engine.make_params(kind, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=a, eps=eps)
            bo = engine.batch_solve(dg, p, seeds)
            binned = int(bo.results["bin_iters"].sum())
            assert (binned == 0) == (bin_min == 0)
            for q in range(12):
                s = int(seeds[q])
                x = bo.dense(q, g.n)
                if kind == _abi.FLSG_PPR:
                    o = O.ppr_gd(g, s, a, eps, local=True)
                else:
                    o = O.katz_gd(g, s, a, eps, local=True)
                    x[s] -= 1.0
                assert bo.results[q]["oper"] == o["oper"], (kind, q)
                assert _rel_l1(x, o["x"]) <= REL_L1, (kind, q)
        p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=1e-6)
        bo = engine.batch_solve(dg, p, seeds)
        _check_batch(g, seeds, bo, 1e-6, n_check=10)
        assert engine.arena_dirty(dg) == 0
    finally:
        dg.close()


@pytest.fixture(scope="module")
def dev_large(large_graph):
    return engine.DeviceGraph(large_graph)


def test_ppr_local_gd_parity_large(large_graph, dev_large, capsys):
    """Full-size graph (2.39M nodes), eps = 1e-7: oracle parity on 3 seeds.
    Frontier sets must be equal in every iteration; the oracle's enqueue tests
    that fell within one ulp of the threshold (where another summation order
    could flip membership) are counted and reported beside the result."""
    g = large_graph
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-7)
    for s in graph.uniform_seeds(g, 3, seed=21):
        o = O.ppr_gd(g, s, ALPHA, 1e-7, local=True, trace=True)
        out = engine.solve(dev_large, p, int(s), trace=True, trace_cap_nodes=64 * g.n)
        diffs = _compare_traces(out.trace, o["trace"])
        with capsys.disabled():
            print(f"\n  seed {s}: {len(o['trace'])} iterations, frontier set differences {len(diffs)}, "
                  f"enqueue tests within 1 ulp of the threshold {o['near_thr']}")
        assert diffs == [], f"seed {s}: {diffs} (near-threshold tests {o['near_thr']})"
        assert out.result["oper"] == o["oper"]
        assert _rel_l1(out.x.cpu().numpy(), o["x"]) <= REL_L1


@pytest.mark.parametrize("eps", [1e-7, 1e-6])
def test_batch_certified_large(large_graph, dev_large, eps):
    """1024 mixed seeds at full size, eps in {1e-7, 1e-6}: every query
    certified on the device; 16 of them re-certified by an independent host
    FP64 recompute."""
    g = large_graph
    seeds = graph.mixed_seeds(g, 1024, seed=0)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=eps)
    bo = engine.batch_solve(dev_large, p, seeds)
    assert all(r["status"] == 0 and r["cert"] < 1.0 for r in bo.results)
    # queries touching more than n/4 records reset by streaming the slot,
    # the others through their touched list: both leave the arena clean
    tn = np.array([r["touched"] for r in bo.results])
    print(f"eps {eps}: touched > n/4 in {(tn * 4 > g.n).sum()} of {len(tn)} queries")
    assert engine.arena_dirty(dev_large) == 0
    _check_batch(g, seeds, bo, eps, n_check=16)


def test_batch_fp64_large(large_graph, dev_large):
    """Full size, eps = 1e-7: 256 mixed seeds through the batched FP64 push,
    every query certified on the device, 3 against the oracle."""
    g = large_graph
    eps = 1e-7
    seeds = graph.mixed_seeds(g, 256, seed=2)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=ALPHA, eps=eps)
    bo = engine.batch_solve(dev_large, p, seeds)
    assert all(r["status"] == 0 and r["cert"] < 1.0 for r in bo.results)
    assert engine.arena_dirty(dev_large) == 0
    for q in range(3):
        o = O.ppr_gd(g, int(seeds[q]), ALPHA, eps, local=True)
        assert bo.results[q]["oper"] == o["oper"]
        assert _rel_l1(bo.dense(q, g.n), o["x"]) <= REL_L1


def test_ppr_local_cheb_exact_small(small_graph, dev_small):
    """Local Chebyshev is order-dependent (residuals change sign); the GPU
    replays the reference's sequential order, so queues match in ORDER and x,
    r are bit-identical to the oracle."""
    g = small_graph
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_CHEB, True, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-6)
    for s in graph.uniform_seeds(g, 10, seed=6):
        o = O.ppr_cheby(g, s, ALPHA, 1e-6, local=True, trace=True)
        out = engine.solve(dev_small, p, int(s), trace=True)
        assert len(out.trace) == len(o["trace"]), f"seed {s}"
        for t, (a, b) in enumerate(zip(out.trace, o["trace"])):
            assert np.array_equal(a, b), f"seed {s}: queue order differs at iteration {t}"
        assert out.result["oper"] == o["oper"]
        assert np.array_equal(out.x.cpu().numpy(), o["x"])
        assert np.array_equal(out.r.cpu().numpy(), o["r"])


def test_katz_local_cheb_exact_small(small_graph, dev_small):
    g = small_graph
    beta = 1.0 / (g.d_max + 1)
    import scipy.sparse.linalg as sla
    A = O.adjacency(g)
    lam_max = float(sla.eigsh(A, 1, which="LA")[0][0])
    lam_min = float(sla.eigsh(A, 1, which="SA")[0][0])
    l_max, mu_min = 1 - beta * lam_min, 1 - beta * lam_max
    p = engine.make_params(_abi.FLSG_KATZ, _abi.FLSG_CHEB, True, _abi.FLSG_FP64, alpha=beta, eps=1e-8,
                           l_max=l_max, mu_min=mu_min)
    for s in graph.uniform_seeds(g, 5, seed=8):
        o = O.katz_cheby(g, s, beta, l_max, mu_min, 1e-8, local=True, trace=True)
        out = engine.solve(dev_small, p, int(s), trace=True)
        assert len(out.trace) == len(o["trace"])
        for a, b in zip(out.trace, o["trace"]):
            assert np.array_equal(a, b)
        assert out.result["oper"] == o["oper"]
        assert np.array_equal(out.x.cpu().numpy(), o["x"])


def test_ppr_local_cheb_exact_large(large_graph, dev_large):
    g = large_graph
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_CHEB, True, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-7)
    for s in graph.uniform_seeds(g, 2, seed=22):
        o = O.ppr_cheby(g, s, ALPHA, 1e-7, local=True, trace=True)
        out = engine.solve(dev_large, p, int(s), trace=True, trace_cap_nodes=64 * g.n)
        assert len(out.trace) == len(o["trace"])
        for a, b in zip(out.trace, o["trace"]):
            assert np.array_equal(a, b)
        assert out.result["oper"] == o["oper"]
        assert np.array_equal(out.x.cpu().numpy(), o["x"])


def test_large_local_vs_fp64_global_solve(large_graph, dev_large):
    """Full size (2.39M nodes, eps = 1e-7): local GD and local Chebyshev
    estimates checked against an FP64 global GD solve to residual 1e-15
    (thresholds 1e-15 * alpha * d), through what the mathematics fixes:
      * GD: r >= 0 and x* - x = Q^-1 r, and 1^T Q^-1 = 1^T / alpha on the
        non-isolated nodes, so ||x* - x||_1 = sum(r) / alpha exactly, and x <= x*;
      * Chebyshev (r of both signs): ||x* - x||_1 <= ||r||_1 / alpha.
    The global solve's own distance to x* is at most ||r_glob||_1 / alpha."""
    g = large_graph
    s = int(graph.uniform_seeds(g, 1, seed=27)[0])
    pg = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, False, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-15)
    ref = engine.solve(dev_large, pg, s)
    xs = ref.x.cpu().numpy()
    err_ref = np.abs(ref.r.cpu().numpy()).sum() / ALPHA
    assert err_ref < 1e-8
    for method in (_abi.FLSG_GD, _abi.FLSG_CHEB):
        p = engine.make_params(_abi.FLSG_PPR, method, True, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-7)
        out = engine.solve(dev_large, p, s)
        x = out.x.cpu().numpy()
        r = out.r.cpu().numpy()
        l1 = np.abs(x - xs).sum()
        if method == _abi.FLSG_GD:
            assert (r >= 0).all()
            expect = r.sum() / ALPHA
            assert abs(l1 - expect) <= 1e-9 * expect + 2 * err_ref, (l1, expect)
            assert (x <= xs + 1e-15).all()
        else:
            assert l1 <= np.abs(r).sum() / ALPHA * (1 + 1e-9) + err_ref


def test_relabelled_graph_parity_small(small_graph):
    """A degree-relabelled device graph returns results in the caller's ids:
    FP64 local GD still matches the oracle's frontier sets and op count, and
    batched outputs map back (certified against the original graph)."""
    g = small_graph
    dgr = engine.DeviceGraph(g, relabel=True)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=ALPHA, eps=1e-6)
    for s in graph.uniform_seeds(g, 4, seed=30):
        o = O.ppr_gd(g, s, ALPHA, 1e-6, local=True, trace=True)
        out = engine.solve(dgr, p, int(s), trace=True)
        assert _compare_traces(out.trace, o["trace"]) == []
        assert out.result["oper"] == o["oper"]
        assert _rel_l1(out.x.cpu().numpy(), o["x"]) <= REL_L1
    seeds = graph.mixed_seeds(g, 48, seed=31)
    p32 = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=1e-6)
    for block in (256, 512, 1024):
        bo = engine.batch_solve(dgr, p32, seeds, block=block)
        _check_batch(g, seeds, bo, 1e-6)
    dgr.close()


def test_autotune_small(small_graph):
    """The autotuner measures its candidates, leaves the best applied, and the
    tuned launches still return certified results."""
    from flsg import autotune
    g = small_graph
    dg = engine.DeviceGraph(g)
    try:
        seeds = graph.mixed_seeds(g, 64, seed=37)
        p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=1e-6)
        out = autotune.tune_batched(dg, p, seeds, rounds=1)
        best = out["best"]
        assert out["qps"] > 0 and len(out["trials"]) >= 5
        assert best["block"] in (256, 512, 1024) and best["split_edges"] >= 64
        # a knob moves only for a gain above 3%: no trial beats the kept
        # setting by more than that, and the kept rate is a measured one
        assert max(t["qps"] for t in out["trials"]) <= out["qps"] * 1.03 * (1 + 1e-12)
        assert any(t["qps"] == out["qps"] and all(t[k] == v for k, v in best.items())
                   for t in out["trials"])
        bo = engine.batch_solve(dg, p, seeds)
        _check_batch(g, seeds, bo, 1e-6, n_check=8)
    finally:
        dg.close()


@pytest.mark.parametrize("hub_bytes", [136 * 1024])
def test_batch_hub_cache_beyond_120k_large(large_graph, dev_large, hub_bytes):
    """A hub cache past 120 KiB (up to the 160 KiB of LDS per workgroup less
    the kernel's static LDS) keeps every query certified; the whole-LDS
    maximum is accepted and one byte more is refused."""
    g = large_graph
    eps = 1e-7
    seeds = graph.mixed_seeds(g, 64, seed=21)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=eps)
    try:
        bo = engine.batch_solve(dev_large, p, seeds, hub_bytes=hub_bytes)
        _check_batch(g, seeds, bo, eps, n_check=6)
        assert dev_large.lib.flsg_batch_set_hub_cache(dev_large.handle, 160 * 1024) != 0
    finally:
        assert dev_large.lib.flsg_batch_set_hub_cache(dev_large.handle, -1) == 0


def test_batch_residual_l1_bounds_the_error(small_graph, dev_small):
    """Each query's r_l1 (device sum of |r_v| plus the FP32 error bounds)
    covers the host FP64 residual's L1 norm, and r_l1 / alpha covers the L1
    distance to the FP64 global solve (residual 1e-15)."""
    g = small_graph
    eps = 1e-6
    seeds = graph.mixed_seeds(g, 32, seed=13)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=eps)
    bo = engine.batch_solve(dev_small, p, seeds)
    for q in range(6):
        s = int(seeds[q])
        x = bo.dense(q, g.n)
        rl1 = float(bo.results[q]["r_l1"])
        rr = O.ppr_residual(g, s, ALPHA, x)
        assert np.abs(rr).sum() <= rl1 * (1 + 1e-9) + 1e-18, (q, np.abs(rr).sum(), rl1)
        xs = O.ppr_global_solve(g, s, ALPHA)
        assert np.abs(x - xs).sum() <= rl1 / ALPHA * (1 + 1e-9) + 1e-14
