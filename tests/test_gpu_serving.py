"""Serving loop: continuous batching on the GPU.

Bars: every PPR / Katz result is certified (|r_v| < thr_v in FP64, recomputed
on the host on the graph current at its launch); a query's result is the same
alone and in the mix -- for PPR / Katz within the two certified residuals
(||x_a - x_b||_1 <= (||r_a||_1 + ||r_b||_1) / alpha), for the heat kernel
within FP32 rounding (relative L1 1e-5); edge-insert snapshots land in place.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from flsg import _abi, engine, graph, serving  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _dense(r, n):
    x = np.zeros(n)
    ids, vals = r.result
    x[ids] = vals
    return x


def _katz_residual(g, s, beta, x):
    A = O.adjacency(g)
    es = np.zeros(g.n)
    es[s] = 1.0
    return es - (x - beta * (A @ x))


@pytest.fixture(scope="module")
def srv_small(small_graph):
    srv = serving.Server(small_graph, slots=24, block=256)
    yield srv
    srv.close()


def test_serving_mix_certified_and_same_alone(small_graph, srv_small):
    g = small_graph
    eps = 1e-6
    reqs = serving.mixed_requests(g, rate=2000.0, count=120, seed=3, eps=eps)
    st = srv_small.run(reqs, collect=True)
    assert st["ok"] == len(reqs) and st["certified"]
    # continuous admission: one persistent launch took requests published
    # while it ran (no launch boundary between them)
    assert st["launches"] == 1 and st["submits"] >= 2 and st["idle_exits"] == 0
    # every launch (PPR, Katz and heat-kernel queries mixed in the same slots)
    # left every slot array in its between-queries state
    srv_small.torch.cuda.synchronize()
    assert engine.arena_dirty(srv_small.dg) == 0
    for r in reqs:
        x = _dense(r, g.n)
        if r.kind == _abi.FLSG_PPR:
            rr = O.ppr_residual(g, r.seed, r.alpha, x)
            assert (np.abs(rr) / (eps * r.alpha * np.maximum(g.degree, 1))).max() < 1 + 1e-9
        elif r.kind == _abi.FLSG_KATZ:
            rr = _katz_residual(g, r.seed, r.alpha, x)
            assert (np.abs(rr) / (eps * np.maximum(g.degree, 1))).max() < 1 + 1e-9
    # alone vs in the mix
    picks = {k: next(r for r in reqs if r.kind == k) for k in (_abi.FLSG_PPR, _abi.FLSG_KATZ, _abi.FLSG_HK)}
    for k, r in picks.items():
        alone = serving.Request(rid=0, kind=r.kind, seed=r.seed, arrival=0.0, alpha=r.alpha, eps=r.eps)
        srv_small.run([alone], collect=True)
        a, b = _dense(alone, g.n), _dense(r, g.n)
        if k == _abi.FLSG_HK:
            assert np.abs(a - b).sum() <= 1e-5 * np.abs(a).sum()
        else:
            res = O.ppr_residual if k == _abi.FLSG_PPR else _katz_residual
            ra = np.abs(res(g, r.seed, r.alpha, a)).sum()
            rb = np.abs(res(g, r.seed, r.alpha, b)).sum()
            scale = r.alpha if k == _abi.FLSG_PPR else 1.0 - r.alpha * g.d_max
            assert np.abs(a - b).sum() <= (ra + rb) / scale * (1 + 1e-9) + 1e-15


def test_serving_rejects_bad_query(small_graph, srv_small):
    """A request the kernel cannot run (PPR alpha outside (0, 1), eps <= 0)
    fails alone with FLSG_EARG; the requests launched beside it are served."""
    g = small_graph
    s0, s1 = (int(v) for v in graph.uniform_seeds(g, 2, seed=43))
    reqs = [serving.Request(rid=0, kind=_abi.FLSG_PPR, seed=s0, arrival=0.0, alpha=1.5, eps=1e-6),
            serving.Request(rid=1, kind=_abi.FLSG_PPR, seed=s1, arrival=0.0, alpha=0.1, eps=1e-6),
            serving.Request(rid=2, kind=_abi.FLSG_KATZ, seed=s1, arrival=0.0, alpha=1e-4, eps=0.0)]
    srv_small.run(reqs, collect=True)
    assert reqs[0].status == _abi.FLSG_EARG and reqs[2].status == _abi.FLSG_EARG
    assert reqs[1].status == _abi.FLSG_OK and reqs[1].cert < 1.0
    x = _dense(reqs[1], g.n)
    rr = O.ppr_residual(g, s1, 0.1, x)
    assert (np.abs(rr) / (1e-6 * 0.1 * np.maximum(g.degree, 1))).max() < 1 + 1e-9


def test_serving_hk_matches_oracle(small_graph, srv_small):
    """Batched FP32 heat kernel vs the oracle's FP64 level push (same N, same
    thresholds): equal up to FP32 rounding."""
    g = small_graph
    reqs = [serving.Request(rid=i, kind=_abi.FLSG_HK, seed=int(s), arrival=0.0)
            for i, s in enumerate(graph.uniform_seeds(g, 6, seed=8))]
    srv_small.run(reqs, collect=True)
    for r in reqs:
        o = O.hk_push(g, r.seed, 5.0, 1e-7, local=True, n_terms=40)
        x = _dense(r, g.n)
        assert np.abs(x - o["x"]).sum() <= 1e-5 * np.abs(o["x"]).sum()
        assert abs(r.oper - o["oper"]) <= 0.02 * o["oper"] + 50


def test_serving_inserts_in_place(small_graph):
    g = small_graph
    srv = serving.Server(g, slots=16, block=256, headroom=6)
    try:
        eps = 1e-6
        ins = serving.random_insert_batches(g, every=0.01, size=300, count=2, seed=5)
        reqs = serving.mixed_requests(g, rate=3000.0, count=60, seed=4, eps=eps, mix=(1.0, 0.0, 0.0))
        st = srv.run(reqs, inserts=ins, collect=True)
        assert st["inserts_applied"] == 2 and st["edges_inserted"] > 0
        # queries served after the last insert are certified on the final graph
        u = np.concatenate([g.indices.astype(np.int64)] + [b.us for b in ins] + [b.vs for b in ins])
        rows = np.concatenate([np.repeat(np.arange(g.n), g.degree)] + [b.vs for b in ins] + [b.us for b in ins])
        g2 = graph.CSRGraph.from_edges(g.n, rows, u)
        last = max(b.applied for b in ins)
        late = [r for r in reqs if r.arrival > last + 1e-3 and r.done > 0]
        for r in late[:10]:
            x = _dense(r, g.n)
            rr = O.ppr_residual(g2, r.seed, r.alpha, x)
            assert (np.abs(rr) / (eps * r.alpha * np.maximum(g2.degree, 1))).max() < 1 + 1e-9
        ind = srv.dg.indices.cpu().numpy()
        deg = srv.dg.degree.cpu().numpy()
        assert int(deg.sum()) == g2.m
    finally:
        srv.close()


@pytest.mark.parametrize("hub_bytes", [0, -1])
def test_serving_hk_and_ppr_with_and_without_hub_cache(small_graph, hub_bytes):
    """Batched heat kernel and PPR with the LDS hub cache off (every level
    residual in the node records) and on (the small graph fits in it whole):
    heat kernel equal to the oracle up to FP32 rounding, PPR certified."""
    g = small_graph
    srv = serving.Server(g, slots=8, block=256)
    try:
        assert srv.lib.flsg_batch_set_hub_cache(srv.dg.handle, hub_bytes) == 0
        seeds = graph.uniform_seeds(g, 4, seed=17)
        reqs = [serving.Request(rid=i, kind=_abi.FLSG_HK, seed=int(s), arrival=0.0) for i, s in enumerate(seeds)]
        reqs += [serving.Request(rid=10 + i, kind=_abi.FLSG_PPR, seed=int(s), arrival=0.0, eps=1e-6)
                 for i, s in enumerate(seeds)]
        st = srv.run(reqs, collect=True)
        assert st["ok"] == len(reqs) and st["certified"]
        for r in reqs:
            x = _dense(r, g.n)
            if r.kind == _abi.FLSG_HK:
                o = O.hk_push(g, r.seed, 5.0, 1e-7, local=True, n_terms=40)
                assert np.abs(x - o["x"]).sum() <= 1e-5 * np.abs(o["x"]).sum()
                assert abs(r.oper - o["oper"]) <= 0.02 * o["oper"] + 50
            else:
                rr = O.ppr_residual(g, r.seed, r.alpha, x)
                assert (np.abs(rr) / (1e-6 * r.alpha * np.maximum(g.degree, 1))).max() < 1 + 1e-9
    finally:
        srv.close()


@pytest.mark.parametrize("precision", [_abi.FLSG_FP32, _abi.FLSG_FP64])
def test_serving_hk_large(large_graph, precision):
    """Full-size graph (2.39M nodes, t = 5, order 40, eps = 1e-7): batched heat
    kernel (hubs in LDS, the rest in node records) against the oracle's FP64
    level push, 3 seeds.

    Bars: FP64 levels -- relative L1 1e-11 (summation order only) and op counts
    within 0.1% (a node whose level residual lands within rounding of its
    threshold may flip).  FP32 levels -- a hub's level residual collects up to
    ~10^3 FP32 additions per level, each rounding by at most 2^-24 of the
    partial sum, so its relative error is bounded by ~10^3 * 6e-8 = 6e-5: the
    bar is relative L1 1e-4 (observed ~1e-5), op counts within 2%."""
    g = large_graph
    srv = serving.Server(g, slots=4, precision=precision)
    # the dense levels take the binned push even with few slots busy
    assert srv.lib.flsg_batch_set_bin_load(srv.dg.handle, 1) == 0
    fp64 = precision == _abi.FLSG_FP64
    try:
        seeds = graph.uniform_seeds(g, 3, seed=23)
        reqs = [serving.Request(rid=i, kind=_abi.FLSG_HK, seed=int(s), arrival=0.0) for i, s in enumerate(seeds)]
        st = srv.run(reqs, collect=True)
        assert st["ok"] == len(reqs)
        for r in reqs:
            o = O.hk_push(g, r.seed, 5.0, 1e-7, local=True, n_terms=40)
            x = _dense(r, g.n)
            rel = np.abs(x - o["x"]).sum() / np.abs(o["x"]).sum()
            assert rel <= (1e-11 if fp64 else 1e-4), f"seed {r.seed}: rel L1 {rel}"
            assert abs(r.oper - o["oper"]) <= (1e-3 if fp64 else 0.02) * o["oper"] + 50
    finally:
        srv.close()


def test_serving_hk_fp64_small(small_graph):
    """Batched FP64 heat kernel on the small graph: equal to the oracle's level
    push up to summation order (relative L1 1e-12), the same operation count."""
    g = small_graph
    srv = serving.Server(g, slots=8, block=256, precision=_abi.FLSG_FP64)
    try:
        seeds = graph.uniform_seeds(g, 6, seed=29)
        reqs = [serving.Request(rid=i, kind=_abi.FLSG_HK, seed=int(s), arrival=0.0) for i, s in enumerate(seeds)]
        st = srv.run(reqs, collect=True)
        assert st["ok"] == len(reqs)
        for r in reqs:
            o = O.hk_push(g, r.seed, 5.0, 1e-7, local=True, n_terms=40)
            x = _dense(r, g.n)
            assert np.abs(x - o["x"]).sum() <= 1e-12 * np.abs(o["x"]).sum()
            assert r.oper == int(o["oper"])
    finally:
        srv.close()


@pytest.mark.slow
def test_serving_mix_large_with_insert(large_graph):
    """The serving mix on the 2.39M-node graph (eps = 1e-7): PPR, Katz and heat
    kernel queries arriving over time, a 10,000-edge insert batch in the
    middle.  Every PPR / Katz result is certified on the graph its session
    ran on (the one before the insert for queries done before it, the one
    after for the rest), heat-kernel results equal the oracle's level push up
    to FP32 rounding, and one query of each kind served alone afterwards
    equals its result in the mix (PPR / Katz within the two certified
    residuals, heat kernel within relative L1 1e-4)."""
    g = large_graph
    eps = 1e-7
    srv = serving.Server(g)
    try:
        reqs = serving.mixed_requests(g, rate=300.0, count=40, seed=11, eps=eps)
        mid = reqs[len(reqs) // 2].arrival
        ins = serving.random_insert_batches(g, every=mid, size=10_000, count=1, seed=13)
        st = srv.run(reqs, inserts=ins, collect=True)
        print(f"\n  large mix: {st}")
        assert st["ok"] == len(reqs) and st["certified"] and st["idle_exits"] == 0
        assert st["inserts_applied"] == 1 and st["edges_inserted"] > 9_000
        b = ins[0]
        u = np.concatenate([g.indices.astype(np.int64), b.us, b.vs])
        rows = np.concatenate([np.repeat(np.arange(g.n), g.degree), b.vs, b.us])
        g2 = graph.CSRGraph.from_edges(g.n, rows, u)
        assert int(srv.dg.degree.cpu().numpy().sum()) == g2.m
        ops = {}

        def residual(gg, r, x):
            key = (id(gg), r.kind, r.alpha)
            if key not in ops:
                ops[key] = O.ppr_operator(gg, r.alpha) if r.kind == _abi.FLSG_PPR else None
            if r.kind == _abi.FLSG_PPR:
                bb = np.zeros(gg.n)
                bb[r.seed] = r.alpha
                return bb - ops[key] @ x
            return _katz_residual(gg, r.seed, r.alpha, x)

        before = [r for r in reqs if r.done < b.applied]
        after = [r for r in reqs if r.done >= b.applied]
        assert before and after
        n_hk = 0
        for r in reqs:
            gg = g if r.done < b.applied else g2
            x = _dense(r, g.n)
            if r.kind == _abi.FLSG_HK:
                if n_hk < 2:
                    o = O.hk_push(gg, r.seed, 5.0, 1e-7, local=True, n_terms=40)
                    assert np.abs(x - o["x"]).sum() <= 1e-4 * np.abs(o["x"]).sum()
                    n_hk += 1
                continue
            rr = residual(gg, r, x)
            thr = eps * (r.alpha if r.kind == _abi.FLSG_PPR else 1.0) * np.maximum(gg.degree, 1)
            assert (np.abs(rr) / thr).max() < 1 + 1e-9, (r.kind, r.seed)
        # alone (on the final graph) vs in the mix
        for k in (_abi.FLSG_PPR, _abi.FLSG_KATZ, _abi.FLSG_HK):
            r = next((q for q in after if q.kind == k), None)
            if r is None:
                continue
            alone = serving.Request(rid=0, kind=r.kind, seed=r.seed, arrival=0.0, alpha=r.alpha, eps=r.eps)
            srv.run([alone], collect=True)
            a, bx = _dense(alone, g.n), _dense(r, g.n)
            if k == _abi.FLSG_HK:
                assert np.abs(a - bx).sum() <= 1e-4 * np.abs(a).sum()
            else:
                ra = np.abs(residual(g2, r, a)).sum()
                rb = np.abs(residual(g2, r, bx)).sum()
                scale = r.alpha if k == _abi.FLSG_PPR else 1.0 - r.alpha * g2.d_max
                assert np.abs(a - bx).sum() <= (ra + rb) / scale * (1 + 1e-9) + 1e-15
        assert engine.arena_dirty(srv.dg) == 0
    finally:
        srv.close()
