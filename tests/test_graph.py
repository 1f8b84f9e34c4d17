"""Synthetic workloads and the reference's harness conventions."""
import numpy as np
import pytest

from flsg import graph


def _check_csr(g):
    assert g.indptr[0] == 0 and g.indptr[-1] == g.indices.size
    assert np.array_equal(np.diff(g.indptr), g.degree)
    rows = np.repeat(np.arange(g.n), g.degree)
    assert not np.any(rows == g.indices), "self-loop"
    # sorted, de-duplicated rows
    same_row = rows[1:] == rows[:-1]
    assert np.all(g.indices[1:][same_row] > g.indices[:-1][same_row])
    # symmetric
    fwd = np.sort(rows.astype(np.int64) * g.n + g.indices)
    bwd = np.sort(g.indices.astype(np.int64) * g.n + rows)
    assert np.array_equal(fwd, bwd)


def test_small_powerlaw_graph(small_graph):
    g = small_graph
    _check_csr(g)
    assert g.n == 10_000
    assert 40_000 <= g.m <= 70_000
    assert g.d_max <= 600
    assert g.degree.dtype == np.int32


def test_small_graph_deterministic():
    a = graph.powerlaw_graph_small()
    b = graph.powerlaw_graph_small()
    assert np.array_equal(a.indices, b.indices) and np.array_equal(a.indptr, b.indptr)


@pytest.mark.slow
def test_large_powerlaw_graph(large_graph):
    g = large_graph
    _check_csr(g)
    assert g.n == 2_394_385
    assert 9.0e6 <= g.m <= 9.6e6
    assert 80_000 <= g.d_max <= 130_000
    assert 0.08 <= (g.degree == 0).mean() <= 0.12


def test_dynamic_graph_snapshots_rebuild_full_graph():
    dg = graph.citation_dynamic_graph(n=3000, m_directed=20000, delnum=4000, snapshots=16, seed=3)
    live = dg.live_csr()
    assert dg.us.shape == (16, 250)
    # live + all snapshot edges == capacity of every row
    deg = dg.degree.astype(np.int64).copy()
    np.add.at(deg, dg.us.ravel(), 1)
    np.add.at(deg, dg.vs.ravel(), 1)
    assert np.array_equal(deg, np.diff(dg.indptr))
    # slots past the live degree hold the sentinel n
    for u in range(0, dg.n, 97):
        tail = dg.indices[dg.indptr[u] + dg.degree[u]:dg.indptr[u + 1]]
        assert np.all(tail == dg.n)
    # the live graph is connected-forest based: no isolated node that had edges
    assert (live.degree[np.diff(dg.indptr) > 0] > 0).all()


def test_demo_seed_draw():
    """demo.py:19,45-46: np.random.seed(17); np.unique(randint(n, 50))."""
    s = graph.demo_seeds(2_388_953)
    np.random.seed(17)
    want = np.unique(np.random.randint(2_388_953, size=50)).astype(np.int32)
    assert np.array_equal(s, want)
    assert s.dtype == np.int32 and np.all(np.diff(s) > 0)


def test_alpha_roundtrip():
    assert graph.alpha_roundtrip(0.1)[0] == 0.10000000000000002
    a, w = graph.alpha_roundtrip(0.2)
    assert a == 0.2 and abs(w - 1.25) < 1e-15
    assert abs(graph.alpha_roundtrip(0.1)[1] - 1.3928644583850192) < 1e-15


def test_mixed_seeds_non_isolated(small_graph):
    s = graph.mixed_seeds(small_graph, 1024, seed=0)
    assert s.size == 1024
    assert (small_graph.degree[s] > 0).all()


def test_npz_roundtrip(tmp_path, small_graph):
    p = str(tmp_path / "g.npz")
    graph.save_npz(p, small_graph)
    h = graph.load_npz(p)
    assert np.array_equal(h.indices, small_graph.indices) and np.array_equal(h.degree, small_graph.degree)


def test_exp_npz_layout_and_roundtrip(tmp_path):
    """The dynamic workload file has the members, dtypes and free-slot sentinel
    the reference's InstantGNN experiment reads (data_process_GNN.py:86-87,
    instantGNN-arxiv-exp.py:174-190), and reads back unchanged."""
    sg = graph.citation_dynamic_graph(n=400, m_directed=2400, delnum=300, snapshots=4, seed=5)
    feat = np.random.default_rng(0).standard_normal((sg.n, 8)).astype(np.float32)
    path = str(tmp_path / "dyn-exp.npz")
    graph.save_exp_npz(path, sg, feat)
    with np.load(path, allow_pickle=False) as d:
        assert set(d.files) == {"indices", "indptr", "degree", "Us", "Vs", "node_feature", "labels", "split_idx"}
        assert d["indices"].dtype == np.int32 and d["degree"].dtype == np.int64
        assert d["Us"].shape == d["Vs"].shape == (4, 75) and d["Us"].dtype == np.int32
        assert d["labels"].shape == (sg.n, 1)
        ip, idx, deg = d["indptr"], d["indices"], d["degree"]
        # slots past the live length hold n, as data_process_GNN.py:117-119 writes them
        for u in range(sg.n):
            assert np.all(idx[ip[u] + deg[u]:ip[u + 1]] == sg.n)
    # the reference reads split_idx through pickle; it is a dict of index arrays
    with np.load(path, allow_pickle=True) as d:
        split = d["split_idx"].reshape(1, -1)[0][0]
        assert sorted(split) == ["test", "train", "valid"]
        assert sum(v.size for v in split.values()) == sg.n
    g2, f2 = graph.load_exp_npz(path)
    assert g2.n == sg.n
    assert np.array_equal(g2.indptr, sg.indptr) and np.array_equal(g2.indices, sg.indices)
    assert np.array_equal(g2.degree, sg.degree)
    assert np.array_equal(g2.us, sg.us) and np.array_equal(g2.vs, sg.vs)
    assert np.array_equal(f2, feat)
