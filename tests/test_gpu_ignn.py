"""GPU InstantGNN propagation and in-place edge insertion.

Bars:
  * graph mutation (integer work): after a snapshot indices and degree equal
    the oracle's updetaGraph byte for byte, and the changed rows come back in
    the reference's insert_neighbor key order;
  * propagation: on return every entry satisfies |R| < eps d^(1-beta) in FP64
    AND R equals the host FP64 recompute X - P/alpha + ((1-alpha)/alpha) M P
    (relative 1e-9 of max|X|), which is the invariant the reference's push
    keeps; on small graphs P is also compared with the dense solution
    pi = alpha (I - (1-alpha) M)^-1 X through pi - P = alpha(...)^-1 R exactly;
  * the reference's FIFO push order is not reproduced, so P is compared with
    the oracle's P through that identity, not bitwise.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from flsg import _abi, graph  # noqa: E402
from flsg.ignn import DynamicPropagation, ignn_residual  # noqa: E402
from oracle import oracle as O  # noqa: E402

ALPHA, BETA = 0.1, 0.5


def _features(n, c, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, c))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    return X.astype(np.float32).astype(np.float64)


def _dense_K(indptr, indices, degree, alpha, beta):
    n = degree.size
    M = np.zeros((n, n))
    for u in range(n):
        for e in range(indptr[u], indptr[u] + degree[u]):
            v = indices[e]
            M[v, u] += 1.0 / (degree[v] ** beta * degree[u] ** (1 - beta))
    return alpha * np.linalg.inv(np.eye(n) - (1 - alpha) * M)


def _certify(prop, eps, X):
    ind, deg = prop.graph_arrays()
    P = prop.P.cpu().numpy()
    R = prop.R.cpu().numpy()
    thr = eps * deg.astype(np.float64) ** (1 - BETA)
    assert (np.abs(R) < thr[:, None]).all(), "certificate violated"
    Rh = ignn_residual(prop.dg.indptr.cpu().numpy().astype(np.int64), ind, deg, X, P, ALPHA, BETA)
    assert np.abs(Rh - R).max() <= 1e-9 * np.abs(X).max()
    return P, R


@pytest.mark.parametrize("precision", [_abi.FLSG_FP64, _abi.FLSG_FP32])
def test_ignn_initial_small_dense(precision):
    sg = graph.citation_dynamic_graph(n=400, m_directed=2400, delnum=300, snapshots=4, seed=5)
    X = _features(sg.n, 8)
    eps = 1e-6
    prop = DynamicPropagation(sg.copy(), X, ALPHA, eps, BETA, precision=precision)
    st = prop.initial()
    assert st.cert < 1.0
    P, R = _certify(prop, eps, X)
    K = _dense_K(sg.indptr, sg.indices, sg.degree, ALPHA, BETA)
    pi = K @ X
    assert np.abs(pi - P - K @ R).max() < 1e-9
    # the oracle's per-column FIFO push satisfies the same identity; the two
    # answers differ by K (R_gpu - R_oracle)
    for c in range(3):
        p_o, r_o, _ = O.ignn_column(sg.indptr, sg.indices, sg.degree, ALPHA, eps, BETA, X[:, c])
        diff = np.abs(P[:, c] - p_o.astype(np.float64) - K @ (r_o - R[:, c])).max()
        assert diff < 1e-5   # p_o is float32 in the reference's initial pass


def _mixed_batch(sg, rng):
    """A snapshot's pairs plus repeats in both orientations and pairs already
    in the graph, shuffled."""
    us, vs = sg.us[0], sg.vs[0]
    rep = rng.choice(us.size, 60, replace=False)
    rows = np.repeat(np.arange(sg.n), sg.degree)
    live = np.concatenate([sg.indices[sg.indptr[u]:sg.indptr[u] + sg.degree[u]] for u in range(sg.n)])
    ex = rng.choice(rows.size, 40, replace=False)
    bu = np.concatenate([us, us[rep[:30]], vs[rep[30:]], rows[ex]]).astype(np.int32)
    bv = np.concatenate([vs, vs[rep[:30]], us[rep[30:]], live[ex]]).astype(np.int32)
    perm = rng.permutation(bu.size)
    return bu[perm], bv[perm]


def test_insert_edges_byte_exact():
    """indices and degree equal the oracle's updetaGraph byte for byte (each
    row's new neighbours appended in batch order, repeats and present pairs
    skipped), and the endpoints come back in insert_neighbor key order."""
    sg = graph.citation_dynamic_graph(n=2000, m_directed=14000, delnum=3000, snapshots=4, seed=7)
    X = _features(sg.n, 4)
    prop = DynamicPropagation(sg.copy(), X, ALPHA, 1e-6, BETA, precision=_abi.FLSG_FP64)
    prop.initial()
    ind_o = sg.indices.copy()
    deg_o = sg.degree.copy()
    us, vs = _mixed_batch(sg, np.random.default_rng(3))
    keys, off, nbr = O.update_graph(sg.indptr, ind_o, deg_o, us, vs)
    st = prop.insert(us, vs)
    ind, deg = prop.graph_arrays()
    assert np.array_equal(deg, deg_o)
    assert np.array_equal(ind, ind_o)
    assert st.inserted == sg.us.shape[1]
    assert np.array_equal(prop.keys[:keys.size].cpu().numpy(), keys)


def test_updetagraph_dropin_returns_reference_dict():
    import InstantGNN
    from flsg.engine import FlsgError
    sg = graph.citation_dynamic_graph(n=1500, m_directed=9000, delnum=2000, snapshots=4, seed=8)
    ind_o, deg_o = sg.indices.copy(), sg.degree.copy()
    ind, deg = sg.indices.copy(), sg.degree.copy()
    us, vs = _mixed_batch(sg, np.random.default_rng(4))
    keys, off, nbr = O.update_graph(sg.indptr, ind_o, deg_o, us, vs)
    ins, dele = InstantGNN.updetaGraph(sg.indptr, ind, deg, us, vs)
    assert dele == {}
    assert list(ins.keys()) == keys.tolist()
    assert all(ins[int(k)] == nbr[off[i]:off[i + 1]].tolist() for i, k in enumerate(keys))
    assert np.array_equal(ind, ind_o) and np.array_equal(deg, deg_o)
    # a self-loop the reference would insert fails and changes nothing
    free = int(np.flatnonzero(deg < np.diff(sg.indptr))[0])
    before = (ind.copy(), deg.copy())
    with pytest.raises(FlsgError):
        InstantGNN.updetaGraph(sg.indptr, ind, deg, np.array([0, free]), np.array([1, free]))
    assert np.array_equal(ind, before[0]) and np.array_equal(deg, before[1])
    with pytest.raises(ValueError):
        InstantGNN.updetaGraph(sg.indptr, ind, deg, np.array([0]), np.array([sg.n]))


@pytest.mark.parametrize("precision", [_abi.FLSG_FP64, _abi.FLSG_FP32])
def test_ignn_snapshots_small_dense(precision):
    """After every snapshot the state is the propagation of X on the new graph."""
    sg = graph.citation_dynamic_graph(n=400, m_directed=2400, delnum=300, snapshots=4, seed=9)
    X = _features(sg.n, 8, seed=1)
    eps = 1e-6
    prop = DynamicPropagation(sg.copy(), X, ALPHA, eps, BETA, precision=precision)
    prop.initial()
    for k in range(sg.us.shape[0]):
        st = prop.insert(sg.us[k], sg.vs[k])
        assert st.cert < 1.0
        P, R = _certify(prop, eps, X)
        ind, deg = prop.graph_arrays()
        K = _dense_K(sg.indptr, ind, deg, ALPHA, BETA)
        assert np.abs(K @ X - P - K @ R).max() < 1e-9


def test_ignn_update_matches_oracle_update():
    """Same snapshot through the oracle's instantGNNUpdateBatch: both end with
    the invariant on the new graph, so their P differ by K (R_a - R_b)."""
    sg = graph.citation_dynamic_graph(n=300, m_directed=1800, delnum=200, snapshots=2, seed=11)
    X = _features(sg.n, 2, seed=2)
    eps = 1e-7
    prop = DynamicPropagation(sg.copy(), X, ALPHA, eps, BETA, precision=_abi.FLSG_FP64)
    prop.initial()
    P0 = prop.P.cpu().numpy().copy()
    R0 = prop.R.cpu().numpy().copy()
    prop.insert(sg.us[0], sg.vs[0])
    P1, R1 = prop.P.cpu().numpy(), prop.R.cpu().numpy()
    ind_o, deg_o = sg.indices.copy(), sg.degree.copy()
    old = deg_o.copy()
    keys, off, nbr = O.update_graph(sg.indptr, ind_o, deg_o, sg.us[0], sg.vs[0])
    K = _dense_K(sg.indptr, ind_o, deg_o, ALPHA, BETA)
    for c in range(2):
        p = np.ascontiguousarray(P0[:, c])
        r = np.ascontiguousarray(R0[:, c])
        O.ignn_update_column(sg.indptr, ind_o, deg_o, old, p, r, X[:, c], ALPHA, eps, BETA, keys, off, nbr)
        assert np.abs(P1[:, c] - p - K @ (r - R1[:, c])).max() < 1e-10


@pytest.mark.slow
def test_ignn_128_columns_full_size():
    """169,343 nodes, 128 columns, eps = 1e-7: the initial propagation and all
    16 insert snapshots of the workload (59,375 pairs each).  After every
    snapshot: indices, degree and the insert_neighbor key order equal the
    oracle's updetaGraph byte for byte, and P / R are certified in FP64 and
    equal to the host recompute."""
    sg = graph.citation_dynamic_graph()
    X = _features(sg.n, 128, seed=3)
    eps = 1e-7
    prop = DynamicPropagation(sg.copy(), X, ALPHA, eps, BETA, precision=_abi.FLSG_FP32)
    st = prop.initial()
    assert st.cert < 1.0
    _certify(prop, eps, X)
    ind_o = sg.indices.copy()
    deg_o = sg.degree.copy()
    for k in range(sg.us.shape[0]):
        keys, _, _ = O.update_graph(sg.indptr, ind_o, deg_o, sg.us[k], sg.vs[k])
        st = prop.insert(sg.us[k], sg.vs[k])
        assert st.cert < 1.0 and st.inserted == sg.us.shape[1], k
        ind, deg = prop.graph_arrays()
        assert np.array_equal(deg, deg_o) and np.array_equal(ind, ind_o), k
        assert np.array_equal(prop.keys[:keys.size].cpu().numpy(), keys), k
        _certify(prop, eps, X)
