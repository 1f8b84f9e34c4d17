"""The result checkers of the oracle (per-query L1 distance to the FP64 local
GD, and the FP64 residual certificate recomputed from the graph) against
dense scipy computations of the same quantities on the 10,000-node graph."""
import numpy as np
import pytest
import scipy.sparse as sp

from flsg import graph
from oracle import oracle as O

ALPHA = 0.1


@pytest.fixture(scope="module")
def g():
    return graph.powerlaw_graph_small()


def _adj(g):
    n = g.n
    rows = np.repeat(np.arange(n), np.diff(g.indptr))
    return sp.csr_matrix((np.ones(rows.size), (rows, g.indices)), shape=(n, n))


def _sparse(xs):
    ids, vals, off, cnt = [], [], [], []
    o = 0
    for x in xs:
        nz = np.flatnonzero(x)
        ids.append(nz.astype(np.int32))
        vals.append(x[nz])
        off.append(o)
        cnt.append(nz.size)
        o += nz.size
    return np.array(off), np.array(cnt), np.concatenate(ids), np.concatenate(vals)


def test_residual_cert_matches_dense(g):
    A = _adj(g)
    d = g.degree.astype(np.float64)
    rng = np.random.default_rng(3)
    seeds = graph.uniform_seeds(g, 4, seed=9)
    eps = 1e-6
    xs = [O.ppr_gd(g, int(s), ALPHA, eps, local=True)["x"] for s in seeds]
    # a perturbed vector too: the checker sees any x, not only converged ones
    xp = xs[0].copy()
    nz = np.flatnonzero(xp)
    xp[nz] *= 1.0 + 0.01 * rng.standard_normal(nz.size)
    xs.append(xp)
    sd = list(seeds) + [seeds[0]]
    cert, rl1 = O.residual_cert(g, 0, ALPHA, eps, sd, *_sparse(xs))
    dinv = np.where(d > 0, 1.0 / np.maximum(d, 1), 0.0)
    for q, (s, x) in enumerate(zip(sd, xs)):
        b = np.zeros(g.n)
        b[s] = ALPHA
        r = b - x + (1 - ALPHA) * (A.T @ (dinv * x))
        thr = eps * ALPHA * d
        m = np.abs(r) > 0
        ref = np.max(np.abs(r[m]) / thr[m])
        assert cert[q] == pytest.approx(ref, rel=1e-9)
        assert rl1[q] == pytest.approx(np.abs(r).sum(), rel=1e-9)
    assert np.all(cert[:4] < 1.0)   # the oracle's own results are certified
    assert cert[4] > 1.0            # the perturbed one is not


def test_residual_cert_katz_matches_dense(g):
    A = _adj(g)
    d = g.degree.astype(np.float64)
    beta = 1.0 / (g.d_max + 1)
    eps = 1e-7
    seeds = graph.uniform_seeds(g, 3, seed=4)
    xs = []
    for s in seeds:
        x = O.katz_gd(g, int(s), beta, eps, local=True)["x"].copy()
        x[s] += 1.0   # the full solution: katz_gd returns x - e_s
        xs.append(x)
    cert, rl1 = O.residual_cert(g, 1, beta, eps, seeds, *_sparse(xs))
    for q, (s, x) in enumerate(zip(seeds, xs)):
        b = np.zeros(g.n)
        b[s] = 1.0
        r = b - x + beta * (A.T @ x)
        m = np.abs(r) > 0
        assert cert[q] == pytest.approx(np.max(np.abs(r[m]) / (eps * d[m])), rel=1e-9)
        assert rl1[q] == pytest.approx(np.abs(r).sum(), rel=1e-9)
    assert np.all(cert < 1.0)


def test_batch_l1_matches_dense(g):
    rng = np.random.default_rng(5)
    seeds = graph.uniform_seeds(g, 5, seed=11)
    eps = 1e-6
    xs, ys = [], []
    for s in seeds:
        x = O.ppr_gd(g, int(s), ALPHA, eps, local=True)["x"]
        y = x.copy()
        nz = np.flatnonzero(y)
        y[nz] *= 1.0 + 1e-3 * rng.standard_normal(nz.size)
        y[nz[: nz.size // 10]] = 0.0                       # entries the result misses
        extra = rng.choice(g.n, 20, replace=False)
        y[extra] += 1e-9                                   # entries only the result has
        xs.append(x)
        ys.append(y)
    l1, norm, opers = O.ppr_batch_l1(g, seeds, ALPHA, eps, *_sparse(ys))
    for q in range(len(seeds)):
        assert l1[q] == pytest.approx(np.abs(xs[q] - ys[q]).sum(), rel=1e-9)
        assert norm[q] == pytest.approx(np.abs(xs[q]).sum(), rel=1e-12)
        assert opers[q] == O.ppr_gd(g, int(seeds[q]), ALPHA, eps, local=True)["oper"]
