"""Pins of the oracle's Chebyshev, Katz GD and heat-kernel restatements.

Each check is chosen so that a plausible slip in the restatement fails it:

* local Chebyshev (PPR and Katz): exact-rational hand traces written as
  literals -- the queue of every iteration, the operation count and exact x
  and r (derived step by step in ``fractions.Fraction`` by
  ``tests/golden/cheb_hand_trace.py``).  The graphs make each of these slips
  change the queues: the previous iteration's coefficients, always or never
  sorting the queue (``rear > n/4``), dropping a node's stale dx when it
  re-enters, testing u after ``r_u -= dx_u``;
* global Chebyshev (PPR and Katz): after K sweeps x equals the Chebyshev
  semi-iterate of order K+1 for the spectral interval, computed from the
  closed form of its residual polynomial, r_k = T_k((theta - Q)/delta) b /
  T_k(theta/delta) (no coefficient recurrence involved);
* Katz GD: a dense Jacobi replay with the thresholds eps*d_v (not
  eps*beta*d_v) gives every iteration's frontier set and the operation count;
* heat-kernel level push: a dense level-by-level replay with psi_j from its
  closed form sum_m t^m j!/(j+m)! and the thresholds e^t eps d/(N psi_j) gives
  every level's pushed set, the operation count and x; and the result meets
  the hk-relax guarantee ||D^-1 (e^{-t} exp(tP) e_s - x)||_inf < eps against
  scipy's expm.
"""
import math
from fractions import Fraction as F

import numpy as np
import pytest
import scipy.linalg as sla

from flsg.graph import CSRGraph
from oracle import oracle as O
from tests.golden import cheb_hand_trace as H
from tests.test_oracle import random_graph


def _graph(n, edges):
    u = np.array([a for a, _ in edges])
    v = np.array([b for _, b in edges])
    return CSRGraph.from_edges(n, u, v)


# ------------------------------------------------------------------ local Chebyshev, hand traces
PPR_TRACE = [[1, 2], [0, 2, 3], [0, 1, 2, 3], [1, 2, 0], [0]]
PPR_OPER = 32
PPR_X = ['147500/255339', '24211/136836', '1499/7602', '3/97', '0', '0', '0', '0', '0', '0', '0', '0']
PPR_R = ['-418279/739461744', '31943/92432718', '-80833609/23909263056', '8567/4424364', '1/194', '1/194', '0',
         '0', '0', '0', '0', '0']
KATZ_TRACE = [[1, 6], [0, 2, 5, 7], [1, 6, 3, 4], [0, 2, 4, 5, 7], [1, 6, 3, 5], [0, 2, 4, 5, 7], [1, 6], [7]]
KATZ_OPER = 54
KATZ_X = ['267405/229411', '51738469/166097659', '41669/458822', '591/17563', '219853/6357962',
          '15667103/166093564', '20136183079/60127352558', '18533976655/224398334524', '0', '0', '0', '0', '0', '0',
          '0', '0']
KATZ_R = ['-2436695384191/606324223194872', '1250896497/478551083816', '-732383077035/161630628558854',
          '-9883023/4294803331', '-32149686573/12437750446576', '-148483086571/75790527899359',
          '15794814848344925/21289862449041540536', '9534119018657/8445006921476216', '0', '0', '0', '0', '0', '0',
          '0', '0']


def _exact(vals):
    return np.array([float(F(v)) for v in vals])


def test_cheb_hand_trace_derivation():
    """The literals are what the exact-rational derivation prints."""
    tr, op, x, r = H.ppr_case()
    assert tr == PPR_TRACE and op == PPR_OPER
    assert [str(v) for v in x] == PPR_X and [str(v) for v in r] == PPR_R
    tr, op, x, r = H.katz_case()
    assert tr == KATZ_TRACE and op == KATZ_OPER
    assert [str(v) for v in x] == KATZ_X and [str(v) for v in r] == KATZ_R


def test_ppr_cheby_local_hand_trace():
    """Two triangles joined by an edge plus six isolated nodes (n = 12), seed 0,
    alpha = 1/2, eps = 1/128.  Queues per iteration: [1,2], [0,2,3],
    [0,1,2,3] (rear 4 > n/4: ascending), [1,2,0] (rear 3: queue order), [0];
    residuals of both signs at the end."""
    g = _graph(H.PPR_N, H.PPR_EDGES)
    o = O.ppr_cheby(g, H.PPR_S, float(H.PPR_ALPHA), float(H.PPR_EPS), local=True, trace=True)
    assert [t.tolist() for t in o["trace"]] == PPR_TRACE
    assert o["oper"] == PPR_OPER
    assert np.abs(o["x"] - _exact(PPR_X)).max() < 1e-15
    assert np.abs(o["r"] - _exact(PPR_R)).max() < 1e-15


def test_katz_cheby_local_hand_trace():
    """A 7-cycle (nodes 0-6) with a pendant node 7 on node 6 and eight isolated
    nodes (n = 16), seed 0, beta = 1/4,
    l_max = 3/2, mu_min = 1/2, eps = 1/256: queues [1,6,3,4] and [1,6,3,5] keep
    their order (rear 4 = n/4), [0,2,4,5,7] is sorted (rear 5 > n/4)."""
    g = _graph(H.KATZ_N, H.KATZ_EDGES)
    o = O.katz_cheby(g, H.KATZ_S, float(H.KATZ_BETA), float(H.KATZ_L), float(H.KATZ_MU), float(H.KATZ_EPS),
                     local=True, trace=True)
    assert [t.tolist() for t in o["trace"]] == KATZ_TRACE
    assert o["oper"] == KATZ_OPER
    assert np.abs(o["x"] - _exact(KATZ_X)).max() < 1e-15
    assert np.abs(o["r"] - _exact(KATZ_R)).max() < 1e-15


# ------------------------------------------------------------------ global Chebyshev, closed form
def _chebyshev_iterate(Q, b, lo, hi, k):
    """x_k of the Chebyshev semi-iteration from x_0 = 0 for Q x = b with the
    spectrum in [lo, hi]: its residual is R_k(Q) b with
    R_k(l) = T_k((theta - l) / delta) / T_k(theta / delta)."""
    theta, delta = (hi + lo) / 2.0, (hi - lo) / 2.0
    Z = (theta * np.eye(Q.shape[0]) - Q) / delta
    sigma = theta / delta
    y0, y1 = b.copy(), Z @ b
    t0, t1 = 1.0, sigma
    for _ in range(k - 1):
        y0, y1 = y1, 2.0 * (Z @ y1) - y0
        t0, t1 = t1, 2.0 * sigma * t1 - t0
    r = y1 / t1 if k >= 1 else b
    return np.linalg.solve(Q, b - r), r


@pytest.mark.parametrize("seed,eps", [(0, 1e-4), (3, 1e-6), (5, 1e-9)])
def test_ppr_cheby_global_is_chebyshev_semi_iteration(seed, eps):
    g = random_graph(seed=seed)
    s, alpha = 2, 0.15
    o = O.ppr_cheby(g, s, alpha, eps, local=False)
    sweeps = o["oper"] // (g.n + g.m)
    assert sweeps * (g.n + g.m) == o["oper"] and sweeps >= 2
    Q = O.ppr_operator(g, alpha).toarray()
    b = np.zeros(g.n)
    b[s] = alpha
    # the initial push is the first Chebyshev step (x_1 = b / theta, theta = 1)
    x, r = _chebyshev_iterate(Q, b, alpha, 2.0 - alpha, sweeps + 1)
    assert np.abs(o["x"] - x).max() <= 1e-13 * np.abs(x).max()
    assert np.abs(o["r"] - r).max() <= 1e-13 * alpha
    # one sweep fewer or more is far off
    for k in (sweeps, sweeps + 2):
        xk, _ = _chebyshev_iterate(Q, b, alpha, 2.0 - alpha, k)
        assert np.abs(o["x"] - xk).max() > 1e3 * 1e-13 * np.abs(x).max()


@pytest.mark.parametrize("seed,eps", [(8, 1e-6), (4, 1e-10)])
def test_katz_cheby_global_is_chebyshev_semi_iteration(seed, eps):
    g = random_graph(seed=seed)
    s = 1
    A = O.adjacency(g).toarray()
    ev = np.linalg.eigvalsh(A)
    beta = 0.5 / np.abs(ev).max()
    l_max, mu_min = 1 - beta * ev.min(), 1 - beta * ev.max()
    o = O.katz_cheby(g, s, beta, l_max, mu_min, eps, local=False)
    sweeps = o["oper"] // (g.n + g.m)
    assert sweeps * (g.n + g.m) == o["oper"] and sweeps >= 2
    Q = np.eye(g.n) - beta * A
    b = np.zeros(g.n)
    b[s] = 1.0
    x, r = _chebyshev_iterate(Q, b, mu_min, l_max, sweeps + 1)
    assert np.abs(o["x"] - x).max() <= 1e-13 * np.abs(x).max()
    assert np.abs(o["r"] - r).max() <= 1e-13
    xk, _ = _chebyshev_iterate(Q, b, mu_min, l_max, sweeps)
    assert np.abs(o["x"] - xk).max() > 1e-10


# ------------------------------------------------------------------ Katz GD, dense Jacobi replay
@pytest.mark.parametrize("seed", [7, 21])
def test_katz_gd_local_frontier_replay(seed):
    """Pushes are >= 0, so the frontier after each sweep is exactly
    {v : r_v >= eps * d_v} (katz_solver.py:179, 216); the first frontier is
    the seed whatever its residual (:185-188)."""
    g = random_graph(n=60, p=0.08, seed=seed)
    A = O.adjacency(g).toarray()
    lam = float(np.abs(np.linalg.eigvalsh(A)).max())
    beta, eps, s = 0.6 / lam, 5e-4, 0
    o = O.katz_gd(g, s, beta, eps, local=True, trace=True)
    d = g.degree.astype(float)
    r = np.zeros(g.n)
    r[s] = 1.0
    x = np.zeros(g.n)
    front = np.array([s])
    total = 0
    for t, fr in enumerate(o["trace"]):
        assert set(fr.tolist()) == set(front.tolist()), t
        total += int(d[front].sum())
        push = np.zeros(g.n)
        push[front] = r[front]
        x[front] += r[front]
        r[front] = 0.0
        r = r + beta * (A @ push)
        front = np.flatnonzero(r >= eps * d)
    assert front.size == 0
    assert total == o["oper"]
    x[s] -= 1.0
    assert np.abs(o["x"] - x).max() < 1e-14
    # the rule matters here: a beta-scaled threshold changes the frontiers
    assert len(o["trace"]) >= 4


# ------------------------------------------------------------------ heat kernel, level replay
def _psi_closed(N, t):
    return [sum(t ** m * math.factorial(j) / math.factorial(j + m) for m in range(N - j + 1)) for j in range(N + 1)]


@pytest.mark.parametrize("seed,t,eps", [(11, 5.0, 3e-3), (13, 5.0, 3e-4), (10, 3.0, 1e-2), (11, 2.0, 3e-3),
                                        (12, 3.0, 1e-2)])
def test_hk_push_level_replay(seed, t, eps):
    """Level residuals only grow (omega = 1), so level j pushes exactly
    {u : r_j,u >= e^t eps d_u / (N psi_j)}; the last level adds r/d to x with
    no t/(j+1) factor, residual left below a level's threshold is dropped, and
    the push ends at the first level none of whose nodes crossed its
    threshold (hk_solver.py:225-291).  The parameters leave many residuals
    between the thresholds of neighbouring levels, and in the last two cases
    a level whose residuals lie between e^t eps d/(N psi_j) and
    e^t eps d/(N psi_{j+1}): a psi index off by one changes the sets or the
    level count."""
    g = random_graph(n=50, p=0.08, seed=seed)
    s = 3
    o = O.hk_push(g, s, t, eps, local=True, trace=True)
    N = o["est_n"]
    assert N == O.estimate_n(t, eps / 2)
    psi = _psi_closed(N, t)
    A = O.adjacency(g).toarray()
    d = g.degree.astype(float)
    thr = math.exp(t) * eps * d / N
    r = np.zeros(g.n)
    r[s] = 1.0
    x = np.zeros(g.n)
    oper = 0.0
    levels = 0
    for j in range(N):
        pushed = np.flatnonzero(r >= thr / psi[j])
        if pushed.size == 0:   # an empty level queue ends the push (hk_solver.py:237)
            break
        assert levels < len(o["trace"]), j
        assert set(o["trace"][levels].tolist()) == set(pushed.tolist()), j
        levels += 1
        oper += d[pushed].sum()
        x[pushed] += r[pushed]
        send = np.zeros(g.n)
        send[pushed] = r[pushed] / d[pushed]
        nxt = A @ send
        if j + 1 == N:
            x += nxt
            break
        r = (t / (j + 1.0)) * nxt   # below-threshold residual of level j is dropped
    assert levels == len(o["trace"])
    assert oper == o["oper"]
    x /= math.exp(t)
    assert np.abs(o["x"] - x).max() < 1e-15
    # hk-relax guarantee against the exact heat kernel
    P = A / d[None, :]
    es = np.zeros(g.n)
    es[s] = 1.0
    h = sla.expm(t * (P - np.eye(g.n))) @ es
    assert (np.abs(h - o["x"]) / d).max() < eps


def test_hk_psi_closed_form_matches_recurrence():
    """psi_N = 1, psi_i = psi_{i+1} t/(i+1) + 1 (hk_solver.py:226-229) is the
    tail sum psi_j = sum_{m=0}^{N-j} t^m j!/(j+m)!."""
    for N, t in ((9, 2.0), (27, 5.0), (40, 5.0)):
        psi = [0.0] * (N + 1)
        psi[N] = 1.0
        for i in range(N - 1, -1, -1):
            psi[i] = psi[i + 1] * t / (i + 1.0) + 1.0
        c = _psi_closed(N, t)
        assert max(abs(a - b) / b for a, b in zip(psi, c)) < 1e-14


@pytest.mark.parametrize("seed,t,n_terms,connected", [(21, 5.0, 0, True), (22, 3.0, 7, True), (23, 5.0, 12, False)])
def test_hk_global_sweep_closed_form(seed, t, n_terms, connected):
    """The global level sweep (hk_solver.py:265-279, omega = 1) is the
    truncated Taylor series with the reference's last level:
        x = e^-t ( sum_{j<N} r_j + P r_{N-1} ),  r_j = (t^j / j!) P^j e_s,
    P = A D^-1 (no t/N factor on the last level's push), one sweep per level
    (oper = N (m + n)), and both level residuals zero on return.  On a graph
    with isolated nodes they are left out of the sweep (the reference never
    ends there)."""
    g = random_graph(n=60, p=0.06, seed=seed, connected=connected)
    d = g.degree.astype(float)
    s = int(np.flatnonzero(d > 0)[0])
    eps = 1e-4
    o = O.hk_push(g, s, t, eps, local=False, n_terms=n_terms)
    N = o["est_n"]
    assert N == (n_terms if n_terms else O.estimate_n(t, eps / 2))
    A = O.adjacency(g).toarray()
    P = A / np.where(d > 0, d, 1.0)[None, :]
    r = np.zeros(g.n)
    r[s] = 1.0
    x = np.zeros(g.n)
    for j in range(N):
        x += r
        if j + 1 == N:
            x += P @ r
        else:
            r = (t / (j + 1.0)) * (P @ r)
    x /= math.exp(t)
    assert np.abs(o["x"] - x).sum() <= 1e-13 * np.abs(x).sum()
    assert o["oper"] == N * (int(g.degree.sum()) + g.n)
    assert not np.any(o["r"])
