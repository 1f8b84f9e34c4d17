"""Exact-rational hand trace of the reference's local Chebyshev push.

Works ``ppr_cheby(local=True)`` (ppr_solver.py:582-601, 619-661) and
``katz_cheby(local=True)`` (katz_solver.py:340-356, 384-427) step by step in
``fractions.Fraction`` on two tiny graphs, following the reference's text:
queue order, the ``rear > n/4`` switch to ascending ids, every frontier node
updated (t = b1 r_u + b2 dx_u, stale dx kept for nodes that left the queue)
before any pushes, each neighbour tested right after its add, and r_u -= dx_u
after u's neighbour loop with no test.  Prints the literals that
tests/test_oracle_pins.py holds (queues per iteration, oper, exact x and r).

The graphs were picked so that each of these restatement slips changes the
queue sequence: using the previous iteration's coefficients, always or never
sorting the queue, dropping a node's stale dx when it re-enters, and testing
u after r_u -= dx_u.

    python tests/golden/cheb_hand_trace.py
"""
from fractions import Fraction as F

# two triangles joined by an edge, 6 isolated nodes (n = 12, n/4 = 3)
PPR_EDGES = [(0, 1), (0, 2), (1, 2), (2, 3), (3, 4), (3, 5), (4, 5)]
PPR_N, PPR_S, PPR_ALPHA, PPR_EPS = 12, 0, F(1, 2), F(1, 128)
# a 7-cycle (nodes 0-6) with a pendant node 7, 8 isolated nodes (n = 16, n/4 = 4)
KATZ_EDGES = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 6), (0, 6), (6, 7)]
KATZ_N, KATZ_S, KATZ_BETA, KATZ_L, KATZ_MU, KATZ_EPS = 16, 0, F(1, 4), F(3, 2), F(1, 2), F(1, 256)


def adjacency(n, edges):
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    return [sorted(a) for a in adj]


def cheb_local(n, adj, s, eps, kind, alpha=None, beta=None, L=None, mu=None):
    deg = [len(a) for a in adj]
    x, dx, r = [F(0)] * n, [F(0)] * n, [F(0)] * n
    if kind == "ppr":
        r[s] = alpha
        x[s] = alpha
        dx[s] = alpha
        thr = [eps * alpha * d for d in deg]
        delta = 1 - alpha
        tmp = [F(0)] * n
        for v in adj[s]:
            tmp[v] += (1 - alpha) * r[s] / deg[s]
        r = tmp                                   # r replaced: r_s = 0, init push not counted
        push = lambda u: (1 - alpha) * dx[u] / deg[u]
    else:
        x[s] = 2 / (L + mu)
        dx[s] = x[s]
        r[s] = 1 - 2 / (L + mu)
        for v in adj[s]:
            r[v] += 2 * beta / (L + mu)
        thr = [eps * d for d in deg]
        delta = (L - mu) / (L + mu)
        push = lambda u: beta * dx[u]
    queue, mark = [], [False] * n
    for v in adj[s]:
        if not mark[v] and thr[v] <= abs(r[v]):
            queue.append(v)
            mark[v] = True
    trace, oper = [], 0
    while queue:
        oper += sum(deg[u] for u in queue)
        if kind == "ppr":
            delta = 1 / (2 / (1 - alpha) - delta)
            b1 = 2 * delta / (1 - alpha)
            b2 = b1 - 1                           # t = beta r + (beta - 1) dx
        else:
            delta = 1 / (2 * ((L + mu) / (L - mu)) - delta)
            b1 = 4 * delta / (L - mu)
            b2 = -(1 - 2 * delta * (L + mu) / (L - mu))   # t = b1 r - b2' dx
        if len(queue) > n / 4:
            queue = [i for i in range(n) if mark[i]]
        trace.append(list(queue))
        for u in queue:
            t = b1 * r[u] + b2 * dx[u]
            x[u] += t
            dx[u] = t
            mark[u] = False
        pre, queue = queue, []
        for u in pre:
            val = push(u)
            for v in adj[u]:
                r[v] += val
                if not mark[v] and thr[v] <= abs(r[v]):
                    queue.append(v)
                    mark[v] = True
            r[u] -= dx[u]
    return trace, oper, x, r


def ppr_case():
    return cheb_local(PPR_N, adjacency(PPR_N, PPR_EDGES), PPR_S, PPR_EPS, "ppr", alpha=PPR_ALPHA)


def katz_case():
    return cheb_local(KATZ_N, adjacency(KATZ_N, KATZ_EDGES), KATZ_S, KATZ_EPS, "katz", beta=KATZ_BETA,
                      L=KATZ_L, mu=KATZ_MU)


if __name__ == "__main__":
    for name, case in (("PPR", ppr_case), ("KATZ", katz_case)):
        trace, oper, x, r = case()
        print(f"{name}_TRACE = {trace}")
        print(f"{name}_OPER = {oper}")
        print(f"{name}_X = {[str(v) for v in x]}")
        print(f"{name}_R = {[str(v) for v in r]}")
