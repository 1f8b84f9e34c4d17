"""Drop-in for the reference's ``InstantGNN.py``: propagation of feature
columns P ~ sum_k alpha (1-alpha)^k (D^-b A D^(b-1))^k X and its update under
edge-insert snapshots (InstantGNN.py:46-71, 137-180, 298-359).

Same signatures; P, R (n x F float64) and the slack-CSR ``indices`` /
``degree`` are mutated in place like the reference.  All F columns run in one
device propagation (FP32 rounds + FP64 certificate, so on return
|R| < eps d^(1-beta) holds in FP64 for every entry).  The push order differs
from the reference's per-column FIFO, so P differs from it by at most what
the two residuals allow (both satisfy the same invariant and certificate).
``runtimes`` are the device time split evenly over the columns.
``instantGNNPush`` / ``instantGNNUpdateBatch`` (one column, the reference's
queue arguments) run the same device propagation in FP64.
``instantGNNSorPush`` / ``instantGNNSorParallel`` and the 'sor' update
algorithms run a graph-coloured SOR push per column in FP64 (colour order, not
the reference's FIFO order; same invariant and certificate).  Edge deletion
is not built (NotImplementedError).
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np

from flsg import _abi
from flsg import dropin as _d
from flsg.graph import SlackGraph
from flsg.ignn import DynamicPropagation

__all__ = ["instantGNNParallel", "instantGNNUpdateBatchParallel", "updetaGraph", "instantGNNPush",
           "instantGNNSorPush", "instantGNNSorParallel", "instantGNNUpdateBatch", "instantGNNUpdateEdgeParallel"]

# device state of the last propagations, for the updates that follow them:
# id(R) -> (DynamicPropagation, key): the key holds alpha, eps, beta, the
# features X the state was built for and the content fingerprint of P, R and
# the graph as the call returned them.  An update reuses an entry only when
# its arguments match that key exactly; otherwise it starts a new
# propagation state from the arrays it is given.
_STATE: "OrderedDict[int, tuple]" = OrderedDict()
_STATE_MAX = 2


def _state_key(alpha, eps, beta, fx, P, R, indptr, indices, degree):
    return (float(alpha), float(eps), float(beta), fx) + _d.fingerprint(P, R, indptr, indices, degree)


def _remember(R, prop, fp):
    old = _STATE.pop(id(R), None)
    if old is not None and old[0] is not prop:
        old[0].close()
    _STATE[id(R)] = (prop, fp)
    while len(_STATE) > _STATE_MAX:
        _, (p_old, _) = _STATE.popitem(last=False)
        p_old.close()


def _slack(indptr, indices, degree):
    n = len(degree)
    return SlackGraph(n=n, indptr=np.asarray(indptr, dtype=np.int64), indices=np.asarray(indices, dtype=np.int32),
                      degree=np.asarray(degree, dtype=np.int32), us=np.zeros((0, 0), np.int32),
                      vs=np.zeros((0, 0), np.int32))


def instantGNNParallel(n, indptr, indices, degree, alpha, eps, P, R, beta, batchsize=80):
    """InstantGNN.py:46-71.  P (zeros) and R (features) are n x F; both are
    updated in place.  Returns per-column runtimes (seconds)."""
    F = R.shape[1]
    fx = _d.fingerprint(R)   # the features the state is built for
    # FP64 rounds: at the reference's eps (0.01/n ~ 6e-8) FP32 rounds leave
    # residuals that many FP64 rounds must then correct, so the all-FP64
    # propagation is the faster certified one (16 snapshots of the
    # 169,343-node citation workload: 3.4 s vs 4.3 s)
    prop = DynamicPropagation(_slack(indptr, indices, degree), X=R, alpha=float(alpha), eps=float(eps),
                              beta=float(beta), precision=_abi.FLSG_FP64, R0=R, P0=P)
    st = prop.initial()
    P[:] = prop.P.cpu().numpy()
    R[:] = prop.R.cpu().numpy()
    _remember(R, prop, _state_key(alpha, eps, beta, fx, P, R, indptr, indices, degree))
    return np.full(F, st.ms * 1e-3 / F)


def instantGNNUpdateBatchParallel(n, indptr, indices, degree, alpha, eps, omega, P, R, X, us, vs, beta, algo,
                                  batchsize=80):
    """InstantGNN.py:332-359 with algo 'fwd...insert': inserts (us, vs) into
    the slack CSR (indices / degree updated in place), corrects P and R and
    pushes.  Returns per-column runtimes."""
    if "insert" not in algo:
        raise NotImplementedError("edge deletion is not built on the GPU (algo must contain 'insert')")
    F = R.shape[1]
    fx = _d.fingerprint(X)
    hit = _STATE.get(id(R))
    prop = None
    if (hit is not None and hit[0].n == n and hit[0].cols == F
            and hit[1] == _state_key(alpha, eps, beta, fx, P, R, indptr, indices, degree)):
        prop = hit[0]
    if prop is None:
        prop = DynamicPropagation(_slack(indptr, indices, degree), X=X, alpha=float(alpha), eps=float(eps),
                                  beta=float(beta), precision=_abi.FLSG_FP64, R0=R, P0=P)
    st = prop.insert(us, vs, sor_omega=float(omega) if "sor" in algo else None)
    P[:] = prop.P.cpu().numpy()
    R[:] = prop.R.cpu().numpy()
    ind, deg = prop.graph_arrays()
    indices[:] = ind.astype(indices.dtype)
    degree[:] = deg.astype(degree.dtype)
    _remember(R, prop, _state_key(alpha, eps, beta, fx, P, R, indptr, indices, degree))
    return np.full(F, st.ms * 1e-3 / F)


def updetaGraph(indptr, indices, degree, us, vs, onlyinsert=True):
    """InstantGNN.py:137-180 (insert only): mutates indices/degree in place
    with the reference's byte layout (each row's new neighbours appended in
    batch order) and returns (insert_neighbor, delete_neighbor) dicts
    {u: [new neighbours]}, keys in first-insertion order.  A self-loop the
    reference would insert raises FlsgError (FLSG_EARG; the reference would
    leave an unwritten entry in the live row) and nothing is inserted;
    endpoints outside [0, n) raise ValueError."""
    if not onlyinsert:
        raise NotImplementedError("edge deletion is not built on the GPU")
    from flsg.engine import DeviceGraph, _check, _ptr, _stream_ptr
    from flsg.graph import edge_ids
    import ctypes as C
    g = _slack(indptr, indices, degree)
    hu, hv = edge_ids(us, vs, g.n)
    dg = DeviceGraph(g)
    torch = dg.torch
    d_us = torch.from_numpy(hu).to(dg.device)
    d_vs = torch.from_numpy(hv).to(dg.device)
    keys = torch.empty(max(2 * d_us.numel(), 1), dtype=torch.int32, device=dg.device)
    old = torch.empty(g.n, dtype=torch.int32, device=dg.device)
    nk, ni = C.c_int64(0), C.c_int64(0)
    try:
        _check(dg.lib.flsg_graph_insert_edges(dg.handle, _ptr(d_us), _ptr(d_vs), int(d_us.numel()), _ptr(keys),
                                              C.byref(nk), _ptr(old), C.byref(ni), _stream_ptr(torch)))
        ind = dg.indices.cpu().numpy()
        deg = dg.degree.cpu().numpy()
        od = old.cpu().numpy()
        # keys in first-insertion order; each row's appended tail is its list
        ins = {}
        for u in keys[:nk.value].cpu().numpy().tolist():
            ins[u] = ind[g.indptr[u] + od[u]:g.indptr[u] + deg[u]].tolist()
    finally:
        dg.close()
    indices[:] = ind.astype(indices.dtype)
    degree[:] = deg.astype(degree.dtype)
    return ins, {}


def _eps_from_vec(degree, eps_vec, beta) -> float:
    """eps of eps_vec = eps * degree**(1-beta), the thresholds InstantGNN.py:57
    builds; other threshold vectors are rejected."""
    d = np.asarray(degree, dtype=np.float64)
    ev = np.asarray(eps_vec, dtype=np.float64)
    pos = d > 0
    if not pos.any():
        return float(ev.max()) if ev.size else 0.0
    w = d[pos] ** (1.0 - beta)
    eps = float(np.median(ev[pos] / w))
    if not np.allclose(ev[pos], eps * w, rtol=1e-6, atol=0.0):
        raise ValueError("eps_vec must be eps * degree**(1-beta) (InstantGNN.py:57)")
    return eps


def _queue_nodes(queue, front, rear, qlen):
    """Entries front+1 .. rear of the circular queue (InstantGNN.py:14-16)."""
    out = []
    f = int(front)
    while f != int(rear):
        f = (f + 1) % int(qlen)
        out.append(int(queue[f]))
    return np.unique(np.asarray(out, dtype=np.int32))


def _column_prop(indptr, indices, degree, x, p, r, alpha, eps, beta):
    cols = lambda a: np.asarray(a, dtype=np.float64).reshape(-1, 1)
    return DynamicPropagation(_slack(indptr, indices, degree), X=cols(x), alpha=float(alpha), eps=float(eps),
                              beta=float(beta), precision=_abi.FLSG_FP64, R0=cols(r), P0=cols(p))


def _push_from(prop, nodes):
    torch = prop.torch
    d_nodes = torch.from_numpy(np.ascontiguousarray(nodes, dtype=np.int32)).to(prop.dg.device)
    return prop._propagate(d_nodes, int(nodes.size)) if nodes.size else None


def instantGNNPush(n, indptr, indices, degree, alpha, eps_vec, p, r, queue, front, rear, q_mark, beta):
    """InstantGNN.py:12-43 for one feature column: pushes from the nodes in
    queue positions front+1 .. rear (the queue is circular with length n,
    which the reference's drivers pass as node count + 2) until every
    |r_u| < eps_vec_u; p and r are updated in place and q_mark cleared.  The
    device pushes in synchronous FP64 rounds, not in FIFO order, so p matches
    the reference's through the propagation invariant and the certificate
    rather than bitwise.  Returns the runtime in seconds."""
    import time
    t0 = time.perf_counter()
    eps = _eps_from_vec(degree, eps_vec, beta)
    nodes = _queue_nodes(queue, front, rear, n)
    prop = _column_prop(indptr, indices, degree, r, p, r, alpha, eps, beta)
    try:
        _push_from(prop, nodes)
        p[:] = prop.P.cpu().numpy()[:, 0].astype(p.dtype)
        r[:] = prop.R.cpu().numpy()[:, 0].astype(r.dtype)
    finally:
        prop.close()
    q_mark[:] = False
    return time.perf_counter() - t0


def instantGNNUpdateBatch(n, indptr, indices, degree, old_degree, p, r, x, alpha, queue, front, rear, q_mark, eps_vec,
                          insert_neighbor, delete_neighbor, algo='fwd', omega=1, beta=0):
    """InstantGNN.py:298-329 for one column, after updetaGraph has appended the
    inserted neighbours to the rows: rescales / corrects p and r of every
    endpoint in insert_neighbor for the new degrees (the sum over a row's new
    neighbours is taken over the appended row tail, the same set), then pushes
    from them and from any nodes already queued.  Edge deletion and SOR pushes
    are not built.  Returns the runtime in seconds."""
    import time
    import ctypes as C
    from flsg.engine import _check, _ptr, _stream_ptr
    if "insert" not in algo and any(len(v) for v in dict(delete_neighbor).values()):
        raise NotImplementedError("edge deletion is not built on the GPU")
    t0 = time.perf_counter()
    eps = _eps_from_vec(degree, eps_vec, beta)
    keys = np.asarray(sorted(int(u) for u in dict(insert_neighbor).keys()), dtype=np.int32)
    prop = _column_prop(indptr, indices, degree, x, p, r, alpha, eps, beta)
    try:
        torch = prop.torch
        dev = prop.dg.device
        d_keys = torch.from_numpy(keys if keys.size else np.zeros(1, np.int32)).to(dev)
        d_old = torch.from_numpy(np.ascontiguousarray(old_degree, dtype=np.int32)).to(dev)
        prop.params.degrees_changed = 1
        _check(prop.dg.lib.flsg_ignn_update(prop.dg.handle, C.byref(prop.params), _ptr(prop.X), _ptr(prop.P),
                                            _ptr(prop.R), _ptr(d_keys), int(keys.size), _ptr(d_old),
                                            _stream_ptr(torch)))
        nodes = np.union1d(keys, _queue_nodes(queue, front, rear, n)).astype(np.int32)
        if "sor" in algo:
            _sor_from(prop, nodes, omega)
        else:
            _push_from(prop, nodes)
        p[:] = prop.P.cpu().numpy()[:, 0].astype(p.dtype)
        r[:] = prop.R.cpu().numpy()[:, 0].astype(r.dtype)
    finally:
        prop.close()
    q_mark[:] = False
    return time.perf_counter() - t0


def _sor_from(prop, nodes, omega):
    from flsg import engine
    p = prop.P[:, 0].contiguous()
    r = prop.R[:, 0].contiguous()
    st = engine.ignn_sor_columns(prop.dg, prop.params.alpha, prop.params.eps, prop.params.beta, float(omega), p, r,
                                 nodes=nodes)
    prop.P[:, 0] = p
    prop.R[:, 0] = r
    return st


def instantGNNSorPush(n, indptr, indices, degree, alpha, eps_vec, omega, p, r, queue, front, rear, q_mark, beta):
    """InstantGNN.py:74-105 for one feature column: SOR push from the nodes in
    queue positions front+1 .. rear (circular, length n = node count + 2 as
    the reference's drivers pass) until every |r_u| < eps_vec_u; p and r are
    updated in place and q_mark cleared.  Graph-coloured on the device (the
    nodes of one colour push at once, colours in order), so p matches the
    reference's through the propagation invariant and the certificate rather
    than bitwise.  Returns the runtime in seconds."""
    import time
    t0 = time.perf_counter()
    eps = _eps_from_vec(degree, eps_vec, beta)
    nodes = _queue_nodes(queue, front, rear, n)
    prop = _column_prop(indptr, indices, degree, r, p, r, alpha, eps, beta)
    try:
        if nodes.size:
            _sor_from(prop, nodes, omega)
        p[:] = prop.P.cpu().numpy()[:, 0].astype(p.dtype)
        r[:] = prop.R.cpu().numpy()[:, 0].astype(r.dtype)
    finally:
        prop.close()
    q_mark[:] = False
    return time.perf_counter() - t0


def instantGNNSorParallel(n, indptr, indices, degree, alpha, eps, omega, P, R, beta, batchsize=80):
    """InstantGNN.py:108-134: per column, p starts at zero and r at R's
    column, the queue holds every node with |r| >= eps d^(1-beta); P and R are
    overwritten in place.  Returns per-column device times (seconds)."""
    from flsg import engine
    F = R.shape[1]
    times = np.zeros(F)
    prop = DynamicPropagation(_slack(indptr, indices, degree), X=R, alpha=float(alpha), eps=float(eps),
                              beta=float(beta), precision=_abi.FLSG_FP64, R0=R, P0=np.zeros_like(R))
    try:
        for c in range(F):
            p = prop.P[:, c].contiguous()
            r = prop.R[:, c].contiguous()
            st = engine.ignn_sor_columns(prop.dg, float(alpha), float(eps), float(beta), float(omega), p, r)
            prop.P[:, c] = p
            prop.R[:, c] = r
            times[c] = st["ms"] * 1e-3
        P[:] = prop.P.cpu().numpy()
        R[:] = prop.R.cpu().numpy()
    finally:
        prop.close()
    return times


def instantGNNUpdateEdgeParallel(*args, **kwargs):
    raise NotImplementedError("dead code in the reference (InstantGNN.py:250-295); not built")
