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
SOR pushes and edge deletion stay CPU reference baselines (NotImplementedError).
"""
from __future__ import annotations

import numpy as np

from flsg import _abi
from flsg import dropin as _d
from flsg.graph import SlackGraph
from flsg.ignn import DynamicPropagation

__all__ = ["instantGNNParallel", "instantGNNUpdateBatchParallel", "updetaGraph", "instantGNNPush",
           "instantGNNSorPush", "instantGNNSorParallel", "instantGNNUpdateBatch", "instantGNNUpdateEdgeParallel"]

_STATE: dict = {}


def _slack(indptr, indices, degree):
    n = len(degree)
    return SlackGraph(n=n, indptr=np.asarray(indptr, dtype=np.int64), indices=np.asarray(indices, dtype=np.int32),
                      degree=np.asarray(degree, dtype=np.int32), us=np.zeros((0, 0), np.int32),
                      vs=np.zeros((0, 0), np.int32))


def instantGNNParallel(n, indptr, indices, degree, alpha, eps, P, R, beta, batchsize=80):
    """InstantGNN.py:46-71.  P (zeros) and R (features) are n x F; both are
    updated in place.  Returns per-column runtimes (seconds)."""
    F = R.shape[1]
    prop = DynamicPropagation(_slack(indptr, indices, degree), X=R, alpha=float(alpha), eps=float(eps),
                              beta=float(beta), precision=_abi.FLSG_FP32, R0=R, P0=P)
    st = prop.initial()
    P[:] = prop.P.cpu().numpy()
    R[:] = prop.R.cpu().numpy()
    _STATE[id(R)] = prop
    return np.full(F, st.ms * 1e-3 / F)


def instantGNNUpdateBatchParallel(n, indptr, indices, degree, alpha, eps, omega, P, R, X, us, vs, beta, algo,
                                  batchsize=80):
    """InstantGNN.py:332-359 with algo 'fwd...insert': inserts (us, vs) into
    the slack CSR (indices / degree updated in place), corrects P and R and
    pushes.  Returns per-column runtimes."""
    if "sor" in algo:
        _d.sor_not_built("instantGNNUpdateBatchParallel(algo=%r)" % algo)
    if "insert" not in algo:
        raise NotImplementedError("edge deletion is not built on the GPU (algo must contain 'insert')")
    F = R.shape[1]
    prop = _STATE.get(id(R))
    if prop is None or prop.n != n or prop.cols != F:
        prop = DynamicPropagation(_slack(indptr, indices, degree), X=X, alpha=float(alpha), eps=float(eps),
                                  beta=float(beta), precision=_abi.FLSG_FP32, R0=R, P0=P)
        _STATE[id(R)] = prop
    st = prop.insert(us, vs)
    P[:] = prop.P.cpu().numpy()
    R[:] = prop.R.cpu().numpy()
    ind, deg = prop.graph_arrays()
    indices[:] = ind.astype(indices.dtype)
    degree[:] = deg.astype(degree.dtype)
    return np.full(F, st.ms * 1e-3 / F)


def updetaGraph(indptr, indices, degree, us, vs, onlyinsert=True):
    """InstantGNN.py:137-180 (insert only): mutates indices/degree in place and
    returns (insert_neighbor, delete_neighbor) dicts {u: [new neighbours]}."""
    if not onlyinsert:
        raise NotImplementedError("edge deletion is not built on the GPU")
    from flsg.engine import DeviceGraph, _check, _ptr, _stream_ptr
    import ctypes as C
    g = _slack(indptr, indices, degree)
    dg = DeviceGraph(g)
    torch = dg.torch
    d_us = torch.from_numpy(np.ascontiguousarray(us, dtype=np.int32)).to(dg.device)
    d_vs = torch.from_numpy(np.ascontiguousarray(vs, dtype=np.int32)).to(dg.device)
    keys = torch.empty(max(2 * d_us.numel(), 1), dtype=torch.int32, device=dg.device)
    old = torch.empty(g.n, dtype=torch.int32, device=dg.device)
    nk, ni = C.c_int64(0), C.c_int64(0)
    _check(dg.lib.flsg_graph_insert_edges(dg.handle, _ptr(d_us), _ptr(d_vs), int(d_us.numel()), _ptr(keys),
                                          C.byref(nk), _ptr(old), C.byref(ni), _stream_ptr(torch)))
    ind = dg.indices.cpu().numpy()
    deg = dg.degree.cpu().numpy()
    od = old.cpu().numpy()
    ins = {}
    for u in sorted(keys[:nk.value].cpu().numpy().tolist()):
        ins[u] = ind[g.indptr[u] + od[u]:g.indptr[u] + deg[u]].tolist()
    indices[:] = ind.astype(indices.dtype)
    degree[:] = deg.astype(degree.dtype)
    dg.close()
    return ins, {}


def instantGNNPush(*args, **kwargs):
    raise NotImplementedError("the per-column FIFO push is replaced by instantGNNParallel's device propagation")


def instantGNNUpdateBatch(*args, **kwargs):
    raise NotImplementedError("use instantGNNUpdateBatchParallel (all columns on the device)")


def instantGNNSorPush(*args, **kwargs):
    _d.sor_not_built("instantGNNSorPush")


def instantGNNSorParallel(*args, **kwargs):
    _d.sor_not_built("instantGNNSorParallel")


def instantGNNUpdateEdgeParallel(*args, **kwargs):
    raise NotImplementedError("dead code in the reference (InstantGNN.py:250-295); not built")
