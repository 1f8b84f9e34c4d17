"""Host side of the push engine: device-resident graphs and solve calls.

PyTorch-ROCm is used for device memory and the current stream only; every
computation runs in libflsg.so's HIP kernels.  Nothing here has a CPU
fallback: without a GPU or without the built library the calls raise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._build import load_lib
from .graph import CSRGraph, SlackGraph


class FlsgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"flsg error {code} ({_abi.ERROR_NAMES.get(code, '?')}): {msg}")
        self.code = code


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("flsg needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _check(rc: int) -> None:
    if rc != 0:
        lib = load_lib()
        msg = lib.flsg_last_error()
        msg = msg.decode() if msg else ""
        if rc == _abi.FLSG_EASSERT:
            raise AssertionError(msg)
        raise FlsgError(rc, msg)


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _stream_ptr(torch) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class DeviceGraph:
    """A CSR (or slack CSR) resident in HBM with a libflsg graph handle.

    The device copy is int32 indptr/indices/degree; the live adjacency of u is
    ``indices[indptr[u] : indptr[u] + degree[u]]``.
    """

    def __init__(self, g: CSRGraph | SlackGraph, device: str = "cuda:0", relabel: bool = False):
        torch = _torch()
        self.torch = torch
        self.device = torch.device(device)
        self.n = int(g.n)
        if int(g.indptr[-1]) >= 2 ** 31:
            raise ValueError("graph too large for int32 CSR")
        self.slack = isinstance(g, SlackGraph)
        self.perm = None          # internal id -> caller id (relabelled graphs)
        self.iperm = None         # caller id -> internal id
        if relabel:
            if self.slack:
                raise ValueError("relabelling is for static CSR graphs")
            g, perm, iperm = relabel_by_degree(g)
            self.perm = torch.from_numpy(perm.astype(np.int32)).to(self.device)
            self.iperm = torch.from_numpy(iperm.astype(np.int64)).to(self.device)
            self.perm_np, self.iperm_np = perm, iperm
        self.indptr = torch.from_numpy(np.ascontiguousarray(g.indptr, dtype=np.int32)).to(self.device)
        self.indices = torch.from_numpy(np.ascontiguousarray(g.indices, dtype=np.int32)).to(self.device)
        self.degree = torch.from_numpy(np.ascontiguousarray(g.degree, dtype=np.int32)).to(self.device)
        self.nnz_cap = int(self.indices.numel())
        self.lib = load_lib()
        h = C.c_void_p()
        torch.cuda.synchronize(self.device)
        _check(self.lib.flsg_graph_create(self.n, self.nnz_cap, _ptr(self.indptr), _ptr(self.indices),
                                          _ptr(self.degree), 1 if self.slack else 0, C.byref(h)))
        self.handle = h
        self.d_max = int(g.degree.max()) if self.n else 0
        if self.perm is not None:
            _check(self.lib.flsg_graph_set_out_map(self.handle, _ptr(self.perm)))

    def to_internal(self, seeds):
        s = np.asarray(seeds, dtype=np.int64)
        return s if self.perm is None else self.iperm_np[s]

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.flsg_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def relabel_by_degree(g: CSRGraph):
    """Relabel nodes by non-increasing degree (ties by id): hubs, where most
    pushes land in a power-law graph, become a contiguous prefix of every
    per-node array.  Returns (graph, perm, iperm) with perm[new] = old."""
    n = g.n
    deg = np.asarray(g.degree, dtype=np.int64)
    perm = np.argsort(-deg, kind="stable").astype(np.int64)
    iperm = np.empty(n, dtype=np.int64)
    iperm[perm] = np.arange(n, dtype=np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    src = iperm[rows]
    dst = iperm[np.asarray(g.indices, dtype=np.int64)]
    order = np.lexsort((dst, src))
    ndeg = deg[perm]
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(ndeg, out=indptr[1:])
    h = CSRGraph(n=n, indptr=indptr, indices=dst[order].astype(np.int32), degree=ndeg.astype(np.int32),
                 name=g.name + "-bydegree")
    return h, perm, iperm


def make_params(kind: int, method: int = _abi.FLSG_GD, local: bool = True, precision: int = _abi.FLSG_FP64,
                alpha: float = 0.1, eps: float = 1e-7, l_max: float = 0.0, mu_min: float = 0.0,
                tau: float = 0.0, n_terms: int = 0, max_iters: int = 0) -> _abi.Params:
    return _abi.Params(kind=kind, method=method, local=1 if local else 0, precision=precision, alpha=alpha,
                       eps=eps, l_max=l_max, mu_min=mu_min, tau=tau, n_terms=n_terms, max_iters=max_iters)


@dataclass
class SolveOut:
    x: object            # torch float64 [n] (device)
    r: object            # torch float64 [n] (HK: [2, n]) (device)
    result: dict
    trace: list | None   # per-iteration frontier node arrays (numpy), if requested
    extra: dict


def solve(dg: DeviceGraph, params: _abi.Params, seed: int, trace: bool = False,
          trace_cap_nodes: int | None = None, trace_cap_iters: int = 100000) -> SolveOut:
    """One single-source solve on the whole GPU (FP64 path)."""
    torch = dg.torch
    n = dg.n
    x = torch.empty(n, dtype=torch.float64, device=dg.device)
    r = torch.empty((2, n) if params.kind == _abi.FLSG_HK else n, dtype=torch.float64, device=dg.device)
    res = _abi.Result()
    tr_ptr = None
    tr = None
    if trace:
        cap = trace_cap_nodes if trace_cap_nodes is not None else max(4 * n, 1 << 20)
        d_nodes = torch.empty(cap, dtype=torch.int32, device=dg.device)
        h_off = np.zeros(trace_cap_iters + 1, dtype=np.int64)
        tr = _abi.Trace(d_nodes=d_nodes.data_ptr(), cap_nodes=cap, h_off=h_off.ctypes.data,
                        cap_iters=trace_cap_iters, n_iters=0, overflow=0)
        tr_ptr = C.byref(tr)
    extra = {}
    seed = int(dg.to_internal([seed])[0])
    if params.kind == _abi.FLSG_HK:
        # Taylor order as the solve computes it (hk_solver.py:225), for the
        # per-level sums' buffer; its capacity goes in through est_n
        n_terms = int(params.n_terms) if params.n_terms > 0 else int(
            dg.lib.flsg_hk_estimate_n(float(params.tau), 0.5 * float(params.eps)))
        lvl = np.zeros(2 * max(n_terms, 1), dtype=np.float64)
        est_n = C.c_int64(lvl.size)
        rc = dg.lib.flsg_hk_solve(dg.handle, C.byref(params), int(seed), _ptr(x), _ptr(r), C.byref(res),
                                  C.c_void_p(lvl.ctypes.data), C.byref(est_n), tr_ptr, _stream_ptr(torch))
        _check(rc)
        extra["est_n"] = int(est_n.value)
        extra["lvl_sums"] = lvl[:2 * int(est_n.value)].copy()
    else:
        rc = dg.lib.flsg_solve(dg.handle, C.byref(params), int(seed), _ptr(x), _ptr(r), C.byref(res),
                               tr_ptr, _stream_ptr(torch))
        _check(rc)
    frontiers = None
    if trace:
        if tr.overflow:
            raise FlsgError(_abi.FLSG_ECAP, "trace capacity exceeded")
        nodes = d_nodes[: int(h_off[tr.n_iters])].cpu().numpy()
        if dg.perm is not None:
            nodes = dg.perm_np[nodes].astype(np.int32)
        frontiers = [nodes[h_off[t]:h_off[t + 1]] for t in range(tr.n_iters)]
    if dg.perm is not None:
        x = x[dg.iperm]
        r = r[..., dg.iperm]
    return SolveOut(x=x, r=r, result=res.as_dict(), trace=frontiers, extra=extra)


def sor_solve(dg: DeviceGraph, alpha: float, eps: float, omega: float, seed: int, local: bool = True,
              max_iters: int = 0, kind: int = _abi.FLSG_PPR) -> SolveOut:
    """Graph-coloured SOR push on the whole GPU: FP64, local (frontier) or
    global (sweeps); PPR (ppr_sor, ppr_solver.py:331-461) or Katz with
    alpha = beta (katz_sor, katz_solver.py:252-337)."""
    torch = dg.torch
    n = dg.n
    x = torch.empty(n, dtype=torch.float64, device=dg.device)
    r = torch.empty(n, dtype=torch.float64, device=dg.device)
    res = _abi.Result()
    p = make_params(kind, _abi.FLSG_GD, local, _abi.FLSG_FP64, alpha=alpha, eps=eps, max_iters=max_iters)
    seed = int(dg.to_internal([seed])[0])
    _check(dg.lib.flsg_sor_solve(dg.handle, C.byref(p), float(omega), seed, _ptr(x), _ptr(r), C.byref(res),
                                 _stream_ptr(torch)))
    if dg.perm is not None:
        x = x[dg.iperm]
        r = r[dg.iperm]
    return SolveOut(x=x, r=r, result=res.as_dict(), trace=None, extra={})


def ignn_sor_columns(dg: DeviceGraph, alpha: float, eps: float, beta: float, omega: float, P, R, nodes=None,
                     max_iters: int = 0) -> dict:
    """Graph-coloured SOR push (instantGNNSorPush, InstantGNN.py:74-105) of
    every column of the FP64 n x C device tensors P, R (updated in place),
    one column at a time.  nodes: optional starting frontier (numpy, distinct
    ids) shared by every column; by default each column starts from its
    nodes above threshold.  Returns the summed device ms, iterations and
    pushed degrees."""
    torch = dg.torch
    deg = dg.degree.to(torch.float64)
    # eps_vec, d^(1-beta), d^beta as the reference builds them (np.power on
    # the host, InstantGNN.py:118)
    d_np = deg.cpu().numpy()
    ev = torch.from_numpy(np.power(d_np, 1 - beta) * eps).to(dg.device)
    dp1 = torch.from_numpy(np.power(d_np, 1 - beta)).to(dg.device)
    dpb = torch.from_numpy(np.power(d_np, beta)).to(dg.device)
    d_nodes = None
    nn = 0
    if nodes is not None:
        nodes = np.unique(np.asarray(nodes, dtype=np.int32))
        d_nodes = torch.from_numpy(nodes).to(dg.device)
        nn = int(nodes.size)
    out = {"ms": 0.0, "iters": 0, "oper": 0}
    cols = 1 if P.dim() == 1 else P.shape[1]
    for c in range(cols):
        p = P if P.dim() == 1 else P[:, c].contiguous()
        r = R if R.dim() == 1 else R[:, c].contiguous()
        res = _abi.Result()
        _check(dg.lib.flsg_sor_ignn_column(dg.handle, float(alpha), float(omega), _ptr(ev), _ptr(dp1), _ptr(dpb),
                                           _ptr(p), _ptr(r), _ptr(d_nodes) if d_nodes is not None else None, nn,
                                           int(max_iters), C.byref(res), _stream_ptr(torch)))
        if P.dim() != 1:
            P[:, c] = p
            R[:, c] = r
        out["ms"] += res.ms
        out["iters"] += res.iters
        out["oper"] += res.oper
    return out


def colouring(dg: DeviceGraph):
    """The SOR colouring of the device graph: (colour[n] numpy int32 in the
    caller's node ids, number of colours, rounds)."""
    torch = dg.torch
    col = torch.empty(dg.n, dtype=torch.int32, device=dg.device)
    k = C.c_int32(0)
    rounds = C.c_int64(0)
    _check(dg.lib.flsg_graph_colouring(dg.handle, _ptr(col), C.byref(k), C.byref(rounds), _stream_ptr(torch)))
    if dg.perm is not None:
        col = col[dg.iperm]
    return col.cpu().numpy(), int(k.value), int(rounds.value)


def arena_dirty(dg: DeviceGraph) -> int:
    """Slot-arena records not back in the between-queries state (0 after
    every completed batched launch)."""
    torch = dg.torch
    cnt = C.c_int64(0)
    _check(dg.lib.flsg_batch_arena_dirty(dg.handle, C.byref(cnt), _stream_ptr(torch)))
    return int(cnt.value)


@dataclass
class BatchOut:
    ids: object          # torch int32 [total] (device)
    vals: object         # torch float64 [total] (device)
    offsets: np.ndarray  # int64 [nq] start of query q in ids/vals
    results: np.ndarray  # per-query result records (structured array, fields of flsg_result)
    ms: float            # wall-clock of the launch (events), milliseconds

    def query(self, q: int):
        """(ids, values) of query q as numpy arrays."""
        o = int(self.offsets[q])
        k = int(self.results[q]["supp"])
        return self.ids[o:o + k].cpu().numpy(), self.vals[o:o + k].cpu().numpy()

    def dense(self, q: int, n: int) -> np.ndarray:
        ids, vals = self.query(q)
        out = np.zeros(n, dtype=np.float64)
        out[ids] = vals
        return out


def batch_solve(dg: DeviceGraph, params: _abi.Params, seeds, out_cap: int | None = None,
                slots: int = 0, block: int = 0, hub_bytes: int | None = None,
                allow_failed: bool = False) -> BatchOut:
    """Many single-source local solves in one persistent launch
    (params.precision FLSG_FP32: FP32 push, FP64 threshold and FP64
    certificate; FLSG_FP64: FP64 push).  A query that fails raises FlsgError
    unless allow_failed, in which case its status is in results[q]."""
    torch = dg.torch
    seeds = np.ascontiguousarray(dg.to_internal(seeds))
    nq = int(seeds.size)
    if out_cap is None:
        # supports of eps=1e-7 PPR queries on the 2.39M-node graph are ~0.4-0.6M
        out_cap = nq * min(dg.n, max(200000, dg.n // 2))
    d_seeds = torch.from_numpy(seeds).to(dg.device)
    ids = torch.empty(out_cap, dtype=torch.int32, device=dg.device)
    vals = torch.empty(out_cap, dtype=torch.float64, device=dg.device)
    offs = torch.empty(max(nq, 1), dtype=torch.int64, device=dg.device)
    res = (_abi.Result * max(nq, 1))()
    if slots:
        _check(dg.lib.flsg_batch_set_slots(dg.handle, int(slots)))
    if block:
        _check(dg.lib.flsg_batch_set_block(dg.handle, int(block)))
    if hub_bytes is not None:
        _check(dg.lib.flsg_batch_set_hub_cache(dg.handle, int(hub_bytes)))
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    rc = dg.lib.flsg_batch_solve(dg.handle, C.byref(params), _ptr(d_seeds), nq, _ptr(ids), _ptr(vals),
                                 _ptr(offs), int(out_cap), res, _stream_ptr(torch))
    e1.record()
    per_query = rc != _abi.FLSG_OK and any(res[i].status != _abi.FLSG_OK for i in range(nq))
    if not (allow_failed and per_query):
        _check(rc)
    torch.cuda.synchronize()
    # the result records as a structured array over the ctypes buffer (no copy)
    return BatchOut(ids=ids, vals=vals, offsets=offs.cpu().numpy()[:nq], results=np.ctypeslib.as_array(res)[:nq],
                    ms=e0.elapsed_time(e1))
