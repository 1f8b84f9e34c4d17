"""InstantGNN feature propagation and dynamic edge inserts on the GPU (host side).

State: P, R (FP64, n x C node-major) and the features X stay resident in HBM
with a slack-CSR graph handle; each snapshot of inserted edges is applied in
place on the device, P and R of the touched endpoints are corrected so the
propagation invariant holds on the new graph, and the push resumes from them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from .engine import DeviceGraph, _check, _ptr, _stream_ptr
from .graph import SlackGraph, edge_ids


class IgnnParams(C.Structure):
    _fields_ = [("precision", C.c_int32), ("degrees_changed", C.c_int32), ("cols", C.c_int64),
                ("alpha", C.c_double), ("eps", C.c_double), ("beta", C.c_double), ("max_rounds", C.c_int64)]


@dataclass
class PropagationStats:
    ms: float
    rounds: int
    rounds_fp64: int
    oper: int
    oper_cert: int
    cert: float
    keys: int = 0
    inserted: int = 0


class DynamicPropagation:
    """P ~ sum_k alpha (1-alpha)^k (D^-b A D^(b-1))^k X for all columns of X,
    maintained under edge-insert snapshots (InstantGNN.py:46-71, 298-359)."""

    def __init__(self, g: SlackGraph, X: np.ndarray, alpha: float, eps: float, beta: float,
                 precision: int = _abi.FLSG_FP32, device: str = "cuda:0", R0: np.ndarray | None = None,
                 P0: np.ndarray | None = None):
        self.dg = DeviceGraph(g, device)
        torch = self.dg.torch
        self.torch = torch
        self.n = g.n
        X = np.ascontiguousarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X[:, None]
        assert X.shape[0] == g.n
        self.cols = X.shape[1]
        dev = self.dg.device
        self.X = torch.from_numpy(X).to(dev)
        self.R = torch.from_numpy(np.ascontiguousarray(R0, dtype=np.float64)).to(dev) if R0 is not None \
            else self.X.clone()
        self.P = torch.from_numpy(np.ascontiguousarray(P0, dtype=np.float64)).to(dev) if P0 is not None \
            else torch.zeros_like(self.X)
        self.params = IgnnParams(precision=precision, degrees_changed=1, cols=self.cols, alpha=alpha, eps=eps,
                                 beta=beta, max_rounds=0)
        self.keys = torch.empty(max(2 * int(g.us.size) if g.us is not None else 0, 1), dtype=torch.int32,
                                device=dev)
        self.old_degree = torch.empty(g.n, dtype=torch.int32, device=dev)

    def _propagate(self, nodes=None, n_nodes=0) -> PropagationStats:
        res = _abi.Result()
        lib = self.dg.lib
        _check(lib.flsg_ignn_propagate(self.dg.handle, C.byref(self.params), _ptr(self.X), _ptr(self.P),
                                       _ptr(self.R), _ptr(nodes) if nodes is not None else None, int(n_nodes),
                                       C.byref(res), _stream_ptr(self.torch)))
        self.params.degrees_changed = 0
        return PropagationStats(ms=res.ms, rounds=res.iters, rounds_fp64=res.n_corr, oper=res.oper,
                                oper_cert=res.oper_cert, cert=res.cert)

    def initial(self) -> PropagationStats:
        """instantGNNParallel: push from every node above threshold."""
        return self._propagate()

    def insert(self, us, vs, sor_omega: float | None = None) -> PropagationStats:
        """One snapshot: in-place edge insertion, invariant correction of the
        endpoints (instantGNNUpdateBatch), then the push from them -- the
        coloured SOR push per column (FP64) when sor_omega is given (the
        reference's 'sor' algorithms, InstantGNN.py:320-322)."""
        torch = self.torch
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        hu, hv = edge_ids(us, vs, self.n)
        d_us = torch.from_numpy(hu).to(self.dg.device)
        d_vs = torch.from_numpy(hv).to(self.dg.device)
        k = int(d_us.numel())
        if self.keys.numel() < 2 * k:
            self.keys = torch.empty(2 * k, dtype=torch.int32, device=self.dg.device)
        nk = C.c_int64(0)
        ni = C.c_int64(0)
        lib = self.dg.lib
        _check(lib.flsg_graph_insert_edges(self.dg.handle, _ptr(d_us), _ptr(d_vs), k, _ptr(self.keys),
                                           C.byref(nk), _ptr(self.old_degree), C.byref(ni),
                                           _stream_ptr(torch)))
        self.params.degrees_changed = 1
        _check(lib.flsg_ignn_update(self.dg.handle, C.byref(self.params), _ptr(self.X), _ptr(self.P),
                                    _ptr(self.R), _ptr(self.keys), nk.value, _ptr(self.old_degree),
                                    _stream_ptr(torch)))
        if sor_omega is None:
            st = self._propagate(self.keys, nk.value)
        else:
            from .engine import ignn_sor_columns
            keys = self.keys[:nk.value].cpu().numpy()
            so = ignn_sor_columns(self.dg, self.params.alpha, self.params.eps, self.params.beta, sor_omega,
                                  self.P, self.R, nodes=keys)
            st = PropagationStats(ms=so["ms"], rounds=so["iters"], rounds_fp64=so["iters"], oper=so["oper"],
                                  oper_cert=0, cert=0.0)
        e1.record()
        torch.cuda.synchronize()
        st.ms = e0.elapsed_time(e1)
        st.keys = nk.value
        st.inserted = ni.value
        return st

    def graph_arrays(self):
        """(indices, degree) of the device slack CSR, on the host."""
        return self.dg.indices.cpu().numpy(), self.dg.degree.cpu().numpy()

    def close(self):
        self.dg.close()


def ignn_residual(indptr, indices, degree, X, P, alpha, beta):
    """Host FP64 recompute R = X - P/alpha + ((1-alpha)/alpha) M P of the
    propagation invariant (M_vu = 1/(d_v^b d_u^(1-b)) over live rows)."""
    import scipy.sparse as sp
    n = degree.size
    rows = np.repeat(np.arange(n), degree)
    cols = np.concatenate([indices[indptr[u]:indptr[u] + degree[u]] for u in range(n)]) if n else np.zeros(0, int)
    d = degree.astype(np.float64)
    with np.errstate(divide="ignore"):
        w = 1.0 / (d[rows] ** beta * d[cols] ** (1 - beta))
    M = sp.csr_matrix((w, (rows, cols)), shape=(n, n))
    R = X - P / alpha + ((1 - alpha) / alpha) * (M @ P)
    return R


def synthetic_workload(snapshots: int = 16, cols: int = 128):
    """The 169,343-node citation graph with 950,000 edges held out in
    `snapshots` insert batches, and X = `cols` row-normalised N(0,1) FP32
    feature columns (held as FP64)."""
    from . import graph as _graph
    sg = _graph.citation_dynamic_graph(snapshots=snapshots)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((sg.n, cols))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    return sg, X.astype(np.float32).astype(np.float64)


def snapshot_bench(snapshots: int = 16, cols: int = 128, alpha: float = 0.1, eps: float = 1e-7, beta: float = 0.5,
                   plain_cols: int = 8, log=print, dataset: str | None = None, sor_cols: int = 0,
                   batched: bool = True) -> list:
    """The dynamic feature-propagation workload: the 169,343-node citation
    graph with 950,000 edges held out and re-inserted in `snapshots` batches,
    X = 128 row-normalised N(0,1) FP32 feature columns.  After the initial
    propagation every snapshot is inserted in place and P updated
    incrementally: all columns at once in FP32 with the FP64 certificate, all
    columns at once in FP64, and -- the per-column form of the reference,
    InstantGNN.py:332-359 -- one FP64 propagation per column for `plain_cols`
    columns (its per-snapshot time scaled to all columns is reported as such).
    With `sor_cols` > 0 the coloured SOR push (instantGNNSorParallel and the
    'sorbatchinsert' updates, omega from the experiment's alpha round-trip)
    runs on that many columns, scaled to all columns the same way.
    With `dataset`, the graph, insert batches and features come from a file
    in the reference's experiment layout (graph.save_exp_npz).
    Returns one dict per (path, snapshot)."""
    from . import graph as _graph
    if dataset:
        sg, feat = _graph.load_exp_npz(dataset)
        X = feat.astype(np.float64)
        cols = X.shape[1]
        snapshots = sg.us.shape[0]
    else:
        sg, X = synthetic_workload(snapshots, cols)
    out = []

    def rec(path, k, st, scale=1.0):
        d = {"path": path, "snapshot": k, "ms": st.ms * scale, "rounds": st.rounds, "rounds_fp64": st.rounds_fp64,
             "oper": st.oper, "oper_cert": st.oper_cert, "cert": st.cert, "keys": st.keys, "inserted": st.inserted}
        out.append(d)
        log(d)

    for prec, name in ((_abi.FLSG_FP32, "batched-fp32-cert"), (_abi.FLSG_FP64, "batched-fp64")) if batched else ():
        prop = DynamicPropagation(sg.copy(), X, alpha, eps, beta, precision=prec)
        try:
            rec(name, -1, prop.initial())
            for k in range(snapshots):
                rec(name, k, prop.insert(sg.us[k], sg.vs[k]))
        finally:
            prop.close()
    if sor_cols > 0:
        from .engine import ignn_sor_columns
        omega = _graph.alpha_roundtrip(alpha)[1]
        prop = DynamicPropagation(sg.copy(), X[:, :sor_cols], alpha, eps, beta, precision=_abi.FLSG_FP64)
        try:
            so = ignn_sor_columns(prop.dg, alpha, eps, beta, omega, prop.P, prop.R)
            st = PropagationStats(ms=so["ms"], rounds=so["iters"], rounds_fp64=so["iters"], oper=so["oper"],
                                  oper_cert=0, cert=0.0)
            rec(f"coloured-sor-fp64 omega {omega:.4f} (x{cols}/{sor_cols})", -1, st, scale=cols / sor_cols)
            for k in range(snapshots):
                rec(f"coloured-sor-fp64 omega {omega:.4f} (x{cols}/{sor_cols})", k,
                    prop.insert(sg.us[k], sg.vs[k], sor_omega=omega), scale=cols / sor_cols)
        finally:
            prop.close()
    if plain_cols > 0:
        props = [DynamicPropagation(sg.copy(), X[:, c:c + 1], alpha, eps, beta, precision=_abi.FLSG_FP64)
                 for c in range(plain_cols)]
        try:
            for k in range(-1, snapshots):
                sts = [p.initial() if k < 0 else p.insert(sg.us[k], sg.vs[k]) for p in props]
                agg = PropagationStats(ms=sum(s.ms for s in sts), rounds=max(s.rounds for s in sts),
                                       rounds_fp64=0, oper=sum(s.oper for s in sts), oper_cert=0,
                                       cert=max(s.cert for s in sts), keys=sts[0].keys, inserted=sts[0].inserted)
                rec(f"per-column-fp64 (x{cols}/{plain_cols})", k, agg, scale=cols / plain_cols)
        finally:
            for p in props:
                p.close()
    return out


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="dynamic feature-propagation snapshot benchmark (one GPU)")
    ap.add_argument("--snapshots", type=int, default=16)
    ap.add_argument("--plain-cols", type=int, default=8)
    ap.add_argument("--sor-cols", type=int, default=0, help="columns of the coloured SOR path (0 = off)")
    ap.add_argument("--skip-batched", action="store_true", help="only the per-column paths")
    ap.add_argument("--out", default=None)
    ap.add_argument("--dataset", default=None, help="workload file in the reference's experiment layout")
    ap.add_argument("--write-dataset", default=None, help="write the synthetic workload to this file and exit")
    args = ap.parse_args()
    if args.write_dataset:
        from . import graph as _graph
        sg, X = synthetic_workload(args.snapshots)
        _graph.save_exp_npz(args.write_dataset, sg, X.astype(np.float32))
        sys.exit(0)
    fh = open(args.out, "w") if args.out else None

    def _log(d):
        line = json.dumps(d)
        print(line, flush=True)
        if fh:
            fh.write(line + "\n")
            fh.flush()

    snapshot_bench(snapshots=args.snapshots, plain_cols=args.plain_cols, log=_log, dataset=args.dataset,
                   sor_cols=args.sor_cols, batched=not args.skip_batched)
    if fh:
        fh.close()
    sys.exit(0)
