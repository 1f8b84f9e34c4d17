"""Serving loop: continuous batching of PPR / Katz / heat-kernel queries on a
graph held resident in HBM, with in-place edge-insert snapshots.

The reference has no serving path; its per-query GPU functions copy the whole
graph to the device on every call (ppr_solver.py:101-107).  Here:

* the graph (relabelled by degree, stored as a slack CSR with per-row headroom
  for inserts) and a slot arena stay resident;
* one persistent batched launch (a serving session, flsg_session_*) runs over
  every slot: each slot takes the next published query as soon as its
  previous one is done, so requests that arrive while others run are
  admitted as slots free up, and PPR, Katz and heat-kernel queries of any
  length share the slots with no launch boundary to wait for;
* admission by memory footprint: a request is published while the session's
  output arena can hold its expected support (a per-kind estimate, raised
  whenever a query overflows; an overflowing query is re-submitted in the
  next session); when the arena or the query ring is used up the session is
  closed and a new one opened;
* edge-insert snapshots are applied between sessions: the loop closes the
  session, lets the published queries finish, inserts on the device and
  opens a new session;
* latency = completion (the host sees the query's done flag, written by the
  kernel into host-coherent memory) minus arrival; p50 / p99 and queries/s
  are reported for the batched path and for the plain FP64 single-source
  path serving the same arrival stream.
"""
from __future__ import annotations

import ctypes as C
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import _abi
from .engine import DeviceGraph, _check, _ptr, make_params, relabel_by_degree
from .graph import CSRGraph, SlackGraph


@dataclass
class Request:
    rid: int
    kind: int                 # FLSG_PPR / FLSG_KATZ / FLSG_HK
    seed: int                 # caller node id
    arrival: float            # seconds from the start of the run
    alpha: float = 0.1        # PPR alpha / Katz beta
    eps: float = 1e-7
    done: float = -1.0
    status: int = -1
    oper: int = 0
    supp: int = 0
    cert: float = 0.0
    device_ms: float = 0.0
    result: tuple | None = None   # (ids, values) when collected


@dataclass
class InsertBatch:
    at: float                 # seconds from the start of the run
    us: np.ndarray
    vs: np.ndarray
    applied: float = -1.0
    inserted: int = 0


def mixed_requests(g: CSRGraph, rate: float, count: int, seed: int = 0, d_max: int | None = None,
                   eps: float = 1e-7, alpha: float = 0.1, mix=(0.7, 0.2, 0.1)) -> list:
    """Poisson arrivals at `rate`/s: PPR (alpha, eps) 70%, Katz
    (beta = 1/(d_max+1), eps) 20%, heat kernel 10%, seeded uniformly at
    non-isolated nodes."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    pool = np.flatnonzero(g.degree >= 1)
    seeds = pool[rng.integers(0, pool.size, size=count)]
    kinds = rng.choice([_abi.FLSG_PPR, _abi.FLSG_KATZ, _abi.FLSG_HK], size=count, p=list(mix))
    beta = 1.0 / ((d_max if d_max is not None else g.d_max) + 1)
    out = []
    for i in range(count):
        k = int(kinds[i])
        out.append(Request(rid=i, kind=k, seed=int(seeds[i]), arrival=float(t[i]),
                           alpha=alpha if k == _abi.FLSG_PPR else beta, eps=eps))
    return out


def random_insert_batches(g: CSRGraph, every: float, size: int, count: int, seed: int = 1) -> list:
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        us = rng.integers(0, g.n, size=size).astype(np.int32)
        vs = rng.integers(0, g.n, size=size).astype(np.int32)
        vs = np.where(vs == us, (vs + 1) % g.n, vs).astype(np.int32)   # no self-loops (the insert rejects them)
        out.append(InsertBatch(at=every * (b + 1), us=us, vs=vs))
    return out


def slack_with_headroom(g: CSRGraph, headroom: int) -> SlackGraph:
    """Slack CSR (data_process_GNN.py:92-119) holding g's rows with `headroom`
    free slots per row for in-place inserts."""
    cap = g.degree.astype(np.int64) + headroom
    indptr = np.zeros(g.n + 1, dtype=np.int64)
    np.cumsum(cap, out=indptr[1:])
    indices = np.full(int(indptr[-1]), g.n, dtype=np.int32)
    rows = np.repeat(np.arange(g.n, dtype=np.int64), g.degree)
    pos = indptr[rows] + (np.arange(rows.size) - np.repeat(g.indptr[:-1], g.degree))
    indices[pos] = g.indices
    return SlackGraph(n=g.n, indptr=indptr, indices=indices, degree=g.degree.astype(np.int32).copy(),
                      us=np.zeros((0, 0), np.int32), vs=np.zeros((0, 0), np.int32), name=g.name + "-slack")


class Server:
    """Continuous-batching server over one resident graph (see the module
    docstring).  `slots` workgroups serve queries (0: one per compute unit);
    `query_cap` queries at most per session; `out_cap` output entries per
    session (default: 3 n per slot)."""

    def __init__(self, g: CSRGraph, slots: int = 0, block: int = 1024, headroom: int = 8,
                 query_cap: int = 1 << 16, out_cap: int | None = None, hk_tau: float = 5.0,
                 hk_eps: float = 1e-7, hk_terms: int = 40, precision: int = _abi.FLSG_FP32,
                 idle_seconds: float = 30.0, device: str = "cuda:0"):
        h, perm, iperm = relabel_by_degree(g)
        self.g = g
        self.precision = precision
        self.iperm = iperm
        self.perm = perm
        sg = slack_with_headroom(h, headroom)
        self.dg = DeviceGraph(sg, device)
        torch = self.dg.torch
        self.torch = torch
        self.d_perm = torch.from_numpy(perm.astype(np.int32)).to(self.dg.device)
        lib = self.dg.lib
        self.lib = lib
        _check(lib.flsg_graph_set_out_map(self.dg.handle, _ptr(self.d_perm)))
        _check(lib.flsg_batch_set_block(self.dg.handle, int(block)))
        if slots <= 0:
            nc = C.c_int32(0)
            free = C.c_int64(0)
            tot = C.c_int64(0)
            _check(lib.flsg_device_info(C.byref(nc), C.byref(free), C.byref(tot)))
            slots = int(nc.value)
        self.slots = slots
        _check(lib.flsg_batch_reserve(self.dg.handle, int(slots), None))
        torch.cuda.synchronize()
        self.query_cap = int(query_cap)
        self.out_cap = int(out_cap) if out_cap else max(1 << 20, int(3 * g.n * slots))
        self.idle_seconds = float(idle_seconds)
        sess = C.c_void_p()
        _check(lib.flsg_session_create(self.dg.handle, self.query_cap, self.out_cap, C.byref(sess)))
        self.sess = sess
        self.hk = _abi.HkConfig(tau=hk_tau, eps=hk_eps, n_terms=hk_terms)
        # per-kind support estimate used for admission (raised on overflow)
        self.est = {_abi.FLSG_PPR: max(1, g.n // 4), _abi.FLSG_KATZ: min(g.n, 50_000),
                    _abi.FLSG_HK: max(1, (g.n * 4) // 5)}
        self.update_stream = torch.cuda.Stream(device=self.dg.device)
        self._done = np.zeros(self.query_cap, dtype=np.int32)
        self._qbuf = (_abi.Query * self.query_cap)()
        self.is_open = False

    # ------------------------------------------------------------------ session plumbing
    def _open(self):
        _check(self.lib.flsg_session_open(self.sess, 0, self.slots, C.byref(self.hk), 0, self.idle_seconds))
        self.is_open = True
        self.ring = []             # ring index -> Request
        self.first_open = 0        # ring indices below are all done
        self.budget = self.out_cap
        self.collected = []        # (Request, out_off, supp) whose output stays in the arena

    def _submit(self, reqs: list) -> None:
        for j, r in enumerate(reqs):
            q = self._qbuf[j]
            q.seed = int(self.iperm[r.seed])
            q.kind = r.kind
            q.precision = self.precision
            q.alpha = r.alpha
            q.eps = r.eps
        first = C.c_int64(0)
        _check(self.lib.flsg_session_submit(self.sess, self._qbuf, len(reqs), C.byref(first)))
        assert first.value == len(self.ring)
        self.ring.extend(reqs)

    def _harvest(self, now: float, requeue: deque) -> int:
        """Record every query of the open session whose done flag is set."""
        lo, hi = self.first_open, len(self.ring)
        if hi <= lo:
            return 0
        _check(self.lib.flsg_session_done(self.sess, lo, hi - lo, self._done.ctypes.data))
        finished = 0
        rr = _abi.Result()
        off = C.c_int64(0)
        for i in range(lo, hi):
            r = self.ring[i]
            if r is None or not self._done[i - lo]:
                continue
            _check(self.lib.flsg_session_result(self.sess, i, C.byref(rr), C.byref(off)))
            self.ring[i] = None
            if rr.status == _abi.FLSG_ECAP:
                self.est[r.kind] = max(self.est[r.kind], int(rr.supp * 1.25) + 1)
                requeue.append(r)
                continue
            finished += 1
            r.done = now
            r.status = rr.status
            r.oper = rr.oper
            r.supp = rr.supp
            r.cert = rr.cert
            r.device_ms = rr.ms
            self.est[r.kind] = max(self.est[r.kind], int(rr.supp * 1.1) + 1)
            if rr.status == _abi.FLSG_OK and off.value >= 0:
                self.collected.append((r, int(off.value), int(rr.supp)))
        while self.first_open < hi and self.ring[self.first_open] is None:
            self.first_open += 1
        return finished

    def _end(self, t0: float, time_scale: float, collect: bool, requeue: deque) -> int:
        """Close the open session, wait for its launch to end (recording the
        queries as they finish), then copy the outputs out if asked and
        re-queue whatever was published but not done (an idle exit)."""
        _check(self.lib.flsg_session_close(self.sess))
        t_close = time.perf_counter()
        finished = 0
        ended = C.c_int32(0)
        idle = C.c_int64(0)
        while True:
            _check(self.lib.flsg_session_poll(self.sess, C.byref(ended), None))
            finished += self._harvest((time.perf_counter() - t0) / time_scale, requeue)
            if ended.value:
                break
            time.sleep(0)
        self.drain_s += time.perf_counter() - t_close
        _check(self.lib.flsg_session_poll(self.sess, C.byref(ended), C.byref(idle)))
        finished += self._harvest((time.perf_counter() - t0) / time_scale, requeue)
        for i in range(self.first_open, len(self.ring)):
            if self.ring[i] is not None:     # published, never run (idle exit)
                requeue.append(self.ring[i])
                self.ring[i] = None
        self.idle_exits = int(idle.value)
        if collect:
            for r, o, k in self.collected:
                ids = np.empty(k, dtype=np.int32)
                vals = np.empty(k, dtype=np.float64)
                _check(self.lib.flsg_session_output(self.sess, o, k, ids.ctypes.data, vals.ctypes.data))
                r.result = (ids, vals)
        self.collected = []
        self.is_open = False
        return finished

    def _apply_insert(self, ins: InsertBatch):
        torch = self.torch
        lib = self.lib
        dev = self.dg.device
        us = torch.from_numpy(self.iperm[ins.us].astype(np.int32)).to(dev)
        vs = torch.from_numpy(self.iperm[ins.vs].astype(np.int32)).to(dev)
        keys = torch.empty(max(2 * us.numel(), 1), dtype=torch.int32, device=dev)
        nk, ni = C.c_int64(0), C.c_int64(0)
        with torch.cuda.stream(self.update_stream):
            _check(lib.flsg_graph_insert_edges(self.dg.handle, _ptr(us), _ptr(vs), int(us.numel()), _ptr(keys),
                                               C.byref(nk), None, C.byref(ni),
                                               C.c_void_p(self.update_stream.cuda_stream)))
            _check(lib.flsg_batch_refresh_degrees(self.dg.handle, _ptr(keys), nk.value,
                                                  C.c_void_p(self.update_stream.cuda_stream)))
        self.update_stream.synchronize()
        ins.inserted = int(ni.value)

    # ------------------------------------------------------------------ loop
    def run(self, requests: list, inserts: list | None = None, collect: bool = False,
            time_scale: float = 1.0) -> dict:
        """Serve `requests` (arrival times in seconds, replayed in real time
        scaled by time_scale) and apply `inserts` at their times."""
        inserts = sorted(inserts or [], key=lambda b: b.at)
        order = sorted(requests, key=lambda r: r.arrival)
        pending: deque = deque()
        requeue: deque = deque()
        nxt = 0
        ins_i = 0
        total = len(order)
        done = 0
        sessions = 0
        submits = 0
        idle_exits = 0
        t0 = time.perf_counter()
        self.drain_s = 0.0
        self._open()
        sessions += 1
        try:
            while True:
                now = (time.perf_counter() - t0) / time_scale
                while nxt < total and order[nxt].arrival <= now:
                    pending.append(order[nxt])
                    nxt += 1
                done += self._harvest(now, requeue)
                while requeue:
                    pending.appendleft(requeue.pop())
                due = ins_i < len(inserts) and inserts[ins_i].at <= now
                finished_all = done >= total and nxt >= total and not pending
                if due or (finished_all and ins_i >= len(inserts)):
                    # drain the session before mutating the graph (or at the end)
                    done += self._end(t0, time_scale, collect, requeue)
                    idle_exits += self.idle_exits
                    if due:
                        self._apply_insert(inserts[ins_i])
                        inserts[ins_i].applied = (time.perf_counter() - t0) / time_scale
                        ins_i += 1
                    while requeue:
                        pending.appendleft(requeue.pop())
                    if done >= total and nxt >= total and not pending and ins_i >= len(inserts):
                        break
                    self._open()
                    sessions += 1
                    continue
                if pending:
                    # admission: publish while the output arena can hold the
                    # expected supports; a full arena or ring starts a new session
                    batch = []
                    room = self.query_cap - len(self.ring)
                    while pending and len(batch) < room:
                        need = self.est[pending[0].kind]
                        if need > self.budget:
                            break
                        self.budget -= need
                        batch.append(pending.popleft())
                    if batch:
                        self._submit(batch)
                        submits += 1
                    elif self.first_open >= len(self.ring):
                        # nothing in flight and the next request does not fit
                        done += self._end(t0, time_scale, collect, requeue)
                        idle_exits += self.idle_exits
                        while requeue:
                            pending.appendleft(requeue.pop())
                        self._open()
                        sessions += 1
                        continue
                time.sleep(0)
        finally:
            if self.is_open:
                self._end(t0, time_scale, collect, requeue)
        wall = time.perf_counter() - t0
        st = self._stats(order, wall, sessions, inserts)
        st["submits"] = submits
        st["idle_exits"] = idle_exits
        st["drain_s"] = self.drain_s
        return st

    @staticmethod
    def _stats(order: list, wall: float, launches: int, inserts: list) -> dict:
        lat = np.array([r.done - r.arrival for r in order if r.done >= 0])
        ok = [r for r in order if r.status == _abi.FLSG_OK]
        certified = all(r.cert < 1.0 for r in ok if r.kind != _abi.FLSG_HK)
        span = max(r.done for r in order) - min(r.arrival for r in order) if order else 0.0
        return {
            "queries": len(order), "ok": len(ok), "launches": launches, "wall_s": wall,
            "qps": len(ok) / span if span > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
            "device_ms_mean": float(np.mean([r.device_ms for r in ok])) if ok else None,
            "certified": certified,
            "inserts_applied": sum(1 for b in inserts if b.applied >= 0),
            "edges_inserted": sum(b.inserted for b in inserts),
            "by_kind": {name: Server._kind_stats([r for r in order if r.kind == k])
                        for k, name in ((_abi.FLSG_PPR, "ppr"), (_abi.FLSG_KATZ, "katz"), (_abi.FLSG_HK, "hk"))},
        }

    @staticmethod
    def _kind_stats(rs: list) -> dict:
        lat = np.array([r.done - r.arrival for r in rs if r.done >= 0])
        dev = np.array([r.device_ms for r in rs if r.done >= 0])
        if not lat.size:
            return {"n": 0}
        return {"n": int(lat.size), "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "device_ms_p50": float(np.percentile(dev, 50))}

    def close(self):
        if self.sess:
            if self.is_open:
                self._end(time.perf_counter(), 1.0, False, deque())
            self.lib.flsg_session_destroy(self.sess)
            self.sess = None
        self.dg.close()


def run_plain(g: CSRGraph, requests: list, hk_tau: float = 5.0, hk_eps: float = 1e-7, hk_terms: int = 40,
              device: str = "cuda:0", dg: DeviceGraph | None = None, time_scale: float = 1.0) -> dict:
    """The same arrival stream served one query at a time by the plain FP64
    single-source path (whole GPU per query, FIFO)."""
    from .engine import solve
    own = dg is None
    dg = dg or DeviceGraph(g, device, relabel=True)
    torch = dg.torch
    order = sorted(requests, key=lambda r: r.arrival)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in order:
        now = (time.perf_counter() - t0) / time_scale
        if r.arrival > now:
            time.sleep((r.arrival - now) * time_scale)
        if r.kind == _abi.FLSG_HK:
            p = make_params(_abi.FLSG_HK, _abi.FLSG_GD, True, _abi.FLSG_FP64, eps=hk_eps, tau=hk_tau,
                            n_terms=hk_terms)
        else:
            p = make_params(r.kind, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=r.alpha, eps=r.eps)
        out = solve(dg, p, r.seed)
        torch.cuda.synchronize()
        r.done = (time.perf_counter() - t0) / time_scale
        r.status = 0
        r.oper = out.result["oper"]
        r.device_ms = out.result["ms"]
    wall = time.perf_counter() - t0
    if own:
        dg.close()
    st = Server._stats(order, wall, len(order), [])
    st["certified"] = True
    return st


def serving_bench(g: CSRGraph | None = None, rates=(1000, 2000, 5000, 10000, 20000, 50000), seconds: float = 2.0,
                  max_queries: int = 4000, plain_queries: int = 600, insert_every: float = 1.0,
                  insert_size: int = 10000, precisions=(_abi.FLSG_FP32, _abi.FLSG_FP64), log=print) -> list:
    """The serving mix on the 2,394,385-node power-law graph: Poisson arrivals
    at each offered rate (PPR 70% alpha=0.1, Katz 20% beta=1/(d_max+1), heat
    kernel 10% t=5 order 40; eps=1e-7), min(rate*seconds, max_queries)
    requests per rate, a batch of insert_size random edge inserts every
    insert_every seconds (at least one per rate), served by the batched path
    at each precision and -- plain_queries requests per rate -- by the plain
    FP64 single-source path.  Returns one dict per (path, rate)."""
    from . import graph as _graph
    g = g if g is not None else _graph.powerlaw_graph_large()
    out = []
    for prec in precisions:
        name = "batched-fp32" if prec == _abi.FLSG_FP32 else "batched-fp64"
        srv = Server(g, precision=prec)
        try:
            srv.run(mixed_requests(g, rate=1e7, count=256, seed=99))      # warm-up (not reported)
            for rate in rates:
                count = int(min(rate * seconds, max_queries))
                reqs = mixed_requests(g, rate=float(rate), count=count, seed=int(rate))
                span = reqs[-1].arrival
                ins = random_insert_batches(g, every=insert_every, size=insert_size,
                                            count=max(1, int(span // insert_every)), seed=int(rate) + 7)
                st = srv.run(reqs, inserts=ins)
                rec = {"path": name, "offered_qps": rate, **st}
                out.append(rec)
                log(rec)
        finally:
            srv.close()
    if plain_queries > 0:
        dg = DeviceGraph(g, relabel=True)
        try:
            run_plain(g, mixed_requests(g, rate=1e7, count=8, seed=98), dg=dg)      # warm-up
            for rate in rates:
                reqs = mixed_requests(g, rate=float(rate), count=plain_queries, seed=int(rate))
                st = run_plain(g, reqs, dg=dg)
                rec = {"path": "plain-fp64", "offered_qps": rate, **st}
                out.append(rec)
                log(rec)
        finally:
            dg.close()
    return out


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="serving-mix benchmark (one GPU)")
    ap.add_argument("--rates", default="1000,2000,5000,10000,20000,50000")
    ap.add_argument("--seconds", type=float, default=2.0)
    ap.add_argument("--max-queries", type=int, default=4000)
    ap.add_argument("--plain-queries", type=int, default=600)
    ap.add_argument("--out", default=None, help="JSON-lines file for the results")
    ap.add_argument("--precisions", default="fp32,fp64", help="batched paths to run: fp32, fp64 or both")
    args = ap.parse_args()
    fh = open(args.out, "w") if args.out else None

    def _log(rec):
        line = json.dumps(rec)
        print(line, flush=True)
        if fh:
            fh.write(line + "\n")
            fh.flush()

    precs = tuple({"fp32": _abi.FLSG_FP32, "fp64": _abi.FLSG_FP64}[p] for p in args.precisions.split(",") if p)
    serving_bench(rates=tuple(int(r) for r in args.rates.split(",")), seconds=args.seconds,
                  max_queries=args.max_queries, plain_queries=args.plain_queries, precisions=precs, log=_log)
    if fh:
        fh.close()
    sys.exit(0)
