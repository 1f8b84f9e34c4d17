"""Serving loop: continuous batching of PPR / Katz / heat-kernel queries on a
graph held resident in HBM, with in-place edge-insert snapshots.

The reference has no serving path; its per-query GPU functions copy the whole
graph to the device on every call (ppr_solver.py:101-107).  Here:

* the graph (relabelled by degree, stored as a slack CSR with per-row headroom
  for inserts) and a slot arena stay resident;
* the slot arena is split into lanes -- each lane has its own HIP stream, slot
  range, query / result / output buffers -- so launches on different lanes run
  concurrently on the device;
* requests that arrive while lanes are busy wait in a queue and are admitted
  into the next free lane's launch as long as the lane's output arena can hold
  their expected footprint (a per-kind support estimate, raised whenever a
  query overflows; an overflowing query is re-queued);
* edge-insert snapshots are applied between launches: the loop stops admitting,
  lets in-flight launches finish, inserts on the device and refreshes the slots'
  degree copies;
* latency = completion (the host observes, by flsg_stream_poll, that the
  lane's stream has drained) minus
  arrival; p50 / p99 and queries/s are reported for the FP32-batched path and
  for the plain FP64 single-source path serving the same arrival stream.
"""
from __future__ import annotations

import ctypes as C
import time
from collections import deque
from dataclasses import dataclass, field

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


class _Lane:
    def __init__(self, srv: "Server", idx: int, slot_base: int, n_slots: int, hk: bool = False):
        torch = srv.torch
        dev = srv.dg.device
        self.idx = idx
        self.hk = hk                 # heat-kernel lane (takes short queries only when no HK query waits)
        self.slot_base = slot_base
        self.n_slots = n_slots
        self.stream = torch.cuda.Stream(device=dev)
        self.qcap = srv.max_batch
        # queries are staged in pinned memory and uploaded asynchronously on the lane's stream
        self.h_qbuf = torch.empty(self.qcap * C.sizeof(_abi.Query), dtype=torch.uint8).pin_memory()
        self.h_q = (_abi.Query * self.qcap).from_address(self.h_qbuf.data_ptr())
        self.d_q = torch.empty(self.qcap * C.sizeof(_abi.Query), dtype=torch.uint8, device=dev)
        self.d_res = torch.empty(self.qcap * C.sizeof(_abi.Result), dtype=torch.uint8, device=dev)
        self.h_res = torch.empty(self.qcap * C.sizeof(_abi.Result), dtype=torch.uint8).pin_memory()
        self.d_off = torch.empty(self.qcap, dtype=torch.int64, device=dev)
        self.h_off = torch.empty(self.qcap, dtype=torch.int64).pin_memory()
        self.d_ctl = torch.zeros(2, dtype=torch.int64, device=dev)
        # output arena in (id, value) entries: sized to the lane's slots and the
        # kind it serves (heat-kernel supports cover most of the graph)
        self.out_cap = srv.out_per_lane or max(1 << 20, int(n_slots * srv.g.n * (0.9 if hk else 0.5)))
        self.ids = torch.empty(self.out_cap, dtype=torch.int32, device=dev)
        self.vals = torch.empty(self.out_cap, dtype=torch.float64, device=dev)
        self.busy = False
        self.event = None
        self.reqs: list = []


class Server:
    def __init__(self, g: CSRGraph, lanes: int = 4, slots: int = 0, block: int = 1024, headroom: int = 8,
                 max_batch: int = 256, out_per_lane: int | None = None, hk_tau: float = 5.0,
                 hk_eps: float = 1e-7, hk_terms: int = 40, precision: int = _abi.FLSG_FP32,
                 hk_lanes: int | None = None, hk_slot_share: float = 0.5, device: str = "cuda:0"):
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
        self.max_batch = max_batch
        self.out_per_lane = out_per_lane
        # A launch ends with its slowest query, and a heat-kernel query does ~7x
        # the work of a PPR query: heat-kernel queries get lanes (and a slot
        # share) of their own so they never hold PPR / Katz slots idle.
        if hk_lanes is None:
            hk_lanes = 1 if lanes >= 2 else 0
        hk_lanes = max(0, min(hk_lanes, lanes - 1))
        short = lanes - hk_lanes
        hk_slots = int(round(slots * hk_slot_share)) if hk_lanes else 0
        hk_slots = min(max(hk_slots, hk_lanes), slots - short)
        sizes = [((slots - hk_slots) // short) + (1 if i < (slots - hk_slots) % short else 0) for i in range(short)]
        if hk_lanes:
            sizes += [(hk_slots // hk_lanes) + (1 if i < hk_slots % hk_lanes else 0) for i in range(hk_lanes)]
        self.lanes = []
        base = 0
        for i, sz in enumerate(sizes):
            self.lanes.append(_Lane(self, i, base, sz, hk=i >= short))
            base += sz
        self.hk = _abi.HkConfig(tau=hk_tau, eps=hk_eps, n_terms=hk_terms)
        # per-kind support estimate used for admission (raised on overflow)
        self.est = {_abi.FLSG_PPR: max(1, g.n // 4), _abi.FLSG_KATZ: min(g.n, 50_000),
                    _abi.FLSG_HK: max(1, (g.n * 4) // 5)}
        self.update_stream = torch.cuda.Stream(device=self.dg.device)

    # ------------------------------------------------------------------ plumbing
    def _launch(self, lane: _Lane, reqs: list, t0: float):
        torch = self.torch
        for j, r in enumerate(reqs):
            q = lane.h_q[j]
            q.seed = int(self.iperm[r.seed])
            q.kind = r.kind
            q.precision = self.precision
            q.alpha = r.alpha
            q.eps = r.eps
        nbytes = len(reqs) * C.sizeof(_abi.Query)
        with torch.cuda.stream(lane.stream):
            lane.d_q[:nbytes].copy_(lane.h_qbuf[:nbytes], non_blocking=True)
            _check(self.lib.flsg_batch_launch(self.dg.handle, _ptr(lane.d_q), len(reqs), lane.slot_base,
                                              min(lane.n_slots, len(reqs)), _ptr(lane.ids), _ptr(lane.vals),
                                              _ptr(lane.d_off), lane.out_cap, _ptr(lane.d_ctl), _ptr(lane.d_res),
                                              0, C.byref(self.hk), C.c_void_p(lane.stream.cuda_stream)))
            rb = len(reqs) * C.sizeof(_abi.Result)
            lane.h_res[:rb].copy_(lane.d_res[:rb], non_blocking=True)
            lane.h_off[:len(reqs)].copy_(lane.d_off[:len(reqs)], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(lane.stream)
        lane.event = ev
        lane.busy = True
        lane.reqs = reqs

    def _poll(self, lane: _Lane) -> bool:
        """Whether the lane's launch and result copies have completed
        (flsg_stream_poll on the lane's stream; never blocks)."""
        done = C.c_int32(0)
        _check(self.lib.flsg_stream_poll(C.c_void_p(lane.stream.cuda_stream), C.byref(done)))
        return bool(done.value)

    def _complete(self, lane: _Lane, now: float, collect: bool, requeue: deque) -> int:
        res = (_abi.Result * len(lane.reqs)).from_buffer_copy(
            lane.h_res[:len(lane.reqs) * C.sizeof(_abi.Result)].numpy().tobytes())
        offs = lane.h_off[:len(lane.reqs)].numpy()
        finished = 0
        for j, r in enumerate(lane.reqs):
            rr = res[j]
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
            if collect and rr.status == _abi.FLSG_OK:
                o, k = int(offs[j]), int(rr.supp)
                r.result = (lane.ids[o:o + k].cpu().numpy(), lane.vals[o:o + k].cpu().numpy())
        lane.busy = False
        lane.reqs = []
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
        pend_short: deque = deque()
        pend_hk: deque = deque()
        requeue: deque = deque()
        has_hk_lane = any(l.hk for l in self.lanes)
        nxt = 0
        ins_i = 0
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        total = len(order)
        launches = 0
        while done < total or ins_i < len(inserts):
            now = (time.perf_counter() - t0) / time_scale
            while nxt < total and order[nxt].arrival <= now:
                r = order[nxt]
                (pend_hk if has_hk_lane and r.kind == _abi.FLSG_HK else pend_short).append(r)
                nxt += 1
            for lane in self.lanes:
                if lane.busy and self._poll(lane):
                    done += self._complete(lane, now, collect, requeue)
            while requeue:
                r = requeue.pop()
                (pend_hk if has_hk_lane and r.kind == _abi.FLSG_HK else pend_short).appendleft(r)
            due = ins_i < len(inserts) and inserts[ins_i].at <= now
            if due:
                if any(l.busy for l in self.lanes):
                    continue          # drain before mutating the graph
                self._apply_insert(inserts[ins_i])
                inserts[ins_i].applied = (time.perf_counter() - t0) / time_scale
                ins_i += 1
                continue
            if not pend_short and not pend_hk:
                if nxt >= total and not any(l.busy for l in self.lanes) and ins_i >= len(inserts):
                    break
                time.sleep(0)
                continue
            for lane in self.lanes:
                if lane.busy:
                    continue
                # heat-kernel lanes fall back to short queries when no HK query waits
                pending = pend_hk if (lane.hk and pend_hk) else pend_short
                if not pending:
                    continue
                batch, budget = [], lane.out_cap
                while pending and len(batch) < lane.qcap:
                    need = self.est[pending[0].kind]
                    if batch and need > budget:
                        break
                    budget -= need
                    batch.append(pending.popleft())
                self._launch(lane, batch, t0)
                launches += 1
        wall = time.perf_counter() - t0
        return self._stats(order, wall, launches, inserts)

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
        }

    def close(self):
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
    args = ap.parse_args()
    fh = open(args.out, "w") if args.out else None

    def _log(rec):
        line = json.dumps(rec)
        print(line, flush=True)
        if fh:
            fh.write(line + "\n")
            fh.flush()

    serving_bench(rates=tuple(int(r) for r in args.rates.split(",")), seconds=args.seconds,
                  max_queries=args.max_queries, plain_queries=args.plain_queries, log=_log)
    if fh:
        fh.close()
    sys.exit(0)
