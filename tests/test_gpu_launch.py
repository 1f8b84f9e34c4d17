"""Asynchronous batched launches through the C ABI (flsg_batch_launch,
flsg_stream_poll): two launches over disjoint slot ranges of one arena, on
two streams at once, each over a caller-held query list that mixes kinds
and precisions.

Bars: every PPR / Katz result certified by an independent host FP64
recompute of its residual (|r_v| < eps alpha d_v for PPR, eps d_v for Katz);
the FP64 PPR queries' operation counts equal those of the same seeds in a
flsg_batch_solve launch (FP64 push: integer work, bit-exact); heat-kernel
results equal to the oracle's level push up to FP32 rounding (relative L1
1e-5); the arena back in its between-queries state afterwards.
"""
import ctypes as C
import time

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from flsg import _abi, engine, graph  # noqa: E402
from oracle import oracle as O  # noqa: E402

ALPHA, EPS = 0.1, 1e-6
TAU, HK_EPS, HK_N = 5.0, 1e-7, 40


def _queries(rows):
    qs = (_abi.Query * len(rows))()
    for i, (seed, kind, prec, a) in enumerate(rows):
        qs[i].seed = int(seed)
        qs[i].kind = kind
        qs[i].precision = prec
        qs[i].alpha = a
        qs[i].eps = EPS
    return qs


def _to_device(torch, obj, device):
    raw = np.frombuffer(bytes(obj), dtype=np.uint8).copy()
    return torch.from_numpy(raw).to(device)


class _Launch:
    def __init__(self, dg, rows, slot_base, n_slots, out_cap, hk=None):
        torch = dg.torch
        self.rows = rows
        self.nq = len(rows)
        self.stream = torch.cuda.Stream(device=dg.device)   # default priority
        self.d_q = _to_device(torch, _queries(rows), dg.device)
        self.d_ids = torch.empty(out_cap, dtype=torch.int32, device=dg.device)
        self.d_vals = torch.empty(out_cap, dtype=torch.float64, device=dg.device)
        self.d_off = torch.empty(self.nq, dtype=torch.int64, device=dg.device)
        self.d_ctl = torch.zeros(2, dtype=torch.int64, device=dg.device)
        self.d_res = torch.zeros(self.nq * C.sizeof(_abi.Result), dtype=torch.uint8, device=dg.device)
        torch.cuda.synchronize(dg.device)
        rc = dg.lib.flsg_batch_launch(dg.handle, engine._ptr(self.d_q), self.nq, slot_base, n_slots,
                                      engine._ptr(self.d_ids), engine._ptr(self.d_vals), engine._ptr(self.d_off),
                                      out_cap, engine._ptr(self.d_ctl), engine._ptr(self.d_res), 0,
                                      C.byref(hk) if hk is not None else None, C.c_void_p(self.stream.cuda_stream))
        assert rc == 0, dg.lib.flsg_last_error()

    def done(self, lib):
        flag = C.c_int32(0)
        assert lib.flsg_stream_poll(C.c_void_p(self.stream.cuda_stream), C.byref(flag)) == 0
        return bool(flag.value)

    def results(self):
        raw = self.d_res.cpu().numpy().tobytes()
        res = (_abi.Result * self.nq).from_buffer_copy(raw)
        return res, self.d_off.cpu().numpy(), self.d_ids.cpu().numpy(), self.d_vals.cpu().numpy()


def test_two_concurrent_launches_mixed_kinds(small_graph):
    g = small_graph
    dg = engine.DeviceGraph(g)
    lib = dg.lib
    try:
        assert lib.flsg_batch_reserve(dg.handle, 16, None) == 0
        beta = 1.0 / (float(g.degree.max()) + 1.0)
        seeds = graph.mixed_seeds(g, 36, seed=71)
        kinds_a = [(_abi.FLSG_PPR, _abi.FLSG_FP64, ALPHA), (_abi.FLSG_PPR, _abi.FLSG_FP32, ALPHA),
                   (_abi.FLSG_KATZ, _abi.FLSG_FP64, beta), (_abi.FLSG_KATZ, _abi.FLSG_FP32, beta)]
        rows_a = [(int(s),) + kinds_a[i % 4] for i, s in enumerate(seeds[:24])]
        rows_b = [(int(s), _abi.FLSG_HK if i % 2 else _abi.FLSG_PPR, _abi.FLSG_FP32, ALPHA)
                  for i, s in enumerate(seeds[24:])]
        hk = _abi.HkConfig(tau=TAU, eps=HK_EPS, n_terms=HK_N)
        la = _Launch(dg, rows_a, 0, 8, 24 * g.n)
        lb = _Launch(dg, rows_b, 8, 8, 12 * g.n, hk=hk)
        t0 = time.time()
        while not (la.done(lib) and lb.done(lib)):
            assert time.time() - t0 < 60.0, "launches did not end"
            time.sleep(0.001)
        for launch in (la, lb):
            res, off, ids, vals = launch.results()
            for q, (seed, kind, prec, a) in enumerate(launch.rows):
                rr = res[q]
                assert rr.status == _abi.FLSG_OK, (q, rr.status)
                k = int(rr.supp)
                o = int(off[q])
                if kind == _abi.FLSG_HK:
                    x = np.zeros(g.n)
                    x[ids[o:o + k]] = vals[o:o + k]
                    ref = O.hk_push(g, seed, TAU, HK_EPS, local=True, n_terms=HK_N)
                    assert np.abs(x - ref["x"]).sum() <= 1e-5 * np.abs(ref["x"]).sum()
                    continue
                assert rr.cert < 1.0
                cert, _ = O.residual_cert(g, 0 if kind == _abi.FLSG_PPR else 1, a, EPS, [seed],
                                          [0], [k], ids[o:o + k], vals[o:o + k])
                assert cert[0] < 1.0 + 1e-9, (q, kind, prec, cert[0])
        # FP64 PPR: the same operation counts as a flsg_batch_solve launch
        ppr64 = [(q, r[0]) for q, r in enumerate(rows_a) if r[1] == _abi.FLSG_PPR and r[2] == _abi.FLSG_FP64]
        p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=ALPHA, eps=EPS)
        bo = engine.batch_solve(dg, p, [s for _, s in ppr64])
        res_a = la.results()[0]
        for i, (q, _) in enumerate(ppr64):
            assert res_a[q].oper == int(bo.results[i]["oper"])
        dg.torch.cuda.synchronize()
        assert engine.arena_dirty(dg) == 0
    finally:
        dg.close()


def test_launch_refuses_bad_slot_range(small_graph):
    g = small_graph
    dg = engine.DeviceGraph(g)
    lib = dg.lib
    try:
        assert lib.flsg_batch_reserve(dg.handle, 8, None) == 0
        torch = dg.torch
        rows = [(int(graph.uniform_seeds(g, 1, seed=72)[0]), _abi.FLSG_PPR, _abi.FLSG_FP64, ALPHA)]
        d_q = _to_device(torch, _queries(rows), dg.device)
        ids = torch.empty(g.n, dtype=torch.int32, device=dg.device)
        vals = torch.empty(g.n, dtype=torch.float64, device=dg.device)
        off = torch.empty(1, dtype=torch.int64, device=dg.device)
        ctl = torch.zeros(2, dtype=torch.int64, device=dg.device)
        res = torch.zeros(C.sizeof(_abi.Result), dtype=torch.uint8, device=dg.device)
        rc = lib.flsg_batch_launch(dg.handle, engine._ptr(d_q), 1, 4, 8, engine._ptr(ids), engine._ptr(vals),
                                   engine._ptr(off), g.n, engine._ptr(ctl), engine._ptr(res), 0, None, None)
        assert rc == _abi.FLSG_EARG
    finally:
        dg.close()
