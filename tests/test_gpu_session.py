"""Serving sessions through the C ABI (flsg_session_*): one persistent batched
launch whose slots take each query as the host publishes it.

Bars: queries published in several submits while the launch runs come back
done with the same certified results (PPR: |r_v| < eps alpha d_v in FP64 on
the host) and the same operation counts as the same seeds in one batched
launch (FP64 push: integer work, bit-exact op counts); the bounded wait of an
idle launch ends it by itself (forced from the host with a short idle bound)
and a new session then serves normally; misuse is refused with FLSG_EARG /
FLSG_ECAP and nothing runs.
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


def _session(dg, qcap, out_cap):
    s = C.c_void_p()
    assert dg.lib.flsg_session_create(dg.handle, qcap, out_cap, C.byref(s)) == 0
    return s


def _wait_ended(lib, s, limit=60.0):
    ended, idle = C.c_int32(0), C.c_int64(0)
    t0 = time.time()
    while True:
        assert lib.flsg_session_poll(s, C.byref(ended), C.byref(idle)) == 0
        if ended.value:
            return idle.value
        assert time.time() - t0 < limit, "session launch did not end"
        time.sleep(0.001)


def _queries(seeds, precision):
    qs = (_abi.Query * len(seeds))()
    for i, sd in enumerate(seeds):
        qs[i].seed = int(sd)
        qs[i].kind = _abi.FLSG_PPR
        qs[i].precision = precision
        qs[i].alpha = ALPHA
        qs[i].eps = EPS
    return qs


@pytest.mark.parametrize("precision", [_abi.FLSG_FP64, _abi.FLSG_FP32])
def test_session_matches_batched_launch(small_graph, precision):
    g = small_graph
    dg = engine.DeviceGraph(g)
    lib = dg.lib
    try:
        assert lib.flsg_batch_reserve(dg.handle, 16, None) == 0
        seeds = graph.mixed_seeds(g, 40, seed=61)
        p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, precision, alpha=ALPHA, eps=EPS)
        ref = engine.batch_solve(dg, p, seeds)
        s = _session(dg, 64, 40 * g.n)
        try:
            assert lib.flsg_session_open(s, 0, 16, None, 0, 30.0) == 0
            qs = _queries(seeds, precision)
            first = C.c_int64(-1)
            # three submits while the launch runs (the first slots are already waiting)
            for lo, hi in ((0, 5), (5, 23), (23, 40)):
                sub = (_abi.Query * (hi - lo))(*qs[lo:hi])
                assert lib.flsg_session_submit(s, sub, hi - lo, C.byref(first)) == 0
                assert first.value == lo
                time.sleep(0.002)
            done = np.zeros(40, dtype=np.int32)
            t0 = time.time()
            while not done.all():
                assert lib.flsg_session_done(s, 0, 40, done.ctypes.data) == 0
                assert time.time() - t0 < 60.0
            assert lib.flsg_session_close(s) == 0
            assert _wait_ended(lib, s) == 0
            thr = EPS * ALPHA * np.maximum(g.degree, 1)
            rr = _abi.Result()
            off = C.c_int64(0)
            for q in range(40):
                assert lib.flsg_session_result(s, q, C.byref(rr), C.byref(off)) == 0
                assert rr.status == _abi.FLSG_OK and rr.cert < 1.0
                if precision == _abi.FLSG_FP64:
                    assert rr.oper == int(ref.results["oper"][q])
                k = int(rr.supp)
                ids = np.empty(k, dtype=np.int32)
                vals = np.empty(k, dtype=np.float64)
                assert lib.flsg_session_output(s, off.value, k, ids.ctypes.data, vals.ctypes.data) == 0
                x = np.zeros(g.n)
                x[ids] = vals
                r = O.ppr_residual(g, int(seeds[q]), ALPHA, x)
                assert (np.abs(r) / thr).max() < 1.0 + 1e-9
                xb = ref.dense(q, g.n)
                rb = O.ppr_residual(g, int(seeds[q]), ALPHA, xb)
                # both certified: within the two residuals (ppr_solver.py:321-328 certificate)
                assert np.abs(x - xb).sum() <= (np.abs(r).sum() + np.abs(rb).sum()) / ALPHA * (1 + 1e-9) + 1e-15
        finally:
            lib.flsg_session_destroy(s)
        dg.torch.cuda.synchronize()
        assert engine.arena_dirty(dg) == 0
    finally:
        dg.close()


def test_session_idle_exit_then_new_session(small_graph):
    """A launch with nothing published ends by itself after its idle bound
    (here 50 ms), reports the idle exit, and the next session serves."""
    g = small_graph
    dg = engine.DeviceGraph(g)
    lib = dg.lib
    try:
        assert lib.flsg_batch_reserve(dg.handle, 8, None) == 0
        s = _session(dg, 16, 16 * g.n)
        try:
            assert lib.flsg_session_open(s, 0, 8, None, 0, 0.05) == 0
            assert _wait_ended(lib, s, limit=30.0) >= 1
            # a new session on the same object serves a query
            assert lib.flsg_session_open(s, 0, 8, None, 0, 30.0) == 0
            seed = int(graph.uniform_seeds(g, 1, seed=62)[0])
            qs = _queries([seed], _abi.FLSG_FP32)
            assert lib.flsg_session_submit(s, qs, 1, None) == 0
            done = np.zeros(1, dtype=np.int32)
            t0 = time.time()
            while not done[0]:
                assert lib.flsg_session_done(s, 0, 1, done.ctypes.data) == 0
                assert time.time() - t0 < 60.0
            assert lib.flsg_session_close(s) == 0
            assert _wait_ended(lib, s) == 0
            rr = _abi.Result()
            assert lib.flsg_session_result(s, 0, C.byref(rr), None) == 0
            assert rr.status == _abi.FLSG_OK and rr.cert < 1.0
        finally:
            lib.flsg_session_destroy(s)
        dg.torch.cuda.synchronize()
        assert engine.arena_dirty(dg) == 0
    finally:
        dg.close()


def test_session_refuses_misuse(small_graph):
    g = small_graph
    dg = engine.DeviceGraph(g)
    lib = dg.lib
    try:
        assert lib.flsg_batch_reserve(dg.handle, 8, None) == 0
        s = _session(dg, 4, 4 * g.n)
        try:
            seed = int(graph.uniform_seeds(g, 1, seed=63)[0])
            qs = _queries([seed] * 5, _abi.FLSG_FP64)
            # not open yet: nothing may be published
            assert lib.flsg_session_submit(s, qs, 1, None) == _abi.FLSG_EARG
            # a slot range outside the reserved arena
            assert lib.flsg_session_open(s, 4, 8, None, 0, 30.0) == _abi.FLSG_EARG
            assert lib.flsg_session_open(s, 0, 8, None, 0, 0.0) == _abi.FLSG_EARG
            assert lib.flsg_session_open(s, 0, 8, None, 0, 30.0) == 0
            # a second open while the launch runs
            assert lib.flsg_session_open(s, 0, 8, None, 0, 30.0) == _abi.FLSG_EARG
            # more queries than the ring holds
            assert lib.flsg_session_submit(s, qs, 5, None) == _abi.FLSG_ECAP
            # outputs only once the launch has ended
            ids = np.empty(1, dtype=np.int32)
            vals = np.empty(1, dtype=np.float64)
            assert lib.flsg_session_output(s, 0, 1, ids.ctypes.data, vals.ctypes.data) == _abi.FLSG_EARG
            assert lib.flsg_session_close(s) == 0
            assert _wait_ended(lib, s) == 0
        finally:
            lib.flsg_session_destroy(s)
    finally:
        dg.close()
