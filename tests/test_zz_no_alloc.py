"""The allocation-free local push (flsg_solve_local_noalloc) against flsg_solve
on the same inputs, through the C ABI.

Bar: the operation count and iteration count equal (integer work, bit-exact);
x and r within relative L1 1e-12 of flsg_solve's (FP64, atomics accumulate in
a different order from run to run).  Two calls in a row share the library's
fixed scratch array; the second runs on a smaller graph, so stale entries
left by the first call must not leak into it.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from flsg import _abi, engine, graph  # noqa: E402
from flsg.graph import CSRGraph  # noqa: E402

REL_L1 = 1e-12


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def _noalloc(dg, params, seed):
    torch = dg.torch
    # the caller's x / r: filled with garbage first, the entry point zeroes them
    x = torch.full((dg.n,), 7.0, dtype=torch.float64, device=dg.device)
    r = torch.full((dg.n,), -3.0, dtype=torch.float64, device=dg.device)
    res = _abi.Result()
    rc = dg.lib.flsg_solve_local_noalloc(dg.handle, C.byref(params), int(seed), C.c_void_p(x.data_ptr()),
                                         C.c_void_p(r.data_ptr()), C.byref(res),
                                         C.c_void_p(torch.cuda.current_stream().cuda_stream))
    engine._check(rc)
    torch.cuda.synchronize()
    return x.cpu().numpy(), r.cpu().numpy(), res


def _compare(dg, params, seed):
    ref = engine.solve(dg, params, seed)
    x, r, res = _noalloc(dg, params, seed)
    print(f"n={dg.n} kind={params.kind} seed={seed} oper={res.oper} iters={res.iters}", flush=True)
    assert res.status == 0
    assert res.oper == ref.result["oper"] and res.iters == ref.result["iters"]
    assert _rel_l1(x, ref.x.cpu().numpy()) <= REL_L1
    assert _rel_l1(r, ref.r.cpu().numpy()) <= REL_L1
    return res


def test_noalloc_matches_solve_two_calls():
    big = graph.powerlaw_graph_small(n=300_000, wmax=6000.0, seed=3)
    small = graph.powerlaw_graph_small()
    dg_big = engine.DeviceGraph(big)
    dg_small = engine.DeviceGraph(small)
    s_big = int(np.argmax(big.degree))       # a hub seed: split hub pieces in the first iterations
    s_small = int(graph.uniform_seeds(small, 1, seed=5)[0])
    ppr = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=0.1, eps=1e-7)
    katz = engine.make_params(_abi.FLSG_KATZ, _abi.FLSG_GD, True, _abi.FLSG_FP64,
                              alpha=1.0 / (small.d_max + 1), eps=1e-8)
    r1 = _compare(dg_big, ppr, s_big)
    r2 = _compare(dg_small, ppr, s_small)
    assert r1.oper > r2.oper > 0
    _compare(dg_small, katz, s_small)
    _compare(dg_big, engine.make_params(_abi.FLSG_KATZ, _abi.FLSG_GD, True, _abi.FLSG_FP64,
                                        alpha=1.0 / (big.d_max + 1), eps=1e-8), s_big)


def test_noalloc_rejects_what_it_cannot_hold():
    lib = engine.load_lib()
    assert lib.flsg_solve_local_noalloc_scratch_bytes() < 64 << 20
    # a path of 2^20 + 1 nodes: beyond the fixed scratch, rejected before any launch
    n = (1 << 20) + 1
    u = np.arange(n - 1, dtype=np.int64)
    dg = engine.DeviceGraph(CSRGraph.from_edges(n, u, u + 1, name="path"))
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=0.1, eps=1e-4)
    with pytest.raises(engine.FlsgError, match="fixed scratch"):
        _noalloc(dg, p, 0)
    small = engine.DeviceGraph(graph.powerlaw_graph_small())
    for q in (engine.make_params(_abi.FLSG_PPR, _abi.FLSG_CHEB, True, _abi.FLSG_FP64, alpha=0.1, eps=1e-6),
              engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, False, _abi.FLSG_FP64, alpha=0.1, eps=1e-6),
              engine.make_params(_abi.FLSG_HK, _abi.FLSG_GD, True, _abi.FLSG_FP64, tau=5.0, eps=1e-6),
              engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=0.1, eps=1e-6)):
        with pytest.raises(engine.FlsgError, match="allocation-free solve"):
            _noalloc(small, q, 0)
    # the reference's assert eps * degree[s] <= 1 holds here too
    s = int(np.argmax(graph.powerlaw_graph_small().degree))
    with pytest.raises(AssertionError):
        _noalloc(small, engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=0.1,
                                           eps=0.5), s)
