"""Parity at the real shape: the 2,394,385-node power-law graph, alpha 0.1,
eps 1e-7 (the drop-in configuration of the plain FP64 path and the batched
configuration), against the oracle run over seeds on every host core.

  * global GD and global Chebyshev: x and r bit-identical to the oracle, and
    the operation count equal, on 5 seeds each;
  * local GD: operation count equal and x within relative
    L1 1e-12 (atomic adds reorder the FP64 sums) on 10 seeds; local
    Chebyshev (sequential replay): x and r bit-identical on 10 seeds;
  * the demo harness's alpha round trip (0.1 -> 0.10000000000000002) on one
    seed per method, to the same bars;
  * the heat kernel's global level sweep (t = 5): x bit-identical, 2 seeds;
  * the batched FP32 launch of the bench's 1024 mixed seeds: every result
    re-certified on the host by an FP64 residual recomputed from the graph
    alone (max |r_v| / (eps alpha d_v) < 1, up to the FP64 rounding of x in
    the residual, 64 ulp(1) absolute), and its residual L1 within the same
    rounding of the device's certified one.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from flsg import _abi, engine, graph  # noqa: E402
from oracle import oracle as O  # noqa: E402

ALPHA = 0.1
EPS = 1e-7
REL_L1 = 1e-12
CORES = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def _oracle(fn, g, seeds, alpha, local):
    """The oracle over seeds in a thread pool (its C calls release the GIL)."""
    with ThreadPoolExecutor(max_workers=CORES) as ex:
        return list(ex.map(lambda s: fn(g, int(s), alpha, EPS, local=local), seeds))


@pytest.fixture(scope="module")
def dev_large(large_graph):
    return engine.DeviceGraph(large_graph)


def _check_exact(dg, p, seeds, outs):
    for s, o in zip(seeds, outs):
        print(f"seed {int(s)} kind={p.kind} method={p.method} local={p.local} alpha={p.alpha!r}", flush=True)
        out = engine.solve(dg, p, int(s))
        assert out.result["oper"] == o["oper"], f"seed {s}"
        assert np.array_equal(out.x.cpu().numpy(), o["x"]), f"seed {s}"
        assert np.array_equal(out.r.cpu().numpy(), o["r"]), f"seed {s}"


def _check_local_gd(dg, p, seeds, outs, g):
    for s, o in zip(seeds, outs):
        print(f"seed {int(s)} local GD alpha={p.alpha!r}", flush=True)
        out = engine.solve(dg, p, int(s))
        assert out.result["oper"] == o["oper"], f"seed {s}"
        assert _rel_l1(out.x.cpu().numpy(), o["x"]) <= REL_L1, f"seed {s}"


@pytest.mark.parametrize("method", ["gd", "cheb"])
def test_global_bit_identical_large(large_graph, dev_large, method):
    g = large_graph
    m = _abi.FLSG_GD if method == "gd" else _abi.FLSG_CHEB
    fn = O.ppr_gd if method == "gd" else O.ppr_cheby
    seeds = graph.uniform_seeds(g, 5, seed=31)
    outs = _oracle(fn, g, seeds, ALPHA, False)
    _check_exact(dev_large, engine.make_params(_abi.FLSG_PPR, m, False, _abi.FLSG_FP64, alpha=ALPHA, eps=EPS),
                 seeds, outs)


def test_local_gd_large(large_graph, dev_large):
    g = large_graph
    seeds = graph.uniform_seeds(g, 10, seed=32)
    outs = _oracle(O.ppr_gd, g, seeds, ALPHA, True)
    _check_local_gd(dev_large, engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=ALPHA,
                                                  eps=EPS), seeds, outs, g)


def test_local_cheb_bit_identical_large(large_graph, dev_large):
    g = large_graph
    seeds = graph.uniform_seeds(g, 10, seed=33)
    outs = _oracle(O.ppr_cheby, g, seeds, ALPHA, True)
    _check_exact(dev_large, engine.make_params(_abi.FLSG_PPR, _abi.FLSG_CHEB, True, _abi.FLSG_FP64, alpha=ALPHA,
                                               eps=EPS), seeds, outs)


def test_alpha_roundtrip_large(large_graph, dev_large):
    g = large_graph
    a, _ = graph.alpha_roundtrip(ALPHA)
    assert a == 0.10000000000000002
    s = graph.uniform_seeds(g, 1, seed=34)
    for fn, m, local in ((O.ppr_gd, _abi.FLSG_GD, False), (O.ppr_cheby, _abi.FLSG_CHEB, False),
                         (O.ppr_cheby, _abi.FLSG_CHEB, True)):
        outs = _oracle(fn, g, s, a, local)
        _check_exact(dev_large, engine.make_params(_abi.FLSG_PPR, m, local, _abi.FLSG_FP64, alpha=a, eps=EPS),
                     s, outs)
    outs = _oracle(O.ppr_gd, g, s, a, True)
    _check_local_gd(dev_large, engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=a,
                                                  eps=EPS), s, outs, g)


def test_batch_recertified_large(large_graph, dev_large):
    g = large_graph
    seeds = graph.mixed_seeds(g, 1024, seed=0)
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=ALPHA, eps=EPS)
    print("batched FP32, 1024 mixed seeds", flush=True)
    bo = engine.batch_solve(dev_large, p, seeds)
    res = bo.results
    assert np.all(res["status"] == 0) and np.all(res["cert"] < 1.0)
    off = bo.offsets.astype(np.int64)
    cnt = res["supp"].astype(np.int64)
    used = int((off + cnt).max())
    cert, rl1 = O.residual_cert(g, 0, ALPHA, EPS, seeds, off, cnt, bo.ids[:used].cpu().numpy(),
                                bo.vals[:used].cpu().numpy(), nthreads=CORES)
    print(f"host certificate: max {cert.max():.12f}, device max {res['cert'].max():.12f}", flush=True)
    ratio = rl1 / res["r_l1"]
    cr = cert / res["cert"]
    for q in np.argsort(-ratio)[:5]:
        print(f"  query {q} seed {seeds[q]}: host r_l1 {rl1[q]:.9e} device {res['r_l1'][q]:.9e} "
              f"(x{ratio[q]:.9f}); host cert {cert[q]:.9f} device {res['cert'][q]:.9f}; supp {cnt[q]} "
              f"iters {res['iters'][q]} corr {res['n_corr'][q]}", flush=True)
    print(f"  host/device cert ratio max {cr.max():.9f}", flush=True)
    # The two FP64 residuals (the device's, tracked through the push and the
    # correction, and the host's, recomputed) differ by rounding at the scale
    # of x (||x||_1 ~ 1): a few 1e-16 absolute, which is up to ~1e-7 of the
    # smallest thresholds (eps alpha d = 1e-8 at d = 1)
    slack = 64 * 2.0 ** -52
    thr_min = EPS * ALPHA
    assert np.all(cert < 1.0 + slack / thr_min), np.flatnonzero(cert >= 1.0 + slack / thr_min)[:10]
    assert np.all(rl1 <= res["r_l1"] + slack * cnt)


def test_hk_global_bit_identical_large(large_graph, dev_large):
    """hk_sor_queue(local=False) at full size (t = 5, eps = 1e-7): x
    bit-identical to the oracle's level sweep and oper = N (m + n), 2 seeds."""
    g = large_graph
    seeds = graph.uniform_seeds(g, 2, seed=35)
    with ThreadPoolExecutor(max_workers=2) as ex:
        outs = list(ex.map(lambda s: O.hk_push(g, int(s), 5.0, EPS, local=False), seeds))
    p = engine.make_params(_abi.FLSG_HK, _abi.FLSG_GD, False, _abi.FLSG_FP64, eps=EPS, tau=5.0)
    for s, o in zip(seeds, outs):
        print(f"seed {int(s)} heat kernel global", flush=True)
        out = engine.solve(dev_large, p, int(s))
        assert out.extra["est_n"] == o["est_n"]
        assert out.result["oper"] == int(o["oper"])
        assert np.array_equal(out.x.cpu().numpy(), o["x"])
