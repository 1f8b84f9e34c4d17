"""The device path against the committed golden fixture (tests/golden/
config1_ppr.npz, the oracle's results on the 10,000-node graph, alpha 0.1,
eps 1e-6, 20 seeds), read from the file alone (no oracle call):
  * local Chebyshev: operation count, iteration count, every queue in order
    (order-sensitive hash) and the SHA-256 of x and r as FP64 bytes equal --
    bit-identical;
  * local GD: operation count, iteration count and every frontier as a set
    (size and id sum) equal; the L1 norm of x within relative 1e-12, and the
    full x of the stored seeds within relative L1 1e-12;
  * Katz GD / Katz Chebyshev / the heat-kernel level push on 5 seeds of the
    same graph (config1_katz_hk.npz) to the same bars."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from flsg import _abi, engine  # noqa: E402
from flsg.graph import CSRGraph  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
MG = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(MG)


@pytest.fixture(scope="module")
def gold():
    with np.load(os.path.join(HERE, "golden", "config1_ppr.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@pytest.fixture(scope="module")
def dev(gold):
    g = CSRGraph(n=gold["degree"].size, indptr=gold["indptr"].astype(np.int64), indices=gold["indices"],
                 degree=gold["degree"], name="golden-config1")
    d = engine.DeviceGraph(g)
    yield d
    d.close()


def test_local_cheb_matches_golden_bitwise(gold, dev):
    alpha, eps = float(gold["alpha"]), float(gold["eps"])
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_CHEB, True, _abi.FLSG_FP64, alpha=alpha, eps=eps)
    off = gold["cheb_front_off"]
    for i, s in enumerate(gold["seeds"]):
        out = engine.solve(dev, p, int(s), trace=True)
        assert out.result["oper"] == gold["cheb_oper"][i], f"seed {s}"
        assert len(out.trace) == gold["cheb_iters"][i]
        assert np.array_equal(MG.frontier_checks(out.trace), gold["cheb_front"][off[i]:off[i + 1]])
        assert np.array_equal(MG.digest(out.x.cpu().numpy()), gold["cheb_x_sha256"][i])
        assert np.array_equal(MG.digest(out.r.cpu().numpy()), gold["cheb_r_sha256"][i])


def test_local_gd_matches_golden(gold, dev):
    alpha, eps = float(gold["alpha"]), float(gold["eps"])
    p = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=alpha, eps=eps)
    off = gold["gd_front_off"]
    n = gold["degree"].size
    for i, s in enumerate(gold["seeds"]):
        out = engine.solve(dev, p, int(s), trace=True)
        assert out.result["oper"] == gold["gd_oper"][i], f"seed {s}"
        assert len(out.trace) == gold["gd_iters"][i]
        checks = MG.frontier_checks([np.sort(t) for t in out.trace])
        assert np.array_equal(checks[:, :2], gold["gd_front"][off[i]:off[i + 1], :2])
        x = out.x.cpu().numpy()
        assert abs(np.abs(x).sum() - gold["gd_x_l1"][i]) <= 1e-12 * gold["gd_x_l1"][i]
        if f"gd_x{i}_ids" in gold:
            xg = np.zeros(n)
            xg[gold[f"gd_x{i}_ids"]] = gold[f"gd_x{i}_vals"]
            assert np.abs(x - xg).sum() <= 1e-12 * np.abs(xg).sum()


@pytest.fixture(scope="module")
def gold_kh():
    with np.load(os.path.join(HERE, "golden", "config1_katz_hk.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("name", ["katz_gd", "katz_cheb", "hk0", "hk40"])
def test_katz_hk_match_golden(gold_kh, dev, name):
    """Katz Chebyshev bit-identical (digest, queues in order); Katz GD and the
    heat kernel: counts and frontier sets equal, x within relative 1e-12."""
    g = gold_kh
    if name == "katz_gd":
        p = engine.make_params(_abi.FLSG_KATZ, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=float(g["beta"]),
                               eps=float(g["katz_eps"]))
    elif name == "katz_cheb":
        p = engine.make_params(_abi.FLSG_KATZ, _abi.FLSG_CHEB, True, _abi.FLSG_FP64, alpha=float(g["beta"]),
                               eps=float(g["katz_eps"]), l_max=float(g["l_max"]), mu_min=float(g["mu_min"]))
    else:
        p = engine.make_params(_abi.FLSG_HK, _abi.FLSG_GD, True, _abi.FLSG_FP64, eps=float(g["hk_eps"]),
                               tau=float(g["hk_tau"]), n_terms=40 if name == "hk40" else 0)
    off = g[f"{name}_front_off"]
    n = dev.n
    for i, s in enumerate(g["seeds"]):
        out = engine.solve(dev, p, int(s), trace=True)
        assert out.result["oper"] == int(g[f"{name}_oper"][i]), f"{name} seed {s}"
        assert len(out.trace) == g[f"{name}_iters"][i]
        if name.startswith("hk"):
            assert out.extra["est_n"] == g[f"{name}_est_n"][i]
        x = out.x.cpu().numpy()
        if name == "katz_cheb":
            assert np.array_equal(MG.frontier_checks(out.trace), g[f"{name}_front"][off[i]:off[i + 1]])
            assert np.array_equal(MG.digest(x), g[f"{name}_x_sha256"][i])
        else:
            checks = MG.frontier_checks([np.sort(t) for t in out.trace])
            assert np.array_equal(checks[:, :2], g[f"{name}_front"][off[i]:off[i + 1], :2])
            assert abs(np.abs(x).sum() - g[f"{name}_x_l1"][i]) <= 1e-12 * g[f"{name}_x_l1"][i]
        if i == 0:
            xg = np.zeros(n)
            xg[g[f"{name}_x0_ids"]] = g[f"{name}_x0_vals"]
            assert np.abs(x - xg).sum() <= 1e-12 * np.abs(xg).sum()
