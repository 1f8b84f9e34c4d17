"""The oracle against the committed golden fixture (tests/golden/
config1_ppr.npz, written by tests/golden/make_golden.py from the oracle):
the graph generator reproduces the stored graph and seeds exactly, and the
oracle reproduces every stored operation count, iteration count, per-
iteration frontier checksum and the SHA-256 of x and r as FP64 bytes --
bit for bit."""
import os

import numpy as np
import pytest

from flsg import graph
from oracle import oracle as O

import importlib.util

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
MG = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(MG)


@pytest.fixture(scope="module")
def gold():
    with np.load(os.path.join(HERE, "golden", "config1_ppr.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_generator_reproduces_the_stored_graph(gold, small_graph):
    assert np.array_equal(small_graph.indptr.astype(np.int32), gold["indptr"])
    assert np.array_equal(small_graph.indices, gold["indices"])
    assert np.array_equal(small_graph.degree, gold["degree"])
    assert np.array_equal(graph.uniform_seeds(small_graph, 20, seed=0), gold["seeds"])


@pytest.mark.parametrize("name", ["gd", "cheb"])
def test_oracle_matches_golden(gold, small_graph, name):
    fn = O.ppr_gd if name == "gd" else O.ppr_cheby
    alpha, eps = float(gold["alpha"]), float(gold["eps"])
    off = gold[f"{name}_front_off"]
    for i, s in enumerate(gold["seeds"]):
        o = fn(small_graph, int(s), alpha, eps, local=True, trace=True)
        assert o["oper"] == gold[f"{name}_oper"][i]
        assert len(o["trace"]) == gold[f"{name}_iters"][i]
        assert np.array_equal(MG.frontier_checks(o["trace"]), gold[f"{name}_front"][off[i]:off[i + 1]])
        assert np.array_equal(MG.digest(o["x"]), gold[f"{name}_x_sha256"][i])
        assert np.array_equal(MG.digest(o["r"]), gold[f"{name}_r_sha256"][i])
        if f"{name}_x{i}_ids" in gold:
            x = np.zeros(small_graph.n)
            x[gold[f"{name}_x{i}_ids"]] = gold[f"{name}_x{i}_vals"]
            assert np.array_equal(o["x"], x)


@pytest.fixture(scope="module")
def gold_kh():
    with np.load(os.path.join(HERE, "golden", "config1_katz_hk.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def _kh_runs(gk, g):
    beta, l_max, mu_min = float(gk["beta"]), float(gk["l_max"]), float(gk["mu_min"])
    ke, tau, he = float(gk["katz_eps"]), float(gk["hk_tau"]), float(gk["hk_eps"])
    return {"katz_gd": lambda s: O.katz_gd(g, s, beta, ke, local=True, trace=True),
            "katz_cheb": lambda s: O.katz_cheby(g, s, beta, l_max, mu_min, ke, local=True, trace=True),
            "hk0": lambda s: O.hk_push(g, s, tau, he, local=True, n_terms=0, trace=True),
            "hk40": lambda s: O.hk_push(g, s, tau, he, local=True, n_terms=40, trace=True)}


@pytest.mark.parametrize("name", ["katz_gd", "katz_cheb", "hk0", "hk40"])
def test_oracle_matches_golden_katz_hk(gold_kh, small_graph, name):
    fn = _kh_runs(gold_kh, small_graph)[name]
    off = gold_kh[f"{name}_front_off"]
    assert np.array_equal(graph.uniform_seeds(small_graph, 5, seed=5), gold_kh["seeds"])
    for i, s in enumerate(gold_kh["seeds"]):
        o = fn(int(s))
        assert o["oper"] == gold_kh[f"{name}_oper"][i]
        assert len(o["trace"]) == gold_kh[f"{name}_iters"][i]
        assert o.get("est_n", 0) == gold_kh[f"{name}_est_n"][i]
        assert np.array_equal(MG.frontier_checks(o["trace"]), gold_kh[f"{name}_front"][off[i]:off[i + 1]])
        assert np.array_equal(MG.digest(o["x"]), gold_kh[f"{name}_x_sha256"][i])
