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
