"""The drop-in modules keep the reference's call signatures: the same
positional parameters in the same order (extra trailing keyword parameters
with defaults are allowed), so a caller of the reference's functions can
switch modules.  Expected lists are the reference's def lines (file:line)."""
import inspect

import pytest

EXPECTED = {
    # ppr_solver.py
    ("ppr_solver", "ppr_gd"): ("ppr_solver.py:209",
                               "n indptr indices degree s_node alpha eps local opt_x debug"),
    ("ppr_solver", "ppr_cheby"): ("ppr_solver.py:583", "n indptr indices degree s_node alpha eps local"),
    ("ppr_solver", "ppr_gd_local_gpu"): ("ppr_solver.py:82", "n indptr indices degree s_node alpha eps"),
    ("ppr_solver", "ppr_gd_standard_gpu"): ("ppr_solver.py:26", "n indptr indices degree s_node alpha eps"),
    ("ppr_solver", "solve_a_graph_gpu"): ("ppr_solver.py:149", "n indptr indices degree alpha eps s_nodes"),
    ("ppr_solver", "solve_a_graph_all"): ("ppr_solver.py:512",
                                          "n indptr indices degree alpha eps omega graph_name s_nodes"),
    ("ppr_solver", "solve_a_graph_cheby"): ("ppr_solver.py:665",
                                            "n indptr indices degree alpha eps omega graph_name s_nodes"),
    ("ppr_solver", "ppr_sor"): ("ppr_solver.py:332",
                                "n indptr indices degree s_node alpha eps omega local opt_x debug"),
    ("ppr_solver", "solve_a_graph_omega"): ("ppr_solver.py:465",
                                            "n indptr indices degree alpha eps omega graph_name s_nodes"),
    # katz_solver.py
    ("katz_solver", "katz_gd"): ("katz_solver.py:168", "n indptr indices degree s_node alpha eps local"),
    ("katz_solver", "katz_sor"): ("katz_solver.py:252", "n indptr indices degree s_node alpha eps omega local"),
    ("katz_solver", "solve_a_graph_all"): ("katz_solver.py:606",
                                           "n indptr indices degree alpha eps omega graph_name s_nodes"),
    ("katz_solver", "solve_a_graph_omega"): ("katz_solver.py:562",
                                             "n indptr indices degree alpha eps omega graph_name s_nodes"),
    ("katz_solver", "katz_cheby"): ("katz_solver.py:341",
                                    "n indptr indices degree s_node alpha l_max mu_min eps local"),
    ("katz_solver", "katz_gd_local_gpu"): ("katz_solver.py:81", "n indptr indices degree s_node alpha eps"),
    ("katz_solver", "katz_gd_standard_gpu"): ("katz_solver.py:24", "n indptr indices degree s_node alpha eps"),
    ("katz_solver", "solve_a_graph_cheby"): ("katz_solver.py:671",
                                             "n indptr indices degree alpha l_max mu_min eps graph_name s_nodes"),
    # hk_solver.py
    ("hk_solver", "estimate_n"): ("hk_solver.py:29", "t eps"),
    ("hk_solver", "hk_sor_queue"): ("hk_solver.py:207", "n indptr indices degree s_node tau eps omega local"),
    # InstantGNN.py
    ("InstantGNN", "instantGNNPush"): ("InstantGNN.py:13",
                                       "n indptr indices degree alpha eps_vec p r queue front rear q_mark beta"),
    ("InstantGNN", "instantGNNParallel"): ("InstantGNN.py:47",
                                           "n indptr indices degree alpha eps P R beta batchsize"),
    ("InstantGNN", "updetaGraph"): ("InstantGNN.py:138", "indptr indices degree us vs onlyinsert"),
    ("InstantGNN", "instantGNNUpdateBatch"): ("InstantGNN.py:299",
                                              "n indptr indices degree old_degree p r x alpha queue front rear "
                                              "q_mark eps_vec insert_neighbor delete_neighbor algo omega beta"),
    ("InstantGNN", "instantGNNUpdateBatchParallel"): ("InstantGNN.py:333",
                                                      "n indptr indices degree alpha eps omega P R X us vs beta "
                                                      "algo batchsize"),
}


@pytest.mark.parametrize("key", sorted(EXPECTED), ids=lambda k: f"{k[0]}.{k[1]}")
def test_dropin_signature(key):
    mod_name, fn_name = key
    where, params = EXPECTED[key]
    mod = __import__(mod_name)
    fn = getattr(mod, fn_name)
    got = list(inspect.signature(fn).parameters.values())
    want = params.split()
    names = [p.name for p in got]
    assert names[:len(want)] == want, f"{mod_name}.{fn_name} vs {where}: {names}"
    for extra in got[len(want):]:
        assert extra.default is not inspect.Parameter.empty, f"{fn_name}: extra parameter {extra.name} needs a default"
