"""Drop-in for the reference's ``ppr_solver.py``: personalized PageRank
Q x = b with Q = I - (1-alpha) A D^{-1}, b = alpha e_s (ppr_solver.py:18-23).

Same call signatures and returns as the reference; every solve runs in
libflsg.so's HIP kernels on the GPU (no CPU fallback).  Differences:
  * the graph is uploaded to HBM once per CSR and stays resident, so
    ``algo_time`` is the solve alone (the reference's GPU times include the
    whole-graph host->device copy, ppr_solver.py:101-107);
  * ``ppr_gd`` / ``ppr_cheby`` run the FP64 parity path (frontier sets and
    operation counts of the reference's numba code);
  * ``ppr_gd_local_gpu`` runs the certified FP32 path (the reference's GPU
    path is FP32 end-to-end); it returns FP64 arrays;
  * ``ppr_sor`` (SOR, and Gauss-Southwell / APPR at omega = 1) runs a
    graph-coloured variant: the nodes of one colour share no edge and push at
    once, colours in order.  It stops on the reference's certificate; its
    push order (colour order, not FIFO) and so its x, oper and iteration
    counts differ from the reference's within that certificate;
  * isolated nodes (degree 0) are never active (the reference's global loops
    divide by zero / never stop on them).
"""
from __future__ import annotations

import numpy as np

from flsg import _abi, engine
from flsg import dropin as _d

__all__ = ["ppr_gd", "ppr_cheby", "ppr_gd_local_gpu", "ppr_gd_standard_gpu", "solve_a_graph_gpu",
           "solve_a_graph_all", "solve_a_graph_cheby", "solve_a_graph_omega", "ppr_sor", "ppr_gd_batch"]


def _params(method, local, alpha, eps, precision=_abi.FLSG_FP64):
    return engine.make_params(_abi.FLSG_PPR, method, local, precision, alpha=float(alpha), eps=float(eps))


def ppr_gd(n, indptr, indices, degree, s_node, alpha, eps, local, opt_x=None, debug=False):
    """Local / global gradient descent (ppr_solver.py:208-328).
    Returns (xt, rt, oper, algo_time)."""
    return _d.single(n, indptr, indices, degree, s_node, _params(_abi.FLSG_GD, local, alpha, eps))


def ppr_cheby(n, indptr, indices, degree, s_node, alpha, eps, local):
    """Local / global Chebyshev (ppr_solver.py:582-661).
    Returns (xt, rt, oper, algo_time)."""
    return _d.single(n, indptr, indices, degree, s_node, _params(_abi.FLSG_CHEB, local, alpha, eps))


def ppr_gd_local_gpu(n, indptr, indices, degree, s_node, alpha, eps):
    """Local GD on the GPU (ppr_solver.py:82-146): certified FP32 path.
    Returns (xt, rt, oper, algo_time); rt is the FP64 residual b - Q xt."""
    import time
    import torch
    dg = _d.device_graph(n, indptr, indices, degree)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bo = engine.batch_solve(dg, _params(_abi.FLSG_GD, True, alpha, eps, _abi.FLSG_FP32),
                            np.array([s_node], dtype=np.int64))
    algo_time = time.perf_counter() - t0
    x = bo.dense(0, int(n))
    r = _residual(n, indptr, indices, degree, s_node, alpha, x)
    return x, r, int(bo.results[0]["oper"]), algo_time


def _residual(n, indptr, indices, degree, s, alpha, x):
    """b - Q x in FP64 on the host (only used to fill the rt return slot)."""
    import scipy.sparse as sp
    ip = np.asarray(indptr, dtype=np.int64)
    A = sp.csr_matrix((np.ones(ip[-1]), np.asarray(indices), ip), shape=(n, n))
    d = np.asarray(degree, dtype=np.float64)
    y = np.where(d > 0, x / np.where(d > 0, d, 1.0), 0.0)
    r = (1.0 - alpha) * (A @ y) - x
    r[s] += alpha
    return r


def ppr_gd_standard_gpu(n, indptr, indices, degree, s_node, alpha, eps):
    """Global GD on the GPU (ppr_solver.py:26-79).  Returns (xt, rt, oper, algo_time)."""
    return _d.single(n, indptr, indices, degree, s_node, _params(_abi.FLSG_GD, False, alpha, eps))


def ppr_gd_batch(n, indptr, indices, degree, s_nodes, alpha, eps):
    """Many seeds in one persistent launch (certified FP32).  Returns the
    engine's BatchOut (sparse per-query estimates, per-query opers)."""
    return _d.batch(n, indptr, indices, degree, s_nodes,
                    _params(_abi.FLSG_GD, True, alpha, eps, _abi.FLSG_FP32))


def solve_a_graph_gpu(n, indptr, indices, degree, alpha, eps, s_nodes):
    """ppr_solver.py:149-164: per-seed global and local GPU GD times."""
    s_nodes = np.asarray(s_nodes)
    _, glob_t = _d.per_seed(n, indptr, indices, degree, s_nodes, _params(_abi.FLSG_GD, False, alpha, eps))
    loc_t = np.zeros(len(s_nodes))
    for i, s in enumerate(s_nodes):
        loc_t[i] = ppr_gd_local_gpu(n, indptr, indices, degree, int(s), alpha, eps)[3]
    return s_nodes, loc_t, glob_t


def solve_a_graph_cheby(n, indptr, indices, degree, alpha, eps, omega, graph_name, s_nodes):
    """ppr_solver.py:664-693: (s_nodes, local_opers, local_algo_times,
    global_opers, global_algo_times)."""
    lo, lt = _d.per_seed(n, indptr, indices, degree, s_nodes, _params(_abi.FLSG_CHEB, True, alpha, eps))
    go, gt = _d.per_seed(n, indptr, indices, degree, s_nodes, _params(_abi.FLSG_CHEB, False, alpha, eps))
    return s_nodes, lo, lt, go, gt


def solve_a_graph_all(n, indptr, indices, degree, alpha, eps, omega, graph_name, s_nodes):
    """ppr_solver.py:511-579: (s_nodes, local_sor_opers, local_sor_algo_times,
    global_sor_*, local_appr_* (omega = 1), global_appr_*, local_gd_*,
    global_gd_*), every slot computed on the GPU (SOR / APPR by the coloured
    variant)."""
    lo, lt = _d.per_seed(n, indptr, indices, degree, s_nodes, _params(_abi.FLSG_GD, True, alpha, eps))
    go, gt = _d.per_seed(n, indptr, indices, degree, s_nodes, _params(_abi.FLSG_GD, False, alpha, eps))
    lso, lst = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, omega, True)
    gso, gst = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, omega, False)
    lao, lat = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, 1.0, True)
    gao, gat = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, 1.0, False)
    return (s_nodes, lso, lst, gso, gst, lao, lat, gao, gat, lo, lt, go, gt)


def ppr_sor(n, indptr, indices, degree, s_node, alpha, eps, omega, local, opt_x=None, debug=False):
    """SOR / Gauss-Southwell push (ppr_solver.py:331-461), graph-coloured on
    the GPU.  Returns (xt, rt, oper, algo_time) with oper a float as the
    reference's."""
    return _d.single_sor(n, indptr, indices, degree, s_node, alpha, eps, omega, local)


def solve_a_graph_omega(n, indptr, indices, degree, alpha, eps, omega, graph_name, s_nodes):
    """ppr_solver.py:463-508: the 13-tuple of solve_a_graph_all with only the
    local SOR and local APPR (omega = 1) slots filled; the others stay 0 as
    the reference leaves them."""
    k = len(s_nodes)
    z_o = np.zeros(k, dtype=np.int64)
    z_t = np.zeros(k)
    lso, lst = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, omega, True)
    lao, lat = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, 1.0, True)
    return (s_nodes, lso, lst, z_o, z_t.copy(), lao, lat, z_o.copy(), z_t.copy(), z_o.copy(), z_t.copy(),
            z_o.copy(), z_t.copy())
