"""Drop-in for the reference's ``katz_solver.py``: Katz (I - alpha A) x = e_s,
``alpha`` being Katz's beta (katz_solver.py:167-248, 340-431).

Same signatures and returns; solves run in libflsg.so on the GPU.
``katz_gd`` returns x - e_s (katz_solver.py:243), ``katz_cheby`` returns x
with the e_s component (katz_solver.py:431), as the reference does.  The
reference's 4-hour wall-clock bail-out is replaced by an iteration cap.
``katz_sor`` runs the graph-coloured SOR push (colour order, not the
reference's FIFO order; same certificate |r_u| < eps*d_u).
"""
from __future__ import annotations

import numpy as np

from flsg import _abi, engine
from flsg import dropin as _d

__all__ = ["katz_gd", "katz_cheby", "katz_gd_local_gpu", "katz_gd_standard_gpu", "solve_a_graph_gpu",
           "solve_a_graph_all", "solve_a_graph_cheby", "solve_a_graph_omega", "katz_sor", "katz_gd_batch",
           "eigen_bounds"]


def _params(method, local, alpha, eps, l_max=0.0, mu_min=0.0, precision=_abi.FLSG_FP64):
    return engine.make_params(_abi.FLSG_KATZ, method, local, precision, alpha=float(alpha), eps=float(eps),
                              l_max=float(l_max), mu_min=float(mu_min))


def katz_gd(n, indptr, indices, degree, s_node, alpha, eps, local):
    """(xt, rt, oper, algo_time); xt = x - e_s."""
    return _d.single(n, indptr, indices, degree, s_node, _params(_abi.FLSG_GD, local, alpha, eps))


def katz_cheby(n, indptr, indices, degree, s_node, alpha, l_max, mu_min, eps, local):
    """(xt, rt, oper, algo_time) with the eigen-bounds l_max = 1 - alpha*lambda_min,
    mu_min = 1 - alpha*lambda_max of I - alpha A."""
    return _d.single(n, indptr, indices, degree, s_node,
                     _params(_abi.FLSG_CHEB, local, alpha, eps, l_max, mu_min))


def katz_gd_local_gpu(n, indptr, indices, degree, s_node, alpha, eps):
    """katz_solver.py:81-146 (GPU local GD): certified FP32 path; xt = x - e_s."""
    import time
    import torch
    dg = _d.device_graph(n, indptr, indices, degree)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bo = engine.batch_solve(dg, _params(_abi.FLSG_GD, True, alpha, eps, precision=_abi.FLSG_FP32),
                            np.array([s_node], dtype=np.int64))
    algo_time = time.perf_counter() - t0
    x = bo.dense(0, int(n))
    ip = np.asarray(indptr, dtype=np.int64)
    import scipy.sparse as sp
    A = sp.csr_matrix((np.ones(ip[-1]), np.asarray(indices), ip), shape=(n, n))
    r = -(x - alpha * (A @ x))
    r[s_node] += 1.0
    x[s_node] -= 1.0
    return x, r, int(bo.results[0]["oper"]), algo_time


def katz_gd_standard_gpu(n, indptr, indices, degree, s_node, alpha, eps):
    return _d.single(n, indptr, indices, degree, s_node, _params(_abi.FLSG_GD, False, alpha, eps))


def katz_gd_batch(n, indptr, indices, degree, s_nodes, alpha, eps):
    """Many seeds in one persistent launch (certified FP32).  Values include
    the e_s component (subtract 1 at the seed for katz_gd's convention)."""
    return _d.batch(n, indptr, indices, degree, s_nodes,
                    _params(_abi.FLSG_GD, True, alpha, eps, precision=_abi.FLSG_FP32))


def solve_a_graph_gpu(n, indptr, indices, degree, alpha, eps, s_nodes):
    s_nodes = np.asarray(s_nodes)
    _, glob_t = _d.per_seed(n, indptr, indices, degree, s_nodes, _params(_abi.FLSG_GD, False, alpha, eps))
    loc_t = np.array([katz_gd_local_gpu(n, indptr, indices, degree, int(s), alpha, eps)[3] for s in s_nodes])
    return s_nodes, loc_t, glob_t


def solve_a_graph_all(n, indptr, indices, degree, alpha, eps, omega, graph_name, s_nodes):
    """katz_solver.py:605-666: (s_nodes, local_sor_*, global_sor_*,
    local_akatz_* (omega = 1), global_akatz_*, local_gd_*, global_gd_*), every
    slot on the GPU (SOR / AKatz by the coloured variant)."""
    K = _abi.FLSG_KATZ
    lso, lst = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, omega, True, K)
    gso, gst = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, omega, False, K)
    lao, lat = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, 1.0, True, K)
    gao, gat = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, 1.0, False, K)
    lo, lt = _d.per_seed(n, indptr, indices, degree, s_nodes, _params(_abi.FLSG_GD, True, alpha, eps))
    go, gt = _d.per_seed(n, indptr, indices, degree, s_nodes, _params(_abi.FLSG_GD, False, alpha, eps))
    return (s_nodes, lso, lst, gso, gst, lao, lat, gao, gat, lo, lt, go, gt)


def solve_a_graph_cheby(n, indptr, indices, degree, alpha, l_max, mu_min, eps, graph_name, s_nodes):
    """katz_solver.py:670-693 (argument order as there: alpha, l_max, mu_min, eps)."""
    lo, lt = _d.per_seed(n, indptr, indices, degree, s_nodes,
                         _params(_abi.FLSG_CHEB, True, alpha, eps, l_max, mu_min))
    go, gt = _d.per_seed(n, indptr, indices, degree, s_nodes,
                         _params(_abi.FLSG_CHEB, False, alpha, eps, l_max, mu_min))
    return s_nodes, lo, lt, go, gt


def eigen_bounds(n, indptr, indices, alpha):
    """(l_max, mu_min) of I - alpha A from the extreme eigenvalues of A
    (katz_exp_dmax.py:20 'LM', katz_exp_dmin.py:20 'SA'), by host Lanczos."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    ip = np.asarray(indptr, dtype=np.int64)
    A = sp.csr_matrix((np.ones(ip[-1]), np.asarray(indices), ip), shape=(n, n))
    lam_max = float(sla.eigsh(A, 1, which="LA")[0][0])
    lam_min = float(sla.eigsh(A, 1, which="SA")[0][0])
    return 1.0 - alpha * lam_min, 1.0 - alpha * lam_max


def solve_a_graph_omega(n, indptr, indices, degree, alpha, eps, omega, graph_name, s_nodes):
    """katz_solver.py:561-603: the 13-tuple with only the local SOR and local
    AKatz (omega = 1) slots filled; the others stay 0 as the reference
    leaves them."""
    K = _abi.FLSG_KATZ
    k = len(s_nodes)
    z_o = np.zeros(k, dtype=np.int64)
    z_t = np.zeros(k)
    lso, lst = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, omega, True, K)
    lao, lat = _d.per_seed_sor(n, indptr, indices, degree, s_nodes, alpha, eps, 1.0, True, K)
    return (s_nodes, lso, lst, z_o, z_t.copy(), lao, lat, z_o.copy(), z_t.copy(), z_o.copy(), z_t.copy(),
            z_o.copy(), z_t.copy())


def katz_sor(n, indptr, indices, degree, s_node, alpha, eps, omega, local):
    """SOR / Gauss-Southwell push for Katz (katz_solver.py:252-337),
    graph-coloured on the GPU.  Returns (xt, rt, oper (float), algo_time),
    xt without the e_s term as the reference's."""
    return _d.single_sor(n, indptr, indices, degree, s_node, alpha, eps, omega, local, _abi.FLSG_KATZ)
