"""Drop-in for the reference's ``hk_solver.py``: heat kernel
h = e^{-t} sum_k (t^k/k!) (A D^-1)^k e_s by the level-synchronous Taylor push
(hk_sor_queue with omega=1, hk_solver.py:206-295).

``hk_sor_queue`` keeps the reference's signature and 7-tuple return
(est_n, xt, rt, oper, beta_t, run_time, algo_time); omega must be 1 (the
SOR-weighted level push stays a CPU baseline); local=False runs the level
sweep over all nodes.  ``estimate_n`` and ``estimate_coefficient`` are the
reference's scalar formulas, computed by the library's host code.
"""
from __future__ import annotations

import time

import numpy as np

from flsg import _abi, engine
from flsg import dropin as _d

__all__ = ["estimate_coefficient", "estimate_n", "hk_sor_queue", "hk_push_batch"]


def estimate_coefficient(t, k):
    """t^{k+1}/(k+1)!  (hk_solver.py:15-25), computed by the library."""
    assert k >= 0
    assert t >= 1.
    return float(_lib().flsg_hk_estimate_coefficient(float(t), int(k)))


def estimate_n(t, eps):
    """Taylor order N (hk_solver.py:28-48), computed by the library."""
    assert t >= 1.
    assert eps > 0.
    return int(_lib().flsg_hk_estimate_n(float(t), float(eps)))


def _lib():
    from flsg._build import load_lib
    return load_lib()


def hk_sor_queue(n, indptr, indices, degree, s_node, tau, eps, omega, local, n_terms=0):
    """Returns (est_n, xt, rt[2, n], oper, beta_t, run_time, algo_time)."""
    if omega != 1.0:
        _d.sor_not_built("hk_sor_queue(omega != 1)")
    import torch
    dg = _d.device_graph(n, indptr, indices, degree)
    p = engine.make_params(_abi.FLSG_HK, _abi.FLSG_GD, bool(local), _abi.FLSG_FP64, eps=float(eps),
                           tau=float(tau), n_terms=int(n_terms))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = engine.solve(dg, p, int(s_node))
    torch.cuda.synchronize()
    run_time = time.perf_counter() - t0
    lvl = out.extra["lvl_sums"].reshape(-1, 2)
    # beta_t (hk_solver.py:235,286-290): ratio of mass pushed at a level to the
    # residual mass that level held
    # (updated only when the next level's queue is non-empty, as in the
    # reference; the global sweep never fills a queue, so there it stays [1.0])
    beta_t = [1.0]
    iters = out.result["iters"] if local else 0
    for j in range(iters):
        pushed, sent = lvl[j]
        if j + 1 < iters:
            if beta_t[-1] != 0.0:
                beta_t[-1] = pushed / beta_t[-1]
            beta_t.append(sent)
    return (out.extra["est_n"], out.x.cpu().numpy(), out.r.cpu().numpy(), float(out.result["oper"]), beta_t,
            run_time, run_time)


def hk_push_batch(n, indptr, indices, degree, s_nodes, tau, eps, n_terms=0):
    """One HK push per seed (FP64); returns list of (xt, oper)."""
    outs = []
    for s in s_nodes:
        est_n, x, r, oper, _, _, _ = hk_sor_queue(n, indptr, indices, degree, int(s), tau, eps, 1.0, True, n_terms)
        outs.append((x, oper))
    return outs
