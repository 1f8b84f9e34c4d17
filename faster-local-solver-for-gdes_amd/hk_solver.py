"""Drop-in for the reference's ``hk_solver.py``: heat kernel
h = e^{-t} sum_k (t^k/k!) (A D^-1)^k e_s by the level-synchronous Taylor push
(hk_sor_queue with omega=1, hk_solver.py:206-295).

``hk_sor_queue`` keeps the reference's signature and 7-tuple return
(est_n, xt, rt, oper, beta_t, run_time, algo_time); omega must be 1 (the
SOR-weighted level push stays a CPU baseline).  ``estimate_n`` and
``estimate_coefficient`` are the reference's scalar formulas (host code).
"""
from __future__ import annotations

import math
import time

import numpy as np

from flsg import _abi, engine
from flsg import dropin as _d

__all__ = ["estimate_coefficient", "estimate_n", "hk_sor_queue", "hk_push_batch"]


def estimate_coefficient(t, k):
    """t^{k+1}/(k+1)!  (hk_solver.py:15-25)."""
    assert k >= 0
    assert t >= 1.
    t0 = t
    if k == 0:
        return t0
    for i in range(int(k)):
        t0 = t0 * (t / (i + 2))
    return t0


def estimate_n(t, eps):
    """Taylor order N (hk_solver.py:28-48; the loop's factor uses the initial k)."""
    assert t >= 1.
    assert eps > 0.
    k = int(math.ceil(t - 1))
    cur_est = estimate_coefficient(t, k) * ((k + 2.) / (k + 2. - t))
    if cur_est < eps * math.exp(t):
        return k
    est_k = k
    for _ in range(k, 1000 * (k + 1)):
        tmp = ((k + 2. - t) * t / ((k + 2.) ** 2.)) * ((k + 3.) / (k + 3. - t))
        cur_est = cur_est * tmp
        est_k += 1
        if cur_est < eps * math.exp(t):
            return est_k
    return est_k


def hk_sor_queue(n, indptr, indices, degree, s_node, tau, eps, omega, local, n_terms=0):
    """Returns (est_n, xt, rt[2, n], oper, beta_t, run_time, algo_time)."""
    if omega != 1.0:
        _d.sor_not_built("hk_sor_queue(omega != 1)")
    if not local:
        raise NotImplementedError("hk_sor_queue(local=False): the global level sweep is not built on the GPU")
    import torch
    dg = _d.device_graph(n, indptr, indices, degree)
    p = engine.make_params(_abi.FLSG_HK, _abi.FLSG_GD, True, _abi.FLSG_FP64, eps=float(eps), tau=float(tau),
                           n_terms=int(n_terms))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = engine.solve(dg, p, int(s_node))
    torch.cuda.synchronize()
    run_time = time.perf_counter() - t0
    lvl = out.extra["lvl_sums"].reshape(-1, 2)
    # beta_t (hk_solver.py:235,286-290): ratio of mass pushed at a level to the
    # residual mass that level held
    # (updated only when the next level's queue is non-empty, as in the reference)
    beta_t = [1.0]
    iters = out.result["iters"]
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
