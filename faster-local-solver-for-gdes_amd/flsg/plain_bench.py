"""The plain FP64 single-source path on the 2,394,385-node power-law graph:
local GD, local Chebyshev, global GD and global Chebyshev (PPR, alpha 0.1,
eps 1e-7) over 50 uniformly random non-isolated seeds, one query at a time on
the whole GPU -- the per-method operation counts and times to set beside the
reference's published CPU / CuPy numbers (README.md:39-48, wiki-talk graph of
the same size).  The coloured SOR push (local / global, omega = 1 and the
demo's optimal omega) runs the same way; its colouring is timed once."""
from __future__ import annotations

import json
import time

import numpy as np

from . import _abi, engine, graph


METHODS = {"local_gd": ("local GD", _abi.FLSG_GD, True),
           "local_cheb": ("local Chebyshev", _abi.FLSG_CHEB, True),
           "global_gd": ("global GD", _abi.FLSG_GD, False),
           "global_cheb": ("global Chebyshev", _abi.FLSG_CHEB, False)}
# coloured SOR: (label, omega (None = the demo's optimal omega), local)
SOR_METHODS = {"local_sor": ("local SOR (coloured)", None, True),
               "local_appr": ("local APPR, omega 1 (coloured)", 1.0, True),
               "global_sor": ("global SOR (coloured)", None, False)}


def plain_bench(seeds: int = 50, eps: float = 1e-7, alpha: float = 0.1, log=print,
                methods=("local_gd", "local_cheb", "global_gd", "global_cheb")) -> list:
    g = graph.powerlaw_graph_large()
    dg = engine.DeviceGraph(g)
    try:
        ss = graph.uniform_seeds(g, seeds, seed=0)
        out = []
        sor = [m for m in methods if m in SOR_METHODS]
        if sor:
            t0 = time.perf_counter()
            _, k, rounds = engine.colouring(dg)
            rec = {"method": "colouring", "colours": k, "rounds": rounds, "ms": 1e3 * (time.perf_counter() - t0)}
            out.append(rec)
            log(rec)
        for m in sor:
            name, omega, local = SOR_METHODS[m]
            if omega is None:
                omega = graph.alpha_roundtrip(alpha)[1]
            engine.sor_solve(dg, alpha, eps, omega, int(ss[0]), local=local)   # warm-up
            opers, ms, iters = [], [], []
            t0 = time.perf_counter()
            for s in ss:
                o = engine.sor_solve(dg, alpha, eps, omega, int(s), local=local)
                opers.append(o.result["oper"])
                ms.append(o.result["ms"])
                iters.append(o.result["iters"])
            wall = time.perf_counter() - t0
            rec = {"method": name, "omega": omega, "seeds": int(len(ss)), "mean_oper": float(np.mean(opers)),
                   "mean_iters": float(np.mean(iters)), "mean_ms": float(np.mean(ms)),
                   "median_ms": float(np.median(ms)), "qps": len(ss) / wall}
            out.append(rec)
            log(rec)
        for name, method, local in (METHODS[m] for m in methods if m in METHODS):
            p = engine.make_params(_abi.FLSG_PPR, method, local, _abi.FLSG_FP64, alpha=alpha, eps=eps)
            engine.solve(dg, p, int(ss[0]))      # warm-up
            opers, ms, iters = [], [], []
            t0 = time.perf_counter()
            for s in ss:
                o = engine.solve(dg, p, int(s))
                opers.append(o.result["oper"])
                ms.append(o.result["ms"])
                iters.append(o.result["iters"])
            wall = time.perf_counter() - t0
            rec = {"method": name, "seeds": int(len(ss)), "mean_oper": float(np.mean(opers)),
                   "mean_iters": float(np.mean(iters)), "mean_ms": float(np.mean(ms)),
                   "median_ms": float(np.median(ms)), "qps": len(ss) / wall,
                   "sweeps_equiv": float(np.mean(opers)) / (g.n + g.m)}
            out.append(rec)
            log(rec)
        return out
    finally:
        dg.close()


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--seeds", type=int, default=50)
    ap.add_argument("--methods", default="local_gd,local_cheb,global_gd,global_cheb",
                    help="comma list of " + ", ".join(list(METHODS) + list(SOR_METHODS)))
    args = ap.parse_args()
    plain_bench(seeds=args.seeds, methods=tuple(args.methods.split(",")),
                log=lambda d: print(json.dumps(d), flush=True))
