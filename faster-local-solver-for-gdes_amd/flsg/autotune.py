"""Autotuner of the batched kernel: workgroup size and slot count, the
dynamic LDS size (hub cache / binned push slice), the binned push's
thresholds and the degree bins of the edge phases, picked by measuring the
throughput of a sample of the caller's own queries on the resident graph.

Coordinate descent: starting from the defaults, each knob in turn is swept
over its candidates with the others fixed, and the fastest value is kept.
Every trial runs the same seeds in one batched launch and is timed with HIP
events around the launch; the tuned setting is left applied on the handle.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _abi
from .engine import DeviceGraph, FlsgError, _check, batch_solve


def _apply(dg: DeviceGraph, cfg: dict) -> None:
    lib = dg.lib
    _check(lib.flsg_batch_set_block(dg.handle, int(cfg["block"])))
    _check(lib.flsg_batch_set_slots(dg.handle, int(cfg["slots"])))
    _check(lib.flsg_batch_set_hub_cache(dg.handle, int(cfg["hub_bytes"])))
    _check(lib.flsg_batch_set_degree_bins(dg.handle, int(cfg["split_edges"]), int(cfg["piece_edges"])))
    _check(lib.flsg_batch_set_binning(dg.handle, int(cfg["bin_min"]), int(cfg["dense_min"])))


def tune_batched(dg: DeviceGraph, params: _abi.Params, seeds, rounds: int = 2, log=None,
                 min_gain: float = 0.03) -> dict:
    """Return {"best": cfg, "qps": q/s of the best, "trials": [...]} for the
    batched solve of `seeds` (a sample of the workload: ~2 queries per slot is
    enough) with `params`; the best setting is applied to `dg`.  A knob moves
    off its current value only for a gain above min_gain (run-to-run noise
    is 2-3%)."""
    nc = C.c_int32(0)
    free = C.c_int64(0)
    tot = C.c_int64(0)
    _check(dg.lib.flsg_device_info(C.byref(nc), C.byref(free), C.byref(tot)))
    cus = max(1, int(nc.value))
    slot_b = int(dg.lib.flsg_batch_slot_bytes(dg.n)) + 16 * int(dg.nnz_cap)   # + the binned push's pairs
    shapes = [(1024, cus)]
    if 2 * cus * slot_b < free.value // 2:
        shapes.append((512, 2 * cus))
    space = [
        ("shape", shapes),
        ("hub_bytes", [-1, 64 * 1024, 128 * 1024]),
        ("bin_min", [0, 16384, 262144]),
        ("dense_min", [16, 256]),
        ("split_edges", [128, 512]),
        ("piece_edges", [0, 4096]),
    ]
    cfg = {"block": 1024, "slots": cus, "hub_bytes": -1, "bin_min": 65536, "dense_min": 64, "split_edges": 256,
           "piece_edges": 0}
    trials = []
    seeds = np.asarray(seeds)

    def measure(c):
        try:
            _apply(dg, c)
        except FlsgError:   # a setting this device cannot take (e.g. a hub cache beyond its LDS)
            _apply(dg, cfg)
            return 0.0
        best = 0.0
        for _ in range(rounds):
            bo = batch_solve(dg, params, seeds)
            ok = all(r["status"] == 0 for r in bo.results)
            if not ok:
                return 0.0
            best = max(best, len(seeds) / (bo.ms * 1e-3))
        return best

    cur = measure(cfg)
    trials.append({**cfg, "qps": cur})
    for name, values in space:
        best_v, best_q = None, cur
        for v in values:
            c = dict(cfg)
            if name == "shape":
                c["block"], c["slots"] = v
                if (c["block"], c["slots"]) == (cfg["block"], cfg["slots"]):
                    continue
            else:
                if c[name] == v:
                    continue
                c[name] = v
            t0 = time.perf_counter()
            q = measure(c)
            trials.append({**c, "qps": q})
            if log:
                log(f"autotune {name}={v}: {q:.1f} q/s ({time.perf_counter() - t0:.1f} s)")
            if q > best_q * (1.0 + min_gain):
                best_v, best_q = v, q
        if best_v is not None:
            if name == "shape":
                cfg["block"], cfg["slots"] = best_v
            else:
                cfg[name] = best_v
            cur = best_q
    _apply(dg, cfg)
    return {"best": cfg, "qps": cur, "trials": trials}
