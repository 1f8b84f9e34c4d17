#!/usr/bin/env python
"""Benchmark: certified PPR queries/s of the batched FP32 local push on one
MI355X, beside the plain FP64 single-source GPU path and the CPU baseline.

Workload (one "step"): 1024 seeds (70% uniform over non-isolated nodes, 30%
degree-proportional) on the synthetic Chung-Lu power-law graph with
n = 2,394,385 nodes (~9.3M CSR entries, max degree ~1e5, ~10% isolated),
alpha = 0.1, eps = 1e-7, in one persistent launch: FP32 push with the FP64
threshold test, FP64 residual recompute and FP64 correction push, so every
query ends with |r_v| < eps*alpha*d_v in FP64.  The graph is resident in HBM
before the timed region starts.

Prints ONE JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "faster-local-solver-for-gdes_amd")
for _p in (PKG, REPO):
    if _p not in sys.path:
        sys.path.insert(0, _p)

# algorithmic bytes of the batched FP32 path, per unit (the model the
# roofline's "achieved" uses -- the contract's per-unit figures for the FP32
# push and its FP64 certificate):
#   push, per frontier entry (node phase): 8 row bounds + 4 r_u read + 4 r_u
#     write + 8 x_u read-modify-write + 4 d_u + 4 frontier id read + 4
#     next-frontier id write                                               = 36
#   push, per edge: 4 column index + 8 r_v read-modify-write + 4 d_v       = 16
#   certificate, per edge of the support's rows (FP64 residual recompute
#     of every support node): 4 column index + 4 x_v gather                 =  8
#   certificate, per support node: 8 row bounds + 4 x_u + 4 d_u + 8 r_u
#     write                                                                 = 24
#   correction push (FP64): 52 per pushed node, 24 per edge
B_NODE_FP32, B_EDGE_FP32, B_CERT_EDGE, B_CERT_ROW, B_CORR_NODE, B_CORR_EDGE = 36, 16, 8, 24, 52, 24
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model() -> str:
    """The host CPU's model name (the lscpu 'Model name' line), if readable."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def l1_against(a, b, n, torch):
    """Per-query L1 distance between two batched outputs of the same seeds
    (sparse ids / values on the device), summed on the GPU through one dense
    scratch vector; returns max / mean absolute and max relative to b."""
    import numpy as np
    dev = a.vals.device
    dense = torch.zeros(n, dtype=torch.float64, device=dev)
    nq = len(a.results)
    d1 = torch.empty(nq, dtype=torch.float64, device=dev)
    nb = torch.empty(nq, dtype=torch.float64, device=dev)
    for q in range(nq):
        oa, ka = int(a.offsets[q]), int(a.results[q]["supp"])
        ob, kb = int(b.offsets[q]), int(b.results[q]["supp"])
        ia, va = a.ids[oa:oa + ka].long(), a.vals[oa:oa + ka]
        ib, vb = b.ids[ob:ob + kb].long(), b.vals[ob:ob + kb]
        dense.index_add_(0, ib, vb)
        dense.index_add_(0, ia, -va)
        d1[q] = dense.abs().sum()   # zero outside the two supports
        nb[q] = vb.abs().sum()
        dense.index_fill_(0, ia, 0.0)
        dense.index_fill_(0, ib, 0.0)
    d1h, nbh = d1.cpu().numpy(), nb.cpu().numpy()
    return {"queries": int(nq), "max_l1": float(d1h.max()), "mean_l1": float(d1h.mean()),
            "max_rel_l1": float((d1h / np.maximum(nbh, 1e-300)).max())}


def log_phases(what, res):
    """Per-query phase times of a batched launch (means over its queries)."""
    log(f"{what}, per query: push {res['ms_push'].mean():.2f} ms, recompute {res['ms_recompute'].mean():.2f}, "
        f"correct {res['ms_correct'].mean():.2f}, total {res['ms'].mean():.2f} ms (times from the query's start); "
        f"push phases: node {res['ms_node'].mean():.2f}, pairs {res['ms_pairs'].mean():.2f}, "
        f"bins {res['ms_bins'].mean():.2f}, direct edges {res['ms_direct'].mean():.2f} ms, "
        f"{res['bin_iters'].mean():.1f} of {res['iters'].mean():.1f} iterations binned")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--eps", type=float, default=1e-7)
    ap.add_argument("--alpha", type=float, default=0.1)
    ap.add_argument("--nq", type=int, default=1024)
    ap.add_argument("--plain-queries", type=int, default=32)
    ap.add_argument("--cpu-sample", type=int, default=0, help="seeds timed on the CPU (0 = 2 per core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-l1", action="store_true", help="skip the L1 of every result against the FP64 oracle")
    ap.add_argument("--fp64-batched", type=int, default=1, help="also time the batched FP64 push (1) or not (0)")
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--block", type=int, default=0, help="batched workgroup size (256/512/1024; 0 = default)")
    ap.add_argument("--no-relabel", action="store_true", help="keep the generator's node ids on the device")
    ap.add_argument("--hub-bytes", type=int, default=None,
                    help="LDS hub cache bytes per workgroup (-1 default, 0 off)")
    ap.add_argument("--bin-min", type=int, default=None,
                    help="binned push for iterations of at least this many edges (0 = never; default: the library's)")
    ap.add_argument("--dense-min", type=int, default=None, help="binned push: pairs below which a bin takes atomics")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (workgroup size / slots / hub cache / degree bins)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (3 times): the known-size reference "
                         "for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 3}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (queries would shard across GPUs "
              "with no exchange). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from flsg import _abi, engine, graph

    t0 = time.time()
    g = graph.powerlaw_graph_large()
    log(f"graph n={g.n} nnz={g.m} dmax={g.d_max} isolated={(g.degree == 0).mean():.3f} ({time.time() - t0:.1f}s)")
    # node ids relabelled by degree on the device (outputs map back to the
    # generator's ids); the plain FP64 path is timed on the same resident graph
    dg = engine.DeviceGraph(g, relabel=not args.no_relabel)
    seeds = graph.mixed_seeds(g, args.nq, seed=0)
    if args.bin_min is not None or args.dense_min is not None:
        from flsg.engine import _check
        _check(dg.lib.flsg_batch_set_binning(dg.handle, 65536 if args.bin_min is None else args.bin_min,
                                             64 if args.dense_min is None else args.dense_min))
    p32 = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP32, alpha=args.alpha, eps=args.eps)
    p64 = engine.make_params(_abi.FLSG_PPR, _abi.FLSG_GD, True, _abi.FLSG_FP64, alpha=args.alpha, eps=args.eps)

    # ------------------------------------------------------------ autotune (untimed)
    tuned = None
    if not args.no_autotune and not (args.slots or args.block or args.hub_bytes is not None
                                     or args.bin_min is not None or args.dense_min is not None):
        from flsg import autotune
        t_at = time.perf_counter()
        tn = autotune.tune_batched(dg, p32, seeds[:min(args.nq, 512)], log=log)
        tuned = {**tn["best"], "sample_qps": tn["qps"], "trials": len(tn["trials"]),
                 "seconds": time.perf_counter() - t_at}
        log(f"autotune: {tuned}")

    # ------------------------------------------------------------ batched FP32 (the measured step)
    bo = None
    for w in range(args.warmup):
        bo = None   # a step's outputs are consumed before the next: their memory is reused
        bo = engine.batch_solve(dg, p32, seeds, slots=args.slots, block=args.block, hub_bytes=args.hub_bytes)
        log(f"warmup {w}: {bo.ms:.1f} ms")
    bo = None
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    kern = []
    for k in range(args.steps):
        bo = None
        bo = engine.batch_solve(dg, p32, seeds, slots=args.slots, block=args.block, hub_bytes=args.hub_bytes)
        kern.append(bo.ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    last = bo
    res = last.results
    for k, ms in enumerate(kern):
        log(f"step {k}: kernel {ms:.1f} ms")
    log_phases("last step", res)
    certified = bool(np.all((res["status"] == 0) & (res["cert"] < 1.0)))
    ms_step = wall * 1e3 / args.steps
    qps = args.nq * args.steps / wall

    oper = int(res["oper"].sum())
    nodes = int(res["nodes"].sum())
    supp = int(res["supp"].sum())
    n_corr = int(res["n_corr"].sum())
    corr_edges = int(res["corr_edges"].sum())
    # volume of every query's support (its ids are caller labels), summed on the device
    deg_dev = torch.from_numpy(np.ascontiguousarray(g.degree, dtype=np.int64)).to(last.ids.device)
    used = int((last.offsets.astype(np.int64) + res["supp"].astype(np.int64)).max())
    supp_vol = int(deg_dev[last.ids[:used].long()].sum().item())
    del deg_dev
    bytes_step = (B_NODE_FP32 * nodes + B_EDGE_FP32 * oper + B_CERT_EDGE * supp_vol + B_CERT_ROW * supp
                  + B_CORR_NODE * n_corr + B_CORR_EDGE * corr_edges)
    pulled = int(res["oper_cert"].sum()) - corr_edges
    log(f"algorithmic bytes per launch: {bytes_step / 1e9:.1f} GB (frontier entries {nodes}, push edges {oper}, "
        f"support nodes {supp}, support volume {supp_vol}, correction nodes {n_corr} / edges {corr_edges}); "
        f"the certificate pulled {pulled} row entries ({pulled / max(supp_vol, 1):.4f} of the support volume)")
    kern_ms = float(np.mean(kern))
    achieved = bytes_step / (kern_ms * 1e-3) / 1e9

    # measured HBM bytes per launch, from the committed rocprofv3 counter runs;
    # only for the workload those runs profiled
    traffic = None
    atomics = None
    # the launch configuration this run timed, in the form the profile records
    import ctypes as C
    ncu = C.c_int32(0)
    engine._check(dg.lib.flsg_device_info(C.byref(ncu), None, None))
    if tuned:
        eff = {k: tuned[k] for k in ("block", "slots", "hub_bytes", "bin_min", "dense_min", "split_edges",
                                     "piece_edges")}
    else:
        eff = {"block": args.block or 1024, "slots": args.slots or int(ncu.value),
               "hub_bytes": -1 if args.hub_bytes is None else args.hub_bytes,
               "bin_min": 65536 if args.bin_min is None else args.bin_min,
               "dense_min": 64 if args.dense_min is None else args.dense_min, "split_edges": 256, "piece_edges": 0}
    if eff["hub_bytes"] == -1:
        eff["hub_bytes"] = {1024: 136 * 1024, 512: 48 * 1024, 256: 24 * 1024}[eff["block"]]
    tpath = os.path.join(REPO, "profiles", "k_batch_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        # only for the workload and launch configuration the counter runs profiled
        if (tj.get("n") == g.n and tj.get("seeds") == len(seeds) and tj.get("eps") == args.eps
                and tj.get("config") == eff):
            traffic = tj.get("bytes_per_launch")
            # the limit that binds: memory-side atomics (counted by rocprofv3 on
            # this workload) per second of this run's kernel time, against the
            # random-atomic probe's rate on the same part
            na = tj.get("memory_atomics_per_launch")
            ceil = tj.get("atomic_probe_ceiling_per_s")
            if na and ceil:
                rate = na / (kern_ms * 1e-3)
                atomics = {"per_launch": na, "per_s": rate, "probe_ceiling_per_s": ceil, "frac": rate / ceil}

    # ------------------------------------------------------------ batched FP64 push (same launch, FP64 records)
    b64_qps = None
    l1_64 = None
    b64_ms = None
    if args.fp64_batched:
        # timed exactly as the FP32 step: the same warmup and step counts
        bo64 = None
        for w in range(max(args.warmup, 1)):
            bo64 = None
            bo64 = engine.batch_solve(dg, p64, seeds, slots=args.slots, block=args.block, hub_bytes=args.hub_bytes)
        bo64 = None
        torch.cuda.synchronize()
        tb = time.perf_counter()
        kern64 = []
        for k in range(args.steps):
            bo64 = None
            bo64 = engine.batch_solve(dg, p64, seeds, slots=args.slots, block=args.block, hub_bytes=args.hub_bytes)
            kern64.append(bo64.ms)
        torch.cuda.synchronize()
        b64_qps = args.nq * args.steps / (time.perf_counter() - tb)
        b64_ms = float(np.mean(kern64))
        ok64 = bool(np.all((bo64.results["status"] == 0) & (bo64.results["cert"] < 1.0)))
        log(f"batched FP64: {b64_qps:.1f} q/s (certified {ok64}, kernel {b64_ms:.1f} ms)")
        log_phases("batched FP64", bo64.results)
        l1_64 = l1_against(bo, bo64, g.n, torch)
        log(f"batched FP32 vs batched FP64, every query: {l1_64}")

    # ------------------------------------------------------------ plain FP64 single-source
    nplain = min(args.plain_queries, args.nq)
    plain_qps = None
    if nplain > 0:
        engine.solve(dg, p64, int(seeds[0]))   # warm
        torch.cuda.synchronize()
        tp = time.perf_counter()
        plain_oper = 0
        for q in range(nplain):
            out = engine.solve(dg, p64, int(seeds[q]))
            plain_oper += out.result["oper"]
        torch.cuda.synchronize()
        plain_qps = nplain / (time.perf_counter() - tp)
        log(f"plain FP64: {plain_qps:.1f} q/s over {nplain} seeds")

    # ------------------------------------------------------------ CPU baseline (oracle, OpenMP over seeds)
    cpu = None
    l1 = None
    if not args.no_cpu:
        from oracle import oracle as O
        # every core this process may use (the environment's OpenMP thread
        # count where it sets one: the host's share for this GPU)
        cores = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
        sample = args.cpu_sample or 2 * cores
        sample = min(sample, args.nq)
        tc = time.perf_counter()
        opers, times = O.ppr_batch(g, seeds[:sample], args.alpha, args.eps, kind=0, nthreads=cores)
        wall_c = time.perf_counter() - tc
        cpu = {"value": sample / wall_c, "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": f"first {sample} of the {args.nq} seeds, PPR local GD FP64 (C restatement, OpenMP over "
                         f"seeds), median {float(np.median(times)):.3f} s/query single-thread",
               "host": {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": _cpu_model()}}
        log(f"cpu: {cpu['value']:.2f} q/s on {cores} threads")
        # L1 error of every batched FP32 result against the FP64 oracle (local
        # GD of the same seed, OpenMP over queries), and the oracle's op counts
        if not args.no_l1:
            tl = time.perf_counter()
            res_o = last.offsets.astype(np.int64)
            res_k = res["supp"].astype(np.int64)
            used = int((res_o + res_k).max())
            h_ids = last.ids[:used].cpu().numpy()
            h_vals = last.vals[:used].cpu().numpy()
            l1s, norms, o_opers = O.ppr_batch_l1(g, seeds, args.alpha, args.eps, res_o, res_k, h_ids, h_vals,
                                                 nthreads=cores)
            rel = l1s / np.maximum(norms, 1e-300)
            l1 = {"queries": int(len(seeds)), "max_l1": float(l1s.max()), "mean_l1": float(l1s.mean()),
                  "max_rel_l1": float(rel.max()), "mean_rel_l1": float(rel.mean()),
                  "oper_equal_frac": float(np.mean(o_opers == res["oper"])),
                  "threads": cores, "seconds": time.perf_counter() - tl}
        log(f"batched FP32 vs FP64 oracle: {l1}")

    line = {
        "metric": "certified PPR queries/s, batched FP32 local GD (FP64 threshold + FP64 certificate), eps="
                  + f"{args.eps:g}".replace("e-0", "e-"),
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 push, f64 threshold/certificate",
        "data": "synthetic",
        "config": {
            "workload": "1024-seed batched PPR local GD on the 2,394,385-node Chung-Lu power-law graph",
            "n": g.n, "nnz": g.m, "d_max": g.d_max, "seeds": args.nq, "alpha": args.alpha, "eps": args.eps,
            "seed_mix": "70% uniform non-isolated, 30% degree-proportional",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "memory_atomics": atomics,
            "kernel": "k_batch", "kernel_ms": kern_ms, "algorithmic_bytes": bytes_step, "launch_config": eff,
        },
        "cpu_baseline": cpu,
        "relabelled": not args.no_relabel,
        "tuned": tuned,
        "certified_all": certified,
        "max_cert": float(res["cert"].max()),
        "plain_fp64_qps": plain_qps,
        "batched_fp64_qps": b64_qps,
        "batched_fp64_kernel_ms": b64_ms,
        "speedup_vs_plain_fp64": (qps / plain_qps) if plain_qps else None,
        "l1_vs_fp64_oracle": l1,
        # every query against the batched FP64 push of the same seeds (same
        # frontier semantics and op counts as the FP64 oracle, tested)
        "l1_vs_fp64_batched_all": l1_64,
        # every query: L1 distance to the exact PPR vector <= ||r||_1 / alpha,
        # r the certified residual (device-summed |r_v| plus each FP32
        # record's error bound); for local GD (r >= 0) the FP64 path's own
        # distance equals its ||r||_1 / alpha, so this is the accuracy of the
        # local push at this eps, not a difference between the two paths
        "l1_to_exact_bound_all": {"queries": int(len(res)), "max": float(res["r_l1"].max()) / args.alpha,
                                  "mean": float(res["r_l1"].mean()) / args.alpha},
        "mean_oper": oper / args.nq,
        "mean_corrections": float(res["n_corr"].sum()) / args.nq,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
