"""Writes tests/golden/config1_ppr.npz: the oracle's local GD and local
Chebyshev results on the 10,000-node power-law graph (alpha 0.1, eps 1e-6,
20 seeds drawn uniformly from non-isolated nodes with rng seed 0), and
tests/golden/config1_katz_hk.npz: local Katz GD and Katz Chebyshev (beta =
1/(d_max+1), eps 1e-8, the eigen-bounds stored with them) and the heat-kernel
level push (t = 5, eps 1e-7, Taylor order from estimate_n and 40) on 5 seeds
(rng seed 5) of the same graph.

Stored: the graph (int32 CSR), the seeds, and per method and seed the
operation count, the iteration count, SHA-256 digests of x and r as FP64
bytes, their L1 norms, per-iteration frontier checksums (size, sum of ids,
and an order-sensitive polynomial hash), and the full sparse x for the first
two seeds.  The oracle is deterministic, so the digests pin it bit for bit;
the device path is checked against the same numbers.  Run from the
repository root: python3 tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))

ALPHA, EPS, NSEEDS, NFULL = 0.1, 1e-6, 20, 2
MOD = (1 << 61) - 1


def digest(a: np.ndarray) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).digest(),
                         dtype=np.uint8)


def frontier_checks(trace):
    """(size, id sum, order hash) per iteration; the hash is sum id_k 1000003^k mod 2^61-1."""
    out = np.zeros((len(trace), 3), dtype=np.int64)
    for t, q in enumerate(trace):
        q = np.asarray(q, dtype=np.int64)
        h = 0
        for v in q.tolist():
            h = (h * 1000003 + v + 1) % MOD
        out[t] = (q.size, int(q.sum()), h)
    return out


def main():
    sys.path.insert(0, os.path.join(REPO, "faster-local-solver-for-gdes_amd"))
    sys.path.insert(0, REPO)
    from flsg import graph
    from oracle import oracle as O
    g = graph.powerlaw_graph_small()
    seeds = graph.uniform_seeds(g, NSEEDS, seed=0)
    rec = {"indptr": g.indptr.astype(np.int32), "indices": g.indices.astype(np.int32),
           "degree": g.degree.astype(np.int32), "seeds": seeds.astype(np.int64),
           "alpha": np.float64(ALPHA), "eps": np.float64(EPS)}
    for name, fn in (("gd", O.ppr_gd), ("cheb", O.ppr_cheby)):
        opers, iters, xd, rd, xl1, rl1, checks, coff = [], [], [], [], [], [], [], [0]
        for i, s in enumerate(seeds):
            o = fn(g, int(s), ALPHA, EPS, local=True, trace=True)
            opers.append(o["oper"])
            iters.append(len(o["trace"]))
            xd.append(digest(o["x"]))
            rd.append(digest(o["r"]))
            xl1.append(float(np.abs(o["x"]).sum()))
            rl1.append(float(np.abs(o["r"]).sum()))
            c = frontier_checks(o["trace"])
            checks.append(c)
            coff.append(coff[-1] + len(c))
            if i < NFULL:
                nz = np.flatnonzero(o["x"])
                rec[f"{name}_x{i}_ids"] = nz.astype(np.int32)
                rec[f"{name}_x{i}_vals"] = o["x"][nz]
        rec[f"{name}_oper"] = np.asarray(opers, dtype=np.int64)
        rec[f"{name}_iters"] = np.asarray(iters, dtype=np.int64)
        rec[f"{name}_x_sha256"] = np.stack(xd)
        rec[f"{name}_r_sha256"] = np.stack(rd)
        rec[f"{name}_x_l1"] = np.asarray(xl1)
        rec[f"{name}_r_l1"] = np.asarray(rl1)
        rec[f"{name}_front"] = np.concatenate(checks)
        rec[f"{name}_front_off"] = np.asarray(coff, dtype=np.int64)
    path = os.path.join(HERE, "config1_ppr.npz")
    np.savez_compressed(path, **rec)
    print(path, os.path.getsize(path), "bytes")

    # Katz and the heat kernel
    import scipy.sparse.linalg as sla
    seeds5 = graph.uniform_seeds(g, 5, seed=5)
    beta = 1.0 / (int(g.degree.max()) + 1)
    A = O.adjacency(g)
    lam_max = float(sla.eigsh(A, 1, which="LA")[0][0])
    lam_min = float(sla.eigsh(A, 1, which="SA")[0][0])
    l_max, mu_min = 1 - beta * lam_min, 1 - beta * lam_max
    rk = {"seeds": seeds5.astype(np.int64), "beta": np.float64(beta), "katz_eps": np.float64(1e-8),
          "l_max": np.float64(l_max), "mu_min": np.float64(mu_min), "hk_tau": np.float64(5.0),
          "hk_eps": np.float64(1e-7)}
    runs = (("katz_gd", lambda s: O.katz_gd(g, s, beta, 1e-8, local=True, trace=True)),
            ("katz_cheb", lambda s: O.katz_cheby(g, s, beta, l_max, mu_min, 1e-8, local=True, trace=True)),
            ("hk0", lambda s: O.hk_push(g, s, 5.0, 1e-7, local=True, n_terms=0, trace=True)),
            ("hk40", lambda s: O.hk_push(g, s, 5.0, 1e-7, local=True, n_terms=40, trace=True)))
    for name, fn in runs:
        opers, iters, xd, xl1, checks, coff, est = [], [], [], [], [], [0], []
        for i, s in enumerate(seeds5):
            o = fn(int(s))
            opers.append(o["oper"])
            iters.append(len(o["trace"]))
            xd.append(digest(o["x"]))
            xl1.append(float(np.abs(o["x"]).sum()))
            est.append(o.get("est_n", 0))
            c = frontier_checks(o["trace"])
            checks.append(c)
            coff.append(coff[-1] + len(c))
            if i == 0:
                nz = np.flatnonzero(o["x"])
                rk[f"{name}_x0_ids"] = nz.astype(np.int32)
                rk[f"{name}_x0_vals"] = o["x"][nz]
        rk[f"{name}_oper"] = np.asarray(opers, dtype=np.float64)
        rk[f"{name}_iters"] = np.asarray(iters, dtype=np.int64)
        rk[f"{name}_est_n"] = np.asarray(est, dtype=np.int64)
        rk[f"{name}_x_sha256"] = np.stack(xd)
        rk[f"{name}_x_l1"] = np.asarray(xl1)
        rk[f"{name}_front"] = np.concatenate(checks)
        rk[f"{name}_front_off"] = np.asarray(coff, dtype=np.int64)
    path = os.path.join(HERE, "config1_katz_hk.npz")
    np.savez_compressed(path, **rk)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
