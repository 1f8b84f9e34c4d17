"""bench.py's host side: the one-GPU guard and the per-query L1 helper that
reports the batched FP32 path's distance to the batched FP64 push."""
import os
import subprocess
import sys
import types

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), ([], {"WORLD_SIZE": "8"})])
def test_more_than_one_gpu_exits_with_message(argv, env):
    e = dict(os.environ, **env)
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py")] + argv, env=e,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""     # no JSON line for a run that did not happen


def _batch(torch, n, supports):
    """A batched result (ids / vals / offsets / per-query supp) from dense rows."""
    ids, vals, offs, res = [], [], [], []
    o = 0
    for idx, v in supports:
        offs.append(o)
        res.append({"supp": len(idx)})
        ids.extend(idx)
        vals.extend(v)
        o += len(idx)
    return types.SimpleNamespace(ids=torch.tensor(ids, dtype=torch.int32),
                                 vals=torch.tensor(vals, dtype=torch.float64),
                                 offsets=offs, results=res)


def test_l1_against_matches_dense_numpy():
    import torch
    sys.path.insert(0, REPO)
    import bench
    n = 50
    rng = np.random.default_rng(3)
    qa, qb, dense_a, dense_b = [], [], [], []
    for q in range(5):
        sa = np.sort(rng.choice(n, size=int(rng.integers(0, 12)), replace=False))
        sb = np.sort(rng.choice(n, size=int(rng.integers(1, 12)), replace=False))
        va, vb = rng.random(len(sa)), rng.random(len(sb))
        qa.append((sa.tolist(), va.tolist()))
        qb.append((sb.tolist(), vb.tolist()))
        da, db = np.zeros(n), np.zeros(n)
        da[sa], db[sb] = va, vb
        dense_a.append(da)
        dense_b.append(db)
    out = bench.l1_against(_batch(torch, n, qa), _batch(torch, n, qb), n, torch)
    d1 = np.array([np.abs(a - b).sum() for a, b in zip(dense_a, dense_b)])
    nb = np.array([np.abs(b).sum() for b in dense_b])
    assert out["queries"] == 5
    assert out["max_l1"] == pytest.approx(d1.max(), rel=1e-12)
    assert out["mean_l1"] == pytest.approx(d1.mean(), rel=1e-12)
    assert out["max_rel_l1"] == pytest.approx((d1 / nb).max(), rel=1e-12)
    # identical outputs are at distance zero
    same = bench.l1_against(_batch(torch, n, qb), _batch(torch, n, qb), n, torch)
    assert same["max_l1"] == 0.0
