"""The rounding-error bound of the batched FP32 certificate.

A boundary node v (reached by the FP32 push, never pushed) keeps its FP32
residual r32 when flsg_boundary_bound -- the function the kernels evaluate --
certifies it.  These tests rebuild what the push does to such a node in exact
arithmetic: each of its d neighbours u pushes `iters` FP32 residuals rho
(x32_u is their FP32 sum, each push value is fl32(fl64(fl64(1 - alpha) * rho) /
d_u)), and r32 is the FP32 sum of the push values, iteration by iteration.
The FP64 residual of the returned x at v is sum_u (1 - alpha) x32_u / d_u,
computed here with fractions.  The bound must cover the difference for random
sequences and for sequences built to round as badly as possible (every add
lands just past half an ulp), and must not be loose by orders of magnitude
(the adversarial error reaches 13-37% of it).  A bound without the degree
factor fails the adversarial cases (error up to 15x such a bound).  The
factor 2 per add, there for FP32 atomics whose rounding mode is not
round-to-nearest, is not pinned: these sums round to nearest (numpy).
"""
from fractions import Fraction

import numpy as np
import pytest

from flsg import _build

ALPHA = 0.1
OM = 1.0 - ALPHA   # the kernels' FP64 1 - alpha


def bound(r32, d, iters, pushes=0):
    return _build.load_lib().flsg_boundary_bound(float(r32), int(d), int(iters), int(pushes))


def push_value(rho, du):
    return np.float32((OM * float(rho)) / du)


def run(d_v, iters, rng, adversarial, active=1.0, with_adds=False):
    """(r32, exact residual of x32) at a boundary node of degree d_v; each
    iteration a neighbour pushes with probability `active` (with_adds: also
    the number of push values the node received)."""
    deg = rng.integers(1, 200, size=d_v)
    x32 = [np.float32(0.0)] * d_v
    r32 = np.float32(0.0)
    adds = 0
    for it in range(iters):
        for k in rng.permutation(d_v):
            if active < 1.0 and rng.random() >= active:
                continue
            adds += 1
            du = int(deg[k])
            if adversarial and not (it == 0 and k == 0 and r32 == 0):
                # a push value just past half an ulp of the running sum: the
                # FP32 add rounds it up by almost half an ulp
                ulp = float(np.spacing(r32)) if r32 > 0 else 1e-30
                target = (0.5 + 2.0 ** -12) * ulp
                rho = np.float32(target * du / OM)
                if rho <= 0:
                    rho = np.float32(np.finfo(np.float32).tiny)
            else:
                rho = np.float32(np.exp(rng.normal(-12.0, 3.0)))
            p = push_value(rho, du)
            x32[k] = np.float32(x32[k] + rho)
            r32 = np.float32(r32 + p)
    exact = sum(Fraction(OM) * Fraction(float(x32[k])) / int(deg[k]) for k in range(d_v))
    return (r32, exact, adds) if with_adds else (r32, exact)


@pytest.mark.parametrize("d_v,iters", [(1, 1), (1, 60), (8, 25), (64, 10), (64, 60)])
def test_bound_covers_random_sums(d_v, iters):
    rng = np.random.default_rng(1000 * d_v + iters)
    for _ in range(20):
        r32, exact = run(d_v, iters, rng, adversarial=False)
        err = abs(Fraction(float(r32)) - exact)
        assert err <= Fraction(bound(r32, d_v, iters)), (d_v, iters, float(err), bound(r32, d_v, iters))


@pytest.mark.parametrize("d_v,iters", [(1, 60), (8, 25), (64, 10), (64, 60)])
def test_bound_covers_adversarial_sums_and_is_not_loose(d_v, iters):
    rng = np.random.default_rng(7 * d_v + iters)
    worst = 0.0
    for _ in range(3):
        r32, exact = run(d_v, iters, rng, adversarial=True)
        err = abs(Fraction(float(r32)) - exact)
        b = bound(r32, d_v, iters)
        assert err <= Fraction(b), (d_v, iters, float(err), b)
        worst = max(worst, float(err) / b)
    # every add rounds by almost half an ulp: the error reaches a fair share of
    # the bound (which allows 2 ulp per add for any rounding mode)
    assert worst > 0.05, worst


def test_bound_special_values():
    assert bound(0.0, 5, 10) == 0.0
    assert bound(-0.0, 5, 10) == 0.0
    # the sign of r32 marks a residual moved into x; the bound keeps its size
    assert bound(-1e-6, 5, 10) == bound(1e-6, 5, 10) > 0.0
    # linear in r32 up to the denormal term, growing with degree and iterations
    assert bound(2e-6, 5, 10) > 1.99 * bound(1e-6, 5, 10)
    assert bound(1e-6, 64, 10) > bound(1e-6, 5, 10)
    assert bound(1e-6, 5, 100) > bound(1e-6, 5, 10)


@pytest.mark.parametrize("d_v,iters,active", [(400, 20, 0.02), (2000, 12, 0.005), (300, 40, 0.05)])
@pytest.mark.parametrize("adversarial", [False, True])
def test_bound_capped_by_received_values(d_v, iters, active, adversarial):
    """A high-degree node reached by few pushes: the bound taken with the cap
    m = the push values the node actually received (the kernels use the
    query's push count, which is at least that on a graph with no repeated
    row entry) still covers the FP32 error, and is far below the degree-based
    bound."""
    rng = np.random.default_rng(d_v + 17 * iters + int(adversarial))
    worst = 0.0
    for _ in range(3):
        r32, exact, adds = run(d_v, iters, rng, adversarial, active=active, with_adds=True)
        if adds == 0:
            continue
        err = abs(Fraction(float(r32)) - exact)
        b = bound(r32, d_v, iters, pushes=adds)
        assert err <= Fraction(b), (d_v, iters, adds, float(err), b)
        # a cap above d iters changes nothing; the cap is far below it here
        assert bound(r32, d_v, iters, pushes=d_v * iters + 5) == bound(r32, d_v, iters)
        assert b < 0.2 * bound(r32, d_v, iters)
        worst = max(worst, float(err) / b)
    if adversarial:
        assert worst > 0.05, worst
