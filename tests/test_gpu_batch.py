"""Batched ensembles (through the C ABI): every system of a batch against the FP64
oracle and against a single-system context, for every work decomposition.

Tolerances as for the single-system products: FP64 <= 1e-12 relative (max-norm),
FP32-compensated <= 1e-6 absolute; solves <= 1e-10 relative FP64 residual."""
import numpy as np
import pytest

import oracle
from cufcm_fil_amd import RpyMobilitySolver
from cufcm_fil_amd import geometry as geo
from cufcm_fil_amd.batch import BatchedMobility

pytestmark = pytest.mark.gpu


def _ensemble(n_sys, lo, hi, seed):
    ens = geo.ensemble(n_sys=n_sys, lo=lo, hi=hi, seed=seed)
    sizes = [n for n, _ in ens]
    x = np.concatenate([xs for _, xs in ens])
    return sizes, x, ens


def _err(got, ref, prec):
    if prec == "fp64":
        return np.abs(got - ref).max() / np.abs(ref).max(), 1e-12
    return np.abs(got - ref).max(), 1e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_batch_apply_matches_oracle_per_system(prec):
    sizes, x, ens = _ensemble(5, 20, 400, seed=11)
    f = geo.loads(sum(sizes), 6, 4)
    with BatchedMobility(sizes) as B:
        B.set_positions(x)
        v = B.apply(f, prec)
        for (n, xs), fs, vs in zip(ens, B.split(f), B.split(v)):
            e, t = _err(vs, oracle.apply(xs, 6, xs, 6, fs), prec)
            assert e < t, (n, e)


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_every_decomposition_gives_the_same_product(prec):
    sizes, x, ens = _ensemble(7, 20, 1200, seed=5)
    f = geo.loads(sum(sizes), 6, 8)
    with BatchedMobility(sizes) as B:
        B.set_positions(x)
        ref = B.apply(f, prec)
        for ms, grid in [(256, 64), (512, 300), (1024, 1024), (0, 256), (2048, 2048)]:
            B.set_tuning(ms, grid)
            got = B.apply(f, prec)
            tol = 1e-13 if prec == "fp64" else 1e-6
            assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max()), (ms, grid)
        # the persistent kernel's queue resets itself: repeated launches agree bitwise
        assert np.array_equal(B.apply(f, prec), B.apply(f, prec))


def test_three_dof_and_free_space_batches():
    sizes, x, ens = _ensemble(3, 20, 200, seed=2)
    f3 = geo.loads(sum(sizes), 3, 1)
    with BatchedMobility(sizes, seg_dof=3) as B:
        B.set_positions(x)
        v = B.apply(f3)
        for (n, xs), fs, vs in zip(ens, B.split(f3), B.split(v)):
            e, t = _err(vs, oracle.apply(xs, 3, xs, 3, fs), "fp64")
            assert e < t, e
    f6 = geo.loads(sum(sizes), 6, 2)
    with BatchedMobility(sizes, wall=False, a=1.3, mu=0.7) as B:
        B.set_positions(x)
        v = B.apply(f6)
        for (n, xs), fs, vs in zip(ens, B.split(f6), B.split(v)):
            ref = oracle.apply(xs, 6, xs, 6, fs, wall=False, a_tgt=1.3, a_src=1.3, mu=0.7)
            e, t = _err(vs, ref, "fp64")
            assert e < t, e


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_batched_gmres_matches_single_system_solves(mode):
    sizes, x, ens = _ensemble(6, 20, 700, seed=9)
    rhs = geo.loads(sum(sizes), 6, 3)
    with BatchedMobility(sizes) as B:
        B.set_positions(x)
        sol, stats, ms = B.solve(rhs, rtol=1e-10, precond="self", mode=mode)
        sol2, stats2, _ = B.solve(rhs, rtol=1e-10, precond="self", mode=mode)   # graph replayed
    assert np.array_equal(sol, sol2)
    assert ms > 0.0
    for (n, xs), b, xk, st in zip(ens, np.split(rhs, np.cumsum(sizes)[:-1] * 6),
                                  np.split(sol, np.cumsum(sizes)[:-1] * 6), stats):
        assert st["converged"] and st["rel_residual"] <= 1e-10, st
        assert 0.0 < st["latency_ms"] <= ms + 1e-9
        res = np.linalg.norm(b - oracle.apply(xs, 6, xs, 6, xk)) / np.linalg.norm(b)
        assert res < 1e-9, (n, res)
        with RpyMobilitySolver(n, 0) as m:
            m.set_positions(xs)
            ref, sr = m.solve("seg", b, rtol=1e-10, precond="self", mode="fp64")
        assert np.abs(xk - ref).max() <= 1e-7 * np.abs(ref).max(), n
        assert st["iterations"] <= sr["iterations"] + 60


def test_autotuner_picks_a_tried_candidate():
    sizes, x, ens = _ensemble(12, 100, 2000, seed=4)
    f = geo.loads(sum(sizes), 6, 6)
    with BatchedMobility(sizes) as B:
        B.set_positions(x)
        ref = B.apply(f, "fp32c")
        best, tried = B.tune("fp32c")
        assert len(tried) >= 8
        assert min(t["ms"] for t in tried) == best["ms"] > 0.0
        assert (best["max_sources"], best["grid"]) in {(t["max_sources"], t["grid"]) for t in tried}
        assert B.tuning()["grid"] == best["grid"]
        assert np.abs(B.apply(f, "fp32c") - ref).max() <= 1e-6


def test_sixty_four_systems_of_uneven_size():
    """The benchmarked ensemble: 64 independent wall systems of 100-4,000 segments
    (121,460 particles), every resistance problem solved together to 1e-10 in mixed
    precision.  Every system's reported FP64 residual is at most 1e-10, and the oracle's
    own product confirms it on eight systems (the smallest, the largest and six drawn
    at random; 1e-13 covers the rounding difference of the two FP64 products)."""
    sizes, x, ens = _ensemble(64, 100, 4000, seed=3)
    assert len(sizes) == 64 and min(sizes) <= 300 and max(sizes) >= 3800
    rhs = geo.loads(sum(sizes), 6, 1)
    with BatchedMobility(sizes) as B:
        B.set_positions(x)
        B.tune("fp32c")
        sol, stats, ms = B.solve(rhs, rtol=1e-10, precond="self", mode="mixed")
    assert ms > 0.0
    for st in stats:
        assert st["converged"] and st["rel_residual"] <= 1e-10, st
        assert 0.0 < st["latency_ms"] <= ms + 1e-9
    rng = np.random.default_rng(0)
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))} | set(rng.choice(64, 6, replace=False).tolist())
    bounds = np.cumsum([0] + [6 * n for n in sizes])
    for s in sorted(pick):
        n, xs = ens[s]
        b, xk = rhs[bounds[s]:bounds[s + 1]], sol[bounds[s]:bounds[s + 1]]
        res = np.linalg.norm(b - oracle.apply(xs, 6, xs, 6, xk)) / np.linalg.norm(b)
        assert res <= 1e-10 + 1e-13, (s, n, res)
