"""GPU parity of the mobility products (through the C ABI) against the FP64 oracle.

Tolerances: FP64 path <= 1e-12 relative (max-norm of the error over
max-norm of the oracle result); FP32-compensated path <= 1e-6 of max(1, |v|_max), i.e.
1e-6 absolute on O(1) velocities and 1e-6 relative where random placements put pairs
at a third of a radius apart and the velocities grow (the wall branch has no overlap
forms, as in the reference, so such pairs carry large, cancelling terms).
"""
import numpy as np
import pytest

import oracle
from cufcm_fil_amd import CFM_BB, CFM_BS, CFM_FULL, CFM_SB, CFM_SS, RpyMobilitySolver
from cufcm_fil_amd import geometry as geo

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


def _system(nfil=12, nblob=150, seed=0):
    xs = geo.tilted_lattice(nx=4, ny=3, seed=seed)
    rng = np.random.default_rng(seed + 1)
    xb = np.column_stack([rng.uniform(-5, 40, nblob), rng.uniform(-5, 30, nblob),
                          rng.uniform(50.0, 70.0, nblob)]).reshape(-1)
    fs = geo.loads(xs.size // 3, 6, seed + 2)
    fb = geo.loads(nblob, 3, seed + 3)
    return xs, xb, fs, fb


def _err(got, ref, prec):
    if prec == "fp64":
        return np.abs(got - ref).max() / np.abs(ref).max(), TOL64
    return np.abs(got - ref).max(), TOL32 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("split", [1, 3])
def test_blocks_match_oracle(prec, split):
    xs, xb, fs, fb = _system()
    ns, nb = xs.size // 3, xb.size // 3
    ids, idb = np.arange(ns), np.arange(ns, ns + nb)
    with RpyMobilitySolver(ns, nb, precision=prec, tile_split=split) as m:
        m.set_positions(xs, xb)
        ref_ss = oracle.apply(xs, 6, xs, 6, fs)
        ref_sb = oracle.apply(xs, 6, xb, 3, fb, id_tgt=ids, id_src=idb)
        ref_bb = oracle.apply(xb, 3, xb, 3, fb)
        ref_bs = oracle.apply(xb, 3, xs, 6, fs, id_tgt=idb, id_src=ids)
        v_ss = m.evaluate_segment_segment_mobility(fs)
        e, t = _err(v_ss, ref_ss, prec)
        assert e < t, ("SS", e)
        v_sb = m.evaluate_segment_blob_mobility(fb)           # adds to v_seg (Msb_mult)
        e, t = _err(v_sb, ref_ss + ref_sb, prec)
        assert e < t, ("SB", e)
        e, t = _err(m.evaluate_blob_blob_mobility(fb), ref_bb, prec)
        assert e < t, ("BB", e)
        e, t = _err(m.evaluate_blob_segment_mobility(fs), ref_bs, prec)
        assert e < t, ("BS", e)
        vs, vb = m.evaluate_full_mobility(fs, fb)
        rs, rb = oracle.apply_full(xs, fs, xb, fb)
        for got, ref in ((vs, rs), (vb, rb)):
            e, t = _err(got, ref, prec)
            assert e < t, ("FULL", e)


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_tilted_lattice_1280(prec):
    xs = geo.tilted_lattice()
    fs = geo.loads(1280, 6, 1)
    with RpyMobilitySolver(1280, 0, precision=prec) as m:
        m.set_positions(xs)
        got = m.evaluate_segment_segment_mobility(fs)
    ref = oracle.apply(xs, 6, xs, 6, fs)
    e, t = _err(got, ref, prec)
    assert e < t, e


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_wave_lattice_12800_sampled(prec):
    xs = geo.wave_lattice()
    fs = geo.loads(12800, 6, 2)
    with RpyMobilitySolver(12800, 0, precision=prec) as m:
        m.set_positions(xs)
        got = m.evaluate_segment_segment_mobility(fs).reshape(-1, 6)
    rows = np.random.default_rng(0).choice(12800, 400, replace=False)
    xt = xs.reshape(-1, 3)[rows].reshape(-1)
    ref = oracle.apply(xt, 6, xs, 6, fs, id_tgt=rows, id_src=np.arange(12800)).reshape(-1, 6)
    e, t = _err(got[rows], ref, prec)
    assert e < t, e


@pytest.mark.parametrize("seg_dof", [3, 6])
def test_unbounded_branch_and_overlaps(seg_dof):
    rng = np.random.default_rng(5)
    n = 300
    xs = rng.uniform(0, 12, 3 * n)          # dense: many overlapping pairs
    fs = rng.standard_normal(seg_dof * n)
    with RpyMobilitySolver(n, 0, seg_dof=seg_dof, wall=False, a_seg=1.3, mu=0.7) as m:
        m.set_positions(xs)
        got = m.evaluate_segment_segment_mobility(fs)
    ref = oracle.apply(xs, seg_dof, xs, seg_dof, fs, wall=False, a_tgt=1.3, a_src=1.3, mu=0.7)
    assert np.abs(got - ref).max() / np.abs(ref).max() < TOL64


def test_mu_scaling_and_linearity():
    xs, xb, fs, fb = _system(seed=4)
    ns, nb = xs.size // 3, xb.size // 3
    with RpyMobilitySolver(ns, nb, mu=2.5) as m:
        m.set_positions(xs, xb)
        a_s, a_b = m.evaluate_full_mobility(fs, fb)
        b_s, b_b = m.evaluate_full_mobility(2 * fs, 2 * fb)
    rs, rb = oracle.apply_full(xs, fs, xb, fb, mu=2.5)
    assert np.abs(a_s - rs).max() / np.abs(rs).max() < TOL64
    assert np.allclose(b_s, 2 * a_s, rtol=1e-13, atol=0) and np.allclose(b_b, 2 * a_b, rtol=1e-13)


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_autotuned_split_matches_oracle(prec):
    """tile_split = -1: the first product of each shape times candidate splits into
    scratch memory; accumulating blocks (Msb adds into v_seg) must add exactly once."""
    xs, xb, fs, fb = _system(seed=4)
    ns, nb = xs.size // 3, xb.size // 3
    ids, idb = np.arange(ns), np.arange(ns, ns + nb)
    with RpyMobilitySolver(ns, nb, precision=prec, tile_split=-1) as m:
        m.set_positions(xs, xb)
        ref_ss = oracle.apply(xs, 6, xs, 6, fs)
        ref_sb = oracle.apply(xs, 6, xb, 3, fb, id_tgt=ids, id_src=idb)
        for _ in range(2):      # tuned on the first call, cached on the second
            v_ss = m.evaluate_segment_segment_mobility(fs)
            e, t = _err(v_ss, ref_ss, prec)
            assert e < t, ("SS", e)
            v_sb = m.evaluate_segment_blob_mobility(fb)
            e, t = _err(v_sb, ref_ss + ref_sb, prec)
            assert e < t, ("SB", e)
        vs, vb = m.evaluate_full_mobility(fs, fb)
        rs, rb = oracle.apply_full(xs, fs, xb, fb)
        for got, ref in ((vs, rs), (vb, rb)):
            e, t = _err(got, ref, prec)
            assert e < t, ("FULL", e)


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("ns,nb", [(1, 0), (2, 0), (63, 1), (65, 0), (257, 130), (0, 5),
                                   (0, 257), (300, 64)])
def test_ragged_sizes_match_oracle(ns, nb, prec):
    """Sizes on either side of the 64-target and 256-source tiles, single particles
    (the self term alone) and single-species contexts, every product against the
    oracle's direct double loop."""
    rng = np.random.default_rng(ns * 1000 + nb)
    xs = np.column_stack([rng.uniform(0, 40, ns), rng.uniform(0, 40, ns),
                          rng.uniform(1.5, 30, ns)]).reshape(-1)
    xb = np.column_stack([rng.uniform(0, 40, nb), rng.uniform(0, 40, nb),
                          rng.uniform(1.5, 30, nb)]).reshape(-1)
    fs, fb = geo.loads(ns, 6, 7) if ns else np.zeros(0), geo.loads(nb, 3, 8) if nb else np.zeros(0)
    ids, idb = np.arange(ns), np.arange(ns, ns + nb)
    with RpyMobilitySolver(ns, nb, precision=prec) as m:
        m.set_positions(xs, xb)
        if ns:
            e, t = _err(m.evaluate_segment_segment_mobility(fs), oracle.apply(xs, 6, xs, 6, fs), prec)
            assert e < t, ("SS", e)
        if nb:
            e, t = _err(m.evaluate_blob_blob_mobility(fb), oracle.apply(xb, 3, xb, 3, fb), prec)
            assert e < t, ("BB", e)
        if ns and nb:
            ref = oracle.apply(xb, 3, xs, 6, fs, id_tgt=idb, id_src=ids)
            e, t = _err(m.evaluate_blob_segment_mobility(fs), ref, prec)
            assert e < t, ("BS", e)
            vs, vb = m.evaluate_full_mobility(fs, fb)
            rs, rb = oracle.apply_full(xs, fs, xb, fb)
            for got, r in ((vs, rs), (vb, rb)):
                e, t = _err(got, r, prec)
                assert e < t, ("FULL", e)
