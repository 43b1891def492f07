"""The wall products of one species from itself that evaluate each unordered pair once
(apply_sym.hip: segments with 6 or 3 DOF, blobs; FP32C and FP64) against the FP64 oracle
and against the one-directional tiled kernel, through the C ABI.

Sizes cover the block-ring cases: two blocks (the d = B/2 offset of an even ring taken
by one row only), odd and even rings, and a ragged last block (placeholder sources and
targets).  Tolerance as every FP32C product: <= 1e-6 of max(1, |v|_max) on O(1) loads --
the close chains here (spacing 2.2a) give velocities up to 3.4, whose worst entries
both FP32C kernels carry to about 1.1e-6 absolute (3e-7 relative) -- and the two FP32C
kernels agree to the same bound; FP64 to 1e-12 relative."""
import numpy as np
import pytest

import oracle
from cufcm_fil_amd import RpyMobilitySolver
from cufcm_fil_amd import geometry as geo

pytestmark = pytest.mark.gpu

TOL32 = 1e-6


def _positions(n, seed):
    """n segments in chains of 20 (spacing 2.2) above the wall, chains on a 4a grid."""
    rng = np.random.default_rng(seed)
    pts = []
    per = 20
    for c in range((n + per - 1) // per):
        base = np.array([4.0 * (c % 8), 4.0 * (c // 8), 1.5 + rng.uniform(0, 0.5)])
        t = np.array([rng.normal(0, 0.2), rng.normal(0, 0.2), 1.0])
        t /= np.linalg.norm(t)
        for k in range(per):
            pts.append(base + 2.2 * k * t)
    return np.array(pts[:n]).reshape(-1)


@pytest.mark.parametrize("nseg", [128, 130, 192, 256, 64 * 7 + 5, 64 * 16])
def test_symmetric_segment_product_matches_oracle(nseg):
    xs = _positions(nseg, nseg)
    fs = geo.loads(nseg, 6, nseg + 1)
    ref = oracle.apply(xs, 6, xs, 6, fs)
    with RpyMobilitySolver(nseg, 0, precision="fp32c") as m:
        m.set_positions(xs)
        m.profile(True)
        v_sym = m.evaluate_segment_segment_mobility(fs)
        st = m.profile_read()
        m.set_pair_symmetry(False)
        m.profile(True)
        v_one = m.evaluate_segment_segment_mobility(fs)
        st_one = m.profile_read()
    assert st["launches_sym"] == 1 and st["launches_fp32"] == 0, st
    assert st_one["launches_sym"] == 0 and st_one["launches_fp32"] == 1, st_one
    tol = TOL32 * max(1.0, np.abs(ref).max())
    assert np.abs(v_sym - ref).max() < tol, np.abs(v_sym - ref).max()
    assert np.abs(v_one - ref).max() < tol
    assert np.abs(v_sym - v_one).max() < tol


def test_symmetric_product_inside_the_full_step_products():
    """Mss + Mbs (all targets from the segments, the step's first product) and the
    accumulating Msb product after it: the segment rows from the symmetric kernel, the
    blob rows from the one-directional kernel in the same call."""
    nseg = 64 * 5 + 17
    xs = _positions(nseg, 3)
    rng = np.random.default_rng(4)
    nb = 90
    xb = np.column_stack([rng.uniform(0, 30, nb), rng.uniform(0, 20, nb), rng.uniform(50, 60, nb)]).reshape(-1)
    fs = geo.loads(nseg, 6, 5)
    fb = geo.loads(nb, 3, 6)
    ids, idb = np.arange(nseg), np.arange(nseg, nseg + nb)
    with RpyMobilitySolver(nseg, nb, precision="fp32c") as m:
        m.set_positions(xs, xb)
        vb = m.evaluate_blob_segment_mobility(fs)
        vs = m.evaluate_segment_segment_mobility(fs)
        vsb = m.evaluate_segment_blob_mobility(fb)        # accumulates onto v_seg
    ref_ss = oracle.apply(xs, 6, xs, 6, fs)
    tol = TOL32 * max(1.0, np.abs(ref_ss).max())
    assert np.abs(vb - oracle.apply(xb, 3, xs, 6, fs, id_tgt=idb, id_src=ids)).max() < tol
    assert np.abs(vs - ref_ss).max() < tol
    ref_sb = oracle.apply(xs, 6, xb, 3, fb, id_tgt=ids, id_src=idb)
    assert np.abs(vsb - (ref_ss + ref_sb)).max() < tol


def _blobs(n, seed):
    """n blobs on a jittered 2.4a lattice above the wall."""
    rng = np.random.default_rng(seed)
    k = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)[:n]
    x = 2.4 * g + rng.uniform(-0.2, 0.2, g.shape)
    x[:, 2] += 1.5
    return x.reshape(-1)


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("n", [128, 130, 453])
def test_symmetric_blob_product_matches_oracle(prec, n):
    """Blob-blob (3 x 3 blocks) in both precisions, against the oracle and the
    one-directional kernel: FP64 to 1e-12 relative, FP32C to 1e-6 absolute."""
    xb = _blobs(n, n)
    fb = geo.loads(n, 3, n + 7)
    ref = oracle.apply(xb, 3, xb, 3, fb)
    with RpyMobilitySolver(0, n, precision=prec) as m:
        m.set_positions(None, xb)
        v_sym = m.evaluate_blob_blob_mobility(fb)
        m.set_pair_symmetry(False)
        v_one = m.evaluate_blob_blob_mobility(fb)
    tol = 1e-12 * np.abs(ref).max() if prec == "fp64" else TOL32 * max(1.0, np.abs(ref).max())
    assert np.abs(v_sym - ref).max() < tol, np.abs(v_sym - ref).max()
    assert np.abs(v_one - ref).max() < tol
    assert np.abs(v_sym - v_one).max() < tol


@pytest.mark.parametrize("prec,dof", [("fp64", 6), ("fp64", 3), ("fp32c", 3)])
def test_symmetric_segment_variants_match_oracle(prec, dof):
    """FP64 6-DOF segments and 3-DOF segments (prescribed cilia) in both precisions."""
    nseg = 64 * 7 + 5
    xs = _positions(nseg, 11)
    fs = geo.loads(nseg, dof, 12)
    ref = oracle.apply(xs, dof, xs, dof, fs)
    with RpyMobilitySolver(nseg, 0, seg_dof=dof, precision=prec) as m:
        m.set_positions(xs)
        v_sym = m.evaluate_segment_segment_mobility(fs)
        m.set_pair_symmetry(False)
        v_one = m.evaluate_segment_segment_mobility(fs)
    tol = 1e-12 * np.abs(ref).max() if prec == "fp64" else TOL32 * max(1.0, np.abs(ref).max())
    assert np.abs(v_sym - ref).max() < tol, np.abs(v_sym - ref).max()
    assert np.abs(v_sym - v_one).max() < tol
