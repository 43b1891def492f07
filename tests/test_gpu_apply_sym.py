"""The FP32C wall product of the segments from the segments that evaluates each unordered
pair once (apply_sym.hip) against the FP64 oracle and against the one-directional tiled
kernel, through the C ABI.

Sizes cover the block-ring cases: two blocks (the d = B/2 offset of an even ring taken
by one row only), odd and even rings, and a ragged last block (placeholder sources and
targets).  Tolerance as every FP32C product: <= 1e-6 absolute on O(1) loads (about
1e-6 relative), and the two FP32C kernels agree to the same bound."""
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
    assert np.abs(v_sym - ref).max() < TOL32, np.abs(v_sym - ref).max()
    assert np.abs(v_one - ref).max() < TOL32
    assert np.abs(v_sym - v_one).max() < TOL32


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
    assert np.abs(vb - oracle.apply(xb, 3, xs, 6, fs, id_tgt=idb, id_src=ids)).max() < TOL32
    ref_ss = oracle.apply(xs, 6, xs, 6, fs)
    assert np.abs(vs - ref_ss).max() < TOL32
    ref_sb = oracle.apply(xs, 6, xb, 3, fb, id_tgt=ids, id_src=idb)
    assert np.abs(vsb - (ref_ss + ref_sb)).max() < TOL32
