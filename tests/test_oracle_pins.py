"""CPU pins of the FP64 oracle (oracle/rpy_oracle.cpp).

Each check is independent of the closed-form coefficients the oracle uses:
  * wall pair blocks vs Blake's image tensor + Faxen operators by finite
    differences (tests/blake_faxen.py),
  * wall self term vs the same construction at coincident points and vs the
    published closed forms (Perkins & Jones 1992; Swan & Brady 2007),
  * free-space branch vs the Faxen construction on the Oseen tensor, and the
    textbook equal-sphere overlap form,
  * grand-matrix symmetry and positive definiteness.
A plausible mistake (a dropped term, a sign flip, a swapped index) fails one.
"""
import numpy as np
import pytest

import blake_faxen as bf
import oracle

E6 = np.eye(6)


def block(fn):
    return np.array([fn(E6[:, c]) for c in range(6)]).T


def test_blake_tensor_is_a_wall_bounded_stokes_flow():
    y = np.array([0.3, -0.2, 2.0])
    assert np.abs(bf.blake([1.0, 0.5, 0.0], y)).max() < 1e-15          # no slip
    assert np.abs(bf.divergence_x(bf.blake, [1.0, 0.5, 1.3], y)).max() < 1e-9
    far = np.array([0.0, 0.0, 1e7])
    g = bf.blake(far + [1.0, 2.0, 0.5], far)
    assert np.allclose(g, bf.oseen([1.0, 2.0, 0.5], [0, 0, 0]), atol=1e-6)


@pytest.mark.parametrize("seed", range(6))
def test_wall_pair_matches_blake_faxen(seed):
    rng = np.random.default_rng(seed)
    xi = np.array([rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(1.1, 5.0)])
    xj = np.array([rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(1.1, 5.0)])
    if np.linalg.norm(xi - xj) < 2.1:
        xj[0] += 3.0
    ref = bf.blocks(bf.blake, xi, xj)
    got = block(lambda F: oracle.wall_pair(xi, xj, F))
    assert np.abs(got - ref).max() < 1e-8 * max(1.0, np.abs(ref).max() * 100)
    # every 3x3 block is exercised (none is identically zero here)
    for r in (slice(0, 3), slice(3, 6)):
        for c in (slice(0, 3), slice(3, 6)):
            assert np.abs(ref[r, c]).max() > 1e-5


@pytest.mark.parametrize("z", [1.1, 1.5, 2.0, 4.0, 10.0])
def test_wall_self_matches_image_limit_and_closed_forms(z):
    x = np.array([0.3, 0.1, z])
    img = bf.blocks(bf.blake_image, x, x)
    img[:3, :3] += np.eye(3) / (6 * np.pi)
    img[3:, 3:] += np.eye(3) / (8 * np.pi)
    got = block(lambda F: oracle.wall_self(z, F))
    assert np.abs(got - img).max() < 1e-8
    h = 1.0 / z
    c6, c8 = 1 / (6 * np.pi), 1 / (8 * np.pi)
    # Perkins & Jones (1992) / Swan & Brady (2007) single-sphere wall mobilities
    assert np.isclose(got[0, 0], c6 * (1 - 9 / 16 * h + 1 / 8 * h**3 - 1 / 16 * h**5), rtol=1e-14)
    assert np.isclose(got[2, 2], c6 * (1 - 9 / 8 * h + 1 / 2 * h**3 - 1 / 8 * h**5), rtol=1e-14)
    assert np.isclose(got[3, 3], c8 * (1 - 5 / 16 * h**3), rtol=1e-14)
    assert np.isclose(got[5, 5], c8 * (1 - 1 / 8 * h**3), rtol=1e-14)
    # a sphere pushed along +x next to a wall below it rotates like a rolling wheel
    assert got[4, 0] > 0 and np.isclose(got[4, 0], c6 * 3 / 32 * h**4, rtol=1e-14)
    assert np.allclose(got, got.T, atol=1e-16)


def test_free_branch_matches_faxen_oseen_and_overlap_form():
    xi, xj = np.zeros(3), np.array([1.5, 2.0, -0.7])
    ref = bf.blocks(bf.oseen, xi, xj)
    got = block(lambda F: oracle.free_pair(xi, 1.0, xj, 1.0, F))
    assert np.abs(got - ref).max() < 1e-8
    # equal spheres overlapping (r < 2a): (1/6 pi a)[(1 - 9r/32a) I + (3r/32a) rr]
    r = 1.3
    d = np.array([0.6, -0.8, 0.0])
    got = block(lambda F: oracle.free_pair(r * d, 1.0, np.zeros(3), 1.0, F))
    tt = ((1 - 9 * r / 32) * np.eye(3) + 3 * r / 32 * np.outer(d, d)) / (6 * np.pi)
    assert np.allclose(got[:3, :3], tt, atol=1e-15)
    # continuity of every block across r = a_i + a_j for unequal radii
    ai, aj = 1.0, 1.7
    u = np.array([0.48, 0.6, 0.64])
    inside = block(lambda F: oracle.free_pair((ai + aj - 1e-9) * u, ai, np.zeros(3), aj, F))
    outside = block(lambda F: oracle.free_pair((ai + aj + 1e-9) * u, ai, np.zeros(3), aj, F))
    assert np.abs(inside - outside).max() < 1e-8


def _configs():
    rng = np.random.default_rng(7)
    pts = []
    while len(pts) < 8:
        p = np.array([rng.uniform(0, 9), rng.uniform(0, 9), rng.uniform(1.2, 6)])
        if all(np.linalg.norm(p - q) > 2.2 for q in pts):
            pts.append(p)
    return np.array(pts)


@pytest.mark.parametrize("wall", [True, False])
def test_grand_mobility_symmetric_positive_definite(wall):
    x = _configs()
    M = oracle.mobility_matrix(x, 6, wall=wall)
    assert np.abs(M - M.T).max() < 1e-15
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_mixed_species_blocks_use_global_ids():
    # segments (6-dof) and blobs (3-dof) stacked; blob k must not get the self term of seg k
    x = _configs()
    xs, xb = x[:4], x[4:]
    fs = np.random.default_rng(1).standard_normal(24)
    fb = np.random.default_rng(2).standard_normal(12)
    vs, vb = oracle.apply_full(xs, fs, xb, fb)
    # same physics from one 6-dof species with zero torques/ignored spins on the blobs
    F = np.zeros(48)
    F[:24] = fs
    for k in range(4):
        F[24 + 6 * k: 27 + 6 * k] = fb[3 * k: 3 * k + 3]
    V = oracle.apply(x, 6, x, 6, F)
    assert np.allclose(vs, V[:24], atol=1e-15, rtol=1e-13)
    assert np.allclose(vb, V[24:].reshape(4, 6)[:, :3].reshape(-1), atol=1e-15, rtol=1e-13)
    # cross blocks are transposes of each other (M_sb = M_bs^T)
    ids, idb = np.arange(4), np.arange(4, 8)
    Msb = np.array([oracle.apply(xs, 6, xb, 3, np.eye(12)[:, c], id_tgt=ids, id_src=idb)
                    for c in range(12)]).T
    Mbs = np.array([oracle.apply(xb, 3, xs, 6, np.eye(24)[:, c], id_tgt=idb, id_src=ids)
                    for c in range(24)]).T
    assert np.abs(Msb - Mbs.T).max() < 1e-16
    # keyed on per-species indices (the reference's i == j test) the cross block
    # would pick up a spurious self mobility and lose its symmetry
    bad = np.array([oracle.apply(xs, 6, xb, 3, np.eye(12)[:, c], id_tgt=ids, id_src=ids)
                    for c in range(12)]).T
    assert np.abs(bad - Mbs.T).max() > 1e-3


def test_gmres_reference_matches_dense_solve():
    x = _configs()
    M = oracle.mobility_matrix(x, 6, wall=True)
    b = np.random.default_rng(3).standard_normal(M.shape[0])
    P = np.diag(1.0 / np.diag(M))
    for restart in (None, 10):
        sol, its, rel = oracle.gmres_reference(lambda v: M @ v, lambda v: P @ v, b, tol=1e-12,
                                               restart=restart, max_iter=500)
        assert rel < 1e-11
        assert np.allclose(sol, np.linalg.solve(M, b), rtol=1e-9, atol=1e-9)
