"""Free rigid bodies made of blobs (through the C ABI) against the FP64 oracle:
the saddle-point product, the three block preconditioners, GMRES solves and the body
reference matrices.

Tolerances: FP64 products <= 1e-12 relative (max-norm), FP32C <= 1e-6 relative;
CFM_PC_DIAG against its numpy restatement <= 1e-12; CFM_PC_DENSE inverts the saddle
matrix to the FP32 rounding of the resident inverse (<= 1e-5); solves reach an FP64
relative residual <= 1e-10 and the dense solution to 1e-8."""
import math

import numpy as np
import pytest

import body_oracle as bo
import oracle
from cufcm_fil_amd import RpyMobilitySolver
from cufcm_fil_amd import _lib as L
from cufcm_fil_amd import geometry as geo

pytestmark = pytest.mark.gpu


def _quat(seed):
    q = np.random.default_rng(seed).normal(size=4)
    return q / np.linalg.norm(q)


def _bodies(n_body=2, nb=80, R=5.0, wall=True, seed=1):
    ref = geo.fibonacci_sphere(nb, R)
    x = np.array([[14.0 * b, 3.0 * b, 9.0 + 2.0 * b] for b in range(n_body)])
    q = np.stack([_quat(seed + b) for b in range(n_body)])
    pos, off = bo.place(ref, x, q)
    K = bo.k_matrix(off, n_body)
    M = oracle.mobility_matrix(pos, 3, wall=wall)
    return ref, x, q, pos, off, K, M


def _solver(ref, x, q, n_body, wall=True):
    m = RpyMobilitySolver(0, len(ref) * n_body, wall=wall)
    m.body_init(ref, n_body)
    m.body_set_state(x, q)
    return m


@pytest.mark.parametrize("wall", [True, False])
def test_saddle_product_matches_oracle(wall):
    ref, x, q, pos, off, K, M = _bodies(wall=wall)
    A = bo.saddle_matrix(M, K)
    v = np.random.default_rng(5).normal(size=A.shape[0])
    ref_y = A @ v
    with _solver(ref, x, q, 2, wall) as m:
        for prec, tol in [("fp64", 1e-12), ("fp32c", 1e-6)]:
            y = m.system_apply("body", v, prec)
            assert np.abs(y - ref_y).max() <= tol * np.abs(ref_y).max(), prec
        # blob positions are the placed ones
        yb = m.evaluate_blob_blob_mobility(v[:3 * len(pos)])
        assert np.abs(yb - M @ v[:3 * len(pos)]).max() <= 1e-12 * np.abs(yb).max()


def test_diag_preconditioner_matches_reference_block_form():
    ref, x, q, pos, off, K, M = _bodies()
    kinv = np.linalg.inv(bo.ktk_reference(ref))
    N = np.random.default_rng(3).normal(size=(6, 6)) * 0.01
    N = N @ N.T + 0.02 * np.eye(6)
    v = np.random.default_rng(6).normal(size=3 * len(pos) + 12)
    with _solver(ref, x, q, 2) as m:
        m.body_set_reference_matrices(kinv, N)
        got = m.apply_preconditioner("body", v, "diag")
        k2, n2 = m.body_reference_matrices()
    want = bo.precondition(v, off, q, kinv, N)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    assert np.allclose(k2, kinv, rtol=0, atol=1e-15 * np.abs(kinv).max()) and np.array_equal(n2, N)


def test_self_equals_diag_without_wall_and_none_is_identity():
    ref, x, q, pos, off, K, M = _bodies(wall=False)
    v = np.random.default_rng(7).normal(size=3 * len(pos) + 12)
    with _solver(ref, x, q, 2, wall=False) as m:
        a = m.apply_preconditioner("body", v, "self")
        b = m.apply_preconditioner("body", v, "diag")
        c = m.apply_preconditioner("body", v, "none")
    assert np.array_equal(a, b)
    assert np.array_equal(c, v)


def test_dense_preconditioner_inverts_the_saddle_matrix():
    ref, x, q, pos, off, K, M = _bodies()
    A = bo.saddle_matrix(M, K)
    v = np.random.default_rng(8).normal(size=A.shape[0])
    with _solver(ref, x, q, 2) as m:
        got = m.apply_preconditioner("body", A @ v, "dense")
    assert np.abs(got - v).max() <= 1e-5 * np.abs(v).max()


@pytest.mark.parametrize("precond,mode", [("diag", "fp64"), ("self", "mixed"), ("dense", "fp64"),
                                          ("dense", "mixed")])
def test_free_body_solve(precond, mode):
    ref, x, q, pos, off, K, M = _bodies()
    A = bo.saddle_matrix(M, K)
    rng = np.random.default_rng(12)
    rhs = np.concatenate([rng.normal(size=3 * len(pos)) * 0.1, rng.normal(size=12)])
    want = np.linalg.solve(A, rhs)
    with _solver(ref, x, q, 2) as m:
        if precond == "diag":
            m.make_body_reference_matrices(0, precond="dense")
        sol, st = m.solve("body", rhs, rtol=1e-10, precond=precond, mode=mode, max_iter=3000)
    assert st["converged"] and st["rel_residual"] <= 1e-10, st
    assert np.linalg.norm(rhs - A @ sol) / np.linalg.norm(rhs) < 1e-9
    assert np.abs(sol - want).max() <= 1e-8 * np.abs(want).max()
    if precond == "dense":
        assert st["iterations"] <= 12, st


@pytest.mark.parametrize("wall", [False, True])
def test_reference_matrices_match_dense_body_mobility(wall):
    ref, x, q, pos, off, K, M = _bodies(n_body=2, wall=wall)
    with _solver(ref, x, q, 2, wall) as m:
        kinv, N, st = m.make_body_reference_matrices(1, precond="dense", rtol=1e-12)
    assert np.allclose(kinv, np.linalg.inv(bo.ktk_reference(ref)), rtol=1e-12, atol=0)
    # body 1 alone, lab frame -> body frame
    K1 = bo.k_matrix(off[len(ref):], 1)
    M1 = oracle.mobility_matrix(pos[len(ref):], 3, wall=wall)
    R = bo.rot_mat(q[1])
    want = bo.rotate6(bo.body_mobility(M1, K1), R.T)
    assert np.abs(N - want).max() <= 1e-9 * np.abs(want).max(), (N, want)
    assert st["rel_residual"] <= 1e-12


def test_blob_sphere_stokes_drag_and_rotation_invariance():
    """The reference's self-check: a 400-blob sphere of radius 8 in unbounded fluid is
    within 5% of Stokes drag (the oracle's figure, 4.7%); its body-frame mobility does
    not depend on the orientation it was computed at."""
    ref = geo.fibonacci_sphere(400, 8.0)
    Ns = []
    for qq in [np.array([1.0, 0, 0, 0]), _quat(21)]:
        with _solver(ref, np.zeros((1, 3)), qq[None], 1, wall=False) as m:
            _, N, _ = m.make_body_reference_matrices(0, precond="dense", rtol=1e-12)
            Ns.append(N)
    Nt = np.diag([1 / (6 * math.pi * 8.0)] * 3 + [1 / (8 * math.pi * 8.0 ** 3)] * 3)
    err = np.linalg.norm(Ns[0] - Nt) / np.linalg.norm(Nt)
    assert 0.04 < err < 0.055, err
    assert np.abs(Ns[0] - Ns[1]).max() <= 1e-9 * np.abs(Ns[0]).max()


def test_set_positions_unplaces_the_bodies():
    ref, x, q, pos, off, K, M = _bodies()
    with _solver(ref, x, q, 2) as m:
        m.set_positions(x_blob=pos)
        with pytest.raises(RuntimeError, match="body state not set"):
            m.solve("body", np.ones(3 * len(pos) + 12), precond="diag")
        m.body_set_state(x, q)
        _, st = m.solve("body", np.ones(3 * len(pos) + 12), precond="dense", rtol=1e-10)
        assert st["converged"]
