"""CPU pins of the free-rigid-body oracle (oracle/body_oracle.py): what the mathematics
fixes about K, the reference's block preconditioner, the body mobility and a blob
sphere's hydrodynamics (Stokes drag, Faxen/Brenner wall corrections)."""
import math

import numpy as np
import pytest

import body_oracle as bo
import oracle
from cufcm_fil_amd import geometry as geo
from filament_oracle import rot_mat


def _quat(seed):
    q = np.random.default_rng(seed).normal(size=4)
    return q / np.linalg.norm(q)


def test_ktk_closed_form_is_the_gram_matrix_of_k():
    ref = np.random.default_rng(1).normal(size=(37, 3)) * 3.0
    K = bo.k_matrix(ref, 1)
    assert np.allclose(bo.ktk_reference(ref), K.T @ K, rtol=0, atol=1e-12)
    # K maps (V, Omega) to V + Omega x r
    u = np.random.default_rng(2).normal(size=6)
    v = (K @ u).reshape(-1, 3)
    assert np.allclose(v, u[:3] + np.cross(u[3:], ref), atol=1e-13)


def test_placement_rotates_offsets_about_the_centre():
    ref = np.random.default_rng(3).normal(size=(10, 3))
    x = np.array([[1.0, -2.0, 7.0], [0.5, 0.5, 9.0]])
    q = np.stack([_quat(4), _quat(5)])
    pos, off = bo.place(ref, x, q)
    for b in range(2):
        R = rot_mat(q[b])
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert np.allclose(off[10 * b:10 * b + 10], ref @ R.T, atol=1e-14)
        assert np.allclose(pos[10 * b:10 * b + 10] - x[b], off[10 * b:10 * b + 10], atol=1e-13)


def test_reference_preconditioner_inverts_the_scalar_saddle_matrix():
    """With M_bb replaced by its scalar approximation m I (m = 1/(6 pi mu a)) and the
    body mobility by m (K^T K)^-1, mobility_solver.cpp:1518-1590 is the exact inverse
    of the saddle matrix -- for rotated bodies too (the reference matrices are body
    frame and rotated by R)."""
    mu, a = 0.7, 1.3
    m = 1.0 / (6.0 * math.pi * mu * a)
    ref = geo.fibonacci_sphere(40, 4.0)
    x = np.array([[0.0, 0.0, 10.0], [12.0, 0.0, 10.0]])
    q = np.stack([_quat(7), _quat(8)])
    pos, off = bo.place(ref, x, q)
    K = bo.k_matrix(off, 2)
    A = bo.saddle_matrix(m * np.eye(3 * len(pos)), K)
    kinv = np.linalg.inv(bo.ktk_reference(ref))
    v = np.random.default_rng(9).normal(size=A.shape[0])
    got = bo.precondition(v, off, q, kinv, m * kinv, mu=mu, a=a)
    assert np.allclose(got, np.linalg.solve(A, v), rtol=0, atol=1e-10 * np.abs(got).max())


@pytest.mark.parametrize("wall", [False, True])
def test_body_mobility_from_six_solves(wall):
    """make_body_reference_matrices' solves (rhs -1 at row rows-1-n -> column 5-n)
    give (K^T M^-1 K)^-1, a symmetric positive-definite 6 x 6."""
    ref = geo.fibonacci_sphere(60, 4.5)
    pos, off = bo.place(ref, np.array([[0.0, 0.0, 9.0]]), _quat(11)[None])
    K = bo.k_matrix(off, 1)
    M = oracle.mobility_matrix(pos, 3, wall=wall)
    N = bo.reference_mobility_by_solves(M, K)
    assert np.allclose(N, bo.body_mobility(M, K), rtol=0, atol=1e-12 * np.abs(N).max())
    assert np.allclose(N, N.T, rtol=0, atol=1e-12 * np.abs(N).max())
    assert np.linalg.eigvalsh(0.5 * (N + N.T)).min() > 0.0


def test_blob_sphere_converges_to_stokes_drag():
    """The reference's sphere check (mobility_solver.cpp:3116-3132): the body mobility
    of a blob sphere of radius R approaches diag(1/(6 pi mu R), 1/(8 pi mu R^3)) as
    the blobs fill its surface (relative error 10.5% at 100 blobs, 4.7% at 400)."""
    errs = []
    for n, R in [(100, 4.0), (400, 8.0)]:
        pos, off = bo.place(geo.fibonacci_sphere(n, R), np.zeros((1, 3)), np.array([[1.0, 0, 0, 0]]))
        N = bo.body_mobility(oracle.mobility_matrix(pos, 3, wall=False), bo.k_matrix(off, 1))
        Nt = np.diag([1 / (6 * math.pi * R)] * 3 + [1 / (8 * math.pi * R ** 3)] * 3)
        errs.append(np.linalg.norm(N - Nt) / np.linalg.norm(Nt))
        # isotropic to the seeding's regularity, no translation-rotation coupling
        assert np.abs(np.diag(N)[:3] / N[0, 0] - 1).max() < 0.01
        assert np.abs(N[:3, 3:]).max() < 0.02 * math.sqrt(N[0, 0] * N[3, 3])
    assert 0.09 < errs[0] < 0.12 and 0.04 < errs[1] < 0.055


def test_blob_sphere_near_the_wall_follows_faxen_and_brenner():
    """Translational mobility of the blob sphere at h = 3R above the wall relative to
    its unbounded value: Faxen's parallel series to 1%, Brenner's normal series to 3%
    (a dropped or mis-signed image term moves these by 20-60%)."""
    R, h = 8.0, 24.0
    ref = geo.fibonacci_sphere(400, R)
    pos, off = bo.place(ref, np.zeros((1, 3)), np.array([[1.0, 0, 0, 0]]))
    K = bo.k_matrix(off, 1)
    N0 = bo.body_mobility(oracle.mobility_matrix(pos, 3, wall=False), K)
    Nw = bo.body_mobility(oracle.mobility_matrix(pos + [0.0, 0.0, h], 3, wall=True), K)
    par, nor = bo.faxen_brenner_ratios(R, h)
    assert abs(Nw[0, 0] / N0[0, 0] / par - 1) < 0.01
    assert abs(Nw[1, 1] / N0[1, 1] / par - 1) < 0.01
    assert abs(Nw[2, 2] / N0[2, 2] / nor - 1) < 0.03


def test_legendre_helpers_match_closed_forms():
    """P_n and P_n' against the closed forms of degrees 2-4, and the squirmer's tangential
    modes V_1 = sin(theta), V_2 = sin(theta) cos(theta)."""
    for x in (-0.9, -0.3, 0.2, 0.75):
        for n, p, dp in ((2, 0.5 * (3 * x * x - 1), 3 * x), (3, 0.5 * (5 * x ** 3 - 3 * x), 0.5 * (15 * x * x - 3)),
                         (4, (35 * x ** 4 - 30 * x * x + 3) / 8, (140 * x ** 3 - 60 * x) / 8)):
            got = bo.legendre(n, x)
            assert np.allclose(got, (p, dp), rtol=1e-13, atol=1e-14)
    th = 0.7
    ref = np.array([[math.sin(th), 0.0, math.cos(th)]])
    u1 = bo.squirmer_slip(ref, [0.0], [1.0])[0]          # u_theta = sin(theta) along e_theta
    e_th = np.array([math.cos(th), 0.0, -math.sin(th)])
    assert np.allclose(u1, math.sin(th) * e_th, atol=1e-15)
    u2 = bo.squirmer_slip(ref, [0.0, 0.0], [0.0, 1.0])[0]
    assert np.allclose(u2, math.sin(th) * math.cos(th) * e_th, atol=1e-15)
    ur = bo.squirmer_slip(ref, [2.0], [0.0])[0]          # u_r = A_1 cos(theta) along e_r
    assert np.allclose(ur, 2.0 * math.cos(th) * ref[0], atol=1e-15)


def test_blob_squirmer_swims_at_two_thirds_b1():
    """A spherical squirmer with tangential slip B_1 sin(theta) swims along its axis at
    2 B_1 / 3 (Lighthill 1952, Blake 1971), without rotating, whatever B_2 is; B_2 alone
    does not swim.  A blob sphere of radius R reproduces this up to a correction of order
    a / R (the slip sits at the blob centres): U = 0.808, 0.757 at R = 8, 12.5 (400, 1,000
    blobs), and the 1/R extrapolation lands within 1% of 2/3 (0.665)."""
    Us = []
    for n, R in [(400, 8.0), (1000, 12.5)]:
        ref = geo.fibonacci_sphere(n, R)
        pos, off = bo.place(ref, np.zeros((1, 3)), np.array([[1.0, 0, 0, 0]]))
        M = oracle.mobility_matrix(pos, 3, wall=False)
        K = bo.k_matrix(off, 1)
        U1, _ = bo.squirmer_velocity(M, K, bo.squirmer_slip(ref, [0.0, 0.0], [1.0, 0.0]))
        U12, _ = bo.squirmer_velocity(M, K, bo.squirmer_slip(ref, [0.0, 0.0], [1.0, 3.0]))
        U2, _ = bo.squirmer_velocity(M, K, bo.squirmer_slip(ref, [0.0, 0.0], [0.0, 1.0]))
        assert abs(U1[0, 2] - U12[0, 2]) < 1e-3
        assert np.abs(U2[0]).max() < 2e-4
        assert np.abs(U1[0, [0, 1, 3, 4, 5]]).max() < 5e-4
        Us.append((R, U1[0, 2]))
    (R1, u1), (R2, u2) = Us
    c = (u1 - u2) / (1.0 / R1 - 1.0 / R2)
    u_inf = u2 - c / R2
    assert abs(u_inf / (2.0 / 3.0) - 1.0) < 0.01, (Us, u_inf)
    # the sign of e_theta is pinned: the body swims towards its north pole (+z)
    assert u1 > u2 > 2.0 / 3.0


def _two_bodies_over_wall():
    ref = geo.fibonacci_sphere(60, 4.5)
    x = np.array([[0.0, 0.0, 8.0], [11.0, 2.0, 9.0]])
    q = np.stack([_quat(21), _quat(22)])
    pos, off = bo.place(ref, x, q)
    return ref, pos, off, q, bo.k_matrix(off, 2), oracle.mobility_matrix(pos, 3, wall=True)


def test_prescribed_rotation_velocity_is_the_clamped_mobility():
    """Under a unit force the prescribed-rotation solve moves a body above the wall with
    ((N^-1)_VV)^-1 e_k, N the body mobility of six free solves -- the Schur complement
    N_VV - N_VO N_OO^-1 N_OV, which differs from the torque-free N_VV by the wall's
    translation-rotation coupling (a solve that kept the Omega columns or the torque
    rows would return N_VV)."""
    ref = geo.fibonacci_sphere(60, 4.5)
    pos, off = bo.place(ref, np.array([[0.0, 0.0, 7.0]]), _quat(11)[None])
    K = bo.k_matrix(off, 1)
    M = oracle.mobility_matrix(pos, 3, wall=True)
    N = bo.reference_mobility_by_solves(M, K)
    Nc = bo.clamped_mobility(N)
    assert np.allclose(Nc, N[:3, :3] - N[:3, 3:] @ np.linalg.solve(N[3:, 3:], N[3:, :3]),
                       rtol=0, atol=1e-12 * np.abs(Nc).max())
    got = np.zeros((3, 3))
    for k in range(3):
        rhs_v = np.zeros(3)
        rhs_v[k] = -1.0                      # -K^T F = -e_k: unit force on the body
        F, U = bo.prescribed_rotation_solve(M, K, np.zeros(3 * len(pos)), rhs_v)
        got[:, k] = U[0, :3]
        assert np.all(U[0, 3:] == 0.0)
    assert np.allclose(got, Nc, rtol=0, atol=1e-10 * np.abs(Nc).max())
    assert np.abs(Nc - N[:3, :3]).max() > 1e-4 * np.abs(Nc).max()     # 3e-4 at this height


def test_prescribed_rotation_equals_free_bodies_held_by_their_torques():
    """Two bodies above the wall with a blob right-hand side: the reduced solve's blob
    torques tau_b = -K_O^T F, given to the free system as its torque rows, reproduce the
    same F and V with Omega = 0; the identity-row product the library uses solves to the
    same answer."""
    ref, pos, off, q, K, M = _two_bodies_over_wall()
    rng = np.random.default_rng(23)
    rf, rv = rng.normal(size=3 * len(pos)), rng.normal(size=6)
    F, U = bo.prescribed_rotation_solve(M, K, rf, rv)
    tau = -K.T @ F
    rhs = np.concatenate([rf, tau])
    for b in range(2):
        rhs[3 * len(pos) + 6 * b:3 * len(pos) + 6 * b + 3] = rv[3 * b:3 * b + 3]
    z = np.linalg.solve(bo.saddle_matrix(M, K), rhs)
    n = 3 * len(pos)
    assert np.abs(z[n:].reshape(2, 6)[:, 3:]).max() < 1e-10 * np.abs(U).max()
    assert np.allclose(z[:n], F, rtol=0, atol=1e-10 * np.abs(F).max())
    assert np.allclose(z[n:].reshape(2, 6)[:, :3], U[:, :3], rtol=1e-9)
    rhs_id = np.concatenate([rf, np.zeros(12)])
    for b in range(2):
        rhs_id[n + 6 * b:n + 6 * b + 3] = rv[3 * b:3 * b + 3]
    zi = np.linalg.solve(bo.prescribed_rotation_matrix(M, K), rhs_id)
    assert np.allclose(zi[:n], F, rtol=0, atol=1e-10 * np.abs(F).max())
    assert np.allclose(zi[n:].reshape(2, 6), U, rtol=0, atol=1e-10 * np.abs(U).max())


def test_prescribed_rotation_preconditioner_inverts_the_reduced_scalar_saddle():
    """With M replaced by m I and a body mobility whose clamped block is m / nb I, the
    reduced block preconditioner is the exact inverse of the identity-row matrix."""
    mu, a = 0.7, 1.3
    m = 1.0 / (6.0 * math.pi * mu * a)
    ref = geo.fibonacci_sphere(40, 4.0)
    x = np.array([[0.0, 0.0, 10.0], [12.0, 0.0, 10.0]])
    q = np.stack([_quat(7), _quat(8)])
    pos, off = bo.place(ref, x, q)
    K = bo.k_matrix(off, 2)
    A = bo.prescribed_rotation_matrix(m * np.eye(3 * len(pos)), K)
    rng = np.random.default_rng(9)
    G = rng.normal(size=(3, 3))
    N = np.zeros((6, 6))
    N[:3, :3] = m / len(ref) * np.eye(3)
    N[3:, 3:] = G @ G.T + np.eye(3)               # the rotation block does not enter
    v = rng.normal(size=A.shape[0])
    got = bo.precondition_prescribed_rotation(v, off, q, N, mu=mu, a=a)
    assert np.allclose(got, np.linalg.solve(A, v), rtol=0, atol=1e-10 * np.abs(got).max())
