"""CPU pins of the prescribed-cilia swimmer oracle (oracle/cilia_swimmer_oracle.py) and
of the frame helper the library offers (cufcm_fil_amd.cilia.cilium_frames): frames
against their defining axes, the block preconditioner against the exact inverse of the
saddle matrix with M replaced by (6 pi mu a)^-1 I, make_precon_mat against the Schur
complement of that matrix, and the solve against rigid-body identities."""
import math

import numpy as np
import pytest

import cilia_oracle as co
import cilia_swimmer_oracle as cso
from cufcm_fil_amd import geometry as geo
from cufcm_fil_amd.cilia import cilium_frames
from filament_oracle import lie_exp, qmul, rot_mat

R_BODY, NB, NSEG = 6.0, 100, 6


def swimmer(n_body=1, per=3, seed=0, dt=0.01, phase_dot=None, body_x=None, body_q=None):
    rng = np.random.default_rng(seed)
    blob_ref = geo.fibonacci_sphere(NB, R_BODY).reshape(-1, 3)
    dirs = rng.normal(size=(per, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    fil_ref = (R_BODY + 1.3) * dirs
    fil_q = np.array([cso.frame_quaternion(d, 0.1 * k) for k, d in enumerate(dirs)])
    bx = np.array([[40.0 * b, 0.0, 0.0] for b in range(n_body)]) if body_x is None else body_x
    bq = np.tile([1.0, 0.0, 0.0, 0.0], (n_body, 1)) if body_q is None else body_q
    nc = n_body * per
    pdot = np.full(nc, 2 * math.pi) if phase_dot is None else phase_dot
    return cso.Swimmers(blob_ref, fil_ref, fil_q, bx, bq, rng.uniform(0, 2 * math.pi, nc), pdot,
                        nseg=NSEG, dt=dt)


def test_frames_follow_their_defining_axes():
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(12, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    tilts = rng.uniform(-0.5, 0.5, 12)
    mine = cilium_frames(dirs, 0.0).reshape(-1, 4)
    for d, t, q in zip(dirs, tilts, mine):
        qr = cso.frame_quaternion(d, t)
        R = rot_mat(qr)
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-12) and np.linalg.det(R) > 0
        assert np.allclose(R[:, 0], d, atol=1e-12)
        p = np.array([0.0, 0.0, -1.0]) + d[2] * d
        p /= np.linalg.norm(p)
        assert np.allclose(R[:, 1], math.cos(t) * p + math.sin(t) * np.cross(d, p), atol=1e-12)
        # the library's helper (built from the axes) is the same rotation at tilt 0
        q0 = cso.frame_quaternion(d, 0.0)
        assert min(np.abs(q - q0).max(), np.abs(q + q0).max()) < 1e-12
        qt = cilium_frames([d], t)
        assert min(np.abs(qt - qr).max(), np.abs(qt + qr).max()) < 1e-12
    # the wall cilium's frame is this construction at d = e_z
    assert np.allclose(rot_mat(cso.frame_quaternion([0.0, 0.0, 1.0])), co.FRAME, atol=1e-15)
    assert np.allclose(rot_mat(cilium_frames([[0.0, 0.0, 1.0]])), co.FRAME, atol=1e-15)


@pytest.mark.parametrize("mode", ["prescribed", "phase", "phase_angle"])
def test_block_preconditioner_is_the_exact_inverse_of_the_scalar_saddle(mode):
    sw = swimmer(n_body=2, per=2)
    dynamic, dyn_angle = mode != "prescribed", mode == "phase_angle"
    xs, xb, vp, va = sw.geometry()
    n3 = 3 * (xs.shape[0] + xb.shape[0])
    m = 6 * math.pi
    A = sw.system_matrix(np.eye(n3) / m, xs, xb, vp, va, dynamic, dyn_angle)
    rng = np.random.default_rng(1)
    for _ in range(3):
        r = rng.normal(size=A.shape[0])
        assert np.allclose(sw.precondition(r, xs, xb, vp, va, dynamic, dyn_angle),
                           np.linalg.solve(A, r), rtol=1e-9, atol=1e-12)
    # make_precon_mat: the inverse Schur complement on the body unknowns
    kinv = sw.precon_mat(xs, xb, vp, va, dynamic, dyn_angle)
    ns, n3_, nu, L = sw.layout(dynamic, dyn_angle)
    Kall = -A[:n3, n3:]                      # [K_B, k_phi, k_theta] columns
    S = Kall.T @ (m * Kall)
    idx_b = np.arange(0, nu - n3)
    idx_c = np.arange(nu - n3, L - n3)
    Sbb = S[np.ix_(idx_b, idx_b)]
    if dynamic:
        Sbc, Scc = S[np.ix_(idx_b, idx_c)], S[np.ix_(idx_c, idx_c)]
        Sbb = Sbb - Sbc @ np.linalg.solve(Scc, Sbc.T)
    for b in range(sw.nbody):
        blk = slice(6 * b, 6 * b + 6)
        assert np.allclose(kinv[b], np.linalg.inv(Sbb[blk, blk]), rtol=1e-9, atol=1e-14)
    # a planted mistake (torque arm sign) breaks the identity
    bad = sw.precondition(r, xs, -xb, vp, va, dynamic, dyn_angle)
    assert not np.allclose(bad, np.linalg.solve(A, r), rtol=1e-3)


def test_solve_is_force_and_torque_free_and_rigid():
    sw = swimmer(per=3)
    z, qp, qa = sw.step(0)
    ns, n3, nu, L = sw.layout(False, False)
    xs, xb = sw.last["xs"], sw.last["xb"]
    F = z[:n3].reshape(-1, 3)
    X = np.vstack([xs, xb])
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()
    assert np.abs(np.cross(X - sw.x[0], F).sum(0)).max() < 1e-9 * np.abs(F).max() * R_BODY
    # blob velocities from the mobility equal the rigid motion V + Omega x r
    M = sw.mobility(xs, xb)
    u = (M @ z[:n3]).reshape(-1, 3)[ns:]
    V, W = sw.V[0, :3], sw.V[0, 3:]
    assert np.allclose(u, V + np.cross(W, xb - sw.x[0]), rtol=1e-9, atol=1e-12 * np.abs(u).max())
    assert np.abs(sw.V).max() > 1e-3        # the beat moves the body


def test_translated_and_rotated_swimmers_move_alike():
    s0 = swimmer(seed=4)
    z0, q0, _ = s0.step(0)
    shift = np.array([[13.0, -7.0, 5.0]])
    s1 = swimmer(seed=4, body_x=shift)
    z1, q1, _ = s1.step(0)
    assert np.allclose(s1.V, s0.V, rtol=1e-9, atol=1e-13) and np.allclose(q1, q0, rtol=1e-9)
    qb = np.array([math.cos(0.4), *(math.sin(0.4) * np.array([1.0, 2.0, -2.0]) / 3.0)])
    R0 = rot_mat(qb)
    s2 = swimmer(seed=4, body_q=qb[None, :])
    s2.step(0)
    assert np.allclose(s2.V[0, :3], R0 @ s0.V[0, :3], rtol=1e-8, atol=1e-12)
    assert np.allclose(s2.V[0, 3:], R0 @ s0.V[0, 3:], rtol=1e-8, atol=1e-12)


def test_explicit_update_moves_blobs_by_the_rigid_velocity():
    sw = swimmer(seed=5, dt=1e-4)
    sw.step(0)
    xb0 = sw.last["xb"].copy()
    V, W = sw.V[0, :3].copy(), sw.V[0, 3:].copy()
    x0 = sw.x[0].copy()
    sw.step(1)
    xb1 = sw.last["xb"]
    pred = xb0 + sw.dt * (V + np.cross(W, xb0 - x0))
    assert np.abs(xb1 - pred).max() < 1e-3 * sw.dt * np.abs(V + np.cross(W, xb0 - x0)).max()
    assert np.allclose(sw.q[0], qmul(lie_exp(sw.dt * W), np.array([1.0, 0.0, 0.0, 0.0])), atol=1e-15)


def test_dynamic_phases_balance_the_driving_force():
    """With driving, each cilium's generalised force equals the interpolated drive."""
    sw = swimmer(n_body=1, per=3, seed=6)
    tab = 100.0 + 20.0 * np.sin(2 * math.pi * np.arange(300) / 300)
    sw.driving = (tab, None, 0.0)
    z, qp, qa = sw.step(0)
    drive = [co.driving_forces(sw.phase[c], 0.0, sw.omega0[c], tab, None)[0] for c in range(sw.nc)]
    assert np.allclose(qp, drive, rtol=1e-9)
    assert np.all(sw.phase_dot > 0)


def test_driving_options_fene_spring_and_seeded_noise():
    # SplitMix64: the first output of the generator seeded with 0
    assert co.mix64(0) == 0xE220A8397B1DCDAF
    # FENE factor: linear up to 0.78, continuous there, odd
    for a in (0.1, -0.5, 0.78):
        assert co.fene_factor(a) == a
    for a in (0.78 + 1e-9, 1.0, 1.5):
        assert abs(co.fene_factor(a) + co.fene_factor(-a)) < 1e-15
    assert abs(co.fene_factor(0.78 + 1e-9) - 0.78) < 1e-8
    assert co.fene_factor(1.0) == 0.5 * (1 + 1 / 0.78 ** 2 * math.exp(0.22))
    # deviates: reproducible, distinct per (seed, step, cilium), standard normal
    assert co.normal_deviate(7, 3, 5) == co.normal_deviate(7, 3, 5)
    assert len({co.normal_deviate(7, 3, c) for c in range(50)}) == 50
    assert co.normal_deviate(7, 3, 5) != co.normal_deviate(8, 3, 5) != co.normal_deviate(7, 4, 5)
    z = np.array([co.normal_deviate(11, nt, c) for nt in range(200) for c in range(250)])
    assert abs(z.mean()) < 0.015 and abs(z.var() - 1.0) < 0.02
    assert abs(np.mean(z ** 4) - 3.0) < 0.08 and abs(np.mean(np.abs(z) < 1.0) - 0.6827) < 0.01


def test_prescribed_rotation_is_the_free_swimmer_held_by_its_torque():
    """With the rotation prescribed (Omega_b = 0, Omega columns and torque rows dropped)
    the fluid's torque on the body, tau = -sum r x F, is not balanced.  Applying -tau to
    the free swimmer as the right-hand side of its torque rows must reproduce the same
    forces and V with Omega = 0 -- a dropped column, a wrong sign or a torque row left
    in the reduced system would break the equality."""
    sw = swimmer(per=3, seed=5)
    sw.rot_prescribed = True
    z, qp, qa = sw.step(0)
    ns, n3, nu, L = sw.layout(False, False)
    assert np.abs(sw.V[:, 3:]).max() == 0.0
    xs, xb, vp, va = sw.last["xs"], sw.last["xb"], sw.last["vp"], sw.last["va"]
    F = z[:n3].reshape(-1, 3)
    tau = np.cross(np.vstack([xs, xb]) - sw.x[0], F).sum(0)
    assert np.linalg.norm(tau) > 1e-3 * np.abs(F).max() * R_BODY   # the hold takes a torque
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()           # still force-free
    free = swimmer(per=3, seed=5)
    A = free.system_matrix(free.mobility(xs, xb), xs, xb, vp, va, False, False)
    rhs = free.rhs(vp, False, False)
    rhs[n3 + 3:n3 + 6] = -tau                                         # -K^T F rows
    zf = np.linalg.solve(A, rhs)
    assert np.abs(zf[n3 + 3:n3 + 6]).max() < 1e-10 * np.abs(zf[n3:n3 + 3]).max()
    assert np.allclose(zf[:n3], z[:n3], rtol=0, atol=1e-9 * np.abs(F).max())
    assert np.allclose(zf[n3:n3 + 3], z[n3:n3 + 3], rtol=1e-9)
