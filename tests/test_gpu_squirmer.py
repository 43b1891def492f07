"""GPU parity (through the C ABI) of squirmer bodies: the body velocity of a blob sphere
with slip modes B_1, B_2 against the oracle's dense solve of the same force- and
torque-free system (unbounded and above the wall, every preconditioner and arithmetic),
its rotation covariance, and the explicit body update between steps."""
import math

import numpy as np
import pytest

import body_oracle as bo
import oracle
from cufcm_fil_amd import RpyMobilitySolver
from cufcm_fil_amd import geometry as geo

pytestmark = pytest.mark.gpu

# 400 blobs on the radius-8 sphere overlap (closest pair 1.24a); the wall-corrected blob
# matrix of that seeding has a negative eigenvalue (-0.012), so above the wall the bodies
# carry 250 blobs (closest pair 1.56a, smallest eigenvalue +0.005)
R, NB, NB_WALL = 8.0, 400, 250
B_MODES = [1.0, 0.5]


def _quat(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([[math.cos(angle / 2)], math.sin(angle / 2) * a])


def _oracle_velocity(ref, x, q, wall):
    pos, off = bo.place(ref, np.asarray(x)[None], np.asarray(q)[None])
    M = oracle.mobility_matrix(pos.reshape(-1), 3, wall=wall)
    K = bo.k_matrix(off, 1)
    Rq = bo.rot_mat(q)
    slip = bo.squirmer_slip(ref, [0.0, 0.0], B_MODES) @ Rq.T
    U, _ = bo.squirmer_velocity(M, K, slip)
    return U[0]


@pytest.mark.parametrize("wall,precond,mode", [(False, "dense", "fp64"), (False, "diag", "fp64"),
                                               (False, "dense", "mixed"), (True, "dense", "fp64"),
                                               (True, "self", "mixed")])
def test_squirmer_velocity_matches_oracle(wall, precond, mode):
    nb = NB_WALL if wall else NB
    ref = geo.fibonacci_sphere(nb, R)
    x = np.array([0.0, 0.0, 30.0])
    q = _quat([0.3, -0.5, 0.8], 0.9)
    with RpyMobilitySolver(0, nb, wall=wall, precision="fp32c" if mode == "mixed" else "fp64") as m:
        m.body_init(ref, 1)
        m.body_set_state(x, q)
        m.squirmer_init(B_MODES, dt=0.01, precond=precond, mode=mode, rtol=1e-11)
        vel, st = m.squirmer_step(0)
    assert st["converged"], st
    ref_u = _oracle_velocity(ref, x, q, wall)
    assert np.abs(vel[0] - ref_u).max() <= 1e-8 * np.abs(ref_u).max(), (vel[0], ref_u)
    # the body swims along its own axis R(q) e_z; free space: no rotation
    axis = bo.rot_mat(q)[:, 2]
    if not wall:
        assert np.abs(vel[0, :3] - (vel[0, :3] @ axis) * axis).max() < 1e-3 * abs(vel[0, :3] @ axis)
        assert vel[0, :3] @ axis > 2.0 / 3.0


def test_squirmer_steps_advance_the_body_explicitly():
    """Above the wall the velocity depends on the body's height and orientation, so the
    second step's velocity is that of a fresh solve at x_1 = x_0 + dt V_0,
    q_1 = exp(dt Omega_0) q_0 (the quaternion update computed here with scipy's rotations)."""
    from scipy.spatial.transform import Rotation
    ref = geo.fibonacci_sphere(NB_WALL, R)
    x0 = np.array([1.0, -2.0, R + 5.0])
    q0 = _quat([1.0, 1.0, 0.5], 0.8)
    dt = 0.5
    with RpyMobilitySolver(0, NB_WALL, wall=True) as m:
        m.body_init(ref, 1)
        m.body_set_state(x0, q0)
        m.squirmer_init(B_MODES, dt=dt, precond="dense", rtol=1e-11)
        v0, _ = m.squirmer_step(0)
        v1, _ = m.squirmer_step(1)
    x1 = x0 + dt * v0[0, :3]
    w, a, b, c = q0
    r1 = Rotation.from_rotvec(dt * v0[0, 3:]) * Rotation.from_quat([a, b, c, w])
    a, b, c, w = r1.as_quat()
    q1 = np.array([w, a, b, c])
    assert np.abs(v1 - v0).max() > 1e-3 * np.abs(v0).max()      # the wall makes it move on
    ref_v1 = _oracle_velocity(ref, x1, q1, True)
    assert np.abs(v1[0] - ref_v1).max() <= 1e-8 * np.abs(ref_v1).max(), (v1[0], ref_v1)
