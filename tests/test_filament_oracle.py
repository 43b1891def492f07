"""CPU pins of oracle/filament_oracle.py (the FP64 restatement of the reference's
elastic-filament mechanics used to check the GPU Broyden step)."""
import math

import numpy as np
import pytest
from scipy.linalg import expm

import filament_oracle as fo


def skew(u):
    return np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])


@pytest.mark.parametrize("scale", [1e-7, 1e-3, 0.5, 2.0])
def test_lie_maps(scale):
    rng = np.random.default_rng(1)
    u = scale * rng.standard_normal(3)
    v = rng.standard_normal(3)
    # dexp(u, dexpinv(u, v)) = v  (identity stated at quaternion.cpp:561-563)
    assert np.allclose(fo.dexp(u, fo.dexpinv(u, v)), v, rtol=1e-12, atol=1e-12)
    # exponential map = Rodrigues / matrix exponential of the skew matrix (the
    # reference returns the identity below |u| = 1e-6, quaternion.cpp:503-509)
    tol = 1e-12 if np.linalg.norm(u) > 1e-6 else 2e-6
    assert np.allclose(fo.rot_mat(fo.lie_exp(u)), expm(skew(u)), atol=tol)
    # rcross(v) w = w x v
    w = rng.standard_normal(3)
    assert np.allclose(fo.rcross(v) @ w, np.cross(w, v))


def test_midpoint_is_geodesic_half():
    rng = np.random.default_rng(2)
    q1 = fo.lie_exp(rng.standard_normal(3))
    q2 = fo.qmul(fo.lie_exp(0.4 * rng.standard_normal(3)), q1)
    m = fo.midpoint(q1, q2)
    half = fo.qmul(m, fo.qconj(q1))
    assert np.allclose(fo.qmul(half, half), fo.qmul(q2, fo.qconj(q1)), atol=1e-14)


def _bent(P, nt=3, seed=0):
    rng = np.random.default_rng(seed)
    n = P.nseg
    t0 = np.array([0.3, 0.1, 1.0])
    q0 = fo.quat_from_tangent(t0)
    assert np.allclose(fo.tangent(q0), t0 / np.linalg.norm(t0))
    qs = np.array([fo.qmul(fo.lie_exp(np.array([0.05 * k, 0.02 * k, 0.01])), q0) for k in range(n)])
    x = np.zeros((n, 3))
    x[0] = [0.0, 0.0, 3.0]
    f = fo.Filament(P, x, qs)
    f.robot_arm()
    f.lam = 2.0 * rng.standard_normal((n, 3))
    f.clamp = rng.standard_normal(3)
    f.xm1 = f.x - 0.01
    f.xm2 = f.x - 0.02
    f.um1 = 0.01 * rng.standard_normal((n, 3))
    return f


def test_internal_loads_balance():
    P = fo.Params(nseg=10)
    fil = _bent(P)
    F = fil.internal_forces(5)
    n = P.nseg
    tip_t = fo.tangent(fil.q[-1])
    assert np.allclose(F[:, :3].sum(0), fil.lam[0] - P.end_force * tip_t, atol=1e-10)
    tq = (np.cross(fil.x, F[:, :3]) + F[:, 3:]).sum(0)
    expect = (np.cross(fil.x[0], fil.lam[0]) + np.cross(fil.x[-1], -P.end_force * tip_t)
              + fo.dexpinv_transpose(fil.u[0], fil.clamp))
    assert np.allclose(tq, expect, atol=1e-9)
    # a straight, unstressed, unloaded filament carries only the end force
    s = fo.Filament(P, np.zeros((n, 3)), np.tile(fo.quat_from_tangent([0, 0, 1.0]), (n, 1)))
    s.robot_arm()
    G = s.internal_forces(5)
    G[-1, :3] += P.end_force * fo.tangent(s.q[-1])
    assert np.abs(G).max() < 1e-10


def test_approx_jacobian_matches_free_draining_fd_jacobian():
    """In the free-draining limit (v = F/6 pi mu a, w = T/8 pi mu a^3) the only
    approximation left in J_approx is the elastic-moment derivative."""
    P = fo.Params(nseg=8, dt=0.05)
    fil = _bent(P)
    n, nt = P.nseg, 3

    def resid(f0):
        F = f0.internal_forces(nt, with_perturbation=False)
        v = np.zeros((n, 6))
        v[:, :3] = F[:, :3] / (6 * math.pi)
        v[:, 3:] = F[:, 3:] / (8 * math.pi)
        return f0.errors(v, nt)

    J = np.zeros((6 * n, 6 * n))
    h = 1e-6
    for c in range(6 * n):
        a, b = fil.copy(), fil.copy()
        d = np.zeros(6 * n)
        d[c] = h
        a.update(d)
        b.update(-d)
        J[:, c] = (resid(a) - resid(b)) / (2 * h)
    Ja = fil.approx_jacobian(nt, fil.internal_forces(nt, with_perturbation=False))

    def rel(r, c):
        return np.linalg.norm(Ja[r, c] - J[r, c]) / np.linalg.norm(J[r, c])

    P3, L3 = slice(0, 3 * n), slice(3 * n, 6 * n)
    assert rel(P3, P3) < 1e-6 and rel(P3, L3) < 1e-6 and rel(L3, P3) < 1e-6
    assert rel(L3, L3) < 1e-2
    # a dropped factor in any block is visible
    Jb = Ja.copy()
    Jb[3 * n:, 3 * n:] *= 0.9
    assert np.linalg.norm(Jb[L3, L3] - J[L3, L3]) / np.linalg.norm(J[L3, L3]) > 5e-2


def test_broyden_converges_on_free_draining_filament():
    P = fo.Params(nseg=10, dt=0.05)
    fil = _bent(P, seed=3)
    fil.lam[:] = 0.0
    fil.clamp[:] = 0.0
    n, nt = P.nseg, 0
    fil.xm1 = fil.x.copy()
    fil.xm2 = fil.x.copy()
    fil.um1[:] = 0.0

    def resid(f0):
        F = f0.internal_forces(nt)
        v = np.zeros((n, 6))
        v[:, :3] = F[:, :3] / (6 * math.pi)
        v[:, 3:] = F[:, 3:] / (8 * math.pi)
        return f0.errors(v, nt)

    Jinv = np.linalg.inv(fil.approx_jacobian(nt, fil.internal_forces(nt)))
    B = fo.Broyden()
    e = resid(fil)
    for it in range(100):
        fil.update(B.find_update(lambda r: Jinv @ r, e))
        e_new = resid(fil)
        big = np.abs(e_new[:3 * n]).max() > P.dl * 1e-10 or np.abs(e_new[3 * n:]).max() > 1e-10
        e = B.end_of_iter(lambda r: Jinv @ r, e, e_new, big)
        if not big:
            break
    assert not big and it < 60
