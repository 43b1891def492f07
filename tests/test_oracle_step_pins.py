"""CPU pins of two oracle functions of the implicit step that the GPU parity tests
compare against: the BDF2 / backward-Euler residual (filament_oracle.Filament.errors,
mobility_solver.cpp:2760-2853) and the barrier forces (filament_oracle.barrier_forces,
cuda_functions.cu:1825-1930).

Neither check retypes the function under test.  The residual must vanish on motion the
scheme integrates exactly: quadratic positions, and rotations whose Lie coordinates are
quadratic in time, with quaternions and angular velocity from scipy's rotations and
finite differences rather than from the oracle's Lie maps.  The barrier must reproduce
hand-computed values at chosen separations and heights, respect its cut-offs and
exclusions, and balance pairwise.  Each group ends with a planted error in the oracle
module that its check detects."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import filament_oracle as fo


def _straight(P, z0=3.0):
    n = P.nseg
    q = np.repeat(fo.quat_from_tangent(np.array([0.2, -0.1, 1.0]))[None], n, axis=0)
    x = np.zeros((n, 3))
    x[:, 2] = z0 + P.dl * np.arange(n)
    return fo.Filament(P, x, q)


def _rows(f, v, nt):
    n = f.P.nseg
    e = f.errors(v, nt)
    return e[:3 * n].reshape(n, 3), e[3 * n:].reshape(n, 3)


# ---------------------------------------------------------------- BDF2 position rows
def _quadratic_motion(P, nt, seed=0, linear=False):
    rng = np.random.default_rng(seed)
    n, dt = P.nseg, P.dt
    a, b, c = rng.standard_normal((3, n, 3))
    if linear:
        c[:] = 0.0
    t = 1.3

    def X(s):
        return a + b * s + c * s * s

    f = _straight(P)
    f.x, f.xm1, f.xm2 = X(t), X(t - dt), X(t - 2 * dt)
    v = np.zeros((n, 6))
    v[:, :3] = b + 2.0 * c * t
    return f, v


def test_bdf2_position_rows_vanish_on_quadratic_motion():
    P = fo.Params(nseg=6, dt=0.07)
    f, v = _quadratic_motion(P, nt=5)
    ex, _ = _rows(f, v, nt=5)
    assert np.abs(ex).max() < 1e-14 * max(1.0, np.abs(f.x).max())
    # backward Euler (steps before num_euler_steps) is exact on linear motion only
    f, v = _quadratic_motion(P, nt=0, linear=True)
    ex, _ = _rows(f, v, nt=0)
    assert np.abs(ex).max() < 1e-14 * max(1.0, np.abs(f.x).max())
    f, v = _quadratic_motion(P, nt=0)
    ex, _ = _rows(f, v, nt=0)
    assert np.abs(ex).max() > 1e-4


@pytest.mark.parametrize("coef,value", [("BDF2_C1", -1.3), ("BDF2_C2", 0.3), ("BDF2_C3", -0.6)])
def test_planted_bdf2_coefficient_error_is_detected(monkeypatch, coef, value):
    P = fo.Params(nseg=6, dt=0.07)
    monkeypatch.setattr(fo, coef, value)
    f, v = _quadratic_motion(P, nt=5)
    ex, eu = _rows(f, v, nt=5)
    if coef == "BDF2_C1":
        assert np.abs(ex).max() > 1e-3
    else:
        # C2 and C3 also weight the Lie-algebra rows: a steady rotation exposes them
        f, v = _steady_rotation(P)
        _, eu = _rows(f, v, nt=5)
        assert max(np.abs(ex).max(), np.abs(eu).max()) > 1e-4


# ---------------------------------------------------------------- BDF2 Lie-algebra rows
def _steady_rotation(P, seed=1):
    """Every segment turns at a constant angular velocity omega about a fixed axis:
    q_n = exp(dt omega) q_{n-1}, so u_n = u_{n-1} = dt omega (segment::update)."""
    rng = np.random.default_rng(seed)
    n, dt = P.nseg, P.dt
    f = _straight(P)
    om = rng.standard_normal((n, 3))
    f.u = dt * om
    f.um1 = dt * om
    v = np.zeros((n, 6))
    v[:, 3:] = om
    return f, v


@pytest.mark.parametrize("nt", [0, 4])
def test_lie_rows_vanish_on_steady_rotation(nt):
    P = fo.Params(nseg=5, dt=0.05)
    f, v = _steady_rotation(P)
    _, eu = _rows(f, v, nt)
    assert np.abs(eu).max() < 1e-15


def _quat(rotvec):
    x, y, z, w = Rotation.from_rotvec(rotvec).as_quat()
    return np.array([w, x, y, z])


def _qmul(p, q):  # Hamilton product, (w, x, y, z)
    return np.concatenate([[p[0] * q[0] - p[1:] @ q[1:]],
                           p[0] * q[1:] + q[0] * p[1:] + np.cross(p[1:], q[1:])])


def _log(q):
    """Rotation vector of a unit quaternion (w, x, y, z)."""
    x, y, z, w = q[1], q[2], q[3], q[0]
    return Rotation.from_quat([x, y, z, w]).as_rotvec()


def _lie_residual(dt, nt=5, seed=3, quadratic=True):
    """Lie-algebra residual of one segment on the rotation q(s) = exp(sigma(s)) q_{n-1},
    sigma(s) = A s + B s^2 (A, B not parallel; s = t - t_{n-1}): u_n = log(q_n q_{n-1}^-1)
    = sigma(dt), u_{n-1} = log(q_{n-1} q_{n-2}^-1) = -sigma(-dt), and omega(t_n) the
    spatial angular velocity 2 vec(dq/dt q^-1) by a fourth-order central difference of
    scipy rotations.  sigma obeys d sigma/dt = dexpinv(sigma, omega), and BDF2 is exact
    for quadratic sigma; backward Euler (nt = 0) is exact for linear sigma."""
    rng = np.random.default_rng(seed)
    A, B = rng.standard_normal((2, 3))
    if not quadratic:
        B[:] = 0.0
    qm1 = _quat(rng.standard_normal(3))

    def q(s):
        return _qmul(_quat(A * s + B * s * s), qm1)

    def conj(p):
        return np.array([p[0], -p[1], -p[2], -p[3]])

    h = 1e-4
    qd = (-q(dt + 2 * h) + 8 * q(dt + h) - 8 * q(dt - h) + q(dt - 2 * h)) / (12 * h)
    om = 2.0 * _qmul(qd, conj(q(dt)))[1:]
    un = _log(_qmul(q(dt), conj(qm1)))
    um1 = _log(_qmul(qm1, conj(q(-dt))))
    assert np.allclose(un, A * dt + B * dt * dt, atol=1e-12)
    P = fo.Params(nseg=2, dt=dt)
    f = _straight(P)
    f.u[:] = un
    f.um1[:] = um1
    v = np.zeros((2, 6))
    v[:, 3:] = om
    _, eu = _rows(f, v, nt=nt)
    return np.abs(eu[0]).max(), np.linalg.norm(un)


@pytest.mark.parametrize("dt", [0.05, 0.3])
def test_lie_rows_vanish_on_exactly_integrated_rotation(dt):
    """BDF2 rows vanish on quadratic Lie coordinates and Euler rows on linear ones, for
    rotations of sizeable angle (|u| up to ~0.5 rad) whose increment is not parallel to
    the angular velocity, so every term of dexpinv takes part."""
    e, th = _lie_residual(dt)
    assert e < 1e-9, (e, th)
    e, th = _lie_residual(dt, nt=0, quadratic=False)
    assert e < 1e-9, (e, th)
    # the Euler row is first order only: not exact on the quadratic rotation
    e, th = _lie_residual(dt, nt=0)
    assert e > 1e-4 * dt


def test_planted_dexpinv_error_is_detected(monkeypatch):
    """Without the dexpinv map (W = omega), or with the opposite sign of its first-order
    term, the exactly integrated rotation leaves a residual."""
    true_dexpinv = fo.dexpinv
    for plant in (lambda u, w: w, lambda u, w: true_dexpinv(-u, w)):
        monkeypatch.setattr(fo, "dexpinv", plant)
        e, th = _lie_residual(0.3)
        assert e > 1e-3, e


# ---------------------------------------------------------------- barrier forces
def _fmag(P):
    return P.repulsive_factor * P.end_force


def _two_filaments(d, P):
    """Filament 0 vertical from (0, 0, 50), filament 1 horizontal from (d, 0, 50): only
    their first segments are within range of each other (d <= 2.2)."""
    n = P.nseg
    f0 = np.array([[0.0, 0.0, 50.0 + P.dl * k] for k in range(n)])
    f1 = np.array([[d + P.dl * k, 0.0, 50.0] for k in range(n)])
    return np.vstack([f0, f1])


@pytest.mark.parametrize("d,frac", [(2.1, 1.0 / 16.0), (2.05, 81.0 / 256.0), (2.19, 0.05 ** 4),
                                    (2.0, 1.0), (1.7, 1.0), (2.2 + 1e-9, 0.0), (3.0, 0.0)])
def test_barrier_between_two_segments(d, frac):
    P = fo.Params(nseg=4)
    x = _two_filaments(d, P)
    F = fo.barrier_forces(x, 4, P, wall=False)
    fm = _fmag(P)
    # repulsion along the separation, equal and opposite, nothing on other segments
    assert np.allclose(F[0], [-fm * frac, 0.0, 0.0], rtol=1e-12, atol=1e-12 * fm)
    assert np.allclose(F[4], -F[0], rtol=1e-14, atol=0.0)
    assert not np.delete(F, [0, 4], axis=0).any()


def test_barrier_skips_chain_neighbours_only():
    """A folded 3-segment filament (spacing 2.0 < 2.2): segments 0 and 2 (2.0 apart)
    repel at full strength, the linked pairs (0, 1) and (1, 2) are skipped."""
    P = fo.Params(nseg=3, dl=2.0)
    x = np.array([[0.0, 0.0, 50.0], [2.0, 0.0, 50.0], [1.0, math.sqrt(3.0), 50.0]])
    F = fo.barrier_forces(x, 3, P, wall=False)
    fm = _fmag(P)
    dirn = (x[0] - x[2]) / 2.0
    assert np.allclose(F[0], fm * dirn, rtol=1e-12)
    assert np.allclose(F[2], -fm * dirn, rtol=1e-12)
    assert np.abs(F[1]).max() == 0.0


@pytest.mark.parametrize("z,frac", [(1.0, 1.0), (0.6, 1.0), (1.05, 1.0 / 16.0), (1.075, 1.0 / 256.0),
                                    (1.1, 0.0), (3.0, 0.0)])
def test_wall_barrier_profile(z, frac):
    """Below BASE_HEIGHT_ABOVE_SURFACE = DL/2 = 1.1 the wall pushes up with
    fmag min(1, 1 - (z - a)/(DL/2 - a))^4: fmag at z = a and below."""
    P = fo.Params(nseg=2)
    x = np.array([[0.0, 0.0, z], [0.0, 0.0, z + 2.2], [40.0, 0.0, 30.0], [40.0, 0.0, 32.2]])
    F = fo.barrier_forces(x, 2, P, wall=True)
    assert np.allclose(F[0], [0.0, 0.0, _fmag(P) * frac], rtol=1e-12, atol=1e-12 * _fmag(P))


def test_blob_repels_a_segment_like_a_segment():
    P = fo.Params(nseg=4)
    xs = _two_filaments(2.1, P)[:4]
    xb = np.array([[2.1, 0.0, 50.0], [60.0, 0.0, 50.0]])
    F = fo.barrier_forces(xs, 4, P, x_blob=xb, wall=False)
    assert np.allclose(F[0], [-_fmag(P) / 16.0, 0.0, 0.0], rtol=1e-12)
    assert not F[1:].any()


def test_barrier_balances_pairwise_in_a_crowded_cluster():
    """Segment-segment repulsion is pairwise equal and opposite: it sums to zero over a
    crowded cluster (no wall, no blobs), and each force points away from the neighbours
    that cause it."""
    P = fo.Params(nseg=5)
    rng = np.random.default_rng(7)
    x = rng.uniform(0.0, 6.0, (15, 3)) + [0.0, 0.0, 20.0]
    F = fo.barrier_forces(x, 5, P, wall=False)
    assert np.abs(F).max() > 0.1 * _fmag(P)
    assert np.abs(F.sum(0)).max() < 1e-12 * np.abs(F).max()


@pytest.mark.parametrize("plant", ["power", "range", "sign"])
def test_planted_barrier_error_is_detected(monkeypatch, plant):
    P = fo.Params(nseg=4)
    x = _two_filaments(2.1, P)
    fn = fo.barrier_forces
    if plant == "power":
        monkeypatch.setattr(fo, "BARRIER_POWER", 3)
    elif plant == "range":
        monkeypatch.setattr(fo, "BARRIER_RANGE", 1.2)
    else:
        def fn(*a, **k):
            return -fo.barrier_forces(*a, **k)
    F = fn(x, 4, P, wall=False)
    ok = np.allclose(F[0], [-_fmag(P) / 16.0, 0.0, 0.0], rtol=1e-12)
    assert not ok
