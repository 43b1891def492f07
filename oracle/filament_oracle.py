"""FP64 numpy restatement of the elastic-filament pieces of the implicit step --
TEST INFRASTRUCTURE ONLY (imported by tests/ and never by the product).

Follows, for tethered filaments with a clamped base and a follower end force
(the reference's INSTABILITY_CILIA model, config.hpp:81-87,459-464):
  * quaternion algebra             src/general/quaternion.cpp:79-470
  * lie_exp / midpoint / dexpinv   src/general/quaternion.cpp:500-636
  * rcross                         src/general/matrix.cpp:665-705
  * segment::initial_guess/update  src/cilia/segment.cpp:38-90
  * filament::robot_arm            src/cilia/filament.cpp:629-651
  * filament::initial_guess        src/cilia/filament.cpp:1336-1663 (elastic branch)
  * filament::body_frame_moment    src/cilia/filament.cpp:1729-1743
  * internal_forces_and_torques    src/cilia/filament.cpp:1745-1853
  * body_frame_moment_lie_derivative  src/cilia/filament.cpp:9-60
  * jacobian_lie_algebra_block     src/cilia/filament.cpp:1855-2037
  * invert_approx_jacobian         src/cilia/filament.cpp:2039-2172
  * filament::update               src/cilia/filament.cpp:2174-2208
  * mobility_solver::compute_errors  src/cilia/mobility/mobility_solver.cpp:2760-2853
  * broyden_solver find_update / end_of_iter  src/cilia/broyden_solver.cpp:71-174
  * barrier_forces                 src/cilia/cuda_functions.cu:1825-1930 (segments)

Parity status: pinned by the identities in tests/test_filament_oracle.py (dexp o
dexpinv = identity, lie_exp vs Rodrigues, finite-difference checks of the
elastic-moment derivative, J_approx against a finite-difference Jacobian of the
BDF residual in the free-draining limit, zero internal force for a straight
unstressed filament, Newton-Euler balance of the internal loads) and in
tests/test_oracle_step_pins.py (the BDF2 / Euler residual vanishes on exactly
integrated motion and is third order on a general rotation; the barrier's known
values at chosen separations and heights, its cut-offs, exclusions and pairwise
balance; planted coefficient errors in this module fail those checks).
"""
from __future__ import annotations

import math

import numpy as np

PI = math.pi
SMALL_ANGLE = 1e-5

# BDF2 residual coefficients (mobility_solver.cpp:2767-2769): x + C1 x_m1 + C2 x_m2 + C3 dt v
BDF2_C1, BDF2_C2, BDF2_C3 = -4.0 / 3.0, 1.0 / 3.0, -2.0 / 3.0
# barrier (cuda_functions.cu:1856-1903): pairs closer than BARRIER_RANGE (a_i + a_j) repel
# with magnitude REPULSIVE_FORCE_FACTOR END_FORCE min(1, 1 - chi (d - a_sum))^BARRIER_POWER,
# chi = 1 / ((BARRIER_RANGE - 1) a_sum) = 10 / a_sum; the wall pushes below BASE_HEIGHT (0.5 DL)
BARRIER_RANGE, BARRIER_POWER = 1.1, 4


# ---------------------------------------------------------------- quaternions
def qmul(p, q):
    s = p[0] * q[0] - np.dot(p[1:], q[1:])
    v = p[0] * q[1:] + q[0] * p[1:] + np.cross(p[1:], q[1:])
    return np.concatenate([[s], v])


def qconj(q):
    return np.array([q[0], -q[1], -q[2], -q[3]])


def qsqrt(q):
    q = np.array(q, dtype=float)
    if q[0] < -0.999999999:
        return np.array([0.0, 0.0, 0.0, 1.0])
    s = math.sqrt(0.5 * (1.0 + q[0]))
    return np.concatenate([[s], q[1:] / (2.0 * s)])


def rot_mat(q):
    s, a, b, c = q
    return np.array([
        [1 - 2 * (b * b + c * c), 2 * (a * b - s * c), 2 * (a * c + s * b)],
        [2 * (a * b + s * c), 1 - 2 * (a * a + c * c), 2 * (b * c - s * a)],
        [2 * (a * c - s * b), 2 * (b * c + s * a), 1 - 2 * (a * a + b * b)],
    ])


def tangent(q):
    return rot_mat(q)[:, 0]


def normal(q):
    return rot_mat(q)[:, 1]


def binormal(q):
    return rot_mat(q)[:, 2]


def lie_exp(u):
    th = math.sqrt(float(np.dot(u, u)))
    if th < 1e-6:
        return np.array([1.0, 0.0, 0.0, 0.0])
    return np.concatenate([[math.cos(0.5 * th)], math.sin(0.5 * th) / th * np.asarray(u)])


def midpoint(q1, q2):
    return qmul(qsqrt(qmul(q2, qconj(q1))), q1)


def rcross(v):
    """The reference's 'right cross' matrix: rcross(v) @ w = w x v."""
    return np.array([[0.0, v[2], -v[1]], [-v[2], 0.0, v[0]], [v[1], -v[0], 0.0]])


def dexpinv(u, v):
    th2 = float(np.dot(u, u))
    th = math.sqrt(th2)
    fac = -1.0 / 12.0 if th < SMALL_ANGLE else (0.5 * th * math.cos(0.5 * th) / math.sin(0.5 * th) - 1.0) / th2
    uxv = np.cross(u, v)
    return v - 0.5 * uxv - fac * np.cross(u, uxv)


def dexpinv_transpose(u, v):
    th2 = float(np.dot(u, u))
    th = math.sqrt(th2)
    fac = -1.0 / 12.0 if th < SMALL_ANGLE else (0.5 * th * math.cos(0.5 * th) / math.sin(0.5 * th) - 1.0) / th2
    uxv = np.cross(u, v)
    return v + 0.5 * uxv - fac * np.cross(u, uxv)


def dexp(u, v):
    th2 = float(np.dot(u, u))
    th = math.sqrt(th2)
    if th < SMALL_ANGLE:
        th4 = th2 * th2
        beta = (1.0 / 3.0 + th2 / 72.0) / (2.0 + th2 / 6.0 + th4 / 72.0)
        alpha = 1.0 / 6.0 + 2.0 * beta * (1.0 - th2 / 12.0)
    else:
        alpha = math.sin(0.5 * th)
        alpha *= 2.0 * alpha / th2
        beta = (th - math.sin(th)) / (th * th2)
    uxv = np.cross(u, v)
    return v + alpha * uxv + beta * np.cross(u, uxv)


def psi_mat(q):
    P = np.zeros((4, 3))
    P[0] = -q[1:]
    P[1:] = rcross(q[1:]) + q[0] * np.eye(3)
    return 0.5 * P


def left_mult_mat(q):
    M = np.zeros((4, 4))
    M[0, 0] = q[0]
    M[0, 1:] = -q[1:]
    M[1:, 0] = q[1:]
    M[1:, 1:] = -rcross(q[1:]) + q[0] * np.eye(3)
    return M


def quat_from_tangent(t):
    """Rotation taking the x axis to the unit vector t, as in filament::initial_setup
    (filament.cpp: qtemp(dir0, 0, -dir2, dir1).sqrt_in_place_ortho)."""
    t = np.asarray(t, float) / np.linalg.norm(t)
    q = np.array([t[0], 0.0, -t[2], t[1]])
    if q[0] < -0.999999:
        q[0] = -0.9999999
        s = math.sqrt(0.5 * (1.0 + q[0]))
        q = np.concatenate([[s], q[1:] / (2.0 * s)])
        nq = np.linalg.norm(q)
        q[0] *= nq
        q[1:] /= nq
    else:
        s = math.sqrt(0.5 * (1.0 + q[0]))
        q = np.concatenate([[s], q[1:] / (2.0 * s)])
    return q / np.linalg.norm(q)


# ---------------------------------------------------------------- parameters
class Params:
    def __init__(self, nseg=20, dl=2.2, dt=36.3833 / 300.0, kb=1800.0, kt=1800.0, mu=1.0,
                 a=1.0, dimensionless_force=220.0, repulsive_factor=2.0, tol=1e-4,
                 num_euler_steps=1, gamma=0.1):
        self.nseg, self.dl, self.dt, self.kb, self.kt = nseg, dl, dt, kb, kt
        self.mu, self.a = mu, a
        self.end_force = dimensionless_force * kb / (dl * dl * nseg * nseg)
        self.repulsive_factor = repulsive_factor
        self.tol = tol
        self.num_euler_steps = num_euler_steps
        self.gamma = gamma


class Filament:
    """State of one tethered, clamped filament (numpy FP64)."""

    def __init__(self, P: Params, x_seg, q_seg, pert=None):
        n = P.nseg
        self.P = P
        self.x = np.array(x_seg, float).reshape(n, 3)
        self.xm1 = self.x.copy()
        self.xm2 = self.x.copy()
        self.q = np.array(q_seg, float).reshape(n, 4)
        self.qm1 = self.q.copy()
        self.u = np.zeros((n, 3))
        self.um1 = np.zeros((n, 3))
        self.lam = np.zeros((n, 3))      # row 0: tether multiplier, row k: link (k-1, k)
        self.lam_m1 = np.zeros((n, 3))
        self.lam_m2 = np.zeros((n, 3))
        self.clamp = np.zeros(3)
        self.clamp_m1 = np.zeros(3)
        self.clamp_m2 = np.zeros(3)
        self.base = self.x[0].copy()
        self.pert = np.zeros(3) if pert is None else np.asarray(pert, float)
        self.strain_twist = np.zeros(3)
        self.Jinv = None

    def copy(self):
        import copy
        return copy.deepcopy(self)

    # -- kinematics ---------------------------------------------------------
    def robot_arm(self):
        dl = self.P.dl
        for i in range(1, self.P.nseg):
            self.x[i] = self.x[i - 1] + 0.5 * dl * (tangent(self.q[i - 1]) + tangent(self.q[i]))

    def initial_guess(self, nt, base=None, ub=None):
        """filament::initial_guess (filament.cpp:1336-1663): extrapolation, then
        accept_state_from_rigid_body (:653-678) with the base position and Lie-algebra
        element of the body (a fixed anchor and zero when not given)."""
        P = self.P
        self.xm2 = self.xm1.copy()
        self.xm1 = self.x.copy()
        for i in range(P.nseg):
            if nt <= P.num_euler_steps:
                self.um1[i] = self.u[i]
            else:
                t = 2.0 * self.u[i] - self.um1[i]
                self.um1[i] = self.u[i]
                self.u[i] = t
            self.qm1[i] = self.q[i]
            self.q[i] = qmul(lie_exp(self.u[i]), self.qm1[i])
        if nt == 1:
            self.lam_m1 = self.lam.copy()
            self.clamp_m1 = self.clamp.copy()
        elif nt == 2:
            self.lam_m2 = self.lam_m1.copy()
            self.lam_m1 = self.lam.copy()
            self.lam = 2.0 * self.lam_m1 - self.lam_m2
            self.clamp_m2 = self.clamp_m1.copy()
            self.clamp_m1 = self.clamp.copy()
            self.clamp = 2.0 * self.clamp_m1 - self.clamp_m2
        elif nt > 2:
            t = 3.0 * (self.lam - self.lam_m1) + self.lam_m2
            self.lam_m2 = self.lam_m1.copy()
            self.lam_m1 = self.lam.copy()
            self.lam = t
            t = 3.0 * (self.clamp - self.clamp_m1) + self.clamp_m2
            self.clamp_m2 = self.clamp_m1.copy()
            self.clamp_m1 = self.clamp.copy()
            self.clamp = t
        self.accept_state_from_rigid_body(base, ub)

    def accept_state_from_rigid_body(self, base=None, ub=None):
        """Segment 0 at the base with the body's Lie-algebra element (filament.cpp:653-678)."""
        self.x[0] = self.base if base is None else np.asarray(base, float)
        self.u[0] = np.zeros(3) if ub is None else np.asarray(ub, float).copy()
        self.q[0] = qmul(lie_exp(self.u[0]), self.qm1[0])
        self.robot_arm()

    def update(self, upd, base=None, ub=None):
        """filament::update (filament.cpp:2174-2208); with a body, first
        accept_state_from_rigid_body (swimmer.cpp:1076-1100)."""
        n = self.P.nseg
        if base is not None:
            self.accept_state_from_rigid_body(base, ub)
        upd = np.asarray(upd, float).reshape(2 * n, 3)
        self.lam[0] += upd[0]
        for i in range(1, n):
            self.lam[i] += upd[i]
            self.u[i] += upd[n + i]
            self.q[i] = qmul(lie_exp(self.u[i]), self.qm1[i])
        self.clamp += upd[n]
        self.robot_arm()

    # -- loads ----------------------------------------------------------------
    def body_frame_moment(self, k):
        P = self.P
        mq = qconj(midpoint(self.q[k], self.q[k + 1]))
        mq = 2.0 * qmul(mq, (self.q[k + 1] - self.q[k]) / P.dl)
        return np.array([P.kt * (mq[1] - self.strain_twist[0]), P.kb * (mq[2] - self.strain_twist[1]),
                         P.kb * (mq[3] - self.strain_twist[2])])

    def internal_forces(self, nt, with_perturbation=True):
        P = self.P
        n = P.nseg
        f = np.zeros((n, 6))
        for i in range(n - 1):
            lam = self.lam[i + 1]
            f[i, :3] -= lam
            f[i + 1, :3] += lam
            f[i, 3:] -= 0.5 * P.dl * np.cross(tangent(self.q[i]), lam)
            f[i + 1, 3:] -= 0.5 * P.dl * np.cross(tangent(self.q[i + 1]), lam)
            M = rot_mat(midpoint(self.q[i], self.q[i + 1])) @ self.body_frame_moment(i)
            f[i, 3:] += M
            f[i + 1, 3:] -= M
        f[0, :3] += self.lam[0]
        f[n - 1, :3] -= P.end_force * tangent(self.q[n - 1])
        if nt == 0 and with_perturbation:
            f[4, :3] += self.pert
        f[0, 3:] += dexpinv_transpose(self.u[0], self.clamp)
        return f

    # -- BDF residual (compute_errors, one filament) ----------------------------
    def errors(self, v, nt):
        """v: (nseg, 6) velocities/angular velocities.  Returns the 6*nseg error in
        the reference layout [positions (3 nseg), Lie algebra (3 nseg)]."""
        P = self.P
        n = P.nseg
        e = np.zeros(6 * n)
        c1, c2, c3 = BDF2_C1, BDF2_C2, BDF2_C3
        for j in range(n):
            if nt < P.num_euler_steps:
                e[3 * j:3 * j + 3] = self.x[j] - self.xm1[j] - P.dt * v[j, :3]
            else:
                e[3 * j:3 * j + 3] = self.x[j] + c1 * self.xm1[j] + c2 * self.xm2[j] + c3 * P.dt * v[j, :3]
            W = dexpinv(self.u[j], v[j, 3:])
            k = 3 * n + 3 * j
            if nt < P.num_euler_steps:
                e[k:k + 3] = self.u[j] - P.dt * W
            else:
                e[k:k + 3] = self.u[j] - c2 * self.um1[j] + c3 * P.dt * W
        return e

    # -- approximate Jacobian -----------------------------------------------------
    def moment_lie_derivative(self, q1, q2, upper):
        P = self.P
        A = psi_mat(q2) / P.dl if upper else psi_mat(q1) / (-P.dl)
        mq = qconj(midpoint(q1, q2))
        dMdu = left_mult_mat(mq) @ A
        dq = (q2 - q1) / P.dl
        A = A * (0.5 * P.dl if upper else -0.5 * P.dl)
        dMdu = dMdu + left_mult_mat(dq) @ A
        out = 2.0 * dMdu[1:4, :]
        out[0] *= P.kt
        out[1] *= P.kb
        out[2] *= P.kb
        return out

    def lie_algebra_block(self, nt, f):
        P = self.P
        n = P.nseg
        E = np.zeros((3 * n, 3 * n))
        tsf = P.dt if nt < P.num_euler_steps else 2.0 * P.dt / 3.0
        Tfac = 1.0 / (8.0 * PI * P.mu * P.a ** 3)
        q = self.q

        def tnb(k):
            return tangent(q[k]), normal(q[k]), binormal(q[k])

        def diag_block(dOmega, k):
            Lie = -rcross(self.u[k])
            Om = rcross(Tfac * f[k, 3:])
            uvec = self.u[k]
            return -tsf * (dOmega - 0.5 * (Om + Lie @ dOmega) +
                           (rcross(Om @ uvec) + Lie @ Om + Lie @ Lie @ dOmega) / 12.0)

        def off_block(dOmega, k):
            Lie = -rcross(self.u[k])
            return -tsf * (dOmega - 0.5 * Lie @ dOmega + Lie @ Lie @ dOmega / 12.0)

        # k = 0
        t, nn, b = tnb(0)
        tp, np1, bp = tnb(1)
        Cp = -0.5 * P.dl * rcross(self.lam[1]) @ rcross(t)
        Kph = self.body_frame_moment(0)
        R = rot_mat(midpoint(q[0], q[1]))
        El = 0.5 * (Kph[0] * rcross(t) + Kph[1] * rcross(nn) + Kph[2] * rcross(b)) + \
            R @ self.moment_lie_derivative(q[0], q[1], False)
        E[0:3, 0:3] = diag_block(Tfac * (El + Cp), 0)
        El = 0.5 * (Kph[0] * rcross(tp) + Kph[1] * rcross(np1) + Kph[2] * rcross(bp)) + \
            R @ self.moment_lie_derivative(q[0], q[1], True)
        E[0:3, 3:6] = off_block(Tfac * El, 0)
        # k = n - 1
        t, nn, b = tnb(n - 1)
        tm, nm1, bm = tnb(n - 2)
        Cp = -0.5 * P.dl * rcross(self.lam[n - 1]) @ rcross(t)
        Kmh = self.body_frame_moment(n - 2)
        R = rot_mat(midpoint(q[n - 2], q[n - 1]))
        El = -0.5 * (Kmh[0] * rcross(t) + Kmh[1] * rcross(nn) + Kmh[2] * rcross(b)) - \
            R @ self.moment_lie_derivative(q[n - 2], q[n - 1], True)
        E[3 * n - 3:, 3 * n - 3:] = diag_block(Tfac * (El + Cp), n - 1)
        El = -0.5 * (Kmh[0] * rcross(tm) + Kmh[1] * rcross(nm1) + Kmh[2] * rcross(bm)) - \
            R @ self.moment_lie_derivative(q[n - 2], q[n - 1], False)
        E[3 * n - 3:, 3 * n - 6:3 * n - 3] = off_block(Tfac * El, n - 1)
        # interior
        for k in range(1, n - 1):
            t, nn, b = tnb(k)
            tm, nm1, bm = tnb(k - 1)
            tp, np1, bp = tnb(k + 1)
            Kph = self.body_frame_moment(k)
            Kmh = self.body_frame_moment(k - 1)
            Rph = rot_mat(midpoint(q[k], q[k + 1]))
            Rmh = rot_mat(midpoint(q[k - 1], q[k]))
            lamvec = self.lam[k] + self.lam[k + 1]
            Cp = -0.5 * P.dl * rcross(lamvec) @ rcross(t)
            El = 0.5 * (Kph[0] * rcross(t) + Kph[1] * rcross(nn) + Kph[2] * rcross(b)) + \
                Rph @ self.moment_lie_derivative(q[k], q[k + 1], False)
            El = El - 0.5 * (Kmh[0] * rcross(t) + Kmh[1] * rcross(nn) + Kmh[2] * rcross(b)) - \
                Rmh @ self.moment_lie_derivative(q[k - 1], q[k], True)
            E[3 * k:3 * k + 3, 3 * k:3 * k + 3] = diag_block(Tfac * (Cp + El), k)
            El = -0.5 * (Kmh[0] * rcross(tm) + Kmh[1] * rcross(nm1) + Kmh[2] * rcross(bm)) - \
                Rmh @ self.moment_lie_derivative(q[k - 1], q[k], False)
            E[3 * k:3 * k + 3, 3 * k - 3:3 * k] = off_block(Tfac * El, k)
            El = 0.5 * (Kph[0] * rcross(tp) + Kph[1] * rcross(np1) + Kph[2] * rcross(bp)) + \
                Rph @ self.moment_lie_derivative(q[k], q[k + 1], True)
            E[3 * k:3 * k + 3, 3 * k + 3:3 * k + 6] = off_block(Tfac * El, k)
        E += np.eye(3 * n)
        return E

    def approx_jacobian(self, nt, f):
        """J_approx of invert_approx_jacobian (before inversion); f = internal loads."""
        P = self.P
        n = P.nseg
        J = np.zeros((6 * n, 6 * n))
        euler = nt < P.num_euler_steps
        fac = -P.dt / (6.0 * PI * P.mu * P.a) if euler else -P.dt / (9.0 * PI * P.mu * P.a)
        for i in range(3 * n):
            J[i, i] = fac
            if i < 3 * n - 3:
                J[i, i + 3] = -fac
        for j in range(1, n):
            jpos = 3 * (n + j)
            temp = P.dl * rcross(tangent(self.q[j]))
            J[3 * j:3 * j + 3, jpos:jpos + 3] = 0.5 * temp
            for i in range(j + 1, n):
                J[3 * i:3 * i + 3, jpos:jpos + 3] = temp
        J[3 * (n - 1):3 * n, 3 * (2 * n - 1):6 * n] -= fac * P.end_force * rcross(tangent(self.q[n - 1]))
        fac2 = P.dl * P.dt / (16.0 * PI * P.mu * P.a ** 3) if euler else P.dl * P.dt / (24.0 * PI * P.mu * P.a ** 3)
        for j in range(1, n):
            jpos = 3 * j
            ipos = 3 * (n + j - 1)
            for kk in (j - 1, j):
                um = rcross(self.u[kk])
                tm = -rcross(tangent(self.q[kk]))
                J[ipos:ipos + 3, jpos:jpos + 3] = fac2 * (tm + 0.5 * um @ tm + um @ um @ tm / 12.0)
                ipos += 3
        J[3 * n:, 3 * n:] = self.lie_algebra_block(nt, f)
        fac3 = -P.dt / (8.0 * PI * P.mu * P.a ** 3) if euler else -P.dt / (12.0 * PI * P.mu * P.a ** 3)
        um = rcross(self.u[0])
        blk = np.eye(3) + 0.5 * um + um @ um / 12.0
        blkT = np.eye(3) - 0.5 * um + um @ um / 12.0
        J[3 * n:3 * n + 3, 3 * n:3 * n + 3] = fac3 * blk @ blkT
        J[3 * n + 3:3 * n + 6, 3 * n:3 * n + 3] = 0.0
        return J


# ---------------------------------------------------------------- barrier forces
def barrier_forces(x_seg, nseg, P: Params, x_blob=None, wall=True):
    """Repulsive barrier forces on segments (cuda_functions.cu:1825-1930): pairs closer
    than 1.1 (a_i + a_j), excluding the same and neighbouring segments of one
    filament, plus the wall repulsion below z = 0.5 DL."""
    xs = np.asarray(x_seg, float).reshape(-1, 3)
    xb = np.zeros((0, 3)) if x_blob is None else np.asarray(x_blob, float).reshape(-1, 3)
    ns = xs.shape[0]
    allx = np.vstack([xs, xb])
    F = np.zeros((ns, 3))
    fmag = P.repulsive_factor * P.end_force
    for i in range(ns):
        fi = i // nseg
        if wall and xs[i, 2] < 0.5 * P.dl:
            fz = min(1.0, 1.0 - (xs[i, 2] - P.a) / (0.5 * P.dl - P.a))
            F[i, 2] += fmag * fz ** BARRIER_POWER
        d = xs[i] - allx
        dist = np.sqrt((d * d).sum(1))
        for j in range(allx.shape[0]):
            a_sum = 2.0 * P.a
            if j < ns and fi == j // nseg and abs(i - j) <= 1:
                continue
            if dist[j] < BARRIER_RANGE * a_sum:
                chi = 1.0 / ((BARRIER_RANGE - 1.0) * a_sum)
                fac = min(1.0, 1.0 - chi * (dist[j] - a_sum))
                F[i] += fmag * fac ** BARRIER_POWER * d[j] / dist[j]
    return F


# ---------------------------------------------------------------- Broyden
class Broyden:
    """Limited-memory 'good Broyden' of broyden_solver.cpp:71-174 (columns in lists)."""

    def __init__(self, gamma=0.1, max_iter=400):
        self.gamma = gamma
        self.max_iter = max_iter
        self.C = []
        self.D = []
        self.iter = 0

    def find_update(self, jinv_mult, error):
        upd = -self.gamma * jinv_mult(error)
        for Ci, Di in zip(self.C, self.D):
            upd -= np.dot(error, Di) * Ci
        return upd

    def end_of_iter(self, jinv_mult, error, new_error, too_large):
        if too_large and self.iter < self.max_iter - 1:
            c = -self.gamma * jinv_mult(new_error)
            for Ci, Di in zip(self.C, self.D):
                c -= np.dot(new_error, Di) * Ci
            d = new_error - error
            nd = np.linalg.norm(d)
            self.C.append(c / nd)
            self.D.append(d / nd)
            error = new_error
        self.iter += 1
        return error


# ---------------------------------------------------------------- full step
def badness(e, P: Params, nfil):
    """max(|e_x| / (DL TOL), |e_u| / TOL) over the reference-layout error vector."""
    E = np.asarray(e).reshape(nfil, 2, -1)
    return max(np.abs(E[:, 0]).max() / (P.dl * P.tol), np.abs(E[:, 1]).max() / P.tol)


def implicit_step(fils, nt, velocity_fn, barrier=True, x_blob=None, max_iter=400):
    """One BDF/Broyden step (cilia_sim_main.cu:405-552) over a list of filaments.
    velocity_fn(x_seg (n,3), F_seg (n,6)) -> (n,6) velocities.  Returns iterations."""
    P = fils[0].P
    n = P.nseg

    def loads():
        F = np.concatenate([f.internal_forces(nt) for f in fils])
        if barrier:
            xs = np.concatenate([f.x for f in fils])
            F[:, :3] += barrier_forces(xs, n, P, x_blob)
        return F

    def resid():
        xs = np.concatenate([f.x for f in fils])
        V = velocity_fn(xs, loads())
        return np.concatenate([f.errors(V[i * n:(i + 1) * n], nt) for i, f in enumerate(fils)])

    for f in fils:
        f.initial_guess(nt)
    e = resid()
    big = badness(e, P, len(fils)) > 1.0
    jinvs = None
    if big:
        jinvs = [np.linalg.inv(f.approx_jacobian(nt, f.internal_forces(nt))) for f in fils]

    def jm(r):
        return np.concatenate([J @ r[i * 6 * n:(i + 1) * 6 * n] for i, J in enumerate(jinvs)])

    B = Broyden(P.gamma, max_iter)
    it = 0
    while big and it < max_iter:
        upd = B.find_update(jm, e)
        for i, f in enumerate(fils):
            f.update(upd[i * 6 * n:(i + 1) * 6 * n])
        e_new = resid()
        if not np.all(np.isfinite(e_new)):
            raise FloatingPointError("Broyden diverged")
        big = badness(e_new, P, len(fils)) > 1.0
        e = B.end_of_iter(jm, e, e_new, big)
        it += 1
    return it
