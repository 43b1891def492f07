"""numpy restatement of the reference's SOLVER_TYPE 0 -- TEST INFRASTRUCTURE ONLY (imported
by tests/ and never by the product).

SOLVER_TYPE 0 (config.hpp:360-364, 423, 443-447, USE_BROYDEN_FOR_EVERYTHING) puts the blob
forces, and for a free body its position and orientation, among the Broyden unknowns
instead of solving the blob system by GMRES at every Broyden iteration:
  * unknowns per swimmer: filaments (6 NSEG each), body (x, u), blob forces (3 NBLOB)
    broyden_solver.cpp:30-40, config.hpp:443-447
  * blob-force guess: BDF extrapolation from step 2 on     rigid_body.cpp:459-490
  * update adds the blob-force increment                   rigid_body.cpp:517-527
  * velocities: every product from the current loads       mobility_solver.cpp:2697-2730
  * residual rows: filaments as before; body: -f + sum of the blob forces and their
    torques about the centre; blobs: V + Omega x (Q r) - v_blob with V, Omega from the
    body's BDF2 (or Euler) relation (dexp)              mobility_solver.cpp:2768-2999
  * approximate Jacobian: the filaments' blocks, the blob rows by the self mobility
    -6 pi mu a, the body Schur complement             swimmer.cpp:561-700, 929-1018
Unknown order here: the filaments of every swimmer, then 6 per body, then 3 per blob.
A body held fixed (prescribed zero velocity) keeps only the blob-force unknowns: its body
rows vanish identically and its block of J0 is the identity (swimmer.cpp:567-573,
940-944), so they are dropped.

Parity status: steps are pinned by the physics they must reach -- the converged state is
the SOLVER_TYPE 1 state (blob forces from the GMRES blob solve) to the Broyden
tolerance, and the blob rows and the force and torque balance hold at convergence
(tests/test_broyden_only_oracle.py); the elimination in jinv_mult is pinned against a
dense solve of the assembled block matrix, and the blob terms of the body's Schur
complement against the elimination of the blob rows.  The remaining rotation-rotation
terms of the body block (swimmer.cpp:582-621) are the reference's approximation of the
Jacobian, which changes only how fast Broyden converges.  The blob-force, tether and
end-force terms are each the exact derivative of a torque (exp(delta) a) x f with the
force held, and the clamp term the derivative of the moment map dexpinv^T(u) lam through
second order in u: both pinned by finite differences (torque_rotation_block, with the
factors' order checked; clamp_block, with the error shrinking as |u|^3).  What is
approximate is where the reference evaluates them (the step's initial orientation, the
forces held), which moves Broyden's iterates but not the converged state, pinned above.
"""
from __future__ import annotations

import math

import numpy as np

from filament_oracle import PI, Broyden, badness, barrier_forces, dexp, rcross, rot_mat, tangent
from swimmer_oracle import Swimmer, blob_barrier_forces


def blob_force_guess(fb, fb_m1, fb_m2, nt):
    """rigid_body::initial_guess blob-force branch (rigid_body.cpp:459-490): returns
    (fb, fb_m1, fb_m2) after the step's shift."""
    if nt == 1:
        return fb, fb.copy(), fb_m2
    if nt == 2:
        fb_m2, fb_m1 = fb_m1.copy(), fb.copy()
        return 2.0 * fb_m1 - fb_m2, fb_m1, fb_m2
    if nt > 2:
        t = 3.0 * (fb - fb_m1) + fb_m2
        return t, fb.copy(), fb_m1.copy()
    return fb, fb_m1, fb_m2


def body_velocity(x, xm1, xm2, u, um1, dt, nt, num_euler):
    """V, Omega the body's position and Lie algebra imply (mobility_solver.cpp:2877-2905)."""
    c1, c2, c3 = -4.0 / 3.0, 1.0 / 3.0, -2.0 / 3.0
    if nt < num_euler:
        return (x - xm1) / dt, u / dt
    V = -(x + c1 * xm1 + c2 * xm2) / (c3 * dt)
    w = (c2 * um1 - u) / (c3 * dt)
    return V, dexp(u, w)


def torque_rotation_block(f, a):
    """d/d(delta) of (exp(delta) a) x f at delta = 0 with f held: rcross(f) rcross(a).  The
    body block's blob-force, tether and end-force terms (swimmer.cpp:588-603,607,614-621)
    all have this form."""
    return rcross(f) @ rcross(a)


def clamp_block(u, lam):
    """The clamp moment's term of the body block (swimmer.cpp:609-612): 0.5 rcross(lam)
    plus the second-order correction (rcross(u x lam) - rcross(u) rcross(lam)) / 12 -- the
    derivative of the moment map dexpinv^T(u) lam through second order in u."""
    return 0.5 * rcross(lam) + (rcross(np.cross(u, lam)) - rcross(u) @ rcross(lam)) / 12.0


class FreeSwimmer(Swimmer):
    """A free swimmer with its blob forces among the Broyden unknowns."""

    def __init__(self, P, body, fils, mob_ref, f_ext=None, a_blob=1.0, mu=1.0):
        super().__init__(P, body, fils, mob_ref, f_ext)
        nb = body.ref.shape[0]
        self.fb = np.zeros((nb, 3))
        self.fb_m1 = np.zeros((nb, 3))
        self.fb_m2 = np.zeros((nb, 3))
        self.a_blob, self.mu = a_blob, mu

    def nbroy(self):
        return 6 * self.P.nseg * len(self.fils) + 6 + self.fb.size

    def initial_guess(self, nt):
        super().initial_guess(nt)
        self.fb, self.fb_m1, self.fb_m2 = blob_force_guess(self.fb, self.fb_m1, self.fb_m2, nt)

    def update(self, upd):
        n0 = 6 * self.P.nseg * len(self.fils)
        super().update(upd[:n0 + 6])
        self.fb = self.fb + upd[n0 + 6:].reshape(-1, 3)

    def errors(self, V, v_blob, f_body, nt):
        """Filament rows, then the body's force and torque balance, then the blob rows."""
        P, b = self.P, self.body
        n = P.nseg
        e = [f.errors(V[i * n:(i + 1) * n], nt) for i, f in enumerate(self.fils)]
        d = self.body.offsets()                      # Q r with the current orientation
        eb = -np.asarray(f_body, float).copy()
        eb[:3] += self.fb.sum(0)
        eb[3:] += np.cross(d, self.fb).sum(0)
        Vb, Om = body_velocity(b.x, b.xm1, b.xm2, b.u, b.um1, P.dt, nt, P.num_euler_steps)
        eg = Vb + np.cross(Om, d) - np.asarray(v_blob, float).reshape(-1, 3)
        return np.concatenate(e + [eb, eg.reshape(-1)])

    def badness(self, e):
        P = self.P
        nf = len(self.fils)
        n0 = 6 * P.nseg * nf
        return max(badness(e[:n0], P, nf), np.abs(e[n0:]).max() / P.tol)

    def prepare_jacobian(self, nt, F_int):
        """The filaments' approximate Jacobians and the body Schur complement of
        SOLVER_TYPE 0 (swimmer.cpp:575-695)."""
        P, b = self.P, self.body
        n = P.nseg
        Q = b.Q_init
        u = b.u
        bmf = (-6.0 if nt < P.num_euler_steps else -9.0) * PI * self.mu * self.a_blob / P.dt
        S = np.zeros((6, 6))
        nb = self.fb.shape[0]
        S[:3, :3] = -nb * bmf * np.eye(3)
        for m in range(nb):
            Rc = rcross(Q @ b.ref[m])
            S[3:, 3:] += torque_rotation_block(self.fb[m], Q @ b.ref[m])
            S[:3, 3:] -= bmf * Rc
            S[3:, :3] += bmf * Rc
            S[3:, 3:] += bmf * Rc @ Rc
        self.jinv, self.JB, self.A, self.D, self.Ja = [], [], [], [], []
        for i, f in enumerate(self.fils):
            d = Q @ self.fref[i]
            S[3:, 3:] += torque_rotation_block(f.lam[0], d)
            S[3:, 3:] += clamp_block(u, f.clamp)
            t0, tl = tangent(f.q[0]), tangent(f.q[n - 1])
            S[3:, 3:] += torque_rotation_block(-P.end_force * tl, 0.5 * P.dl * t0)
            Dm = np.zeros((6, 6 * n))
            Dm[0:3, 0:3] = np.eye(3)
            Dm[3:6, 0:3] = rcross(-d)
            Dm[3:6, 3 * n:3 * n + 3] = np.eye(3) + rcross(u) @ rcross(u) / 12.0 - 0.5 * rcross(u)
            E = rcross(-P.end_force * tl)
            Dm[0:3, 6 * n - 3:6 * n] = E
            diff = f.x[n - 1] - b.x
            Dm[3:6, 6 * n - 3:6 * n] = rcross(0.5 * P.dl * tl - diff) @ E
            for m in range(1, n - 1):
                Dm[3:6, 3 * (n + m):3 * (n + m) + 3] = E @ rcross(P.dl * tangent(f.q[m]))
            Ja = f.approx_jacobian(nt, F_int[i * n:(i + 1) * n])
            lab = f.lie_algebra_block(nt, F_int[i * n:(i + 1) * n])
            Jinv = np.linalg.inv(Ja)
            A = np.zeros((6 * n, 6))
            dn = np.linalg.norm(d)
            Du = rcross((0.5 * P.dl + dn) * d / dn)
            for m in range(n):
                A[3 * m:3 * m + 3, 0:3] = np.eye(3)
                A[3 * m:3 * m + 3, 3:6] = Du
            A[3 * n:3 * n + 3, 3:6] = lab[0:3, 0:3]
            A[3 * n + 3:3 * n + 6, 3:6] = lab[3:6, 0:3]
            JB = Dm @ Jinv
            S -= JB @ A
            self.jinv.append(Jinv)
            self.JB.append(JB)
            self.A.append(A)
            self.D.append(Dm)
            self.Ja.append(Ja)
        self.S = S
        self.Sinv = np.linalg.inv(S)
        self.dt_fac = (1.0 if nt < P.num_euler_steps else 1.5) / P.dt
        return self

    def jinv_mult(self, r):
        """J0^-1 r (swimmer.cpp:929-1018)."""
        P, b = self.P, self.body
        n6 = 6 * P.nseg
        nf = len(self.fils)
        n0 = nf * n6
        bmf0 = -6.0 * PI * self.mu * self.a_blob
        b3 = r[n0:n0 + 6].copy()
        b2 = r[n0 + 6:].reshape(-1, 3).copy()
        for i in range(nf):
            b3 -= self.JB[i] @ r[i * n6:(i + 1) * n6]
        d = b.ref @ b.Q_init.T
        b3[:3] -= bmf0 * b2.sum(0)
        b3[3:] -= bmf0 * np.cross(d, b2).sum(0)
        ob = self.Sinv @ b3
        b2 -= self.dt_fac * (ob[:3] + np.cross(ob[3:], d))
        out = [self.jinv[i] @ (r[i * n6:(i + 1) * n6] - self.A[i] @ ob) for i in range(nf)]
        return np.concatenate(out + [ob, (bmf0 * b2).reshape(-1)])


def swimmer_step(sw: FreeSwimmer, nt, velocity_fn, barrier=True, wall=False, max_iter=400):
    """One BDF/Broyden step of a free swimmer with SOLVER_TYPE 0.
    velocity_fn(sw, F_seg (n,6), f_blob (nb,3)) -> (V_seg (n,6), v_blob (nb,3)), every
    product of the current loads.  Returns the Broyden iterations."""
    P = sw.P
    n = P.nseg

    def resid():
        Fint = np.concatenate([f.internal_forces(nt) for f in sw.fils])
        F = Fint.copy()
        rep = None
        if barrier:
            xs = np.concatenate([f.x for f in sw.fils])
            xb = sw.body.blobs()
            F[:, :3] += barrier_forces(xs, n, P, xb, wall=wall)
            rep = blob_barrier_forces(xs, xb, P, xb.shape[0])
        fbody = sw.body_force(Fint, rep)
        V, vb = velocity_fn(sw, F, sw.fb)
        sw.last = (V, vb)
        return sw.errors(V, vb, fbody, nt)

    sw.initial_guess(nt)
    e = resid()
    big = sw.badness(e) > 1.0
    if big:
        sw.prepare_jacobian(nt, np.concatenate([f.internal_forces(nt) for f in sw.fils]))
    B = Broyden(P.gamma, max_iter)
    it = 0
    while big and it < max_iter:
        upd = B.find_update(sw.jinv_mult, e)
        sw.update(upd)
        e_new = resid()
        if not np.all(np.isfinite(e_new)):
            raise FloatingPointError("Broyden diverged")
        big = sw.badness(e_new) > 1.0
        e = B.end_of_iter(sw.jinv_mult, e, e_new, big)
        it += 1
    return it


class FixedBlobs:
    """Filaments beside a body held fixed, its blob forces among the Broyden unknowns
    (PRESCRIBED_BODY_VELOCITIES with zero velocity).  Blob rows: -v_blob; their block of
    J0^-1: -6 pi mu a (swimmer.cpp:940-944)."""

    def __init__(self, P, fils, x_blob, a_blob=1.0, mu=1.0):
        self.P, self.fils = P, fils
        self.x_blob = np.asarray(x_blob, float).reshape(-1, 3)
        nb = self.x_blob.shape[0]
        self.fb = np.zeros((nb, 3))
        self.fb_m1 = np.zeros((nb, 3))
        self.fb_m2 = np.zeros((nb, 3))
        self.a_blob, self.mu = a_blob, mu

    def initial_guess(self, nt):
        for f in self.fils:
            f.initial_guess(nt)
        self.fb, self.fb_m1, self.fb_m2 = blob_force_guess(self.fb, self.fb_m1, self.fb_m2, nt)

    def update(self, upd):
        n6 = 6 * self.P.nseg
        for i, f in enumerate(self.fils):
            f.update(upd[i * n6:(i + 1) * n6])
        self.fb = self.fb + upd[len(self.fils) * n6:].reshape(-1, 3)

    def errors(self, V, v_blob, nt):
        n = self.P.nseg
        e = [f.errors(V[i * n:(i + 1) * n], nt) for i, f in enumerate(self.fils)]
        return np.concatenate(e + [-np.asarray(v_blob, float).reshape(-1)])

    def badness(self, e):
        P = self.P
        nf = len(self.fils)
        n0 = 6 * P.nseg * nf
        return max(badness(e[:n0], P, nf), np.abs(e[n0:]).max() / P.tol)

    def prepare_jacobian(self, nt):
        self.jinv = [np.linalg.inv(f.approx_jacobian(nt, f.internal_forces(nt))) for f in self.fils]

    def jinv_mult(self, r):
        n6 = 6 * self.P.nseg
        nf = len(self.fils)
        out = [J @ r[i * n6:(i + 1) * n6] for i, J in enumerate(self.jinv)]
        return np.concatenate(out + [-6.0 * PI * self.mu * self.a_blob * r[nf * n6:]])


def fixed_step(sys_: FixedBlobs, nt, velocity_fn, barrier=True, wall=True, max_iter=400):
    """One step with the fixed body's blob forces among the Broyden unknowns.
    velocity_fn(x_seg (n,3), F_seg (n,6), f_blob (nb,3)) -> (V_seg (n,6), v_blob (nb,3))."""
    P = sys_.P
    n = P.nseg

    def resid():
        xs = np.concatenate([f.x for f in sys_.fils])
        F = np.concatenate([f.internal_forces(nt) for f in sys_.fils])
        if barrier:
            F[:, :3] += barrier_forces(xs, n, P, sys_.x_blob, wall=wall)
        V, vb = velocity_fn(xs, F, sys_.fb)
        return sys_.errors(V, vb, nt)

    sys_.initial_guess(nt)
    e = resid()
    big = sys_.badness(e) > 1.0
    if big:
        sys_.prepare_jacobian(nt)
    B = Broyden(P.gamma, max_iter)
    it = 0
    while big and it < max_iter:
        upd = B.find_update(sys_.jinv_mult, e)
        sys_.update(upd)
        e_new = resid()
        if not np.all(np.isfinite(e_new)):
            raise FloatingPointError("Broyden diverged")
        big = sys_.badness(e_new) > 1.0
        e = B.end_of_iter(sys_.jinv_mult, e, e_new, big)
        it += 1
    return it
