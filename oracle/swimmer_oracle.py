"""numpy restatement of the free swimmer -- TEST INFRASTRUCTURE ONLY (imported by tests/
and never by the product).

A rigid body of blobs carrying elastic filaments, moving freely (BODY_VELOCITY_TYPE 0,
SOLVER_TYPE 1: blob forces and body velocities by GMRES, everything else by Broyden):
  * rigid_body::initial_guess / update      rigid_body.cpp:410-458, 517-544
  * filament bases attached to the body     swimmer.cpp:434-455, 1076-1100,
                                            filament.cpp:653-678
  * swimmer::forces_and_torques             swimmer.cpp:470-535 (external load given)
  * blob repulsion in the body balance      mobility_solver.cpp:178-238,
                                            cuda_functions.cu:1932-2012
  * free-body blob system                   mobility_solver.cpp:455-632, 1822-1880
  * body rows of compute_errors             mobility_solver.cpp:3003-3040
  * approximate Jacobian with the body      swimmer.cpp:688-917 (no sphere correction)
    Schur complement and its inverse apply  swimmer.cpp:1022-1068
Unknown order: the filaments' 6 NSEG blocks, then 6 per body (x, u).

Parity status: the Schur-complement inverse is pinned against a dense solve with the
assembled block Jacobian; the step against its own residual (BDF2 rows of the body
and filaments below TOL at convergence) and, for the body, against a body without
filaments sedimenting at the velocity the body mobility gives (tests/test_swimmer_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

from filament_oracle import (PI, Broyden, badness, barrier_forces, dexpinv, lie_exp, qmul,
                             rcross, rot_mat, tangent)


class Body:
    """Rigid body: centre x, orientation q = exp(u) qm1 (rigid_body.cpp)."""

    def __init__(self, x, q, blob_ref):
        self.x = np.array(x, float)
        self.xm1 = self.x.copy()
        self.xm2 = self.x.copy()
        self.u = np.zeros(3)
        self.um1 = np.zeros(3)
        self.q = np.array(q, float)
        self.qm1 = self.q.copy()
        self.Q_init = rot_mat(self.q)
        self.ref = np.asarray(blob_ref, float).reshape(-1, 3)

    def initial_guess(self, nt, num_euler):
        if nt <= num_euler:
            gx, gu = 2.0 * self.x - self.xm1, self.u.copy()
        else:
            gx, gu = 3.0 * (self.x - self.xm1) + self.xm2, 2.0 * self.u - self.um1
        self.xm2, self.xm1, self.x = self.xm1.copy(), self.x.copy(), gx
        self.um1, self.u = self.u.copy(), gu
        self.qm1 = self.q.copy()
        self.q = qmul(lie_exp(self.u), self.qm1)
        self.Q_init = rot_mat(self.q)

    def update(self, upd):
        self.x = self.x + upd[:3]
        self.u = self.u + upd[3:6]
        self.q = qmul(lie_exp(self.u), self.qm1)

    def offsets(self):
        return self.ref @ rot_mat(self.q).T

    def blobs(self):
        return self.x + self.offsets()


def sphere_blocks(rv, R, a=1.0, mu=1.0):
    """(Mtt, Mtr, Mrr) of a sphere of radius R and a segment of radius a at separation
    rv = x_seg - x_body (swimmer.cpp:809-840)."""
    r = np.linalg.norm(rv)
    e = rv / r
    rr = np.outer(e, e)
    s2 = R * R + a * a
    Mtt = ((1 + s2 / (3 * r * r)) * np.eye(3) + (1 - s2 / (r * r)) * rr) / (8 * math.pi * mu * r)
    Mtr = rcross(e) / (8 * math.pi * mu * r * r)
    Mrr = (3 * rr - np.eye(3)) / (16 * math.pi * mu * r ** 3)
    return Mtt, Mtr, Mrr


def sphere_coupling(f, xb, R, P, DinvT, a=1.0, mu=1.0):
    """Body rows of B added for a spherical body (swimmer.cpp:800-877): for every segment
    the RPY blocks of a sphere of radius R and a segment of radius a at separation r,
        Mtt = ((1 + (R^2 + a^2) / 3 r^2) I + (1 - (R^2 + a^2) / r^2) r r) / (8 pi mu r),
        Mtr = rcross(r_hat) / (8 pi mu r^2),   Mrr = (3 r r - I) / (16 pi mu r^3),
    placed at the columns of the constraint forces and Lie-algebra elements the segment
    depends on."""
    n = P.nseg
    C = np.zeros((6, 6 * n))
    for m in range(n):
        Mtt, Mtr, Mrr = sphere_blocks(f.x[m] - xb, R, a, mu)
        tc = rcross(tangent(f.q[m]))
        ta, tb = 0.5 * P.dl * Mtr @ tc, 0.5 * P.dl * Mrr @ tc
        if m == 0:
            C[0:3, 0:3] += Mtt
            C[3:6, 0:3] += Mtr
            C[0:3, 3 * n:3 * n + 3] += Mtr @ DinvT
            C[3:6, 3 * n:3 * n + 3] += Mrr @ DinvT
            C[0:3, 3:6] += ta - Mtt
            C[3:6, 3:6] += tb - Mtr
        elif m == n - 1:
            C[0:3, 6 * n - 3:6 * n] += -P.end_force * Mtt @ tc
            C[3:6, 6 * n - 3:6 * n] += -P.end_force * Mtr @ tc
            C[0:3, 3 * n - 3:3 * n] += ta + Mtt
            C[3:6, 3 * n - 3:3 * n] += tb + Mtr
        else:
            C[0:3, 3 * m:3 * m + 3] += ta + Mtt
            C[3:6, 3 * m:3 * m + 3] += tb + Mtr
            C[0:3, 3 * (m + 1):3 * (m + 1) + 3] += ta - Mtt
            C[3:6, 3 * (m + 1):3 * (m + 1) + 3] += tb - Mtr
    return C


def blob_barrier_forces(x_seg, x_blob, P, nb_per_body):
    """Repulsion on each blob from the segments and from the blobs of other bodies
    (cuda_functions.cu:1932-2012; the blob pairs of one body cancel in its total force
    and torque, so they are left out)."""
    xs = np.asarray(x_seg, float).reshape(-1, 3)
    xb = np.asarray(x_blob, float).reshape(-1, 3)
    fmag = P.repulsive_factor * P.end_force
    a_sum = 2.0 * P.a
    F = np.zeros_like(xb)
    for i in range(xb.shape[0]):
        others = [xs]
        bi = i // nb_per_body
        mask = np.arange(xb.shape[0]) // nb_per_body != bi
        others.append(xb[mask])
        allx = np.vstack(others)
        d = xb[i] - allx
        dist = np.sqrt((d * d).sum(1))
        for j in np.nonzero(dist < 1.1 * a_sum)[0]:
            fac = min(1.0, 1.0 - (10.0 / a_sum) * (dist[j] - a_sum))
            F[i] += fmag * fac ** 4 * d[j] / dist[j]
    return F


class Swimmer:
    """One free body with filaments (filament f's base offset fref[f] in the body frame)."""

    def __init__(self, P, body: Body, fils, mob_ref, f_ext=None):
        self.P = P
        self.body = body
        self.fils = fils
        self.mob_ref = np.asarray(mob_ref, float)    # body-frame 6 x 6 body mobility
        self.f_ext = np.zeros(6) if f_ext is None else np.asarray(f_ext, float)
        R = rot_mat(body.q)
        self.fref = [R.T @ (f.x[0] - body.x) for f in fils]
        self.f = np.zeros(6)
        self.vel = np.zeros(6)

    def base(self, i):
        return self.body.x + rot_mat(self.body.q) @ self.fref[i]

    def initial_guess(self, nt):
        self.body.initial_guess(nt, self.P.num_euler_steps)
        for i, f in enumerate(self.fils):
            f.initial_guess(nt, self.base(i), self.body.u)

    def update(self, upd):
        n6 = 6 * self.P.nseg
        self.body.update(upd[len(self.fils) * n6:])
        for i, f in enumerate(self.fils):
            f.update(upd[i * n6:(i + 1) * n6], self.base(i), self.body.u)

    def body_force(self, F_int, blob_rep=None):
        """swimmer::forces_and_torques (minus the segment loads about the centre, plus the
        external load) and the blob repulsion of read_positions_and_forces."""
        f = np.zeros(6)
        xs = np.concatenate([fil.x for fil in self.fils])
        d = xs - self.body.x
        f[:3] -= F_int[:, :3].sum(0)
        f[3:] -= (F_int[:, 3:] + np.cross(d, F_int[:, :3])).sum(0)
        f += self.f_ext
        if blob_rep is not None:
            r = self.body.offsets()
            f[:3] += blob_rep.sum(0)
            f[3:] += np.cross(r, blob_rep).sum(0)
        return f

    def errors(self, V, U, nt):
        """Filament rows, then the body rows of compute_errors (x, u; BDF2 / Euler)."""
        P = self.P
        n = P.nseg
        e = [f.errors(V[i * n:(i + 1) * n], nt) for i, f in enumerate(self.fils)]
        b = self.body
        c1, c2, c3 = -4.0 / 3.0, 1.0 / 3.0, -2.0 / 3.0
        w = dexpinv(b.u, U[3:])
        if nt < P.num_euler_steps:
            ex = b.x - b.xm1 - P.dt * U[:3]
            eu = b.u - P.dt * w
        else:
            ex = b.x + c1 * b.xm1 + c2 * b.xm2 + c3 * P.dt * U[:3]
            eu = b.u - c2 * b.um1 + c3 * P.dt * w
        return np.concatenate(e + [ex, eu])

    def badness(self, e):
        P = self.P
        nf = len(self.fils)
        bf = badness(e[:6 * P.nseg * nf], P, nf)
        eb = e[6 * P.nseg * nf:]
        return max(bf, np.abs(eb[:3]).max() / (P.dl * P.tol), np.abs(eb[3:]).max() / P.tol)

    # -- approximate Jacobian with the body Schur complement (swimmer.cpp:688-917) --
    sphere_R = 0.0   # > 0: the sphere correction of the body rows of B

    def prepare_jacobian(self, nt, F_int):
        P = self.P
        n = P.nseg
        b = self.body
        dt_fac = -P.dt if nt < P.num_euler_steps else -2.0 * P.dt / 3.0
        Q = b.Q_init
        u = b.u
        blc = -rcross(u)
        cbo = -blc / dt_fac
        S = np.eye(6)
        S[3:, 3:] += dt_fac * (blc @ cbo / 12.0 - 0.5 * cbo)
        N = self.mob_ref.copy()
        for i in (0, 3):
            for j in (0, 3):
                N[i:i + 3, j:j + 3] = Q @ self.mob_ref[i:i + 3, j:j + 3] @ Q.T
        old = np.zeros((3, 3))
        for i, f in enumerate(self.fils):
            d = Q @ self.fref[i]
            old -= rcross(f.lam[0]) @ rcross(d)
            uxl = np.cross(u, f.clamp)
            old += 0.5 * rcross(f.clamp) - (rcross(uxl) + blc @ rcross(f.clamp)) / 12.0
            t = tangent(f.q[n - 1])
            dn = np.linalg.norm(d)
            old += rcross(P.end_force * t) @ rcross((0.5 * P.dl + dn) * d / dn)
        old = N[3:, 3:] @ old
        S[3:, 3:] += dt_fac * (old - 0.5 * blc @ old + blc @ blc @ old / 12.0)
        Dinv = np.eye(3) + blc @ blc / 12.0 - 0.5 * blc
        DinvT = np.eye(3) + blc @ blc / 12.0 + 0.5 * blc
        self.S0 = S.copy()   # body block of the approximate Jacobian
        self.jinv, self.BJ, self.A, self.B, self.Ja = [], [], [], [], []
        for i, f in enumerate(self.fils):
            Ja = f.approx_jacobian(nt, F_int[i * n:(i + 1) * n])
            lab = f.lie_algebra_block(nt, F_int[i * n:(i + 1) * n])
            clamp1, clamp2 = lab[0:3, 0:3].copy(), lab[3:6, 0:3].copy()
            Jinv = np.linalg.inv(Ja)
            d = Q @ self.fref[i]
            B = np.zeros((6, 6 * n))
            B[0:3, 0:3] = -np.eye(3)
            B[3:6, 0:3] = rcross(d)
            B[3:6, 3 * n:3 * n + 3] = -DinvT
            t = tangent(f.q[n - 1])
            B[0:3, 6 * n - 3:6 * n] = rcross(P.end_force * t)
            diff = f.x[n - 1] - b.x
            B[3:6, 6 * n - 3:6 * n] = rcross(0.5 * P.dl * t - diff) @ B[0:3, 6 * n - 3:6 * n]
            for m in range(1, n - 1):
                tm = tangent(f.q[m])
                B[3:6, 3 * (n + m):3 * (n + m) + 3] = B[0:3, 6 * n - 3:6 * n] @ rcross(P.dl * tm)
            B = N @ B
            if self.sphere_R > 0.0:
                B += sphere_coupling(f, b.x, self.sphere_R, P, DinvT)
            B[3:6, :] = Dinv @ B[3:6, :]
            B *= dt_fac
            A = np.zeros((6 * n, 6))
            dn = np.linalg.norm(d)
            Dm = rcross((0.5 * P.dl + dn) * d / dn)
            for m in range(n):
                A[3 * m:3 * m + 3, 0:3] = np.eye(3)
                A[3 * m:3 * m + 3, 3:6] = Dm
            A[3 * n:3 * n + 3, 3:6] = clamp1
            A[3 * n + 3:3 * n + 6, 3:6] = clamp2
            BJ = B @ Jinv
            S -= BJ @ A
            self.jinv.append(Jinv)
            self.BJ.append(BJ)
            self.A.append(A)
            self.B.append(B)
            self.Ja.append(Ja)
        self.Sinv = np.linalg.inv(S)
        self.S = S
        return self

    def jinv_mult(self, r):
        n6 = 6 * self.P.nseg
        nf = len(self.fils)
        b2 = r[nf * n6:].copy()
        for i in range(nf):
            b2 -= self.BJ[i] @ r[i * n6:(i + 1) * n6]
        ob = self.Sinv @ b2
        out = [self.jinv[i] @ (r[i * n6:(i + 1) * n6] - self.A[i] @ ob) for i in range(nf)]
        return np.concatenate(out + [ob])


def implicit_step(sw: Swimmer, nt, velocity_fn, barrier=True, wall=False, max_iter=400):
    """One BDF/Broyden step of a free swimmer.  velocity_fn(sw, F_seg (n,6), f_body (6))
    -> (V_seg (n,6), U_body (6)).  Returns the Broyden iterations."""
    P = sw.P
    n = P.nseg

    def loads():
        F = np.concatenate([f.internal_forces(nt) for f in sw.fils])
        rep = None
        if barrier:
            xs = np.concatenate([f.x for f in sw.fils])
            xb = sw.body.blobs()
            F[:, :3] += barrier_forces(xs, n, P, xb, wall=wall)
            rep = blob_barrier_forces(xs, xb, P, xb.shape[0])
        Fint = np.concatenate([f.internal_forces(nt) for f in sw.fils])
        return F, sw.body_force(Fint, rep)

    def resid():
        F, fb = loads()
        V, U = velocity_fn(sw, F, fb)
        sw.vel = U
        return sw.errors(V, U, nt)

    sw.initial_guess(nt)
    e = resid()
    big = sw.badness(e) > 1.0
    if big:
        Fint = np.concatenate([f.internal_forces(nt) for f in sw.fils])
        sw.prepare_jacobian(nt, Fint)
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
