"""numpy restatement of prescribed cilia carried by free rigid bodies -- TEST
INFRASTRUCTURE ONLY (imported by tests/ and never by the product).

Follows the reference's PRESCRIBED_CILIA swimmers with free bodies
(PRESCRIBED_BODY_VELOCITIES 0, unbounded fluid):
  * cilium frames on a body (filament.cpp:370-432): qtemp = sqrt((d_x, 0, -d_z, d_y))
    turns e_x onto the base direction d; q2 = sqrt((p . n, n x p)) turns qtemp's second
    axis n onto the posterior polar direction p (-e_z projected off d, normalised); a
    tilt turns about d; body_q_ref = q_tilt q2 qtemp.  Cilium c of body b sits in the
    frame R(q_b) R(body_q_ref) with its base at x_b + R(q_b) ref_c (swimmer.cpp:448-462,
    filament.cpp:653-663), its segments placed by the fitted beat (filament.cpp:1336-1490);
  * the system of mobility_solver.cpp:1891-2115, unknowns in the reference's order
    [F_seg (3 per segment) | F_blob (3 per blob) | V, Omega (6 per body) | phase speeds |
    angle speeds]:
        M F - K_B U - sum_c (phi_dot_c k_phi,c + theta_dot_c k_theta,c)
        -K_B^T F,   -k_phi,c . F,   -k_theta,c . F
    with (K_B U)_i = V_b + Omega_b x r_i, r_i = x_i - x_b;
  * right-hand side (mobility_solver.cpp:301-458): prescribed phase speeds
    [phase_dot k_phi | 0 | 0]; dynamic phases [0 | 0 | 0 | -q_phase | -q_angle];
  * the block preconditioner with 6 pi mu a on every particle
    (mobility_solver.cpp:1104-1515) and make_precon_mat's 6 x 6 matrix
    (swimmer.cpp:1175-1347), restated step by step;
  * prescribed rotation (BODY_VELOCITY_TYPE 2, rot_prescribed): Omega_b = 0, its
    columns and the torque rows dropped (mobility_solver.cpp:1947-1994, 2177-2181),
    written as identity rows on Omega_b with a zero right-hand side;
  * the explicit time integration (filament.cpp:1340-1356, cilia_sim_main.cu:724-740):
    phase += dt phase_dot before a step, x_b += dt V_b and q_b = exp(dt Omega_b) q_b after.

Parity status: pinned in tests/test_cilia_swimmer_oracle.py by what the mathematics
fixes -- the frames against an independent construction from their defining axes
(orthonormal, right-handed, first axis d, second axis the tilted polar direction), the
preconditioner against the exact inverse of the saddle matrix with M replaced by
(6 pi mu a)^-1 I, make_precon_mat against that matrix's Schur complement, the solve
against force- and torque-free balance and the rigid-body identities (a swimmer
translated moves identically; rotated, its velocities rotate), and the dynamic-phase
solve by each cilium's generalised force equalling its interpolated drive.  The
prescribed-rotation solve is pinned against the free system: applying to the free swimmer
the torque that holds its rotation reproduces the same forces and V with Omega = 0.
"""
from __future__ import annotations

import math

import numpy as np

import cilia_oracle as co
import oracle
from filament_oracle import lie_exp, qmul, qsqrt, rcross, rot_mat


# ---- frames ----------------------------------------------------------------------------

def frame_quaternion(direction, tilt=0.0):
    """body_q_ref of a cilium whose base points along `direction` (body frame), built
    as the reference composes it (filament.cpp:370-432).  Not defined for d = -e_x."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    qtemp = qsqrt(np.array([d[0], 0.0, -d[2], d[1]]))
    n = rot_mat(qtemp)[:, 1]
    p = np.array([0.0, 0.0, -1.0]) + d[2] * d
    q2 = np.array([1.0, 0.0, 0.0, 0.0])
    if np.linalg.norm(p) > 0.0:
        p = p / np.linalg.norm(p)
        q2 = qsqrt(np.concatenate([[p @ n], np.cross(n, p)]))
    qt = np.array([math.cos(0.5 * tilt), *(math.sin(0.5 * tilt) * d)])
    return qmul(qmul(qt, q2), qtemp)


# ---- geometry --------------------------------------------------------------------------

class Swimmers:
    """n_body bodies sharing blob offsets `blob_ref` (nb x 3) and one arrangement of
    cilia (base offsets fil_ref, frame quaternions fil_q); state: body centres and
    quaternions, phases, shape angles and their speeds, body velocities."""

    def __init__(self, blob_ref, fil_ref, fil_q, body_x, body_q, phase, phase_dot, *, nseg=10,
                 dl=2.6, dt=0.01, beat="fulford_blake", s_tol=1e-12, mu=1.0, a=1.0):
        self.blob_ref = np.asarray(blob_ref, dtype=np.float64).reshape(-1, 3)
        self.fil_ref = np.asarray(fil_ref, dtype=np.float64).reshape(-1, 3)
        self.fil_q = np.asarray(fil_q, dtype=np.float64).reshape(-1, 4)
        self.x = np.array(body_x, dtype=np.float64).reshape(-1, 3)
        self.q = np.array(body_q, dtype=np.float64).reshape(-1, 4)
        self.nbody, self.per, self.nb = self.x.shape[0], self.fil_ref.shape[0], self.blob_ref.shape[0]
        self.nc = self.nbody * self.per
        self.phase = np.array(phase, dtype=np.float64).reshape(-1)
        self.phase_dot = np.array(phase_dot, dtype=np.float64).reshape(-1)
        self.omega0 = self.phase_dot.copy()
        self.angle = np.zeros(self.nc)
        self.angle_dot = np.zeros(self.nc)
        self.V = np.zeros((self.nbody, 6))
        self.nseg, self.dl, self.dt, self.beat, self.s_tol = nseg, dl, dt, beat, s_tol
        self.mu, self.a = mu, a
        self.pair_dp = 0.0    # PAIR_DP of the paired beats (BICILIA, BICILIA_LONGT)
        self.driving = None   # (q_phase_table, q_angle_table or None, spring)
        self.fene, self.noise, self.seed, self.nt = False, 0.0, 0, 0
        # BODY_VELOCITY_TYPE 2: Omega_b held at zero, the Omega columns and torque rows
        # dropped (mobility_solver.cpp:1947-1994, 2177-2181), written with identity rows
        self.rot_prescribed = False

    def geometry(self):
        """Segment positions (n_seg x 3), blob positions (n_blob x 3) and per-cilium
        velocity directions (nc x 3 nseg)."""
        xs, vp, va = [], [], []
        for b in range(self.nbody):
            Rb = rot_mat(self.q[b])
            for l in range(self.per):
                c = b * self.per + l
                R = Rb @ rot_mat(self.fil_q[l])
                base = self.x[b] + Rb @ self.fil_ref[l]
                if self.beat in ("bicilia", "bicilia_longt"):   # pairs of nseg / 2
                    longt = self.beat == "bicilia_longt"
                    m = self.nseg // 2
                    psi2 = co.pair_phase(self.phase[c], self.pair_dp, longt)
                    s = np.concatenate([co.s_values(self.phase[c], m, "volvox", tol=self.s_tol),
                                        co.s_values(psi2, m, "volvox", tol=self.s_tol)])
                    x, p, a, _ = co.pair_placement(self.phase[c], base, self.nseg, self.dl,
                                                   self.angle[c], self.pair_dp, longt, self.a,
                                                   frame=R, s=s)
                else:
                    s = co.s_values(self.phase[c], self.nseg, self.beat, tol=self.s_tol)
                    x, p, a, _ = co.placement(self.phase[c], base, self.nseg, self.dl,
                                              self.angle[c], self.beat, s=s, frame=R)
                xs.append(x.reshape(-1, 3))
                vp.append(p)
                va.append(a)
        xb = np.concatenate([self.x[b] + self.blob_ref @ rot_mat(self.q[b]).T
                             for b in range(self.nbody)])
        return np.concatenate(xs), xb, np.array(vp), np.array(va)

    # ---- the system -------------------------------------------------------------------
    def layout(self, dynamic, dyn_angle):
        ns = self.nc * self.nseg
        n3 = 3 * (ns + self.nbody * self.nb)
        nx = self.nc * ((2 if dyn_angle else 1) if dynamic else 0)
        return ns, n3, n3 + 6 * self.nbody, n3 + 6 * self.nbody + nx

    def k_columns(self, xs, xb):
        """Dense K_B (3 M x 6 n_body)."""
        ns = xs.shape[0]
        X = np.vstack([xs, xb])
        K = np.zeros((3 * X.shape[0], 6 * self.nbody))
        for i in range(X.shape[0]):
            b = (i // self.nseg) // self.per if i < ns else (i - ns) // self.nb
            r = X[i] - self.x[b]
            K[3 * i:3 * i + 3, 6 * b:6 * b + 3] = np.eye(3)
            K[3 * i:3 * i + 3, 6 * b + 3:6 * b + 6] = rcross(r)   # Omega x r
        return K

    def system_matrix(self, M, xs, xb, vp, va, dynamic, dyn_angle):
        """The dense system matrix for the mobility M (3 M x 3 M, AoS)."""
        ns, n3, nu, L = self.layout(dynamic, dyn_angle)
        A = np.zeros((L, L))
        A[:n3, :n3] = M
        K = self.k_columns(xs, xb)
        A[:n3, n3:nu] = -K
        A[n3:nu, :n3] = -K.T
        if dynamic:
            cols = list(range(self.nc)) + (list(range(self.nc)) if dyn_angle else [])
            for j, c in enumerate(cols):
                v = np.zeros(n3)
                v[3 * c * self.nseg:3 * (c + 1) * self.nseg] = vp[c] if j < self.nc else va[c]
                A[:n3, nu + j] = -v
                A[nu + j, :n3] = -v
        if self.rot_prescribed:
            for b in range(self.nbody):
                for k in range(3, 6):
                    j = n3 + 6 * b + k
                    A[:, j] = 0.0
                    A[j, :] = 0.0
                    A[j, j] = 1.0
        return A

    def mobility(self, xs, xb):
        return oracle.mobility_matrix(np.vstack([xs, xb]), dof=3, wall=False, a=self.a, mu=self.mu)

    def rhs(self, vp, dynamic, dyn_angle):
        ns, n3, nu, L = self.layout(dynamic, dyn_angle)
        r = np.zeros(L)
        if not dynamic:
            for c in range(self.nc):
                r[3 * c * self.nseg:3 * (c + 1) * self.nseg] = self.phase_dot[c] * vp[c]
            return r
        qt, qa, spring = self.driving
        for c in range(self.nc):
            qp_, qa_ = co.driving_forces(self.phase[c], self.angle[c], self.omega0[c], qt, qa,
                                         spring, self.dl, self.nseg, self.fene, self.noise,
                                         self.seed, self.nt, c)
            r[nu + c] = -qp_
            if dyn_angle:
                r[nu + self.nc + c] = -qa_
        return r

    # ---- the reference's block preconditioner ----------------------------------------
    def precon_mat(self, xs, xb, vp, va, dynamic, dyn_angle):
        """make_precon_mat (swimmer.cpp:1175-1347): per body the inverse of
        seg_fac sum_seg K^T K + blob_fac sum_blob K^T K - seg_fac C_Ainv_B."""
        sf = bf = 6.0 * math.pi * self.mu * self.a
        out = []
        for b in range(self.nbody):
            A = np.zeros((6, 6))
            eye = (self.per * self.nseg * sf + self.nb * bf)
            A[:3, :3] = eye * np.eye(3)
            segs = range(b * self.per * self.nseg, (b + 1) * self.per * self.nseg)
            for i in segs:
                C = rcross(xs[i] - self.x[b])
                A[:3, 3:] += sf * C
                A[3:, 3:] -= sf * C @ C
            for i in range(self.nb):
                C = rcross(xb[b * self.nb + i] - self.x[b])
                A[:3, 3:] += bf * C
                A[3:, 3:] -= bf * C @ C
            A[3:, :3] = -A[:3, 3:]
            if dynamic:
                CAB = np.zeros((6, 6))
                for l in range(self.per):
                    c = b * self.per + l
                    kp = vp[c].reshape(-1, 3)
                    ka = va[c].reshape(-1, 3)
                    r = xs[c * self.nseg:(c + 1) * self.nseg] - self.x[b]
                    vphi, uphi = kp.sum(0), np.cross(r, kp).sum(0)
                    vth, uth = ka.sum(0), np.cross(r, ka).sum(0)
                    c11, c22, c12 = np.sum(kp * kp), np.sum(ka * ka), np.sum(kp * ka)
                    if dyn_angle:
                        det = c11 * c22 - c12 * c12
                        i11, i22, i12 = c22 / det, c11 / det, -c12 / det
                        B11 = (np.outer(vphi, vphi) * i11 + np.outer(vphi, vth) * i12
                               + np.outer(vth, vphi) * i12 + np.outer(vth, vth) * i22)
                        B21 = (np.outer(uphi, vphi) * i11 + np.outer(uphi, vth) * i12
                               + np.outer(uth, vphi) * i12 + np.outer(uth, vth) * i22)
                        B22 = (np.outer(uphi, uphi) * i11 + np.outer(uphi, uth) * i12
                               + np.outer(uth, uphi) * i12 + np.outer(uth, uth) * i22)
                    else:
                        B11 = np.outer(vphi, vphi) / c11
                        B21 = np.outer(uphi, vphi) / c11
                        B22 = np.outer(uphi, uphi) / c11
                    CAB[:3, :3] += B11
                    CAB[3:, :3] += B21
                    CAB[3:, 3:] += B22
                CAB[:3, 3:] = CAB[3:, :3].T
                A -= sf * CAB
            out.append(np.linalg.inv(A))
        return out

    def precondition(self, r, xs, xb, vp, va, dynamic, dyn_angle, kinv=None):
        """apply_preconditioner_new, prescribed cilia on free bodies
        (mobility_solver.cpp:1104-1515), in the reference's order of operations."""
        sm = bm = 6.0 * math.pi * self.mu * self.a
        ns, n3, nu, L = self.layout(dynamic, dyn_angle)
        nc, nseg = self.nc, self.nseg
        if kinv is None:
            kinv = self.precon_mat(xs, xb, vp, va, dynamic, dyn_angle)
        out = np.array(r, dtype=np.float64)
        out[:3 * ns] *= sm
        out[3 * ns:n3] *= bm
        out[n3:nu] *= -1.0
        ph = nu
        for b in range(self.nbody):
            oid = n3 + 6 * b
            segs = range(b * self.per * nseg, (b + 1) * self.per * nseg)
            for i in segs:
                f = out[3 * i:3 * i + 3].copy()
                d = xs[i] - self.x[b]
                out[oid:oid + 3] -= f
                out[oid + 3:oid + 6] -= np.cross(d, f)
            for i in range(self.nb):
                j = ns + b * self.nb + i
                f = out[3 * j:3 * j + 3].copy()
                d = xb[b * self.nb + i] - self.x[b]
                out[oid:oid + 3] -= f
                out[oid + 3:oid + 6] -= np.cross(d, f)
            sums = []
            if dynamic:
                cs = range(b * self.per, (b + 1) * self.per)
                out[ph + cs.start:ph + cs.stop] *= -1.0
                if dyn_angle:
                    out[ph + nc + cs.start:ph + nc + cs.stop] *= -1.0
                for c in cs:
                    kp = vp[c].reshape(-1, 3)
                    ka = va[c].reshape(-1, 3)
                    f = out[3 * c * nseg:3 * (c + 1) * nseg].reshape(-1, 3)
                    out[ph + c] -= np.sum(kp * f)
                    if dyn_angle:
                        out[ph + nc + c] -= np.sum(ka * f)
                    d = xs[c * nseg:(c + 1) * nseg] - self.x[b]
                    vphi, uphi = kp.sum(0), np.cross(d, kp).sum(0)
                    vth, uth = ka.sum(0), np.cross(d, ka).sum(0)
                    c11, c22, c12 = np.sum(kp * kp), np.sum(ka * ka), np.sum(kp * ka)
                    sums.append((c, vphi, uphi, vth, uth, c11, c22, c12))
                    if not dyn_angle:
                        out[ph + c] /= c11
                        out[oid:oid + 3] -= out[ph + c] * vphi
                        out[oid + 3:oid + 6] -= out[ph + c] * uphi
                    else:
                        det = c11 * c22 - c12 * c12
                        i11, i22, i12 = c22 / det, c11 / det, -c12 / det
                        nphi = i11 * out[ph + c] + i12 * out[ph + nc + c]
                        nth = i12 * out[ph + c] + i22 * out[ph + nc + c]
                        out[ph + c], out[ph + nc + c] = nphi, nth
                        out[oid:oid + 3] -= out[ph + c] * vphi + out[ph + nc + c] * vth
                        out[oid + 3:oid + 6] -= out[ph + c] * uphi + out[ph + nc + c] * uth
            out[oid:oid + 6] = kinv[b] @ out[oid:oid + 6]
            if dynamic:
                for (c, vphi, uphi, vth, uth, c11, c22, c12) in sums:
                    out[ph + c] /= sm
                    if dyn_angle:
                        out[ph + nc + c] /= sm
                    V, W = out[oid:oid + 3], out[oid + 3:oid + 6]
                    if not dyn_angle:
                        out[ph + c] -= (vphi @ V + uphi @ W) / c11
                    else:
                        det = c11 * c22 - c12 * c12
                        i11, i22, i12 = c22 / det, c11 / det, -c12 / det
                        out[ph + c] -= (i11 * (vphi @ V) + i12 * (vth @ V) + i11 * (uphi @ W)
                                        + i12 * (uth @ W))
                        out[ph + nc + c] -= (i12 * (vphi @ V) + i22 * (vth @ V) + i12 * (uphi @ W)
                                             + i22 * (uth @ W))
            V, W = out[oid:oid + 3], out[oid + 3:oid + 6]
            for i in segs:
                c = i // nseg
                d = xs[i] - self.x[b]
                add = V + np.cross(W, d)
                if dynamic:
                    add = add + out[ph + c] * vp[c][3 * (i % nseg):3 * (i % nseg) + 3]
                    if dyn_angle:
                        add = add + out[ph + nc + c] * va[c][3 * (i % nseg):3 * (i % nseg) + 3]
                out[3 * i:3 * i + 3] += sm * add
            for i in range(self.nb):
                j = ns + b * self.nb + i
                d = xb[b * self.nb + i] - self.x[b]
                out[3 * j:3 * j + 3] += bm * (V + np.cross(W, d))
        return out

    # ---- time stepping --------------------------------------------------------------
    def step(self, nt):
        """One step (explicit update first for nt > 0); returns (solution, q_phase, q_angle)."""
        dynamic = self.driving is not None
        dyn_angle = dynamic and self.driving[1] is not None
        self.nt = nt
        if nt > 0:
            self.phase += self.dt * self.phase_dot
            if dyn_angle:
                self.angle += self.dt * self.angle_dot
            for b in range(self.nbody):
                self.x[b] += self.dt * self.V[b, :3]
                self.q[b] = qmul(lie_exp(self.dt * self.V[b, 3:]), self.q[b])
        xs, xb, vp, va = self.geometry()
        A = self.system_matrix(self.mobility(xs, xb), xs, xb, vp, va, dynamic, dyn_angle)
        z = np.linalg.solve(A, self.rhs(vp, dynamic, dyn_angle))
        ns, n3, nu, L = self.layout(dynamic, dyn_angle)
        self.V = z[n3:nu].reshape(-1, 6).copy()
        if dynamic:
            self.phase_dot = z[nu:nu + self.nc].copy()
            if dyn_angle:
                self.angle_dot = z[nu + self.nc:].copy()
        f = z[:3 * ns].reshape(self.nc, -1)
        qp = np.array([vp[c] @ f[c] for c in range(self.nc)])
        qa = np.array([va[c] @ f[c] for c in range(self.nc)])
        self.last = dict(z=z, xs=xs, xb=xb, vp=vp, va=va, A=A)
        return z, qp, qa
