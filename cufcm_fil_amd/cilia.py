"""Prescribed-shape cilia above the wall (the reference's CILIA_TYPE 3 path in its
generalised-force mode): fitted Fulford-Blake beat at prescribed phase speeds, one
wall-RPY resistance solve per step, phase and angle generalised forces per cilium.

Mirrors filament::initial_setup / initial_guess (src/cilia/filament.cpp:225-470,
1336-1490) and the time-loop body of src/cilia/cilia_sim_main.cu:405-438.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as L
from .mobility import RpyMobilitySolver


def wall_lattice(nx, ny, spacing, dl=2.6):
    """Bases on a rectangular lattice centred at the origin, DL/2 above the wall
    (the reference's rectangular seeding, swimmer.cpp:173-200, config.hpp:397)."""
    im, jm = 0.5 * (nx - 1), 0.5 * (ny - 1)
    b = [((i - im) * spacing, (j - jm) * spacing, 0.5 * dl) for i in range(nx) for j in range(ny)]
    return np.array(b, dtype=np.float64).reshape(-1)


BEATS = ("fulford_blake", "fulford_blake_original", "volvox", "coral_larvae", "build_a_beat",
         "bicilia", "bicilia_longt")


def _beat(name):
    """Beat name -> CFM_BEAT_* (the reference's SHAPE_SEQUENCE 1, 4, 3, 2, 0, 5 and 6;
    the last two are paired cilia: each unit of nseg segments is two cilia of nseg / 2)."""
    if name not in BEATS:
        raise ValueError(f"unknown beat {name!r}; one of {BEATS}")
    return BEATS.index(name) + 1


def pair_phase_speeds(n_cilia, omega0=2.0 * math.pi, pair_dp=1.0):
    """Intrinsic phase speeds of the reference's PAIR 1 mode (filament.cpp:240-244): the
    second half of the filaments beats at pair_dp times the frequency of the first."""
    w = np.full(int(n_cilia), float(omega0))
    npair = int(n_cilia) // 2           # NPAIR = NFIL / 2; filament id with id / NPAIR == 1
    w[npair:2 * npair] *= pair_dp
    return w


class PrescribedCilia:
    """n cilia of nseg segments; base (3 n), phase and phase_dot (n)."""

    def __init__(self, base, phase=None, phase_dot=None, *, nseg=20, dl=2.6, dt=1.0 / 300,
                 precision="fp64", solve_mode="fp64", precond="self", rtol=1e-10, restart=50,
                 max_iter=1000, device=0, tile_split=0, beat="fulford_blake", s_tol=0.0):
        base = L.f64(base)
        n = base.size // 3
        self.n_cilia, self.nseg = n, nseg
        phase = np.zeros(n) if phase is None else L.f64(phase, n)
        phase_dot = np.full(n, 2.0 * math.pi) if phase_dot is None else L.f64(phase_dot, n)
        self.mob = RpyMobilitySolver(n * nseg, 0, seg_dof=3, wall=True, precision=precision,
                                     device=device, tile_split=tile_split)
        o = L.cfm_solve_opts(system=L.CFM_SYS_SEG, precond=RpyMobilitySolver._PC[precond],
                             mode=RpyMobilitySolver._MODE[solve_mode], restart=restart,
                             max_iter=max_iter, use_guess=1, rtol=rtol)
        self.prm = L.cfm_cilia_params(n_cilia=n, nseg=nseg, beat=_beat(beat), dl=dl, dt=dt,
                                      s_tol=s_tol, solve=o)
        L.check(self.mob._h, L.lib().cfm_cilia_init(self.mob._h, ctypes.byref(self.prm),
                                                    L.ptr(base), L.ptr(phase),
                                                    L.ptr(phase_dot)))

    def close(self):
        self.mob.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_driving(self, q_phase_table, q_angle_table=None, spring=0.0):
        """Dynamic phase (and shape-angle) evolution driven by recorded generalised
        forces of one beat (tables sampled at phase 2 pi k / n)."""
        qp = L.f64(q_phase_table)
        qa = None if q_angle_table is None else L.f64(q_angle_table)
        L.check(self.mob._h, L.lib().cfm_cilia_set_driving(
            self.mob._h, L.ptr(qp), qp.size, L.ptr(qa), 0 if qa is None else qa.size,
            float(spring)))

    def set_driving_options(self, fene=False, noise=0.0, seed=0):
        """FENE form of the shape-angle spring and seeded noise on the phase drive
        (FENE_MODEL, FORCE_NOISE_MAG)."""
        L.check(self.mob._h, L.lib().cfm_cilia_set_driving_options(self.mob._h, int(bool(fene)),
                                                                   float(noise), int(seed)))

    def set_s_table(self, table=None):
        """Segment parameters from a recorded table (rows: phases 2 pi k / rows), as the
        reference reads its s_values file; None returns to the bisection."""
        if table is None:
            L.check(self.mob._h, L.lib().cfm_cilia_set_s_table(self.mob._h, None, 0, 0))
            return
        t = np.ascontiguousarray(table, dtype=np.float64)
        rows, cols = t.shape
        L.check(self.mob._h, L.lib().cfm_cilia_set_s_table(self.mob._h, L.ptr(t.reshape(-1)),
                                                           rows, cols))

    def set_phases(self, phase=None, angle=None):
        """Replace phases and / or shape angles (e.g. from a psi.dat state)."""
        n = self.n_cilia
        ph = None if phase is None else L.f64(phase, n)
        an = None if angle is None else L.f64(angle, n)
        L.check(self.mob._h, L.lib().cfm_cilia_set_phases(self.mob._h, L.ptr(ph), L.ptr(an)))

    def set_pair_dp(self, pair_dp):
        """PAIR_DP of the paired beats (globals.ini pair_dp): the second cilium's phase
        offset in cycles ("bicilia") or phase factor ("bicilia_longt")."""
        L.check(self.mob._h, L.lib().cfm_cilia_set_pair_dp(self.mob._h, float(pair_dp)))

    def write_outputs(self, writer, step, period=1.0):
        """One line of the reference's _seg_states.dat (positions) and _true_states.dat
        (period, phases, shape angles) through a simio.OutputWriter."""
        st = self.state()
        sp = self.speeds()
        writer.states(step, st["x"].reshape(-1, 3))
        writer.true_states(step, period, sp["phase"], sp["angle"])

    def speeds(self):
        n = self.n_cilia
        ph, pd, an, ad = (np.zeros(n) for _ in range(4))
        L.check(self.mob._h, L.lib().cfm_cilia_get_speeds(self.mob._h, L.ptr(ph), L.ptr(pd),
                                                          L.ptr(an), L.ptr(ad)))
        return {"phase": ph, "phase_dot": pd, "angle": an, "angle_dot": ad}

    def step(self, nt):
        """One time step; returns (q_phase, q_angle, solve stats)."""
        qp, qa = np.zeros(self.n_cilia), np.zeros(self.n_cilia)
        st = L.cfm_solve_stats()
        L.check(self.mob._h, L.lib().cfm_cilia_step(self.mob._h, int(nt), L.ptr(qp), L.ptr(qa),
                                                    ctypes.byref(st)))
        return qp, qa, st.as_dict()

    def state(self):
        ns = self.n_cilia * self.nseg
        x, vp, va, f = (np.zeros(3 * ns) for _ in range(4))
        s, ph = np.zeros(ns), np.zeros(self.n_cilia)
        L.check(self.mob._h, L.lib().cfm_cilia_get_state(self.mob._h, L.ptr(x), L.ptr(vp),
                                                         L.ptr(va), L.ptr(s), L.ptr(ph), L.ptr(f)))
        return {"x": x, "vel_dir_phase": vp, "vel_dir_angle": va,
                "s_values": s.reshape(self.n_cilia, self.nseg), "phase": ph, "f": f}


def _quat_from_matrix(R):
    """Unit quaternion (w, x, y, z), w >= 0, of a rotation matrix (Shepperd's method)."""
    t = np.trace(R)
    if t > 0.0:
        w = 0.5 * math.sqrt(1.0 + t)
        q = np.array([w, (R[2, 1] - R[1, 2]) / (4 * w), (R[0, 2] - R[2, 0]) / (4 * w),
                      (R[1, 0] - R[0, 1]) / (4 * w)])
    else:
        i = int(np.argmax(np.diag(R)))
        j, k = (i + 1) % 3, (i + 2) % 3
        r = math.sqrt(max(1.0 + R[i, i] - R[j, j] - R[k, k], 0.0))
        v = np.zeros(3)
        v[i] = 0.5 * r
        v[j] = (R[j, i] + R[i, j]) / (2 * r)
        v[k] = (R[k, i] + R[i, k]) / (2 * r)
        w = (R[k, j] - R[j, k]) / (2 * r)
        q = np.array([w, *v])
    q /= np.linalg.norm(q)
    return q if q[0] >= 0.0 else -q


def cilium_frames(directions, tilt=0.0):
    """Frame quaternions (w, x, y, z) of cilia whose bases point along `directions`
    (body frame, n x 3): the first axis along the direction (the filament axis), the
    second along the posterior polar direction -- the projection of -e_z -- turned by
    `tilt` about the first (filament.cpp:370-432).  At the poles, where the polar
    direction vanishes, the second axis is e_y."""
    out = []
    for d in np.asarray(directions, dtype=np.float64).reshape(-1, 3):
        d = d / np.linalg.norm(d)
        polar = np.array([0.0, 0.0, -1.0]) + d[2] * d
        n = np.linalg.norm(polar)
        e2 = polar / n if n > 0.0 else np.array([0.0, 1.0, 0.0])
        e2 = math.cos(tilt) * e2 + math.sin(tilt) * np.cross(d, e2)
        R = np.column_stack([d, e2, np.cross(d, e2)])
        out.append(_quat_from_matrix(R))
    return np.array(out).reshape(-1)


class CiliatedBodies:
    """n_body free rigid bodies of blobs, each carrying the same arrangement of
    prescribed cilia (the reference's PRESCRIBED_CILIA swimmers with free bodies).

    blob_ref: body-frame blob offsets of one body (3 nb); fil_ref, fil_q: base offsets
    (3 per cilium) and frame quaternions (4 per cilium, see cilium_frames) of one body's
    cilia; body_x, body_q: centres and unit quaternions (w, x, y, z).  Phases and phase
    speeds cover all n_body x per cilia.  Every step solves for the forces, the bodies'
    velocities and, with set_driving, the phase (and shape-angle) speeds; later steps
    first move phases and bodies explicitly with the previous step's speeds.
    body_motion "prescribed_rotation" holds the bodies' rotation at zero (the reference's
    BODY_VELOCITY_TYPE 2)."""

    def __init__(self, blob_ref, fil_ref, fil_q, body_x, body_q, phase=None, phase_dot=None, *,
                 nseg=20, dl=2.6, dt=1.0 / 300, wall=False, precision="fp64", solve_mode="fp64",
                 precond="diag", rtol=1e-10, restart=50, max_iter=2000, device=0, tile_split=0,
                 beat="fulford_blake", s_tol=0.0, body_motion="free"):
        blob_ref = L.f64(blob_ref)
        fil_ref, fil_q = L.f64(fil_ref), L.f64(fil_q)
        body_x, body_q = L.f64(body_x), L.f64(body_q)
        nbody, per, nb = body_x.size // 3, fil_ref.size // 3, blob_ref.size // 3
        n = nbody * per
        self.n_body, self.per_body, self.n_cilia, self.nseg, self.nb = nbody, per, n, nseg, nb
        phase = np.zeros(n) if phase is None else L.f64(phase, n)
        phase_dot = np.full(n, 2.0 * math.pi) if phase_dot is None else L.f64(phase_dot, n)
        self.mob = RpyMobilitySolver(n * nseg, nbody * nb, seg_dof=3, wall=wall, precision=precision,
                                     device=device, tile_split=tile_split)
        # provisional positions: segments at the bases' bodies, blobs placed
        R = [_rot(body_q[4 * b:4 * b + 4]) for b in range(nbody)]
        x_blob = np.concatenate([body_x[3 * b:3 * b + 3] + blob_ref.reshape(-1, 3) @ R[b].T
                                 for b in range(nbody)])
        x_seg = np.repeat(body_x.reshape(-1, 3), per * nseg, axis=0)
        self.mob.set_positions(x_seg, x_blob)
        self.mob.body_init(blob_ref, nbody)
        self.mob.body_set_state(body_x, body_q)
        self.mob.body_set_motion(body_motion)
        o = L.cfm_solve_opts(system=L.CFM_SYS_SEG, precond=RpyMobilitySolver._PC[precond],
                             mode=RpyMobilitySolver._MODE[solve_mode], restart=restart,
                             max_iter=max_iter, use_guess=1, rtol=rtol)
        self.prm = L.cfm_cilia_params(n_cilia=n, nseg=nseg, beat=_beat(beat), dl=dl, dt=dt,
                                      s_tol=s_tol, solve=o)
        L.check(self.mob._h, L.lib().cfm_cilia_init_bodies(self.mob._h, ctypes.byref(self.prm),
                                                           L.ptr(fil_ref), L.ptr(fil_q),
                                                           L.ptr(phase), L.ptr(phase_dot)))

    def close(self):
        self.mob.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    set_driving = PrescribedCilia.set_driving
    set_driving_options = PrescribedCilia.set_driving_options
    set_phases = PrescribedCilia.set_phases
    set_pair_dp = PrescribedCilia.set_pair_dp
    speeds = PrescribedCilia.speeds
    step = PrescribedCilia.step
    state = PrescribedCilia.state

    def bodies(self):
        """Body centres, quaternions, velocities (V, Omega) of the last step, blob forces."""
        nb = self.n_body
        x, q, v = np.zeros(3 * nb), np.zeros(4 * nb), np.zeros(6 * nb)
        fb = np.zeros(3 * nb * self.nb)
        L.check(self.mob._h, L.lib().cfm_cilia_get_bodies(self.mob._h, L.ptr(x), L.ptr(q),
                                                          L.ptr(v), L.ptr(fb)))
        return {"x": x, "q": q, "vel": v, "f_blob": fb}

    def write_outputs(self, writer, step, period=1.0):
        """_seg_states.dat (segment positions), _body_states.dat (centre, quaternion) and
        _true_states.dat lines through a simio.OutputWriter."""
        st, sp, b = self.state(), self.speeds(), self.bodies()
        writer.states(step, st["x"].reshape(-1, 3), b["x"], b["q"])
        writer.true_states(step, period, sp["phase"], sp["angle"])


def _rot(q):
    s, a, b, c = q
    return np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - s * c), 2 * (a * c + s * b)],
                     [2 * (a * b + s * c), 1 - 2 * (a * a + c * c), 2 * (b * c - s * a)],
                     [2 * (a * c - s * b), 2 * (b * c + s * a), 1 - 2 * (a * a + b * b)]])


class CiliaEnsemble:
    """n_members copies of one lattice of dynamic-phase cilia above the wall, stepped
    together (cfm_cilia_batch_*): the same bases, beat and driving tables, each member
    with its own phases, natural frequencies, shape angles and time step.  One step
    advances every member as PrescribedCilia.step does for one, with one product launch
    and one lockstep GMRES (a captured hipGraph per Arnoldi step) for all of them."""

    def __init__(self, base, n_members, q_phase_table, q_angle_table=None, *, spring=0.0,
                 nseg=20, dl=2.6, precond="diag", rtol=1e-10, restart=50, max_iter=1000,
                 beat="fulford_blake", s_tol=0.0, mu=1.0, fene=False, noise=0.0, seed=0,
                 device=0):
        base = L.f64(base)
        self.n_cilia, self.nseg, self.n_members = base.size // 3, nseg, int(n_members)
        o = L.cfm_solve_opts(system=L.CFM_SYS_SEG, precond=RpyMobilitySolver._PC[precond],
                             mode=L.CFM_SOLVE_FP64, restart=restart, max_iter=max_iter,
                             use_guess=1, rtol=rtol)
        prm = L.cfm_cilia_params(n_cilia=self.n_cilia, nseg=nseg, beat=_beat(beat), dl=dl,
                                 dt=0.0, s_tol=s_tol, solve=o)
        cfg = L.cfm_cilia_batch_config(n_members=self.n_members, device=device,
                                       spring_model=int(bool(fene)), mu=mu, spring=float(spring),
                                       noise=float(noise), seed=int(seed), prm=prm)
        qp = L.f64(q_phase_table)
        qa = None if q_angle_table is None else L.f64(q_angle_table)
        self.angles = qa is not None
        h = ctypes.c_void_p()
        rc = L.lib().cfm_cilia_batch_create(ctypes.byref(cfg), L.ptr(base), L.ptr(qp), qp.size,
                                            L.ptr(qa), 0 if qa is None else qa.size, ctypes.byref(h))
        L.check_cilia_batch(None, rc)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().cfm_cilia_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()

    def set_pair_dp(self, pair_dp):
        """PAIR_DP of the paired beats, for every member."""
        L.check_cilia_batch(self._h, L.lib().cfm_cilia_batch_set_pair_dp(self._h, float(pair_dp)))

    def set_state(self, phase, omega0, dt, angle=None):
        """phase, omega0, angle: (n_members, n_cilia); dt: n_members time steps."""
        n = self.n_members * self.n_cilia
        ph, om = L.f64(phase, n), L.f64(omega0, n)
        an = None if angle is None else L.f64(angle, n)
        d = L.f64(dt, self.n_members)
        L.check_cilia_batch(self._h, L.lib().cfm_cilia_batch_set_state(
            self._h, L.ptr(ph), L.ptr(om), L.ptr(an), L.ptr(d)))

    def step(self, nt):
        """One step of every member; returns the generalised forces (q_phase, q_angle),
        each (n_members, n_cilia), and per-member solve statistics."""
        n = self.n_members * self.n_cilia
        qp, qa = np.zeros(n), np.zeros(n)
        st = (L.cfm_batch_sys_stats * self.n_members)()
        L.check_cilia_batch(self._h, L.lib().cfm_cilia_batch_step(self._h, int(nt), L.ptr(qp), L.ptr(qa), st))
        stats = [{"iterations": s.iterations, "converged": bool(s.converged),
                  "rel_residual": s.rel_residual, "guess_rejected": bool(s.reserved0)} for s in st]
        shape = (self.n_members, self.n_cilia)
        return qp.reshape(shape), qa.reshape(shape), stats

    def speeds(self):
        n = self.n_members * self.n_cilia
        ph, pd, an, ad = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        L.check_cilia_batch(self._h, L.lib().cfm_cilia_batch_get_state(
            self._h, L.ptr(ph), L.ptr(pd), L.ptr(an), L.ptr(ad)))
        shape = (self.n_members, self.n_cilia)
        return {"phase": ph.reshape(shape), "phase_dot": pd.reshape(shape),
                "angle": an.reshape(shape), "angle_dot": ad.reshape(shape)}

    def solution(self):
        """Per member: segment forces (n_cilia nseg, 3), phase speeds, angle speeds."""
        ns = self.n_cilia * self.nseg
        lm = 3 * ns + self.n_cilia * (2 if self.angles else 1)
        z = np.zeros(self.n_members * lm)
        L.check_cilia_batch(self._h, L.lib().cfm_cilia_batch_get_solution(self._h, L.ptr(z)))
        z = z.reshape(self.n_members, lm)
        return z[:, :3 * ns].reshape(self.n_members, ns, 3), z[:, 3 * ns:]

    def positions(self):
        x = np.zeros(3 * self.n_members * self.n_cilia * self.nseg)
        L.check_cilia_batch(self._h, L.lib().cfm_cilia_batch_get_positions(self._h, L.ptr(x)))
        return x.reshape(self.n_members, -1, 3)
