"""Host-side driver of the implicit filament step (swimmer/filament/broyden layer).

``FilamentSystem`` plays the part of the reference's time loop around
``swimmer``, ``filament`` and ``broyden_solver`` (src/cilia/cilia_sim_main.cu:405-552):
it owns a mobility context, initialises the filament state and advances it one
implicit step at a time; every step runs on the GPU through the C ABI.

Physical parameters default to the reference's elastic (instability-driven)
filaments: KB = KT = 1800, DIMENSIONLESS_FORCE = 220 with
END_FORCE_MAGNITUDE = DIMENSIONLESS_FORCE KB / (DL^2 NSEG^2) (globals.cu:18),
REPULSIVE_FORCE_FACTOR = 2, TOL = 1e-4, JACOBIAN_CONFIDENCE_FACTOR = 0.1,
MAX_BROYDEN_ITER = 400, NUM_EULER_STEPS = 1, DT = 36.3833 / 300 (the single-
filament beat period over 300 steps, config.hpp:463).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from . import geometry as geo
from .mobility import RpyMobilitySolver

DEFAULTS = dict(nseg=20, max_broyden=400, num_euler_steps=1, barrier=1, dl=2.2,
                dt=36.3833 / 300.0, kb=1800.0, kt=1800.0, dimensionless_force=220.0,
                repulsive_factor=2.0, tol=1e-4, gamma=0.1)


def end_force(dimensionless_force, kb, dl, nseg):
    return dimensionless_force * kb / (dl * dl * nseg * nseg)


class FilamentSystem:
    """Filaments (n_fil x nseg segments) plus an optional fixed rigid body of blobs."""

    def __init__(self, x_seg, tangents, *, x_blob=None, precision="fp64", solve_mode="fp64",
                 blob_rtol=1e-10, blob_restart=50, blob_max_iter=2000, blob_precond="diag",
                 guess="previous", pert=None, device=0, tile_split=0, wall=True, overlap=False,
                 broyden_only=False, **params):
        prm = dict(DEFAULTS)
        prm.update(params)
        self.params = prm
        nseg = prm["nseg"]
        x_seg = L.f64(x_seg)
        n_seg = x_seg.size // 3
        n_blob = 0 if x_blob is None else L.f64(x_blob).size // 3
        self.n_seg, self.n_blob, self.nseg, self.nfil = n_seg, n_blob, nseg, n_seg // nseg
        self.mob = RpyMobilitySolver(n_seg, n_blob, seg_dof=6, wall=wall, precision=precision,
                                     device=device, tile_split=tile_split)
        self.mob.set_positions(x_seg, None if x_blob is None else L.f64(x_blob))
        q = geo.segment_quaternions(tangents, nseg)
        P = L.cfm_fil_params()
        P.nseg = nseg
        P.max_broyden = prm["max_broyden"]
        P.num_euler_steps = prm["num_euler_steps"]
        P.barrier = int(prm["barrier"])
        P.dl, P.dt, P.kb, P.kt = prm["dl"], prm["dt"], prm["kb"], prm["kt"]
        P.end_force = prm.get("end_force") or end_force(prm["dimensionless_force"], prm["kb"],
                                                        prm["dl"], nseg)
        P.repulsive_factor = prm["repulsive_factor"]
        P.tol = prm["tol"]
        P.gamma = prm["gamma"]
        P.pair_precision = RpyMobilitySolver._prec(precision)
        P.guess = {"previous": L.CFM_GUESS_PREVIOUS, "learned": L.CFM_GUESS_LEARNED,
                   "projected": L.CFM_GUESS_PROJECTED}[guess]
        P.blob_solve = L.cfm_solve_opts(system=L.CFM_SYS_BLOB,
                                        precond=RpyMobilitySolver._PC[blob_precond],
                                        mode=RpyMobilitySolver._MODE[solve_mode],
                                        restart=blob_restart, max_iter=blob_max_iter,
                                        use_guess=1, rtol=blob_rtol)
        P.overlap = int(bool(overlap))
        # blob forces (and attached bodies' states) among the Broyden unknowns: the
        # reference's SOLVER_TYPE 0
        P.broyden_only = int(bool(broyden_only))
        self.broyden_only = bool(broyden_only) and n_blob > 0
        self._P = P
        self._x0 = (x_seg, None if x_blob is None else L.f64(x_blob))
        self._q0 = q
        pert_arr = None if pert is None else L.f64(pert, 3 * self.nfil)
        L.check(self.h, L.lib().cfm_fil_init(self.h, ctypes.byref(P), L.ptr(q), L.ptr(pert_arr)))
        self.nt = 0
        self.members = 1

    def set_members(self, members):
        """The filaments as `members` equal independent wall-bounded systems stepped
        together (cfm_fil_set_members): member e owns filaments [e F, (e+1) F) of the
        stacked configuration, F = n_fil / members, and each evolves as if alone."""
        L.check(self.h, L.lib().cfm_fil_set_members(self.h, int(members)))
        self.members = int(members)

    def member_stats(self):
        """Per-member statistics of the last step (cfm_fil_member_stats)."""
        st = (L.cfm_step_stats * self.members)()
        L.check(self.h, L.lib().cfm_fil_member_stats(self.h, st))
        return [s.as_dict() for s in st]

    def restart(self, pert=None):
        """Back to the initial configuration (positions, orientations, zero multipliers)
        with a new tip perturbation; everything else the context holds (blob
        preconditioners, recorded samples, the network) is kept."""
        self.mob.set_positions(*self._x0)
        pert_arr = None if pert is None else L.f64(pert, 3 * self.nfil)
        L.check(self.h, L.lib().cfm_fil_init(self.h, ctypes.byref(self._P), L.ptr(self._q0),
                                              L.ptr(pert_arr)))
        self.nt = 0
        if self.members > 1:
            L.check(self.h, L.lib().cfm_fil_set_members(self.h, self.members))

    @property
    def h(self):
        return self.mob.handle

    # -- free swimmers (swimmer.cpp, rigid_body.cpp) ---------------------------------
    def attach_bodies(self, blob_ref, x_body, q_body, *, f_ext=None, mobility=None,
                      ref_precond="dense", ref_rtol=1e-10, body_motion="free"):
        """The blobs become n_body free rigid bodies (blob_ref: body-frame offsets of one
        body, x_body 3 n_body, q_body 4 n_body) carrying the filaments (n_fil / n_body
        per body); their positions and orientations join the Broyden unknowns.  The body
        mobility for the approximate Jacobian is computed (make_body_reference_matrices)
        unless given (6 x 6, body frame).  body_motion "prescribed_rotation" holds the
        bodies' rotation at zero (the reference's BODY_VELOCITY_TYPE 2)."""
        x_body = L.f64(x_body)
        n_body = x_body.size // 3
        self.mob.body_init(blob_ref, n_body)
        self.mob.body_set_state(x_body, q_body)
        self.mob.body_set_motion(body_motion)
        if mobility is None:
            self.mob.make_body_reference_matrices(0, precond=ref_precond, rtol=ref_rtol)
        else:
            self.mob.body_set_reference_matrices(None, L.f64(mobility, 36))
        fe = None if f_ext is None else L.f64(f_ext, 6 * n_body)
        L.check(self.h, L.lib().cfm_fil_attach_bodies(self.h, L.ptr(fe)))
        self.n_body = n_body

    def set_body_sphere(self, radius):
        """Treat the bodies as spheres of this radius in the approximate Jacobian (the
        reference's sphere correction); 0 turns it off."""
        L.check(self.h, L.lib().cfm_fil_set_body_sphere(self.h, float(radius)))

    def write_outputs(self, writer, step):
        """One line of the reference's _seg_states.dat (segment quaternions) and, with
        bodies attached, _body_states.dat (centre, quaternion) through a simio.OutputWriter."""
        st = self.state()
        if getattr(self, "n_body", 0):
            b = self.body_state()
            writer.states(step, st["q"].reshape(-1, 4), b["x"], b["q"])
        else:
            writer.states(step, st["q"].reshape(-1, 4))

    def body_state(self):
        nb = self.n_body
        x, q, u, v = np.zeros(3 * nb), np.zeros(4 * nb), np.zeros(3 * nb), np.zeros(6 * nb)
        L.check(self.h, L.lib().cfm_fil_get_body_state(self.h, L.ptr(x), L.ptr(q), L.ptr(u),
                                                        L.ptr(v)))
        return {"x": x, "q": q, "u": u, "vel": v}

    def close(self):
        self.mob.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- time stepping ----------------------------------------------------------
    def step(self, nt=None):
        nt = self.nt if nt is None else nt
        st = L.cfm_step_stats()
        L.check(self.h, L.lib().cfm_fil_step(self.h, int(nt), ctypes.byref(st)))
        self.nt = nt + 1
        return st.as_dict()

    def state(self):
        ns, nf = self.n_seg, self.nfil
        x, q, u, lam, cl = (np.zeros(3 * ns), np.zeros(4 * ns), np.zeros(3 * ns),
                            np.zeros(3 * ns), np.zeros(3 * nf))
        L.check(self.h, L.lib().cfm_fil_get_state(self.h, L.ptr(x), L.ptr(q), L.ptr(u),
                                                   L.ptr(lam), L.ptr(cl)))
        return dict(x=x, q=q, u=u, lam=lam, clamp=cl)

    def set_state(self, *, q=None, qm1=None, u=None, um1=None, lam=None, clamp=None, xm1=None,
                  xm2=None):
        a = [None if v is None else L.f64(v) for v in (q, qm1, u, um1, lam, clamp, xm1, xm2)]
        L.check(self.h, L.lib().cfm_fil_set_state(self.h, *[L.ptr(v) for v in a]))

    def blob_rhs_solution_pairs(self):
        """Number of (rhs, solution) pairs recorded by the learned correction."""
        return self.mob.gnn_samples()

    # -- pieces (parity checks) ---------------------------------------------------
    def internal_forces(self, nt):
        f = np.zeros(6 * self.n_seg)
        L.check(self.h, L.lib().cfm_fil_internal_forces(self.h, int(nt), L.ptr(f)))
        return f

    def barrier_forces(self):
        f = np.zeros(3 * self.n_seg)
        L.check(self.h, L.lib().cfm_barrier_forces(self.h, L.ptr(f)))
        return f

    def jacobian_inverse(self, nt):
        N = 6 * self.nseg
        J = np.zeros(self.nfil * N * N)
        L.check(self.h, L.lib().cfm_fil_jacobian_inverse(self.h, int(nt), L.ptr(J)))
        return J.reshape(self.nfil, N, N)

    @property
    def n_unknowns(self):
        """Length of the Broyden vectors: 6 per segment, plus 6 per attached body, plus 3
        per blob with broyden_only."""
        return 6 * self.n_seg + 6 * getattr(self, "n_body", 0) + (3 * self.n_blob if self.broyden_only else 0)

    def blob_forces(self):
        """Blob forces (n_blob, 3): the Broyden state with broyden_only, else the last
        blob solve's."""
        f = np.zeros(3 * self.n_blob)
        L.check(self.h, L.lib().cfm_fil_get_blob_forces(self.h, L.ptr(f)))
        return f.reshape(-1, 3)

    def errors(self, nt, v_seg):
        e = np.zeros(self.n_unknowns)
        bad = ctypes.c_double()
        L.check(self.h, L.lib().cfm_fil_errors(self.h, int(nt), L.ptr(L.f64(v_seg, 6 * self.n_seg)),
                                                L.ptr(e), ctypes.byref(bad)))
        return e, bad.value

    def update(self, upd):
        L.check(self.h, L.lib().cfm_fil_update(self.h, L.ptr(L.f64(upd, self.n_unknowns))))

    def initial_guess(self, nt):
        L.check(self.h, L.lib().cfm_fil_initial_guess(self.h, int(nt)))
