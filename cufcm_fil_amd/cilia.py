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
        beats = {"fulford_blake": L.CFM_BEAT_FULFORD_BLAKE,
                 "fulford_blake_original": L.CFM_BEAT_FULFORD_BLAKE_ORIGINAL}
        self.prm = L.cfm_cilia_params(n_cilia=n, nseg=nseg, beat=beats[beat], dl=dl, dt=dt,
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
