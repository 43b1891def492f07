"""The reference's on-disk formats, so that its drivers and analysis scripts can read
this library's runs and hand it their inputs.

* ``globals.ini``: sections ``[Parameters]`` (numbers) and ``[Filenames]`` (paths), read
  key by key by ``data_from_ini`` (cilia_sim_main.cu:57-94).
* ``psi.dat``: one line, ``k T phase_1 .. phase_NFIL [angle_1 .. angle_NFIL]`` -- the
  initial state the periodic-orbit and bisection drivers write with ``np.savetxt(...,
  newline=" ")`` (find_periodic/newton_solver.py:49-74, bisection/bisection.py:150-162).
* ``<sim>.par``: ``value %% NAME`` lines in the order of cilia_sim_main.cu:215-257.
* per-step output files (globals.cu:51-62), each line ``step`` then the values in
  scientific notation: ``_body_states.dat`` (x, q per swimmer, rigid_body.cpp:565-569),
  ``_seg_states.dat`` (per segment: position for prescribed cilia, quaternion for
  elastic filaments, filament.cpp:2210-2226), ``_true_states.dat`` (``step PERIOD``,
  phases, then shape angles, cilia_sim_main.cu:670-700), and with OUTPUT_FORCES
  ``_seg_vels.dat``, ``_seg_forces.dat`` (6 per segment), ``_blob_forces.dat`` (3 per
  blob), ``_body_vels.dat`` (6 per swimmer) (mobility_solver.cpp:3150-3278).
"""
from __future__ import annotations

import configparser
import os

import numpy as np

# .par entries in the order cilia_sim_main.cu:218-257 writes them
PAR_ORDER = ["NFIL", "NSEG", "NSWIM", "NBLOB", "MU", "RSEG", "DL", "FIL_LENGTH", "RBLOB", "KB",
             "KT", "TOL", "TOTAL_TIME_STEPS", "DT", "PLOT_FREQUENCY_IN_STEPS", "STEPS_PER_PERIOD",
             "PRESCRIBED_CILIA"]


def read_globals(path):
    """{'Parameters': {name: float}, 'Filenames': {name: str}} of a globals.ini."""
    cp = configparser.ConfigParser()
    cp.optionxform = str
    with open(path) as fh:
        cp.read_file(fh)
    out = {"Parameters": {}, "Filenames": {}}
    if cp.has_section("Parameters"):
        for k, v in cp.items("Parameters"):
            out["Parameters"][k] = float(v)
    if cp.has_section("Filenames"):
        out["Filenames"] = dict(cp.items("Filenames"))
    return out


def write_globals(path, parameters, filenames=None):
    """A globals.ini the reference executable reads (other sections left empty)."""
    with open(path, "w") as fh:
        fh.write("[Parameters]\n")
        for k, v in parameters.items():
            fh.write(f"{k}={float(v)!r}\n")
        fh.write("\n[Filenames]\n")
        for k, v in (filenames or {}).items():
            fh.write(f"{k}={v}\n")
        fh.write("\n[Box]\n\n[Hex]\n\n[Concentric]\n\n[Seeding_util]\n")


def read_psi(path, nfil, with_angles=None):
    """(k, T, phases, angles or None) of a psi.dat initial state."""
    v = np.loadtxt(path, ndmin=1)
    k, T = float(v[0]), float(v[1])
    rest = v[2:]
    if with_angles is None:
        with_angles = rest.size >= 2 * nfil
    phases = rest[:nfil].copy()
    angles = rest[nfil:2 * nfil].copy() if with_angles else None
    return k, T, phases, angles


def write_psi(path, k, T, phases, angles=None):
    x = np.concatenate([[k, T], np.asarray(phases, float)]
                       + ([np.asarray(angles, float)] if angles is not None else []))
    np.savetxt(path, x, newline=" ")


def _fmt(values, digits):
    return " ".join(f"{v:.{digits}e}" for v in np.asarray(values, float).reshape(-1))


class OutputWriter:
    """Appends the reference's per-step output lines under ``sim_dir + sim_file``."""

    SUFFIX = {"par": ".par", "body_states": "_body_states.dat", "seg_states": "_seg_states.dat",
              "body_vels": "_body_vels.dat", "seg_vels": "_seg_vels.dat",
              "blob_forces": "_blob_forces.dat", "seg_forces": "_seg_forces.dat",
              "true_states": "_true_states.dat"}

    def __init__(self, sim_dir, sim_file, output_digit=15, force_digit=10):
        self.name = os.path.join(sim_dir, sim_file)
        os.makedirs(sim_dir or ".", exist_ok=True)
        self.digits = output_digit
        self.force_digits = force_digit

    def path(self, kind):
        return self.name + self.SUFFIX[kind]

    def _append(self, kind, step, values, digits, head=()):
        parts = [str(int(step))] + [f"{h:.{digits}e}" for h in head]
        body = _fmt(values, digits)
        with open(self.path(kind), "a") as fh:
            fh.write(" ".join(parts) + " " + body + " \n")

    def write_par(self, values):
        """``value %% NAME`` lines: PAR_ORDER first, then any further entries."""
        keys = [k for k in PAR_ORDER if k in values] + [k for k in values if k not in PAR_ORDER]
        with open(self.path("par"), "w") as fh:
            for k in keys:
                fh.write(f"{values[k]} %% {k}\n")

    def states(self, step, seg_states, body_x=None, body_q=None):
        """seg_states: positions (prescribed cilia) or quaternions (elastic filaments),
        one row per segment; body_x / body_q: one row per swimmer."""
        self._append("seg_states", step, seg_states, self.digits)
        if body_x is not None:
            bx = np.asarray(body_x, float).reshape(-1, 3)
            bq = np.asarray(body_q, float).reshape(-1, 4)
            self._append("body_states", step, np.hstack([bx, bq]), self.digits)

    def true_states(self, step, period, phases, angles=None):
        vals = np.concatenate([np.asarray(phases, float)]
                              + ([np.asarray(angles, float)] if angles is not None else []))
        self._append("true_states", step, vals, self.digits, head=(period,))

    def velocities_forces(self, step, v_seg, f_seg, f_blob=None, v_body=None):
        self._append("seg_vels", step, v_seg, self.force_digits)
        self._append("seg_forces", step, f_seg, self.force_digits)
        if f_blob is not None:
            self._append("blob_forces", step, f_blob, self.force_digits)
        if v_body is not None:
            self._append("body_vels", step, v_body, self.force_digits)


def read_lines(path):
    """(steps, values) of a per-step output file."""
    steps, rows = [], []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            steps.append(int(tok[0]))
            rows.append(np.array(tok[1:], dtype=float))
    return np.array(steps), rows


def read_par(path):
    out = {}
    with open(path) as fh:
        for line in fh:
            if " %% " in line:
                v, k = line.rstrip("\n").split(" %% ", 1)
                out[k.strip()] = float(v)
    return out
