"""numpy restatement of the prescribed-cilium path -- TEST INFRASTRUCTURE ONLY.

Follows the reference's prescribed-shape cilium (CILIA_TYPE 3): the fitted beats of
SHAPE_SEQUENCE 1 (described below), 4 (unaltered coefficients), 3 (volvox, four
polynomial degrees) and 2 (coral larvae), and the build-a-beat tangent-angle model of
SHAPE_SEQUENCE 0 (build_a_beat_angle, build_a_beat_placement).  For SHAPE_SEQUENCE 1:
  * Fourier/polynomial beat of Fulford & Blake as fitted by Ito, Omori & Ishikawa
    (2019, Table 1), rescaled to unit mean length with the zero mode halved
    (filament.cpp:684-758 coefficients, :964-1045 shape, tangent and phase derivative);
  * arc-length parameters s_n of the segments: trapezoid curve length on 10 NSEG
    panels and bisection to within 0.1/NSEG of n/(NSEG-1) (filament.cpp:1047-1160);
  * placement above the wall: base at (0, 0, DL/2), reference x axis along the wall
    normal (the quaternion of filament.cpp:225-226 for dir = e_z), segment n at
    base + FIL_LENGTH R shape(s_n, psi) with FIL_LENGTH = DL (NSEG-1), velocity
    directions FIL_LENGTH R d shape/d psi and FIL_LENGTH R (e_z x shape)
    (filament.cpp:1336-1490, config.hpp:397,579);
  * the resistance solve M F = phase_dot vel_dir_phase over the segments (3-DOF,
    wall-corrected RPY) and the generalised forces sum_n vel_dir . F_n
    (mobility_solver.cpp:301-420, cilia_sim_main.cu:418-438);
  * the coupled system with phase (and shape-angle) speeds as unknowns, driven by the
    recorded generalised forces through periodic cubic Hermite interpolation
    (mobility_solver.cpp:242-356, 1891-2115).

Parity status: pinned by the reference's own output files
input/forcing/fulford_and_blake_reference_{phase,angle}_generalised_forces_NSEG=20_SEP=2.600000.dat
and, for the unaltered coefficients (SHAPE_SEQUENCE 4),
input/forcing/fulford_and_blake_original_reference_{phase,angle}_generalised_forces_...dat
and its recorded segment parameters ..._s_values_NSEG=20_SEP=2.600000.dat
(copied under tests/golden/) -- both directly and in closed loop (driving a cilium
with its own recorded forces returns its beat speed) -- by the arc-length property of
the s-values, by finite differences of the shape in the phase, and by node, periodicity
and slope properties of the interpolation (tests/test_cilia_oracle.py).
The other beats have no reference output: their shape, tangent and phase direction
agree by finite differences, the coral-larvae table has the unit period-mean length its
source states, and the build-a-beat angle is pinned by continuity at its switch points,
its range [theta_s, pi - theta_s], finite differences in the phase, the chord length of
two unit tangents and the derivatives of the placement in phase and angle.  Every
coefficient table equals, value for value, the assignments of filament.cpp:684-960 as
tests/golden/extract_beat_tables.py reads them into tests/golden/beat_tables.json.
Paired cilia (pair_placement) are two single-cilium placements; the pairing is pinned by
the reference's phase rules at the values that make the two cilia coincide in shape
(pair_dp 0 or 1 for BICILIA, 1 for BICILIA_LONGT: the second cilium is the first moved
by 2.2 a), by the second cilium equalling a lone volvox cilium at the second phase, and
by the offset being 2.2 a along the wall frame's third axis, rebuilt from the
quaternion of filament.cpp:225-226.
"""
from __future__ import annotations

import math

import numpy as np

# Ito, Omori & Ishikawa (2019) Table 1 as used by the reference: rows = Fourier
# mode, columns = polynomial degree (s, s^2, s^3); x is the filament axis, y the
# beat direction.
AX = np.array([[9.7204e-01, -2.8315e-01, 4.9243e-02],
               [-1.8466e-02, -1.2926e-01, 2.6981e-01],
               [1.6209e-01, -3.4983e-01, 1.9082e-01],
               [1.0259e-02, 3.5907e-02, -6.8736e-02]])
AY = np.array([[-3.3547e-01, 4.0369e-01, 1.0362e-01],
               [4.0318e-01, -1.5553e+00, 7.3455e-01],
               [-9.9513e-02, 3.2829e-02, -1.2106e-01],
               [8.1046e-02, -3.0982e-01, 1.4568e-01]])
BX = np.array([[1.9697e-01, -5.1193e-01, 3.4778e-01],
               [-5.1295e-02, 4.3396e-01, -3.3547e-01],
               [1.2311e-02, 1.4157e-01, -1.1695e-01]])
BY = np.array([[2.9136e-01, 1.0721e+00, -1.0433e+00],
               [6.1554e-03, 3.2521e-01, -2.8315e-01],
               [-6.0528e-02, 2.3185e-01, -2.0108e-01]])

# The unaltered Table 1 values (SHAPE_SEQUENCE 4, filament.cpp:813-868): the zero
# cosine mode enters with weight 0.5 (filament.cpp:977-979, 1006-1008) and the mean
# length is 0.97475 rather than 1.
AX0 = np.array([[1.895, -0.552, 0.096], [-0.018, -0.126, 0.263], [0.158, -0.341, 0.186],
                [0.010, 0.035, -0.067]])
AY0 = np.array([[-0.654, 0.787, 0.202], [0.393, -1.516, 0.716], [-0.097, 0.032, -0.118],
                [0.079, -0.302, 0.142]])
BX0 = np.array([[0.192, -0.499, 0.339], [-0.050, 0.423, -0.327], [0.012, 0.138, -0.114]])
BY0 = np.array([[0.284, 1.045, -1.017], [0.006, 0.317, -0.276], [-0.059, 0.226, -0.196]])

# The volvox fit (SHAPE_SEQUENCE 3, filament.cpp:748-812): four polynomial degrees.
AXV = np.array([[1.00490080, -0.19958656, -0.20422343, 0.10887946],
                [0.00314934, -0.22239636, 1.18659976, -0.87409016],
                [0.18942274, -1.24525818, 1.94063304, -0.78632572],
                [-0.11162123, 1.07262966, -2.15142914, 1.20544741]])
AYV = np.array([[-0.01665535, -0.27457190, 1.08406999, -0.56102820],
                [0.17505739, -2.09412076, 1.99556763, -0.50179439],
                [-0.12408600, 1.62438155, -3.21015214, 1.70178281],
                [-0.16672661, 0.90896038, -1.50867276, 0.74387795]])
BXV = np.array([[0.09140608, -0.54804089, 1.40643137, -0.60109684],
                [-0.09670739, 1.14634486, -2.60675971, 1.55525152],
                [-0.16451150, 0.98846449, -1.49147276, 0.64896090]])
BYV = np.array([[-0.06343445, 2.15509751, -4.19609063, 2.05489100],
                [-0.28828336, 1.45223183, -1.56580487, 0.44541361],
                [0.02187088, -0.33070891, 0.94906955, -0.66157093]])

# The coral-larvae fit (SHAPE_SEQUENCE 2, filament.cpp:869-893), scaled to unit mean length.
AXC = np.array([[9.034276e-01, -6.634681e-01, 2.016123e-01],
                [-2.468656e-01, 1.215869e+00, -5.867236e-01],
                [3.075836e-01, -6.966322e-01, 4.554729e-01],
                [2.378489e-02, -1.355582e-01, 7.315831e-02]])
AYC = np.array([[-2.766076e-01, 6.368662e-01, -8.513947e-02],
                [3.091205e-01, -1.993712e+00, 1.091847e+00],
                [-2.952120e-01, 7.189021e-01, -6.328589e-01],
                [-9.868018e-02, 3.312812e-01, -2.578206e-01]])
BXC = np.array([[1.336807e-01, -6.807852e-01, 6.008592e-01],
                [8.146824e-02, 3.472676e-01, -2.744220e-01],
                [3.615272e-02, 8.619119e-02, -6.122992e-02]])
BYC = np.array([[9.165546e-01, 8.848315e-02, -4.838832e-01],
                [-2.946211e-01, 1.057964e+00, -6.335721e-01],
                [6.122744e-02, -2.319094e-01, 1.906292e-01]])

# beat name -> (Ax, Ay, Bx, By, weight of the zero cosine mode); columns = degrees
BEATS = {"fulford_blake": (AX, AY, BX, BY, 1.0),
         "fulford_blake_original": (AX0, AY0, BX0, BY0, 0.5),
         "volvox": (AXV, AYV, BXV, BYV, 1.0),
         "coral_larvae": (AXC, AYC, BXC, BYC, 1.0)}


def _modes(psi, c0=1.0):
    cv = np.array([c0] + [math.cos(n * psi) for n in range(1, 4)])
    sv = np.array([math.sin(n * psi) for n in range(1, 4)])
    return cv, sv


def tangent(s, psi, beat="fulford_blake"):
    """d shape / d s in the beat plane (filament.cpp:964-988)."""
    ax, ay, bx, by, c0 = BEATS[beat]
    cv, sv = _modes(psi, c0)
    p = np.array([(d + 1) * s ** d for d in range(ax.shape[1])])
    return (cv @ ax + sv @ bx) @ p, (cv @ ay + sv @ by) @ p


def shape(s, psi, beat="fulford_blake"):
    """Reference-frame position at parameter s (filament.cpp:990-1019)."""
    ax, ay, bx, by, c0 = BEATS[beat]
    cv, sv = _modes(psi, c0)
    p = np.array([s ** (d + 1) for d in range(ax.shape[1])])
    return np.array([(cv @ ax + sv @ bx) @ p, (cv @ ay + sv @ by) @ p, 0.0])


def shape_phase_derivative(s, psi, beat="fulford_blake"):
    """d shape / d psi (filament.cpp:1021-1045)."""
    ax, ay, bx, by, _ = BEATS[beat]
    dc = np.array([0.0] + [-n * math.sin(n * psi) for n in range(1, 4)])
    ds = np.array([n * math.cos(n * psi) for n in range(1, 4)])
    p = np.array([s ** (d + 1) for d in range(ax.shape[1])])
    return np.array([(dc @ ax + ds @ bx) @ p, (dc @ ay + ds @ by) @ p, 0.0])


# Build-a-beat (SHAPE_SEQUENCE 0) with the reference's constants (config.hpp:133-139).
BUILD_AMPLITUDE, BUILD_RECOVERY, BUILD_EFFECTIVE, BUILD_SHIFT = 1.99, 1.0 / 12.0, 0.75, 0.05


def build_a_beat_angle(s, psi):
    """Tangent angle theta(s, psi) of the build-a-beat model and d theta / d psi: theta
    moves between theta_s = acos(amplitude / 2) and pi - theta_s through
    delta = t - sin t cos t on the effective stroke and the recovery window, whose start
    travels along the filament (filament.cpp:1240-1330)."""
    th_s = math.acos(0.5 * BUILD_AMPLITUDE)
    w_r, w_e = 2 * math.pi * BUILD_RECOVERY, 2 * math.pi * BUILD_EFFECTIVE
    trans = w_e + s * (2 * math.pi - w_e - w_r)
    sp = psi - 2 * math.pi * BUILD_SHIFT * s
    sp -= 2 * math.pi * math.floor(sp / (2 * math.pi))
    if sp < w_e:
        t = math.pi * (w_e - sp) / w_e
        delta, d_delta = t - math.sin(t) * math.cos(t), -2 * math.sin(t) ** 2 / w_e
    elif sp < trans:
        delta, d_delta = 0.0, 0.0
    elif sp < trans + w_r:
        t = math.pi * (sp - trans) / w_r
        delta, d_delta = t - math.sin(t) * math.cos(t), 2 * math.sin(t) ** 2 / w_r
    else:
        delta, d_delta = math.pi, 0.0
    span = math.pi - 2 * th_s
    return th_s + span * delta / math.pi, span * d_delta


def build_a_beat_placement(psi, base, nseg, dl, angle, frame):
    """Segments at s = n / (nseg - 1) joined by trapezoid steps of the unit tangent
    (sin theta, cos theta, 0) turned by frame Rshape; the phase direction the same sum
    of d tangent / d psi (filament.cpp:1401-1546).  Two departures from the reference,
    where its placement applies the frame twice to the base tangent and turns the lab
    displacement by Rshape for the angle direction: the frame enters once, and the angle
    direction is frame (e_z x frame^T (x_n - x_0)), as for the fitted beats."""
    ca, sa = math.cos(angle), math.sin(angle)
    R = frame @ np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])

    def tk(s):
        th, dth = build_a_beat_angle(s, psi)
        return (R @ np.array([math.sin(th), math.cos(th), 0.0]),
                R @ (dth * np.array([math.cos(th), -math.sin(th), 0.0])))

    x = np.zeros((nseg, 3))
    vphase = np.zeros((nseg, 3))
    vangle = np.zeros((nseg, 3))
    x[0] = base
    t1, k1 = tk(0.0)
    for n in range(1, nseg):
        t2, k2 = tk(n / (nseg - 1))
        x[n] = x[n - 1] + 0.5 * dl * (t1 + t2)
        vphase[n] = vphase[n - 1] + 0.5 * dl * (k1 + k2)
        ref = frame.T @ (x[n] - x[0])
        vangle[n] = frame @ np.array([-ref[1], ref[0], 0.0])
        t1, k1 = t2, k2
    return x, vphase, vangle, np.arange(nseg) / (nseg - 1)


class _Panels:
    """Tangent norms at the panel nodes k / (10 nseg) of one phase and their running
    trapezoid sums, accumulated in the same order as a loop over the panels, so that
    curve_length gives the same bits with or without the table."""

    def __init__(self, psi, nseg, beat):
        self.psi, self.beat = psi, beat
        self.dl = 1.0 / (10 * nseg)
        self.f = [math.hypot(*tangent(0.0, psi, beat))]
        self.P = [0.0]

    def upto(self, k):
        while len(self.f) <= k:
            n = len(self.f)
            self.f.append(math.hypot(*tangent(n * self.dl, psi=self.psi, beat=self.beat)))
            self.P.append(self.P[-1] + 0.5 * self.dl * (self.f[n - 1] + self.f[n]))


def curve_length(s, psi, nseg=20, beat="fulford_blake", panels=None):
    """Trapezoid arc length from 0 to s on 10 nseg panels, the last panel
    interpolated (filament.cpp:1047-1094).  panels: a _Panels of the same phase, reused
    between calls."""
    if s <= 0.0:
        return 0.0
    t = _Panels(psi, nseg, beat) if panels is None else panels
    dl = t.dl
    fl = math.floor(s / dl)
    ce = math.ceil(s / dl)
    t.upto(max(fl, ce))
    length = t.P[fl]
    if ce > fl:
        length += ((s / dl - fl) / (ce - fl)) * 0.5 * dl * (t.f[fl] + t.f[ce])
    return length


def s_values(psi, nseg=20, beat="fulford_blake", tol=None, margins=False):
    """Segment parameters by bisection on the arc-length fraction to within tol
    (default 0.1/nseg) (filament.cpp:1096-1160).  With margins=True also returns, per
    segment, how close any of its bisection comparisons came to flipping (the smaller
    of | |frac - target| - tol | and |frac - target|)."""
    tol = 0.1 / nseg if tol is None else tol
    if beat == "build_a_beat":      # uniform arc fraction, no search
        s = np.arange(nseg) / (nseg - 1)
        return (s, np.full(nseg, np.inf)) if margins else s
    s = [0.0] * nseg
    s[-1] = 1.0
    m = np.full(nseg, np.inf)
    panels = _Panels(psi, nseg, beat)
    total = curve_length(1.0, psi, nseg, beat, panels)
    for n in range(1, nseg - 1):
        target = n / (nseg - 1)
        lo, hi = s[n - 1], 1.0
        est = 0.5 * (lo + hi)
        frac = curve_length(est, psi, nseg, beat, panels) / total
        m[n] = min(abs(abs(frac - target) - tol), abs(frac - target))
        while abs(frac - target) > tol:
            if frac > target:
                hi = est
            else:
                lo = est
            est = 0.5 * (lo + hi)
            frac = curve_length(est, psi, nseg, beat, panels) / total
            m[n] = min(m[n], abs(abs(frac - target) - tol), abs(frac - target))
        s[n] = est
    return (np.array(s), m) if margins else np.array(s)


def s_values_from_table(table, psi, nseg):
    """Segment parameters read from a recorded table (rows: phases 2 pi k / rows,
    columns: the recorded segments), bilinear in phase and in segment position
    (filament.cpp:1190-1235)."""
    table = np.asarray(table, dtype=np.float64)
    rows, cols = table.shape
    x = 0.5 * psi / math.pi
    x = (x - math.floor(x)) * rows
    lo = int(x)
    if lo == rows:
        svec = table[0]
    else:
        up = table[0] if lo == rows - 1 else table[lo + 1]
        svec = table[lo] + (up - table[lo]) * (x - lo)
    s = np.zeros(nseg)
    s[-1] = 1.0
    for n in range(1, nseg - 1):
        si = (cols - 1) * n / (nseg - 1)
        k = int(si)
        s[n] = svec[k] + (si - k) * (svec[k + 1] - svec[k])
    return s


# Rotation taking the reference x axis to the wall normal e_z: the square root of
# the quaternion (0, 0, -1, 0) (filament.cpp:225-226 with dir = e_z).
FRAME = np.array([[0.0, 0.0, -1.0],
                  [0.0, 1.0, 0.0],
                  [1.0, 0.0, 0.0]])


def placement(psi, base, nseg=20, dl=2.6, angle=0.0, beat="fulford_blake", s=None, frame=None):
    """Positions, phase and angle velocity directions of one cilium (AoS, 3 nseg);
    segment parameters by bisection unless given; frame: the cilium's rotation (3 x 3,
    default the wall frame FRAME)."""
    FRAME_ = FRAME if frame is None else np.asarray(frame, dtype=np.float64)
    if beat == "build_a_beat":
        x, vp, va, s = build_a_beat_placement(psi, np.asarray(base, dtype=np.float64), nseg,
                                              dl, angle, FRAME_)
        return x.reshape(-1), vp.reshape(-1), va.reshape(-1), s
    L = dl * (nseg - 1)
    s = s_values(psi, nseg, beat) if s is None else np.asarray(s)
    ca, sa = math.cos(angle), math.sin(angle)
    Rshape = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    base = np.asarray(base, dtype=np.float64)
    x = np.zeros((nseg, 3))
    vphase = np.zeros((nseg, 3))
    vangle = np.zeros((nseg, 3))
    x[0] = base
    for n in range(1, nseg):
        ref = L * (Rshape @ shape(s[n], psi, beat))
        x[n] = base + FRAME_ @ ref
        vangle[n] = FRAME_ @ np.array([-ref[1], ref[0], 0.0])   # e_z x ref
        vphase[n] = FRAME_ @ (L * (Rshape @ shape_phase_derivative(s[n], psi, beat)))
    return x.reshape(-1), vphase.reshape(-1), vangle.reshape(-1), s


# ---- paired cilia (SHAPE_SEQUENCE 5 / 6, BICILIA / BICILIA_LONGT) ---------------------
# Each filament of NSEG segments is two cilia of NSEG_PER_CILIA = NSEG / 2
# (cilia_sim_main.cu:110-116) beating the volvox fit (filament.cpp:895-960 holds the
# same table as :748-812); the second cilium at phase psi2 (filament.cpp:371-375,
# 1344-1348) is placed by fitted_shape(s, psi2, z_dis / FIL_LENGTH), whose third
# component z_dis = 2.2 RSEG moves it along the frame's z axis (filament.cpp:990-1019,
# 1429-1466); FIL_LENGTH, the bisection panels and tolerance use NSEG_PER_CILIA
# (filament.cpp:1057,1145, config.hpp:579).
PAIR_SEP = 2.2


def pair_phase(psi, pair_dp, longt=False):
    """Phase of the second cilium: psi + 2 pi pair_dp (BICILIA) or pair_dp psi (LONGT)."""
    return pair_dp * psi if longt else psi + pair_dp * 2.0 * math.pi


def pair_placement(psi, base, nseg=20, dl=2.6, angle=0.0, pair_dp=0.0, longt=False, a=1.0,
                   frame=None, s=None):
    """Positions, phase and angle velocity directions (AoS, 3 nseg) and segment parameters
    (nseg) of one pair of cilia of nseg // 2 segments: the first at phase psi from base,
    the second at pair_phase(psi) from base + PAIR_SEP a frame[:, 2]; s: both cilia's
    parameters (nseg), by bisection when None."""
    m = nseg // 2
    F = FRAME if frame is None else np.asarray(frame, dtype=np.float64)
    base = np.asarray(base, dtype=np.float64)
    s1 = None if s is None else np.asarray(s)[:m]
    s2 = None if s is None else np.asarray(s)[m:]
    x1, vp1, va1, s1 = placement(psi, base, m, dl, angle, "volvox", s=s1, frame=F)
    x2, vp2, va2, s2 = placement(pair_phase(psi, pair_dp, longt), base + PAIR_SEP * a * F[:, 2],
                                 m, dl, angle, "volvox", s=s2, frame=F)
    return (np.concatenate([x1, x2]), np.concatenate([vp1, vp2]), np.concatenate([va1, va2]),
            np.concatenate([s1, s2]))


def generalised_forces(mobility, psi, phase_dot=2.0 * math.pi, base=(0.0, 0.0, 1.3), nseg=20,
                       dl=2.6, beat="fulford_blake"):
    """Phase and angle generalised forces of one cilium: F = M^-1 (phase_dot vel_dir_phase)
    with M = mobility(x) the 3-DOF wall mobility matrix of its segments."""
    x, vp, va, _ = placement(psi, base, nseg, dl, beat=beat)
    f = np.linalg.solve(mobility(x), phase_dot * vp)
    return float(vp @ f), float(va @ f)


# ---- dynamic phase / shape-angle evolution (the coupled prescribed system) -----------

def cyclic_hermite(table, phase):
    """Periodic cubic Hermite interpolation of a table sampled at phase = 2 pi k / n
    with central-difference slopes (a one-sided slope at node 0), as
    mobility_solver.cpp:242-281 does."""
    n = len(table)
    x = 0.5 * phase / math.pi
    x = (x - math.floor(x)) * n
    lo = int(x)
    if lo == n:
        return float(table[0])
    hi = 0 if lo == n - 1 else lo + 1
    d_lo = (table[1] - table[0]) if lo == 0 else 0.5 * (table[hi] - table[lo - 1])
    d_hi = (table[0] - table[n - 1]) if hi == 0 else 0.5 * (table[0 if hi == n - 1 else hi + 1] - table[lo])
    t = x - lo
    t2, t3 = t * t, t * t * t
    return float(table[lo] * (2 * t3 - 3 * t2 + 1) + d_lo * (t3 - 2 * t2 + t)
                 + table[hi] * (3 * t2 - 2 * t3) + d_hi * (t3 - t2))


_M64 = (1 << 64) - 1


def mix64(z):
    """SplitMix64 finaliser (Steele, Lea & Flood 2014) on 64-bit integers."""
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def normal_deviate(seed, nt, c):
    """N(0, 1) deviate of (seed, step, cilium): two 53-bit uniforms on (0, 1) from the
    mixed key, then Box-Muller."""
    key = mix64((seed ^ mix64(((nt & 0xFFFFFFFF) << 32) | (c & 0xFFFFFFFF))) & _M64)
    h1, h2 = mix64(key), mix64((key + 1) & _M64)
    u1 = ((h1 >> 11) + 0.5) * 2.0 ** -53
    u2 = ((h2 >> 11) + 0.5) * 2.0 ** -53
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def fene_factor(angle, max_angle=0.78):
    """Angle factor of the FENE torsional spring (mobility_solver.cpp:320-330)."""
    if abs(angle) > max_angle:
        return 0.5 * angle * (1.0 + angle ** 2 / max_angle ** 2 * math.exp(abs(angle) - max_angle))
    return angle


def driving_forces(phase, angle, omega0, q_phase_table, q_angle_table, spring=0.0, dl=2.6,
                   nseg=20, fene=False, noise=0.0, seed=0, nt=0, c=0):
    """Generalised driving forces of cilium c at step nt (mobility_solver.cpp:306-356):
    the interpolated recorded forces scaled by omega0 / 2 pi, the torsional spring
    (linear, or the FENE form) and noise * N(0, 1) on the phase drive."""
    L = dl * (nseg - 1)
    qp = cyclic_hermite(q_phase_table, phase) * 0.5 * omega0 / math.pi
    if noise != 0.0:
        qp += noise * normal_deviate(seed, nt, c)
    qa = None
    if q_angle_table is not None:
        qa = cyclic_hermite(q_angle_table, phase) * 0.5 * omega0 / math.pi
        qa -= spring * omega0 * L ** 3 * (fene_factor(angle) if fene else angle)
    return qp, qa


def coupled_solve(M, vphase, vangle, q_phase, q_angle=None):
    """Dense solve of the coupled system of mobility_solver.cpp:1891-2115 for cilia
    above the wall (body velocities prescribed):
        M F - sum_c (phi_dot_c K_c + psi_dot_c A_c) = 0,
        -K_c . F = -q_phase_c,   -A_c . F = -q_angle_c,
    with K_c, A_c the phase and angle velocity directions of cilium c (3 nseg each,
    zero outside its own segments).  Returns (F, phi_dot, psi_dot or None)."""
    n = M.shape[0]
    nc = len(vphase)
    cols = list(vphase) + (list(vangle) if q_angle is not None else [])
    k = len(cols)
    A = np.zeros((n + k, n + k))
    A[:n, :n] = M
    for j, v in enumerate(cols):
        A[:n, n + j] = -v
        A[n + j, :n] = -v
    rhs = np.zeros(n + k)
    rhs[n:n + nc] = -np.asarray(q_phase)
    if q_angle is not None:
        rhs[n + nc:] = -np.asarray(q_angle)
    z = np.linalg.solve(A, rhs)
    return z[:n], z[n:n + nc], (z[n + nc:] if q_angle is not None else None)
