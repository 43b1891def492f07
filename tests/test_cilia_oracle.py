"""CPU pins of the prescribed-cilium oracle: the reference's own generalised-force
files (tests/golden), the arc-length property of the segment parameters and finite
differences of the beat shape in the phase."""
import math
import os

import numpy as np
import pytest

import cilia_oracle as co
import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _golden(name):
    v = np.loadtxt(os.path.join(GOLD, name))
    assert int(v[0]) == 300 and v.size == 301
    return v[1:]


def test_generalised_forces_match_reference_files():
    """The reference wrote these in single precision with a GMRES tolerance of 1e-4;
    the FP64 oracle agrees to 1e-4 of the phase force at every sampled step and to
    1e-4 of the largest angle force."""
    ph = _golden("fb_phase_generalised_forces_nseg20_sep2.6.dat")
    an = _golden("fb_angle_generalised_forces_nseg20_sep2.6.dat")
    mob = lambda x: oracle.mobility_matrix(x, 3)
    for nt in range(0, 300, 7):
        qp, qa = co.generalised_forces(mob, 2.0 * math.pi * nt / 300)
        assert abs(qp - ph[nt]) <= 1e-4 * abs(ph[nt]), (nt, qp, ph[nt])
        assert abs(qa - an[nt]) <= 1e-4 * np.abs(an).max(), (nt, qa, an[nt])


def test_a_planted_geometry_error_is_detected():
    """A 1% longer filament moves the phase force well outside the tolerance."""
    ph = _golden("fb_phase_generalised_forces_nseg20_sep2.6.dat")
    x, vp, va, _ = co.placement(0.7, (0.0, 0.0, 1.3), dl=2.6 * 1.01)
    f = np.linalg.solve(oracle.mobility_matrix(x, 3), 2.0 * math.pi * vp)
    nt = round(0.7 * 300 / (2.0 * math.pi))
    qp_good, _ = co.generalised_forces(lambda y: oracle.mobility_matrix(y, 3),
                                       2.0 * math.pi * nt / 300)
    assert abs(qp_good - ph[nt]) <= 1e-4 * abs(ph[nt])
    assert abs(vp @ f - ph[nt]) > 1e-2 * abs(ph[nt])


@pytest.mark.parametrize("psi", [0.0, 1.1, 2.9, 4.4])
def test_s_values_split_the_curve_into_equal_arcs(psi):
    s = co.s_values(psi)
    total = co.curve_length(1.0, psi)
    frac = np.array([co.curve_length(v, psi) / total for v in s])
    target = np.arange(20) / 19.0
    assert np.all(np.abs(frac - target) <= 0.1 / 20 + 1e-15)
    assert np.all(np.diff(s) > 0)


@pytest.mark.parametrize("psi", [0.3, 2.0, 5.5])
def test_phase_velocity_direction_is_the_shape_derivative(psi):
    h = 1e-5
    for s in (0.2, 0.55, 1.0):
        fd = (co.shape(s, psi + h) - co.shape(s, psi - h)) / (2 * h)
        assert np.allclose(co.shape_phase_derivative(s, psi), fd, rtol=0, atol=1e-8)
        tfd = (co.shape(s + h, psi) - co.shape(s - h, psi)) / (2 * h)
        assert np.allclose(co.tangent(s, psi), tfd[:2], rtol=0, atol=1e-8)


def test_placement_geometry():
    x, vp, va, s = co.placement(1.0, (0.0, 0.0, 1.3))
    p = x.reshape(-1, 3)
    assert np.allclose(p[0], [0.0, 0.0, 1.3])
    assert np.all(p[:, 2] > 1.0)   # the beat stays above the wall
    # velocity directions vanish at the base; the angle direction is a rotation of
    # (x - base) about the reference frame's z axis, which is the lab's -x axis
    assert not vp[:3].any() and not va[:3].any()
    r = p - p[0]
    assert np.allclose(va.reshape(-1, 3), np.cross([-1.0, 0.0, 0.0], r))


def test_cyclic_hermite_interpolation():
    rng = np.random.default_rng(1)
    table = rng.standard_normal(40)
    n = table.size
    for k in (0, 1, 17, n - 1):
        assert co.cyclic_hermite(table, 2 * math.pi * k / n) == pytest.approx(table[k], abs=1e-14)
    # periodic in the phase, and exact for a constant table
    assert co.cyclic_hermite(table, 0.3) == pytest.approx(co.cyclic_hermite(table, 0.3 + 2 * math.pi), abs=1e-12)
    assert co.cyclic_hermite(np.full(7, 2.5), 1.234) == pytest.approx(2.5, abs=1e-14)
    # interior slopes are central differences: the derivative at node 5 is continuous
    h = 1e-7 * 2 * math.pi / n
    x5 = 2 * math.pi * 5 / n
    left = (co.cyclic_hermite(table, x5) - co.cyclic_hermite(table, x5 - h)) / h
    right = (co.cyclic_hermite(table, x5 + h) - co.cyclic_hermite(table, x5)) / h
    assert left == pytest.approx(right, rel=1e-4)
    assert right == pytest.approx(0.5 * (table[6] - table[4]) * n / (2 * math.pi), rel=1e-4)


@pytest.mark.parametrize("with_angle", [False, True])
def test_driving_a_cilium_with_its_own_forces_recovers_the_beat_speed(with_angle):
    """Closed loop through the reference's own files: they hold the generalised forces
    of a cilium beating at 2 pi with a fixed shape angle, so driving the coupled system
    with them must return phase speed 2 pi (and zero angle speed) to the files'
    single-precision / 1e-4 accuracy."""
    qp_tab = _golden("fb_phase_generalised_forces_nseg20_sep2.6.dat")
    qa_tab = _golden("fb_angle_generalised_forces_nseg20_sep2.6.dat") if with_angle else None
    for nt in (0, 61, 150, 233):
        psi = 2 * math.pi * nt / 300
        x, vp, va, _ = co.placement(psi, (0.0, 0.0, 1.3))
        qp, qa = co.driving_forces(psi, 0.0, 2 * math.pi, qp_tab, qa_tab)
        M = oracle.mobility_matrix(x, 3)
        F, phid, psid = co.coupled_solve(M, [vp], [va], [qp], None if qa is None else [qa])
        assert phid[0] == pytest.approx(2 * math.pi, rel=2e-4)
        if with_angle:
            assert abs(psid[0]) <= 2e-4 * 2 * math.pi


def _s_table():
    v = np.loadtxt(os.path.join(GOLD, "fb_original_s_values_nseg20_sep2.6.dat"))
    assert int(v[0]) == 300 and int(v[1]) == 20 and v.size == 2 + 300 * 20
    return v[2:].reshape(300, 20)


def test_original_beat_s_values_match_reference_file():
    """The reference's recorded segment parameters of the unaltered coefficient set
    (SHAPE_SEQUENCE 4) at sampled steps, with the bisection tolerance 0.001/NSEG the
    file was written with: within 1e-6 (six printed digits plus the single-precision
    rounding of the reference's midpoints; one bisection decision apart is 1.5e-5 or
    more) except where the reference's single-precision arithmetic took the other side
    of a comparison that FP64 puts within 1e-6 of its threshold (steps 20 and 68 here:
    9 of all 300 steps); the present source's 0.1/NSEG stops earlier (several 1e-3)."""
    tab = _s_table()
    flips = 0
    for k in list(range(0, 300, 23)) + [20, 68]:
        psi = 2.0 * math.pi * k / 300
        s, m = co.s_values(psi, 20, "fulford_blake_original", tol=0.001 / 20, margins=True)
        d = np.abs(s - tab[k])
        if d.max() > 1e-6:
            first = int(np.argmax(d > 1e-6))
            assert m[first] < 1e-6 and d[first] > 1e-5, (k, first, m[first])
            flips += 1
    assert flips == 2
    coarse = co.s_values(0.0, 20, "fulford_blake_original")
    assert np.abs(coarse - tab[0]).max() > 1e-3


def test_rescaled_table_is_the_original_over_its_mean_length():
    """filament.cpp:684-758 divides Table 1 by 0.97475 and halves the zero cosine mode
    into the table; both transcriptions agree to the five digits they carry."""
    w = np.ones((4, 1))
    w[0] = 0.5
    for a, a0 in [(co.AX, co.AX0), (co.AY, co.AY0)]:
        assert np.allclose(a, a0 * w / 0.97475, rtol=6e-3, atol=2e-4)
    for b, b0 in [(co.BX, co.BX0), (co.BY, co.BY0)]:
        assert np.allclose(b, b0 / 0.97475, rtol=6e-3, atol=2e-4)
    # same shape everywhere up to that rescaling
    for psi in (0.3, 2.0, 4.4):
        for s in (0.2, 0.7, 1.0):
            assert np.allclose(co.shape(s, psi), co.shape(s, psi, "fulford_blake_original") / 0.97475,
                               atol=2e-3)


def test_s_table_interpolation_rule():
    """filament.cpp:1190-1235: rows at phase 2 pi k / rows (periodic), columns at the
    recorded segments; a table with as many columns as segments returns its rows at the
    sampled phases, their mean half way between, and row 0 after the last row."""
    tab = _s_table()
    for k in (0, 17, 299):
        assert np.allclose(co.s_values_from_table(tab, 2 * math.pi * k / 300, 20), tab[k], atol=1e-12)
    mid = co.s_values_from_table(tab, 2 * math.pi * 10.5 / 300, 20)
    assert np.allclose(mid, 0.5 * (tab[10] + tab[11]), atol=1e-12)
    wrap = co.s_values_from_table(tab, 2 * math.pi * 299.5 / 300, 20)
    assert np.allclose(wrap, 0.5 * (tab[299] + tab[0]), atol=1e-12)
    # a 39-column table (every other column a midpoint) gives back the even columns
    fine = np.zeros((300, 39))
    fine[:, ::2] = tab
    fine[:, 1::2] = 0.5 * (tab[:, :-1] + tab[:, 1:])
    assert np.allclose(co.s_values_from_table(fine, 2 * math.pi * 5 / 300, 20), tab[5], atol=1e-12)


FITTED = ["fulford_blake", "fulford_blake_original", "volvox", "coral_larvae"]


@pytest.mark.parametrize("beat", FITTED)
def test_fitted_beats_evaluate_their_tables_consistently(beat):
    """Every fitted beat, the four-degree volvox one included: the shape starts at the
    base, its s-derivative is the tangent and its phase derivative the phase direction
    (a dropped or mis-weighted polynomial degree breaks the first or second)."""
    h = 1e-5
    for psi in (0.4, 2.5, 5.1):
        assert np.allclose(co.shape(0.0, psi, beat), 0.0)
        for s in (0.15, 0.5, 0.95):
            fd = (co.shape(s, psi + h, beat) - co.shape(s, psi - h, beat)) / (2 * h)
            assert np.allclose(co.shape_phase_derivative(s, psi, beat), fd, atol=1e-8)
            tfd = (co.shape(s + h, psi, beat) - co.shape(s - h, psi, beat)) / (2 * h)
            assert np.allclose(co.tangent(s, psi, beat), tfd[:2], atol=1e-8)


def test_coral_larvae_fit_has_unit_mean_length():
    """The reference scales the coral-larvae coefficients to a period-mean length of 1
    (filament.cpp:871); a sign error in any one coefficient moves it by more than 1e-3."""
    phases = 2 * math.pi * np.arange(64) / 64
    mean = lambda b: np.mean([co.curve_length(1.0, p, 100, b) for p in phases])
    assert abs(mean("coral_larvae") - 1.0) < 2e-4
    ax, ay, bx, by, c0 = co.BEATS["coral_larvae"]
    try:
        for k in range(3):
            bad = ay.copy()
            bad[k + 1, 1] *= -1.0
            co.BEATS["_planted"] = (ax, bad, bx, by, c0)
            assert abs(mean("_planted") - 1.0) > 1e-3
    finally:
        co.BEATS.pop("_planted", None)


def _bab_boundaries(s):
    w_r, w_e = 2 * math.pi * co.BUILD_RECOVERY, 2 * math.pi * co.BUILD_EFFECTIVE
    trans = w_e + s * (2 * math.pi - w_e - w_r)
    shift = 2 * math.pi * co.BUILD_SHIFT * s
    return [b + shift for b in (0.0, w_e, trans, trans + w_r)]


@pytest.mark.parametrize("s", [0.0, 0.3, 0.77, 1.0])
def test_build_a_beat_angle_is_continuous_and_its_phase_derivative_matches(s):
    th_s = math.acos(0.5 * co.BUILD_AMPLITUDE)
    eps = 1e-9
    for b in _bab_boundaries(s):
        lo, _ = co.build_a_beat_angle(s, b - eps)
        hi, _ = co.build_a_beat_angle(s, b + eps)
        assert abs(lo - hi) < 1e-7, (s, b)
    h = 1e-6
    for psi in np.linspace(0.05, 2 * math.pi - 0.05, 57):
        th, dth = co.build_a_beat_angle(s, psi)
        assert th_s - 1e-15 <= th <= math.pi - th_s + 1e-15
        fd = (co.build_a_beat_angle(s, psi + h)[0] - co.build_a_beat_angle(s, psi - h)[0]) / (2 * h)
        assert abs(fd - dth) < 1e-5 * (1 + abs(dth)), (s, psi)
    # the stroke sweeps the full range: theta_s at the end of the effective stroke,
    # pi - theta_s between recovery and the next effective stroke
    b = _bab_boundaries(s)
    assert abs(co.build_a_beat_angle(s, b[1])[0] - th_s) < 1e-12
    if s < 1.0:
        assert abs(co.build_a_beat_angle(s, b[3] + 1e-3)[0] - (math.pi - th_s)) < 1e-12


@pytest.mark.parametrize("psi", [0.5, 3.0, 5.0, 5.9])
def test_build_a_beat_placement_derivatives(psi):
    """Segments are joined by the chord of two unit tangents (length dl cos of half
    their angle); the phase and angle directions are the derivatives of the placement in
    the phase and in the shape angle."""
    base, nseg, dl, ang = np.array([0.3, -0.2, 1.3]), 20, 2.6, 0.4
    x, vp, va, s = co.placement(psi, base, nseg, dl, angle=ang, beat="build_a_beat")
    p = x.reshape(-1, 3)
    assert np.allclose(s, np.arange(nseg) / (nseg - 1)) and np.allclose(p[0], base)
    th = np.array([co.build_a_beat_angle(v, psi)[0] for v in s])
    steps = np.linalg.norm(np.diff(p, axis=0), axis=1)
    assert np.allclose(steps, dl * np.cos(0.5 * np.diff(th)), rtol=1e-12)
    h = 1e-6
    xp = co.placement(psi + h, base, nseg, dl, angle=ang, beat="build_a_beat")[0]
    xm = co.placement(psi - h, base, nseg, dl, angle=ang, beat="build_a_beat")[0]
    assert np.allclose((xp - xm) / (2 * h), vp, atol=1e-5 * np.abs(vp).max())
    xp = co.placement(psi, base, nseg, dl, angle=ang + h, beat="build_a_beat")[0]
    xm = co.placement(psi, base, nseg, dl, angle=ang - h, beat="build_a_beat")[0]
    assert np.allclose((xp - xm) / (2 * h), va, atol=1e-6 * np.abs(va).max())
    # the first chord bisects the base tangent (sin theta, cos theta) of the reference
    # frame, turned once by the shape angle and once by the frame, and the next one
    ca, sa = math.cos(ang), math.sin(ang)
    t = [co.FRAME @ np.array([ca * math.sin(v) - sa * math.cos(v),
                              sa * math.sin(v) + ca * math.cos(v), 0.0]) for v in th[:2]]
    d1 = (p[1] - p[0]) / steps[0]
    assert np.allclose(d1, (t[0] + t[1]) / np.linalg.norm(t[0] + t[1]), atol=1e-12)


def test_build_a_beat_tip_sweeps_the_amplitude():
    """Straight filament at tangent angle theta: the tip lies at L (sin theta, cos theta)
    in the reference frame; between theta_s and pi - theta_s the tip's beat-direction
    coordinate moves by amplitude x L (1.99 filament lengths)."""
    th_s = math.acos(0.5 * co.BUILD_AMPLITUDE)
    L = 1.0
    y = [L * math.cos(t) for t in (th_s, math.pi - th_s)]
    assert abs((y[0] - y[1]) - co.BUILD_AMPLITUDE * L) < 1e-12
    # the model's tip over a cycle stays within that sweep and reaches most of it
    nseg, dl = 20, 1.0 / 19
    tips = []
    for psi in np.linspace(0, 2 * math.pi, 200, endpoint=False):
        x, _, _, _ = co.placement(psi, np.zeros(3), nseg, dl, beat="build_a_beat",
                                  frame=np.eye(3))
        tips.append(x.reshape(-1, 3)[-1])
    tips = np.array(tips)
    assert tips[:, 1].max() <= y[0] + 1e-12 and tips[:, 1].min() >= y[1] - 1e-12
    assert tips[:, 1].max() - tips[:, 1].min() > 0.8 * co.BUILD_AMPLITUDE * L
    assert np.all(np.linalg.norm(tips, axis=1) <= L + 1e-12)


def _wall_frame_from_quaternion():
    """The wall cilium's frame rebuilt from filament.cpp:225-226 for dir = e_z: the square
    root of the quaternion (0, 0, -1, 0) (quaternion.cpp:197-223), as a rotation matrix."""
    w, v = 0.0, np.array([0.0, -1.0, 0.0])
    w = math.sqrt(0.5 * (1.0 + w))
    v = v / (2.0 * w)
    n = math.sqrt(w * w + v @ v)
    w, (a, b, c) = w / n, v / n
    return np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - w * c), 2 * (a * c + w * b)],
                     [2 * (a * b + w * c), 1 - 2 * (a * a + c * c), 2 * (b * c - w * a)],
                     [2 * (a * c - w * b), 2 * (b * c + w * a), 1 - 2 * (a * a + b * b)]])


@pytest.mark.parametrize("pair_dp,longt", [(0.0, False), (1.0, False), (1.0, True)])
def test_pair_whose_phases_coincide_is_one_cilium_moved_by_the_separation(pair_dp, longt):
    """BICILIA at pair_dp 0 or 1 and BICILIA_LONGT at 1 give the second cilium the first
    one's phase (modulo 2 pi): its segments are the first cilium's moved by 2.2 a along
    the wall frame's third axis (lab -x), with the same velocity directions."""
    F = _wall_frame_from_quaternion()
    assert np.allclose(F, co.FRAME, atol=1e-15)
    assert np.allclose(F[:, 2], [-1.0, 0.0, 0.0], atol=1e-15)
    base, a = np.array([3.0, -2.0, 1.3]), 1.25
    x, vp, va, s = co.pair_placement(1.7, base, nseg=16, angle=0.4, pair_dp=pair_dp,
                                     longt=longt, a=a)
    p, vp, va = x.reshape(2, 8, 3), vp.reshape(2, 8, 3), va.reshape(2, 8, 3)
    assert np.allclose(p[1] - p[0], 2.2 * a * np.array([-1.0, 0.0, 0.0]), atol=1e-12)
    assert np.allclose(vp[1], vp[0], atol=1e-12) and np.allclose(va[1], va[0], atol=1e-12)
    assert np.allclose(s[8:], s[:8], atol=1e-12)
    # each cilium is one volvox cilium of nseg / 2 segments (panels and tolerance of 8)
    x1, _, _, s1 = co.placement(1.7, base, 8, 2.6, 0.4, "volvox")
    assert np.array_equal(s[:8], s1) and np.array_equal(x[:24], x1)


@pytest.mark.parametrize("longt", [False, True])
def test_second_cilium_is_a_lone_cilium_at_the_second_phase(longt):
    psi, dp = 2.3, 0.37
    psi2 = dp * psi if longt else psi + 2.0 * math.pi * dp
    assert co.pair_phase(psi, dp, longt) == psi2
    base = np.array([0.0, 5.0, 1.3])
    x, vp, va, s = co.pair_placement(psi, base, nseg=20, pair_dp=dp, longt=longt)
    x2, vp2, va2, s2 = co.placement(psi2, base + np.array([-2.2, 0.0, 0.0]), 10, 2.6, 0.0, "volvox")
    assert np.array_equal(s[10:], s2)
    assert np.allclose(x[30:], x2, atol=1e-13) and np.allclose(vp[30:], vp2, atol=1e-13)
    assert np.allclose(va[30:], va2, atol=1e-13)
    # a planted error: the second cilium at the first cilium's phase is far off
    x_bad, _, _, _ = co.placement(psi, base + np.array([-2.2, 0.0, 0.0]), 10, 2.6, 0.0, "volvox")
    assert np.abs(x_bad - x[30:]).max() > 1.0


@pytest.mark.parametrize("beat", ["fulford_blake", "fulford_blake_original", "volvox",
                                  "coral_larvae", "bicilia"])
def test_oracle_tables_equal_the_reference_assignments(beat):
    """Every coefficient of the oracle's hand-typed beat tables equals the value the
    reference assigns (filament.cpp:684-960), read mechanically into
    tests/golden/beat_tables.json by tests/golden/extract_beat_tables.py; the paired
    beats' table is the volvox table."""
    import json
    with open(os.path.join(GOLD, "beat_tables.json")) as fh:
        ref = json.load(fh)[beat]
    ax, ay, bx, by, _ = co.BEATS["volvox" if beat == "bicilia" else beat]
    for name, mine in (("Ax", ax), ("Ay", ay), ("Bx", bx), ("By", by)):
        assert np.array_equal(np.asarray(ref[name]), mine), (beat, name)
