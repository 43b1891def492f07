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
