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
