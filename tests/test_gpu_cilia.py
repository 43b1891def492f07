"""Prescribed Fulford-Blake cilia on the GPU (through the C ABI): geometry against the
oracle, generalised forces against the reference's own output files for all 300
steps of one beat, and a lattice of cilia against a dense solve."""
import math
import os

import numpy as np
import pytest

import cilia_oracle as co
import oracle
from cufcm_fil_amd import _lib as L
from cufcm_fil_amd.cilia import PrescribedCilia, wall_lattice

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _golden(name):
    return np.loadtxt(os.path.join(GOLD, name))[1:]


def test_geometry_matches_oracle():
    with PrescribedCilia([0.0, 0.0, 1.3], phase=[2.2]) as C:
        st = C.state()
    x, vp, va, s = co.placement(2.2, (0.0, 0.0, 1.3))
    assert np.array_equal(st["s_values"][0], s)          # same bisection decisions
    for got, ref in ((st["x"], x), (st["vel_dir_phase"], vp), (st["vel_dir_angle"], va)):
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_generalised_forces_match_reference_files(mode):
    """One 20-segment cilium, phase speed 2 pi, 300 steps (dt = 1/300): the reference
    wrote these values in single precision with a GMRES tolerance of 1e-4."""
    ph = _golden("fb_phase_generalised_forces_nseg20_sep2.6.dat")
    an = _golden("fb_angle_generalised_forces_nseg20_sep2.6.dat")
    prec = "fp32c" if mode == "mixed" else "fp64"
    qp, qa = np.zeros(300), np.zeros(300)
    with PrescribedCilia([0.0, 0.0, 1.3], precision=prec, solve_mode=mode) as C:
        for nt in range(300):
            p, a, st = C.step(nt)
            assert st["converged"] and st["rel_residual"] <= 1e-10
            qp[nt], qa[nt] = p[0], a[0]
        assert abs(C.state()["phase"][0] - 2.0 * math.pi * 299 / 300) < 1e-12
    assert np.all(np.abs(qp - ph) <= 1e-4 * np.abs(ph)), np.abs(qp / ph - 1).max()
    assert np.abs(qa - an).max() <= 1e-4 * np.abs(an).max()


def test_lattice_matches_dense_solve():
    base = wall_lattice(3, 3, spacing=30.0)
    n = base.size // 3
    rng = np.random.default_rng(5)
    phase = rng.uniform(0.0, 2.0 * math.pi, n)
    pdot = 2.0 * math.pi * (1.0 + 0.1 * rng.standard_normal(n))
    with PrescribedCilia(base, phase, pdot, dt=0.01) as C:
        C.step(0)
        qp, qa, st = C.step(1)
        state = C.state()
    xs, vps, vas = [], [], []
    for c in range(n):
        x, vp, va, _ = co.placement(phase[c] + pdot[c] * 0.01, base[3 * c:3 * c + 3])
        xs.append(x), vps.append(vp), vas.append(va)
    x, vp, va = np.concatenate(xs), np.concatenate(vps), np.concatenate(vas)
    assert np.abs(state["x"] - x).max() <= 1e-11
    f = np.linalg.solve(oracle.mobility_matrix(x, 3), np.repeat(pdot, 60) * vp)
    assert np.abs(state["f"] - f).max() <= 1e-8 * np.abs(f).max()
    ref_qp = np.array([vp[60 * c:60 * c + 60] @ f[60 * c:60 * c + 60] for c in range(n)])
    ref_qa = np.array([va[60 * c:60 * c + 60] @ f[60 * c:60 * c + 60] for c in range(n)])
    assert np.allclose(qp, ref_qp, rtol=1e-8, atol=0) and np.allclose(qa, ref_qa, rtol=1e-8,
                                                                        atol=1e-8 * np.abs(ref_qa).max())


def test_init_rejects_wrong_contexts():
    from cufcm_fil_amd import RpyMobilitySolver
    import ctypes
    prm = L.cfm_cilia_params(n_cilia=1, nseg=20, beat=L.CFM_BEAT_FULFORD_BLAKE, dl=2.6, dt=0.01,
                             solve=L.cfm_solve_opts(precond=L.CFM_PC_SELF, rtol=1e-10))
    b, p, w = np.array([0.0, 0.0, 1.3]), np.zeros(1), np.ones(1)
    for kw in (dict(seg_dof=6), dict(seg_dof=3, wall=False)):
        with RpyMobilitySolver(20, 0, **kw) as m:
            assert L.lib().cfm_cilia_init(m._h, ctypes.byref(prm), L.ptr(b), L.ptr(p), L.ptr(w)) == L.CFM_EINVAL
    with RpyMobilitySolver(40, 0, seg_dof=3) as m:   # 2 x 20 segments, but n_cilia = 1
        assert L.lib().cfm_cilia_init(m._h, ctypes.byref(prm), L.ptr(b), L.ptr(p), L.ptr(w)) == L.CFM_EINVAL
    with RpyMobilitySolver(20, 0, seg_dof=3) as m:
        assert L.lib().cfm_cilia_step(m._h, 0, None, None, None) == L.CFM_ESTATE
