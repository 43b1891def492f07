"""Paired cilia (the reference's deprecated BICILIA / BICILIA_LONGT, SHAPE_SEQUENCE 5 and
6) on the GPU through the C ABI: geometry of both cilia of each pair against the
oracle's pair placement, the wall lattice's forces and generalised forces against a
dense solve, pairs carried by a free body against the body oracle, and the parameter
checks."""
import ctypes
import math

import numpy as np
import pytest

import cilia_oracle as co
import oracle
from cufcm_fil_amd import _lib as L
from cufcm_fil_amd import RpyMobilitySolver
from cufcm_fil_amd.cilia import PrescribedCilia, wall_lattice
from test_cilia_swimmer_oracle import NSEG, swimmer
from test_gpu_cilia_swimmer import _compare, _gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("beat,pair_dp", [("bicilia", 0.3), ("bicilia_longt", 0.8)])
def test_pair_geometry_matches_oracle(beat, pair_dp):
    """Three pairs of two 8-segment cilia at different phases and shape angles.  Segment
    parameters equal the oracle's where no bisection comparison came within 1e-12 of
    flipping; the placement against the oracle's at the library's parameters."""
    base = [0.0, 0.0, 1.3, 40.0, 0.0, 1.3, 0.0, 40.0, 1.3]
    phases, angles = [0.4, 2.9, 5.6], [0.0, 0.3, -0.7]
    longt = beat == "bicilia_longt"
    with PrescribedCilia(base, phase=phases, nseg=16, beat=beat) as C:
        C.set_pair_dp(pair_dp)
        C.set_phases(angle=angles)
        st = C.state()
    for c in range(3):
        s_gpu = st["s_values"][c]
        for h, psi in enumerate((phases[c], co.pair_phase(phases[c], pair_dp, longt))):
            s_ref, margin = co.s_values(psi, 8, "volvox", margins=True)
            safe = margin > 1e-12
            assert np.array_equal(s_gpu[8 * h:8 * h + 8][safe], s_ref[safe]), (c, h)
        x, vp, va, _ = co.pair_placement(phases[c], base[3 * c:3 * c + 3], 16, 2.6, angles[c],
                                         pair_dp, longt, s=s_gpu)
        sl = slice(48 * c, 48 * c + 48)
        for got, ref in ((st["x"], x), (st["vel_dir_phase"], vp), (st["vel_dir_angle"], va)):
            assert np.abs(got[sl] - ref).max() <= 1e-11 * np.abs(ref).max(), c


def test_pair_lattice_matches_dense_solve():
    """A 2 x 2 wall lattice of pairs (two 10-segment cilia each) at prescribed phase
    speeds: after one advance, positions, segment forces and each pair's generalised
    forces (summed over both cilia) against the oracle's placement and dense solve."""
    base = wall_lattice(2, 2, spacing=30.0)
    n = base.size // 3
    rng = np.random.default_rng(8)
    phase = rng.uniform(0.0, 2.0 * math.pi, n)
    pdot = 2.0 * math.pi * (1.0 + 0.1 * rng.standard_normal(n))
    with PrescribedCilia(base, phase, pdot, dt=0.01, beat="bicilia") as C:
        C.set_pair_dp(0.25)
        C.step(0)
        qp, qa, st = C.step(1)
        state = C.state()
    assert st["converged"] and st["rel_residual"] <= 1e-10
    xs, vps, vas = [], [], []
    for c in range(n):
        x, vp, va, _ = co.pair_placement(phase[c] + pdot[c] * 0.01, base[3 * c:3 * c + 3], 20,
                                         pair_dp=0.25, s=state["s_values"][c])
        xs.append(x), vps.append(vp), vas.append(va)
    x, vp, va = np.concatenate(xs), np.concatenate(vps), np.concatenate(vas)
    assert np.abs(state["x"] - x).max() <= 1e-11
    f = np.linalg.solve(oracle.mobility_matrix(x, 3), np.repeat(pdot, 60) * vp)
    assert np.abs(state["f"] - f).max() <= 1e-8 * np.abs(f).max()
    ref_qp = np.array([vp[60 * c:60 * c + 60] @ f[60 * c:60 * c + 60] for c in range(n)])
    ref_qa = np.array([va[60 * c:60 * c + 60] @ f[60 * c:60 * c + 60] for c in range(n)])
    assert np.allclose(qp, ref_qp, rtol=1e-8, atol=0)
    assert np.allclose(qa, ref_qa, rtol=1e-8, atol=1e-8 * np.abs(ref_qa).max())


@pytest.mark.parametrize("beat", ["bicilia", "bicilia_longt"])
def test_pairs_on_a_free_body_match_oracle(beat):
    sw = swimmer(n_body=1, per=3, seed=21, dt=0.02)
    sw.beat, sw.pair_dp = beat, 0.4
    with _gpu(sw, "diag", beat=beat) as G:
        G.set_pair_dp(0.4)
        for nt in range(2):
            qp_g, qa_g, stats = G.step(nt)
            _, qp_o, qa_o = sw.step(nt)
            assert stats["converged"], stats
            _compare(G, sw, nt, qp_o, qa_o, qp_g, qa_g)
    assert NSEG % 2 == 0


def test_pair_beats_take_an_even_number_of_segments():
    prm = L.cfm_cilia_params(n_cilia=1, nseg=9, beat=L.CFM_BEAT_BICILIA, dl=2.6, dt=0.01,
                             solve=L.cfm_solve_opts(precond=L.CFM_PC_SELF, rtol=1e-10))
    b, p, w = np.array([0.0, 0.0, 1.3]), np.zeros(1), np.ones(1)
    with RpyMobilitySolver(9, 0, seg_dof=3, wall=True) as m:
        rc = L.lib().cfm_cilia_init(m._h, ctypes.byref(prm), L.ptr(b), L.ptr(p), L.ptr(w))
        assert rc == L.CFM_EINVAL
        assert b"even nseg" in L.lib().cfm_last_error(m._h)


def test_batched_pairs_match_single_contexts():
    """Two members of a 2 x 2 lattice of dynamic-phase pairs (two 10-segment cilia each)
    stepped together by the ensemble: each member's positions, phases and phase speeds
    after two steps against the same member in its own context."""
    import os
    from cufcm_fil_amd.cilia import CiliaEnsemble
    gold = os.path.join(os.path.dirname(__file__), "golden")
    qp_tab = np.loadtxt(os.path.join(gold, "fb_phase_generalised_forces_nseg20_sep2.6.dat"))[1:]
    base = wall_lattice(2, 2, spacing=30.0)
    n = base.size // 3
    rng = np.random.default_rng(12)
    phase = rng.uniform(0.0, 2.0 * math.pi, (2, n))
    omega = np.full((2, n), 2.0 * math.pi)
    with CiliaEnsemble(base, 2, qp_tab, beat="bicilia", precond="diag", rtol=1e-11,
                       s_tol=1e-12) as E:
        E.set_pair_dp(0.3)
        E.set_state(phase, omega, np.full(2, 1.0 / 40))
        for nt in range(2):
            E.step(nt)
        xe = E.positions()
        st_e = E.speeds()
    for m in range(2):
        with PrescribedCilia(base, phase[m], omega[m], dt=1.0 / 40, beat="bicilia",
                             precond="diag", rtol=1e-11, s_tol=1e-12) as C:
            C.set_pair_dp(0.3)
            C.set_driving(qp_tab)
            for nt in range(2):
                C.step(nt)
            x1 = C.state()["x"]
            sp = C.speeds()
        # both solve to rtol 1e-11 with different reduction orders: speeds agree to the
        # solve tolerance times the conditioning, positions through the phases they move
        assert np.abs(xe[m].reshape(-1) - x1).max() <= 1e-7, m
        assert np.allclose(st_e["phase_dot"][m], sp["phase_dot"], rtol=1e-7, atol=0), m
        assert np.allclose(st_e["phase"][m], sp["phase"], rtol=0, atol=1e-9), m
