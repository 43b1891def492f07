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


@pytest.mark.parametrize("beat", ["volvox", "coral_larvae", "build_a_beat",
                                  "fulford_blake_original"])
def test_other_beats_geometry_matches_oracle(beat):
    """Three cilia at different phases and shape angles.  Segment parameters: equal to
    the oracle's where no bisection comparison came within 1e-12 of flipping, and in any
    case splitting the curve into equal arcs to within the tolerance; the placement
    against the oracle's at the library's parameters."""
    base = [0.0, 0.0, 1.3, 40.0, 0.0, 1.3, 0.0, 40.0, 1.3]
    phases, angles = [0.4, 2.9, 5.6], [0.0, 0.3, -0.7]
    with PrescribedCilia(base, phase=phases, nseg=12, beat=beat) as C:
        C.set_phases(angle=angles)
        st = C.state()
    tol = 0.1 / 12
    for c in range(3):
        s_gpu = st["s_values"][c]
        s_ref, margin = co.s_values(phases[c], 12, beat, margins=True)
        safe = margin > 1e-12
        assert np.array_equal(s_gpu[safe], s_ref[safe]), (beat, c)
        if beat != "build_a_beat":
            total = co.curve_length(1.0, phases[c], 12, beat)
            frac = np.array([co.curve_length(v, phases[c], 12, beat) / total for v in s_gpu])
            assert np.all(np.abs(frac - np.arange(12) / 11) <= tol + 1e-12)
        x, vp, va, _ = co.placement(phases[c], base[3 * c:3 * c + 3], 12, 2.6, angles[c], beat,
                                    s=s_gpu)
        sl = slice(36 * c, 36 * c + 36)
        for got, ref in ((st["x"], x), (st["vel_dir_phase"], vp), (st["vel_dir_angle"], va)):
            assert np.abs(got[sl] - ref).max() <= 1e-11 * np.abs(ref).max(), (beat, c)


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


@pytest.mark.parametrize("with_angle,precond", [(False, "diag"), (True, "diag"), (True, "self")])
def test_driving_with_recorded_forces_recovers_the_beat(with_angle, precond):
    """Closed loop through the reference's files: a cilium driven by its own recorded
    generalised forces beats at 2 pi with no shape rotation, step after step (the
    files hold single-precision values from solves to 1e-4)."""
    qp = _golden("fb_phase_generalised_forces_nseg20_sep2.6.dat")
    qa = _golden("fb_angle_generalised_forces_nseg20_sep2.6.dat") if with_angle else None
    with PrescribedCilia([0.0, 0.0, 1.3], precond=precond) as C:
        C.set_driving(qp, qa, spring=0.0)
        for nt in range(60):
            p, a, st = C.step(nt)
            assert st["converged"] and st["rel_residual"] <= 1e-10
            sp = C.speeds()
            assert sp["phase_dot"][0] == pytest.approx(2 * math.pi, rel=3e-4), nt
            assert abs(sp["angle_dot"][0]) <= 3e-4 * 2 * math.pi
            # the generalised force balances the drive
            assert p[0] == pytest.approx(np.interp(sp["phase"][0], 2 * math.pi * np.arange(300) / 300, qp), rel=1e-3)


def test_coupled_lattice_matches_dense_solve():
    base = wall_lattice(2, 3, spacing=30.0)
    n = base.size // 3
    rng = np.random.default_rng(2)
    phase = rng.uniform(0.0, 2.0 * math.pi, n)
    omega0 = 2.0 * math.pi * (1.0 + 0.05 * rng.standard_normal(n))
    qp_tab = _golden("fb_phase_generalised_forces_nseg20_sep2.6.dat")
    qa_tab = _golden("fb_angle_generalised_forces_nseg20_sep2.6.dat")
    spring = 0.01
    with PrescribedCilia(base, phase, omega0, precond="self") as C:
        C.set_driving(qp_tab, qa_tab, spring=spring)
        C.step(0)
        state, sp = C.state(), C.speeds()
    xs, vps, vas, qps, qas = [], [], [], [], []
    L = 2.6 * 19
    for c in range(n):
        x, vp, va, _ = co.placement(phase[c], base[3 * c:3 * c + 3])
        z = np.zeros(60 * n)
        xs.append(x)
        zp, za = np.zeros(60 * n), np.zeros(60 * n)
        zp[60 * c:60 * c + 60], za[60 * c:60 * c + 60] = vp, va
        vps.append(zp), vas.append(za)
        q1, q2 = co.driving_forces(phase[c], 0.0, omega0[c], qp_tab, qa_tab, spring=spring)
        qps.append(q1), qas.append(q2)
    M = oracle.mobility_matrix(np.concatenate(xs), 3)
    F, phid, psid = co.coupled_solve(M, vps, vas, qps, qas)
    # residual of the GPU solution in the oracle's coupled operator (the solve's own
    # criterion, rtol 1e-10), then the solution itself to the saddle system's conditioning
    m = M.shape[0]
    K = np.stack(vps + vas, axis=1)
    A = np.block([[M, -K], [-K.T, np.zeros((2 * n, 2 * n))]])
    b = np.concatenate([np.zeros(m), -np.asarray(qps), -np.asarray(qas)])
    zg = np.concatenate([state["f"], sp["phase_dot"], sp["angle_dot"]])
    assert np.linalg.norm(A @ zg - b) <= 2e-10 * np.linalg.norm(b)
    assert np.abs(state["f"] - F).max() <= 1e-6 * np.abs(F).max()
    assert np.allclose(sp["phase_dot"], phid, rtol=1e-6, atol=0)
    assert np.allclose(sp["angle_dot"], psid, rtol=0, atol=1e-6 * np.abs(phid).max())


def _s_table():
    v = np.loadtxt(os.path.join(GOLD, "fb_original_s_values_nseg20_sep2.6.dat"))
    return v[2:].reshape(300, 20)


def test_original_beat_s_values_match_reference_file_every_step():
    """The reference's recorded segment parameters of the unaltered coefficients
    (SHAPE_SEQUENCE 4), all 300 steps, with the bisection tolerance 0.001/NSEG the file
    was written with: every step equals the FP64 oracle's bisection exactly, and the
    file within 1e-6 (six printed digits plus the single-precision rounding of the
    reference's midpoints; one decision apart is 1.5e-5 or more) except at steps where
    the reference's single-precision arithmetic took the other side of a comparison
    that FP64 puts within 1e-6 of its threshold (9 of 300)."""
    tab = _s_table()
    flips = []
    with PrescribedCilia([0.0, 0.0, 1.3], beat="fulford_blake_original", s_tol=0.001 / 20) as C:
        for nt in range(300):
            C.step(nt)
            st = C.state()
            d = np.abs(st["s_values"][0] - tab[nt])
            if d.max() > 1e-6 or nt % 60 == 0:
                psi = 2.0 * math.pi * nt / 300
                s, m = co.s_values(psi, 20, "fulford_blake_original", 0.001 / 20, margins=True)
                assert np.array_equal(st["s_values"][0], s), nt
                if d.max() > 1e-6:
                    first = int(np.argmax(d > 1e-6))
                    assert m[first] < 1e-6 and d[first] > 1e-5, (nt, first, m[first])
                    flips.append(nt)
                x, vp, va, _ = co.placement(psi, (0.0, 0.0, 1.3), beat="fulford_blake_original", s=s)
                assert np.abs(st["x"] - x).max() <= 1e-12 * np.abs(x).max()
                assert np.abs(st["vel_dir_phase"] - vp).max() <= 1e-12 * np.abs(vp).max()
    assert len(flips) <= 9, flips


def test_s_table_mode_interpolates_the_recorded_parameters():
    """With the recorded table loaded (the reference's mode whenever it does not write
    its reference files) the parameters at the sampled phases are the table's rows and
    in between follow the oracle's interpolation; clearing it returns to bisection."""
    tab = _s_table()
    phases = [2.0 * math.pi * 37 / 300, 1.234, 2.0 * math.pi * 299.5 / 300]
    with PrescribedCilia([0.0, 0.0, 1.3, 40.0, 0.0, 1.3, 0.0, 40.0, 1.3], phase=phases,
                         beat="fulford_blake_original") as C:
        C.set_s_table(tab)
        st = C.state()
        assert np.abs(st["s_values"][0] - tab[37]).max() <= 1e-12
        for c, psi in enumerate(phases):
            ref = co.s_values_from_table(tab, psi, 20)
            assert np.abs(st["s_values"][c] - ref).max() <= 1e-14, c
            x, _, _, _ = co.placement(psi, (40.0 * (c == 1), 40.0 * (c == 2), 1.3),
                                      beat="fulford_blake_original", s=ref)
            assert np.abs(st["x"][60 * c:60 * c + 60] - x).max() <= 1e-12 * np.abs(x).max()
        _, _, sst = C.step(0)
        assert sst["converged"]
        C.set_s_table(None)
        s_bis = C.state()["s_values"][1]
    assert np.array_equal(s_bis, co.s_values(1.234, 20, "fulford_blake_original"))


def test_psi_restart_and_reference_output_files(tmp_path):
    """A run restarted from a psi.dat state (phases and angles) writes the reference's
    _seg_states.dat / _true_states.dat lines; they hold the library's own state."""
    from cufcm_fil_amd import simio
    base = wall_lattice(2, 2, spacing=30.0)
    rng = np.random.default_rng(3)
    ph, an = rng.uniform(0, 2 * math.pi, 4), 0.1 * rng.normal(size=4)
    simio.write_psi(tmp_path / "psi.dat", 0.0, 1.0, ph, an)
    k, T, ph2, an2 = simio.read_psi(tmp_path / "psi.dat", 4)
    w = simio.OutputWriter(str(tmp_path) + "/", "run")
    with PrescribedCilia(base, np.zeros(4)) as C:
        C.set_phases(ph2, an2)
        sp = C.speeds()
        assert np.array_equal(sp["phase"], ph) and np.array_equal(sp["angle"], an)
        x0, _, _, _ = co.placement(ph[1], base[3:6], angle=an[1])
        assert np.abs(C.state()["x"][60:120] - x0).max() <= 1e-12 * np.abs(x0).max()
        for nt in range(3):
            C.step(nt)
            C.write_outputs(w, nt + 1, period=T)
        x = C.state()["x"]
    steps, rows = simio.read_lines(w.path("seg_states"))
    assert list(steps) == [1, 2, 3] and np.allclose(rows[-1], x, rtol=1e-14, atol=1e-14)
    steps, rows = simio.read_lines(w.path("true_states"))
    assert rows[-1][0] == T and np.allclose(rows[-1][1:5], ph + 2 * math.pi * 2 / 300, atol=1e-12)
