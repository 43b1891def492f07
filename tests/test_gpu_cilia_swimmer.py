"""Prescribed cilia on free rigid bodies on the GPU (through the C ABI) against the
numpy oracle (oracle/cilia_swimmer_oracle.py): prescribed phase speeds and dynamic
phases with shape rotation driven by the reference's recorded generalised forces, three
explicit steps each; body state within 1e-9, forces and speeds within 1e-8 of the
oracle's dense solve.  At a larger size, where the oracle is too slow, the rigid-body
identities: a rotated swimmer's velocities are the rotated velocities."""
import math
import os

import numpy as np
import pytest

from cufcm_fil_amd import geometry as geo
from cufcm_fil_amd.cilia import CiliatedBodies, cilium_frames
from filament_oracle import rot_mat
from test_cilia_swimmer_oracle import NSEG, swimmer

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gpu(sw, precond="diag", rtol=1e-11, **kw):
    return CiliatedBodies(sw.blob_ref.reshape(-1), sw.fil_ref.reshape(-1), sw.fil_q.reshape(-1),
                          sw.x.reshape(-1), sw.q.reshape(-1), sw.phase, sw.phase_dot, nseg=NSEG,
                          dl=sw.dl, dt=sw.dt, wall=False, precond=precond, rtol=rtol,
                          s_tol=sw.s_tol, **kw)


def _compare(G, sw, nt, qp_o, qa_o, qp_g, qa_g, tol=1e-8):
    """tol: relative bound on velocities, forces and speeds (the GMRES tolerance times
    the conditioning); positions and phases to 1e-9 / 1e-12 absolute."""
    st, b, sp = G.state(), G.bodies(), G.speeds()
    z = sw.last["z"]
    ns = sw.nc * NSEG
    assert np.abs(st["x"] - sw.last["xs"].reshape(-1)).max() < 1e-9, nt
    assert np.abs(b["x"] - sw.x.reshape(-1)).max() < 1e-9, nt
    assert np.abs(b["q"] - sw.q.reshape(-1)).max() < 1e-9, nt
    vmax = np.abs(sw.V).max()
    assert np.abs(b["vel"] - sw.V.reshape(-1)).max() < tol * vmax, (nt, b["vel"], sw.V)
    f = z[:3 * ns]
    assert np.abs(st["f"] - f).max() < tol * np.abs(f).max(), nt
    fb = z[3 * ns:3 * ns + 3 * sw.nbody * sw.nb]
    assert np.abs(b["f_blob"] - fb).max() < tol * np.abs(fb).max(), nt
    assert np.allclose(qp_g, qp_o, rtol=tol, atol=tol * np.abs(qp_o).max()), nt
    assert np.allclose(sp["phase"], sw.phase, atol=1e-12), nt
    assert np.allclose(sp["phase_dot"], sw.phase_dot, rtol=tol), nt


@pytest.mark.parametrize("precond", ["diag", "none", "near", "dense"])
def test_prescribed_phase_speeds_match_oracle(precond):
    sw = swimmer(n_body=2, per=3, seed=11, dt=0.02)
    with _gpu(sw, precond) as G:
        for nt in range(3):
            qp_g, qa_g, stats = G.step(nt)
            _, qp_o, qa_o = sw.step(nt)
            assert stats["converged"], stats
            _compare(G, sw, nt, qp_o, qa_o, qp_g, qa_g)


@pytest.mark.parametrize("beat", ["volvox", "coral_larvae", "build_a_beat"])
def test_other_beats_on_bodies_match_oracle(beat):
    sw = swimmer(n_body=1, per=3, seed=14, dt=0.02)
    sw.beat = beat
    with _gpu(sw, "diag", beat=beat) as G:
        for nt in range(2):
            qp_g, qa_g, stats = G.step(nt)
            _, qp_o, qa_o = sw.step(nt)
            assert stats["converged"], stats
            _compare(G, sw, nt, qp_o, qa_o, qp_g, qa_g)


@pytest.mark.parametrize("precond", ["diag", "near", "dense"])
def test_dynamic_phases_and_angles_match_oracle(precond):
    qp = np.loadtxt(os.path.join(GOLD, "fb_phase_generalised_forces_nseg20_sep2.6.dat"))[1:]
    qa = np.loadtxt(os.path.join(GOLD, "fb_angle_generalised_forces_nseg20_sep2.6.dat"))[1:]
    # tables recorded for 20 segments, scaled to these 6-segment cilia
    qp, qa = qp * (NSEG / 20.0) ** 3, qa * (NSEG / 20.0) ** 3
    sw = swimmer(n_body=1, per=4, seed=12, dt=0.02)
    sw.driving = (qp, qa, 0.05)
    # the right-hand side holds only the driving forces: a tighter GMRES tolerance, and
    # forces within 1e-7 of the dense solve
    with _gpu(sw, precond, rtol=1e-13) as G:
        G.set_driving(qp, qa, spring=0.05)
        for nt in range(3):
            qp_g, qa_g, stats = G.step(nt)
            _, qp_o, qa_o = sw.step(nt)
            assert stats["converged"], stats
            _compare(G, sw, nt, qp_o, qa_o, qp_g, qa_g, tol=1e-7)
            sp = G.speeds()
            assert np.allclose(sp["angle_dot"], sw.angle_dot, rtol=1e-7,
                               atol=1e-7 * np.abs(sw.angle_dot).max())
            assert np.allclose(sp["angle"], sw.angle, atol=1e-12)


def test_fene_spring_and_noise_match_oracle():
    qp = np.loadtxt(os.path.join(GOLD, "fb_phase_generalised_forces_nseg20_sep2.6.dat"))[1:]
    qa = np.loadtxt(os.path.join(GOLD, "fb_angle_generalised_forces_nseg20_sep2.6.dat"))[1:]
    qp, qa = qp * (NSEG / 20.0) ** 3, qa * (NSEG / 20.0) ** 3
    sw = swimmer(n_body=1, per=4, seed=13, dt=0.02)
    sw.driving = (qp, qa, 0.05)
    sw.fene, sw.noise, sw.seed = True, 3.0, 12345
    sw.angle[:] = [0.9, -1.1, 0.3, 0.8]     # two beyond the FENE threshold
    with _gpu(sw, "diag", rtol=1e-13) as G:
        G.set_driving(qp, qa, spring=0.05)
        G.set_driving_options(fene=True, noise=3.0, seed=12345)
        G.set_phases(angle=sw.angle)
        for nt in range(3):
            qp_g, qa_g, stats = G.step(nt)
            _, qp_o, qa_o = sw.step(nt)
            assert stats["converged"], stats
            _compare(G, sw, nt, qp_o, qa_o, qp_g, qa_g, tol=1e-7)
            assert np.allclose(G.speeds()["angle_dot"], sw.angle_dot, rtol=1e-7,
                               atol=1e-7 * np.abs(sw.angle_dot).max())


def _big(body_q, phase, precond="diag"):
    R = 20.0
    blob_ref = geo.fibonacci_sphere(1500, R)
    dirs = geo.fibonacci_sphere(40, 1.0).reshape(-1, 3)
    fil_ref = (R + 1.3) * dirs
    fil_q = cilium_frames(dirs, 0.0)
    return CiliatedBodies(blob_ref, fil_ref, fil_q, [5.0, -3.0, 2.0], body_q, phase,
                          np.full(40, 2 * math.pi), nseg=20, dl=2.6, dt=0.01, wall=False,
                          precond=precond, rtol=1e-10)


def test_rotated_swimmer_moves_rotated():
    phase = np.random.default_rng(2).uniform(0, 2 * math.pi, 40)
    with _big([1.0, 0.0, 0.0, 0.0], phase) as G:
        _, _, s0 = G.step(0)
        v0 = G.bodies()["vel"]
    qb = np.array([math.cos(0.6), *(math.sin(0.6) * np.array([2.0, -1.0, 2.0]) / 3.0)])
    R0 = rot_mat(qb)
    with _big(qb, phase) as G:
        _, _, s1 = G.step(0)
        v1 = G.bodies()["vel"]
    assert s0["converged"] and s1["converged"]
    scale = np.abs(v0).max()
    assert np.abs(v1[:3] - R0 @ v0[:3]).max() < 1e-7 * scale
    assert np.abs(v1[3:] - R0 @ v0[3:]).max() < 1e-7 * scale
    # the resident inverse of the blob block (frozen in the body frame) against the
    # reference's 6 pi mu a: same velocities, a fraction of the GMRES steps, also after
    # the body has moved
    with _big(qb, phase, "dense") as G:
        _, _, s2 = G.step(0)
        v2 = G.bodies()["vel"]
        for nt in range(1, 3):
            _, _, s3 = G.step(nt)
            assert s3["converged"] and s3["iterations"] <= s2["iterations"] + 2, s3
    assert np.abs(v2 - v1).max() < 1e-7 * scale
    print("gmres iterations diag", s0["iterations"], "dense", s2["iterations"], s3["iterations"])
    assert s2["iterations"] * 1.5 < s1["iterations"]
