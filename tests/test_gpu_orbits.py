"""In-process outer loops on the GPU (cufcm_fil_amd/orbits.py): flow maps of prescribed
cilia driven by the reference's recorded generalised forces, run one at a time and
concurrently (separate contexts and streams: identical results), and a Newton-Krylov
hookstep search for the periodic orbit of two cilia."""
import math
import os

import numpy as np
import pytest

from cufcm_fil_amd import orbits
from cufcm_fil_amd.cilia import wall_lattice

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _tables():
    ph = np.loadtxt(os.path.join(GOLD, "fb_phase_generalised_forces_nseg20_sep2.6.dat"))[1:]
    an = np.loadtxt(os.path.join(GOLD, "fb_angle_generalised_forces_nseg20_sep2.6.dat"))[1:]
    return ph, an


def test_concurrent_flow_maps_equal_sequential_runs():
    ph, _ = _tables()
    flow = orbits.CiliaFlow(wall_lattice(2, 2, spacing=40.0), ph, steps_per_period=40)
    rng = np.random.default_rng(1)
    xs = [np.concatenate([[1.0], rng.uniform(0, 2 * math.pi, 4)]) for _ in range(4)]
    seq = [flow.run(x, 20) for x in xs]
    par = flow.ensemble(xs, 20, workers=4, batched=False)
    for a, b in zip(seq, par):
        assert np.array_equal(a, b)
    # an isolated cilium driven by its own recorded forces advances by 2 pi per period
    one = orbits.CiliaFlow(wall_lattice(1, 1, spacing=0.0), ph, steps_per_period=40)
    y = one.run([1.0, 0.7], 40)
    assert abs(y[1] - 0.7 - 2 * math.pi) < 0.01


def test_newton_hookstep_and_floquet_exponents_of_a_two_cilium_orbit():
    ph, _ = _tables()
    # two cilia 30 apart in line along the beat direction (the beat plane is y-z): they
    # lock with a small phase lag of the downstream cilium
    flow = orbits.CiliaFlow(wall_lattice(1, 2, spacing=30.0), ph, steps_per_period=40,
                            s_tol=1e-12)
    x0 = np.array([1.0, 0.0, 0.3])
    N = orbits.NewtonHook(flow, x0, eps_j=1e-5)
    x, res, status = N.solve(m=3, gtol=1e-3, tol=1e-7, nits=8)
    assert res < 1e-7 or res < 1e-3 * N.history[0], (status, N.history)
    assert abs(x[0] - 1.0) < 0.25
    # the orbit closes: one more period returns the phases advanced by 2 pi
    y = flow.run(x, flow.spp)
    assert np.abs(y[1:] - x[1:] - 2 * math.pi).max() < max(1e-6, 10 * res)
    # Floquet exponents: the Arnoldi estimate (forward differences, step 1e-2) agrees with
    # the full monodromy matrix (central differences, one concurrent ensemble)
    M, ev_m = orbits.monodromy(flow, x, eps=1e-4, workers=4)
    ev_a, _, k = orbits.floquet_arnoldi(flow, x, n_evals=2, eps=1e-2, tol=1e-4)
    assert k == 2 and np.allclose(ev_a, ev_m, atol=1e-2), (ev_a, ev_m)
    # the exponent nearest zero belongs to the shift along the orbit: its eigenvector is
    # the phase velocity at x
    mu, vec = np.linalg.eig(M)
    s_idx = int(np.argmin(np.abs(np.log(mu.astype(complex)))))
    shift = math.log(mu[s_idx].real) / x[0]
    diff = math.log(mu[1 - s_idx].real) / x[0]
    vel = flow.run(x, 1)[1:] - x[1:]
    cosang = abs(vec[:, s_idx].real @ vel) / np.linalg.norm(vec[:, s_idx].real) / np.linalg.norm(vel)
    assert cosang > 0.999, cosang
    # the shift is neutral for the time-continuous flow; the phases advance by explicit
    # Euler, so the map over spp steps keeps it only to O(T / spp): halving the step
    # halves the exponent
    fine = orbits.CiliaFlow(wall_lattice(1, 2, spacing=30.0), ph, steps_per_period=80,
                            s_tol=1e-12)
    N2 = orbits.NewtonHook(fine, x, eps_j=1e-5)
    x2, res2, _ = N2.solve(m=3, gtol=1e-3, tol=1e-7, nits=8)
    assert res2 < 1e-6
    M2, _ = orbits.monodromy(fine, x2, eps=1e-4, workers=4)
    mu2 = np.linalg.eigvals(M2)
    order = np.argsort(np.abs(np.log(mu2.astype(complex))))
    shift2, diff2 = (math.log(mu2[order[0]].real) / x2[0], math.log(mu2[order[1]].real) / x2[0])
    assert abs(shift2) < 0.6 * abs(shift) and abs(shift) < 0.1, (shift, shift2)
    # the locked phase difference is stable, and the step changes its exponent little
    assert diff < 0 and diff2 < 0 and abs(diff2 - diff) < 0.2 * abs(diff), (diff, diff2)
