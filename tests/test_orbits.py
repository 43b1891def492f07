"""CPU pins of the in-process outer loops (cufcm_fil_amd/orbits.py) on flow maps with
known answers: two phase-locked oscillators (period 2 pi / mean frequency, locked phase
difference asin((w1 - w2) / 2 eps)), a chain of identical oscillators in sync (Floquet
exponents -2 eps (1 - cos(pi k / n)), the spectrum of the path-graph Laplacian), a
linear system for the hookstep GMRES, and the interpolation rule of the bisection."""
import math

import numpy as np

from cufcm_fil_amd import orbits


class OscillatorFlow:
    """dphi_i/dt = w_i + eps sin(phi_j - phi_i), RK4 with T / spp steps; the CiliaFlow
    interface (nfil, spp, run, steporbit, getrhs)."""

    def __init__(self, w=(1.0, 1.2), eps=0.5, spp=200):
        self.w = np.array(w)
        self.eps = eps
        self.nfil = 2
        self.spp = spp
        self.angles = False

    def rhs(self, p):
        return self.w + self.eps * np.sin(p[::-1] - p)

    def run(self, x, ndts, record=False):
        T = x[0]
        dt = T / self.spp
        p = np.array(x[1:3], float)
        rows = []
        for _ in range(ndts):
            k1 = self.rhs(p)
            k2 = self.rhs(p + 0.5 * dt * k1)
            k3 = self.rhs(p + 0.5 * dt * k2)
            k4 = self.rhs(p + dt * k3)
            p = p + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
            rows.append(p.copy())
        y = np.concatenate([[T], p])
        return (y, np.array(rows)) if record else y

    steporbit = orbits.CiliaFlow.steporbit
    getrhs = orbits.CiliaFlow.getrhs
    ensemble = orbits.CiliaFlow.ensemble


class ChainFlow(OscillatorFlow):
    """dphi_i/dt = w + eps sum over chain neighbours j of sin(phi_j - phi_i)."""

    def __init__(self, n=5, w=1.0, eps=0.3, spp=200):
        self.nfil, self.w, self.eps, self.spp, self.angles = n, w, eps, spp, False

    def rhs(self, p):
        d = np.zeros_like(p)
        d[:-1] += np.sin(p[1:] - p[:-1])
        d[1:] += np.sin(p[:-1] - p[1:])
        return self.w + self.eps * d

    def run(self, x, ndts, record=False):
        T = x[0]
        dt = T / self.spp
        p = np.array(x[1:1 + self.nfil], float)
        for _ in range(ndts):
            k1 = self.rhs(p)
            k2 = self.rhs(p + 0.5 * dt * k1)
            k3 = self.rhs(p + 0.5 * dt * k2)
            k4 = self.rhs(p + dt * k3)
            p = p + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        return np.concatenate([[T], p])


def test_floquet_exponents_of_a_synchronised_chain():
    n, eps = 5, 0.3
    flow = ChainFlow(n=n, eps=eps)
    x_star = np.concatenate([[2 * math.pi], np.full(n, 0.4)])
    exact = np.sort(-2 * eps * (1 - np.cos(math.pi * np.arange(n) / n)))[::-1]
    ev, vec, k = orbits.floquet_arnoldi(flow, x_star, n_evals=3, eps=1e-2, tol=1e-6)
    assert np.allclose(ev.real, exact[:3], atol=1e-4) and np.allclose(ev.imag, 0, atol=1e-8)
    # the neutral mode is the shift along the orbit: all phases equally
    v0 = vec[:, 0].real
    assert np.allclose(v0 / v0[0], 1.0, atol=1e-3)
    M, ev_m = orbits.monodromy(flow, x_star, eps=1e-4, workers=2)
    assert np.allclose(ev_m.real, exact, atol=1e-6)
    # a planted mistake in the map (one coupling lost) moves the spectrum
    class Broken(ChainFlow):
        def rhs(self, p):
            d = np.zeros_like(p)
            d[:-1] += np.sin(p[1:] - p[:-1])
            return self.w + self.eps * d
    _, ev_b = orbits.monodromy(Broken(n=n, eps=eps), x_star, eps=1e-4, workers=2)
    assert not np.allclose(ev_b.real, exact, atol=1e-3)


def test_hook_gmres_solves_and_respects_the_trust_radius():
    rng = np.random.default_rng(0)
    A = np.eye(8) + 0.2 * rng.normal(size=(8, 8))
    b = rng.normal(size=8)
    dot = lambda n, a, c: float(a @ c)
    g = orbits.HookGMRES()
    x, res, _, its, info = g.solve(8, 8, np.zeros(8), b, lambda v: A @ v, lambda v: v, dot,
                                   1e-12, 0.0, 100, 0)
    assert info == 0 and np.allclose(x, np.linalg.solve(A, b), atol=1e-10)
    full = np.linalg.norm(np.linalg.solve(A, b))
    g = orbits.HookGMRES()
    _, _, pdel, _, _ = g.solve(8, 8, np.zeros(8), b, lambda v: A @ v, lambda v: v, dot,
                               1e-12, 0.3 * full, 100, 0)
    y, _ = g.hook(0.3 * full)
    assert np.linalg.norm(y) <= 0.3 * full * 1.0001


def test_newton_finds_the_phase_locked_orbit():
    flow = OscillatorFlow()
    om = flow.w.mean()
    x0 = np.array([2 * math.pi / flow.w[0], 0.0, 1.0])
    N = orbits.NewtonHook(flow, x0, eps_j=1e-7)
    x, res, status = N.solve(m=3, gtol=1e-4, tol=1e-9, nits=30)
    assert status == 0 and res < 1e-9, (status, res, N.history)
    assert abs(x[0] - 2 * math.pi / om) < 1e-6
    lock = math.asin((flow.w[0] - flow.w[1]) / (2 * flow.eps))
    d = (x[1] - x[2] - lock + math.pi) % (2 * math.pi) - math.pi
    assert abs(d) < 1e-6
    assert N.history[-1] < N.history[0]


def test_interpolation_and_bisection_bracket_the_edge():
    rng = np.random.default_rng(2)
    left, right = rng.uniform(0, 6, 6), rng.uniform(0, 6, 6)
    assert np.allclose(orbits.interpolate_state(left, right, 1.0, 3)[:3], orbits.box(left[:3], 2 * math.pi))
    assert np.allclose(orbits.interpolate_state(left, right, 0.0, 3)[3:], right[3:])
    mid = orbits.interpolate_state(left, right, 0.5, 3)[:3]
    ref = np.angle(np.exp(1j * left[:3]) + np.exp(1j * right[:3])) % (2 * math.pi)
    assert np.allclose(mid, ref)

    class StillFlow(OscillatorFlow):   # states do not move: the outcome is the initial r
        def run(self, x, ndts, record=False):
            rows = np.tile(np.array(x[1:3]), (ndts, 1))
            return (np.array(x), rows) if record else np.array(x)

    flow = StillFlow()
    L = np.array([1.0, 0.0, 0.0])       # in phase: r = 1
    R = np.array([1.0, 0.0, 2.0])       # r = cos(1)
    classify = lambda r: r[-1] > 0.8    # "left" while r stays above 0.8
    lo, hi = orbits.bisection(flow, L, R, 2, classify, iters=12, workers=2)
    # the interpolated second phase reaches 2 acos(0.8) at the edge
    f = lambda a: orbits.interpolate_state(L[1:], R[1:], a, 2)[1] - 2 * math.acos(0.8)
    assert f(lo) > 0.0 > f(hi) and hi - lo <= 2.0 ** -12
    assert orbits.order_parameter([0.3, 0.3]) == 1.0 or abs(orbits.order_parameter([0.3, 0.3]) - 1.0) < 1e-15
