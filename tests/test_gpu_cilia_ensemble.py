"""Ensembles of dynamic-phase cilia lattices stepped together (cfm_cilia_batch_*, through
the C ABI) against the numpy oracle, single-system contexts and themselves.

* every member's solution at every step, in the oracle's coupled operator at that
  member's state (placement, phase and angle directions and the driving forces of
  oracle/cilia_oracle.py from the member's own phases and angles): the residual is within
  the solve's criterion, ||A z - b|| <= 2 rtol ||b|| -- with and without shape rotation,
  both preconditioners, per-member time steps; the ensemble's positions are the oracle's
  placement (1e-12);
* the advance between steps: phase(nt) = phase(nt-1) + dt_e phase_dot(nt-1) (and the
  angle) to rounding (the device fuses the multiply-add);
* the first solve of each member and of a single-system context (PrescribedCilia) at
  the same state against the oracle's direct solve z* of that member's coupled system:
  ||z - z*|| <= ||A^-1||_2 ||A z - b|| (the residual bound, with ||A^-1||_2 from the
  oracle matrix's singular values) -- the two solvers may differ by the system's
  conditioning (condition numbers of 2e4 to 1e5 here) times the residual, no more;
* a member's results do not depend on the rest of the ensemble: run alone, in an
  ensemble of one, it reproduces its ensemble results bit for bit (the products, dot
  products and Givens rotations of one member never mix with another's);
* the orbit layer's batched flow maps and its concurrent single-system flow maps both
  follow the oracle's flow maps (direct solves) within each member's first-order bound
  from the solve tolerance."""
import math
import os

import numpy as np
import pytest

import cilia_oracle as co
import oracle
from cufcm_fil_amd import orbits
from cufcm_fil_amd.cilia import CiliaEnsemble, PrescribedCilia, wall_lattice

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _tables():
    ph = np.loadtxt(os.path.join(GOLD, "fb_phase_generalised_forces_nseg20_sep2.6.dat"))[1:]
    an = np.loadtxt(os.path.join(GOLD, "fb_angle_generalised_forces_nseg20_sep2.6.dat"))[1:]
    return ph, an


def _oracle_system(base, phase, angle, omega, qp_tab, qa_tab, spring):
    """One member's coupled operator A, right-hand side b and segment positions, all from
    the oracle's placement and driving forces at the member's phases and angles."""
    n = phase.size
    xs, vps, vas, qps, qas = [], [], [], [], []
    for c in range(n):
        s = co.s_values(phase[c], 20, tol=1e-12)
        x, vp, va, _ = co.placement(phase[c], base[3 * c:3 * c + 3], angle=angle[c], s=s)
        xs.append(x)
        zp, za = np.zeros(60 * n), np.zeros(60 * n)
        zp[60 * c:60 * c + 60], za[60 * c:60 * c + 60] = vp, va
        vps.append(zp)
        vas.append(za)
        q1, q2 = co.driving_forces(phase[c], angle[c], omega, qp_tab, qa_tab, spring=spring)
        qps.append(q1)
        qas.append(q2)
    M = oracle.mobility_matrix(np.concatenate(xs), 3)
    cols = vps + (vas if qa_tab is not None else [])
    K = np.stack(cols, axis=1)
    A = np.block([[M, -K], [-K.T, np.zeros((len(cols), len(cols)))]])
    b = np.concatenate([np.zeros(M.shape[0]), -np.asarray(qps)] + ([-np.asarray(qas)] if qa_tab is not None else []))
    return A, b, np.concatenate(xs)


def _oracle_residual(base, phase, angle, omega, qp_tab, qa_tab, spring, F, speeds):
    """||A z - b|| / ||b|| of one member's solution in the oracle's coupled operator."""
    A, b, x = _oracle_system(base, phase, angle, omega, qp_tab, qa_tab, spring)
    z = np.concatenate([F.reshape(-1), speeds])
    return np.linalg.norm(A @ z - b) / np.linalg.norm(b), x


@pytest.mark.parametrize("angles,precond", [(False, "diag"), (True, "diag"), (True, "self")])
def test_ensemble_solves_satisfy_the_oracle_system(angles, precond):
    qp_tab, qa_tab = _tables()
    if not angles:
        qa_tab = None
    base = wall_lattice(3, 3, spacing=20.0)
    n = base.size // 3
    E, steps, spring, rtol = 4, 4, 0.2, 1e-10
    rng = np.random.default_rng(3)
    phase = rng.uniform(0, 2 * math.pi, (E, n))
    angle = rng.uniform(-0.2, 0.2, (E, n)) if angles else None
    dt = np.array([1.0, 0.9, 1.1, 1.05]) / 40.0
    omega = 2 * math.pi
    with CiliaEnsemble(base, E, qp_tab, qa_tab, spring=spring, precond=precond, rtol=rtol,
                       max_iter=4000, s_tol=1e-12) as C:
        C.set_state(phase, np.full((E, n), omega), dt, angle=angle)
        prev = None
        for nt in range(steps + 1):
            qp, qa, st = C.step(nt)
            assert all(s["converged"] and s["rel_residual"] <= rtol for s in st), st
            sp = C.speeds()
            F, speeds = C.solution()
            x = C.positions()
            if prev is not None:  # the advance before this step
                assert np.abs(sp["phase"] - (prev["phase"] + prev["phase_dot"] * dt[:, None])).max() < 1e-14
                if angles:
                    assert np.abs(sp["angle"] - (prev["angle"] + prev["angle_dot"] * dt[:, None])).max() < 1e-15
            for e in range(E):
                ang = sp["angle"][e] if angles else np.zeros(n)
                res, xo = _oracle_residual(base, sp["phase"][e], ang, omega, qp_tab, qa_tab, spring,
                                           F[e], speeds[e])
                assert res <= 2 * rtol, (e, nt, res)
                assert np.abs(x[e].reshape(-1) - xo).max() < 1e-12, (e, nt)
            prev = sp


def test_first_solves_match_single_system_contexts():
    qp_tab, qa_tab = _tables()
    base = wall_lattice(3, 3, spacing=20.0)
    n = base.size // 3
    E = 3
    rng = np.random.default_rng(5)
    phase = rng.uniform(0, 2 * math.pi, (E, n))
    angle = rng.uniform(-0.2, 0.2, (E, n))
    omega = 2 * math.pi
    with CiliaEnsemble(base, E, qp_tab, qa_tab, spring=0.2, precond="diag", rtol=1e-10,
                       s_tol=1e-12) as C:
        C.set_state(phase, np.full((E, n), omega), np.full(E, 1.0 / 40), angle=angle)
        qp, qa, _ = C.step(0)
        F, speeds = C.solution()
    for e in range(E):
        with PrescribedCilia(base, phase[e], np.full(n, omega), dt=1.0 / 40, precond="diag",
                             rtol=1e-10, s_tol=1e-12) as C1:
            C1.set_phases(angle=angle[e])
            C1.set_driving(qp_tab, qa_tab, 0.2)
            q1, a1, _ = C1.step(0)
            s1 = C1.speeds()
            F1 = C1.state()["f"]
        A, b, _ = _oracle_system(base, phase[e], angle[e], omega, qp_tab, qa_tab, 0.2)
        zs = np.linalg.solve(A, b)
        inv_norm = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
        ns3 = 3 * 20 * n
        for z, q_p, q_a in ((np.concatenate([F[e].reshape(-1), speeds[e]]), qp[e], qa[e]),
                            (np.concatenate([F1.reshape(-1), s1["phase_dot"], s1["angle_dot"]]), q1, a1)):
            r = np.linalg.norm(A @ z - b)
            assert r <= 2e-10 * np.linalg.norm(b)
            assert np.linalg.norm(z - zs) <= 1.01 * inv_norm * r + 1e-12 * np.linalg.norm(zs), e
            # the generalised forces K^T F are the oracle's driving forces up to the
            # constraint rows of the residual
            q = np.concatenate([q_p, q_a])
            assert np.linalg.norm(q + b[ns3:]) <= r * (1 + 1e-6) + 1e-14 * np.linalg.norm(b), e


def test_members_do_not_depend_on_the_ensemble():
    qp_tab, qa_tab = _tables()
    base = wall_lattice(2, 3, spacing=20.0)
    n = base.size // 3
    E = 5
    rng = np.random.default_rng(8)
    phase = rng.uniform(0, 2 * math.pi, (E, n))
    angle = rng.uniform(-0.2, 0.2, (E, n))
    dt = np.linspace(0.9, 1.1, E) / 40.0
    kw = dict(spring=0.2, precond="diag", rtol=1e-10, s_tol=1e-12)

    def run(idx):
        with CiliaEnsemble(base, len(idx), qp_tab, qa_tab, **kw) as C:
            C.set_state(phase[idx], np.full((len(idx), n), 2 * math.pi), dt[idx], angle=angle[idx])
            for nt in range(4):
                qp, qa, _ = C.step(nt)
            return C.speeds(), qp, qa, C.positions()

    sp, qp, qa, x = run(np.arange(E))
    for e in (0, 3):
        s1, q1, a1, x1 = run(np.array([e]))
        for key in ("phase", "phase_dot", "angle", "angle_dot"):
            assert np.array_equal(s1[key][0], sp[key][e]), (e, key)
        assert np.array_equal(q1[0], qp[e]) and np.array_equal(a1[0], qa[e])
        assert np.array_equal(x1[0], x[e])


def _oracle_flow(base, x, ndts, spp, qp_tab, rtol, omega=2 * math.pi):
    """Flow map of one state by direct solves of the oracle's coupled system: a solve at
    the initial phases, then ndts steps of phase += (T / spp) phase_dot and a solve.
    Also returns the first-order bound on the phases of a flow whose solves stop at
    relative residual rtol: sum over the solves that feed an advance of
    dt ||A^-1||_2 rtol ||b||, doubled for the propagation of earlier errors."""
    T, ph = x[0], np.array(x[1:], float)
    n = ph.size
    dt = T / spp

    def speeds(p):
        A, b, _ = _oracle_system(base, p, np.zeros(n), omega, qp_tab, None, 0.0)
        smin = np.linalg.svd(A, compute_uv=False)[-1]
        return np.linalg.solve(A, b)[-n:], dt * rtol * np.linalg.norm(b) / smin

    pd, bound = speeds(ph)
    for nt in range(ndts):
        ph = ph + dt * pd
        if nt + 1 < ndts:
            pd, e = speeds(ph)
            bound += e
    return ph, 2.0 * bound


def test_orbit_layer_batched_flow_maps_match_the_oracle():
    """The batched flow maps and the concurrent single-context ones against the oracle's
    flow maps (direct solves): 20 steps of dt = T / 40 at rtol 1e-10, segment parameters
    bisected to 1e-12 so that the map is smooth.  Each solve's speeds may differ from the
    direct solve by ||A^-1|| rtol ||b|| (condition numbers 1e4 to 1e6 at these states), so
    each member is held to its own first-order bound (5e-3 to 1.5e-2 here; measured errors
    2e-7 to 2e-5).  What the test catches is the orbit layer's wiring -- the time step
    T / steps_per_period, the solve before the first advance, the step count, the order
    of members -- each of which moves the phases by 0.05 or more; each step's solve is
    held to 2 rtol in the oracle's operator by the first test of this file."""
    qp_tab, _ = _tables()
    base = wall_lattice(2, 2, spacing=40.0)
    rtol = 1e-10
    flow = orbits.CiliaFlow(base, qp_tab, steps_per_period=40, s_tol=1e-12, rtol=rtol)
    rng = np.random.default_rng(3)
    xs = [np.concatenate([[1.0 + 0.05 * k], rng.uniform(0, 2 * math.pi, 4)]) for k in range(6)]
    par = flow.ensemble(xs, 20, workers=4, batched=False)
    bat = flow.ensemble(xs, 20)
    for x, a, b in zip(xs, par, bat):
        assert a[0] == x[0] and b[0] == x[0]
        ref, bound = _oracle_flow(base, x, 20, 40, qp_tab, rtol)
        assert bound < 0.02
        assert np.abs(b[1:] - ref).max() < bound, (np.abs(b[1:] - ref).max(), bound)
        assert np.abs(a[1:] - ref).max() < bound, (np.abs(a[1:] - ref).max(), bound)
