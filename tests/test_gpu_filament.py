"""GPU parity of the implicit filament step (through the C ABI) against the
numpy oracle (oracle/filament_oracle.py) and the FP64 mobility oracle."""
import numpy as np
import pytest

import filament_oracle as fo
import oracle
from cufcm_fil_amd import geometry as geo
from cufcm_fil_amd.filaments import DEFAULTS, FilamentSystem, end_force

pytestmark = pytest.mark.gpu


def _params(nseg=20, dt=DEFAULTS["dt"]):
    return fo.Params(nseg=nseg, dl=2.2, dt=dt, dimensionless_force=220.0)


def _oracle_fils(x, t, nseg, P, pert=None):
    q = geo.segment_quaternions(t, nseg).reshape(-1, nseg, 4)
    X = x.reshape(-1, nseg, 3)
    fils = []
    for f in range(X.shape[0]):
        fil = fo.Filament(P, X[f], q[f], None if pert is None else pert[3 * f:3 * f + 3])
        fils.append(fil)
    return fils


def _random_state(sysm, fils, seed):
    rng = np.random.default_rng(seed)
    n = sysm.n_seg
    u = 0.05 * rng.standard_normal((n, 3))
    u[::sysm.nseg] = 0.0
    lam = 3.0 * rng.standard_normal((n, 3))
    clamp = rng.standard_normal((sysm.nfil, 3))
    um1 = 0.02 * rng.standard_normal((n, 3))
    q0 = sysm.state()["q"].reshape(n, 4)
    q = np.array([fo.qmul(fo.lie_exp(u[i]), q0[i]) for i in range(n)])
    sysm.set_state(q=q, qm1=q0, u=u, um1=um1, lam=lam, clamp=clamp)
    for f, fil in enumerate(fils):
        s = slice(f * sysm.nseg, (f + 1) * sysm.nseg)
        fil.u, fil.um1, fil.lam, fil.clamp = u[s].copy(), um1[s].copy(), lam[s].copy(), clamp[f].copy()
        fil.qm1, fil.q = q0[s].copy(), q[s].copy()
    return q


def _system(nx=2, ny=2, seed=0):
    x, t = geo.wave_lattice(nx=nx, ny=ny, spacing=12.0, return_tangents=True)
    return x, t


def test_internal_loads_and_jacobian_match_oracle():
    x, t = _system()
    P = _params()
    with FilamentSystem(x, t) as S:
        fils = _oracle_fils(x, t, 20, P)
        _random_state(S, fils, 1)
        for nt in (0, 3):
            got = S.internal_forces(nt).reshape(-1, 6)
            ref = np.concatenate([f.internal_forces(nt) for f in fils])
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
            Jg = S.jacobian_inverse(nt)
            for f, fil in enumerate(fils):
                Ja = fil.approx_jacobian(nt, fil.internal_forces(nt))
                assert np.abs(Ja @ Jg[f] - np.eye(120)).max() < 1e-9
                Jr = np.linalg.inv(Ja)
                assert np.abs(Jg[f] - Jr).max() <= 1e-9 * np.abs(Jr).max()


def test_residual_update_and_guess_match_oracle():
    x, t = _system()
    P = _params()
    rng = np.random.default_rng(4)
    with FilamentSystem(x, t) as S:
        fils = _oracle_fils(x, t, 20, P)
        _random_state(S, fils, 2)
        v = 0.1 * rng.standard_normal((S.n_seg, 6))
        for nt in (0, 2):
            e, bad = S.errors(nt, v.reshape(-1))
            ref = np.concatenate([f.errors(v[i * 20:(i + 1) * 20], nt) for i, f in enumerate(fils)])
            assert np.allclose(e, ref, rtol=1e-13, atol=1e-13)
            assert np.isclose(bad, fo.badness(ref, P, len(fils)), rtol=1e-12)
        upd = 1e-3 * rng.standard_normal(6 * S.n_seg)
        S.update(upd)
        for i, f in enumerate(fils):
            f.update(upd[i * 120:(i + 1) * 120])
        st = S.state()
        assert np.allclose(st["x"].reshape(-1, 3), np.concatenate([f.x for f in fils]), atol=1e-12)
        assert np.allclose(st["lam"].reshape(-1, 3), np.concatenate([f.lam for f in fils]), atol=1e-13)
        assert np.allclose(st["clamp"].reshape(-1, 3), np.array([f.clamp for f in fils]), atol=1e-13)
        for nt in (1, 2, 3):
            S.initial_guess(nt)
            for f in fils:
                f.initial_guess(nt)
        st = S.state()
        assert np.allclose(st["x"].reshape(-1, 3), np.concatenate([f.x for f in fils]), atol=1e-12)
        assert np.allclose(st["q"].reshape(-1, 4), np.concatenate([f.q for f in fils]), atol=1e-13)
        assert np.allclose(st["lam"].reshape(-1, 3), np.concatenate([f.lam for f in fils]), atol=1e-12)


def test_barrier_forces_match_oracle():
    # two filaments crossing each other plus the wall repulsion
    x, t = geo.tilted_lattice(nx=2, ny=1, spacing=2.5, nseg=20, z0=0.9, max_polar=0.5, seed=5,
                              return_tangents=True)
    P = _params()
    with FilamentSystem(x, t) as S:
        got = S.barrier_forces().reshape(-1, 3)
    ref = fo.barrier_forces(x, 20, P)
    assert np.abs(ref).max() > 1.0
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-9)


def _oracle_velocity(xs, F):
    return oracle.apply(xs.reshape(-1), 6, xs.reshape(-1), 6, F.reshape(-1)).reshape(-1, 6)


@pytest.mark.parametrize("nsteps", [3])
def test_implicit_steps_match_oracle(nsteps):
    x, t = _system(nx=2, ny=1)
    P = _params()
    pert = 0.01 * P.end_force * np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0])
    with FilamentSystem(x, t, pert=pert) as S:
        fils = _oracle_fils(x, t, 20, P, pert)
        for nt in range(nsteps):
            st = S.step(nt)
            it = fo.implicit_step(fils, nt, _oracle_velocity)
            assert st["converged"] and st["broyden_iterations"] == it, (nt, st, it)
            g = S.state()
            assert np.abs(g["x"].reshape(-1, 3) - np.concatenate([f.x for f in fils])).max() < 1e-8
            assert np.abs(g["lam"].reshape(-1, 3) - np.concatenate([f.lam for f in fils])).max() < 1e-6


@pytest.mark.parametrize("cap", [1, 2, 3])
def test_capped_broyden_step_matches_oracle(cap):
    """A step stopped by the Broyden iteration cap (MAX_BROYDEN_ITER) before it converges:
    the library forms each next low-rank column before it reads the residual, except on
    the last allowed iteration; the capped state must still be the oracle's after the same
    number of iterations (broyden_solver.cpp:69-125)."""
    x, t = _system(nx=2, ny=1)
    P = _params()
    pert = 0.01 * P.end_force * np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0])
    with FilamentSystem(x, t, pert=pert, max_broyden=cap) as S:
        fils = _oracle_fils(x, t, 20, P, pert)
        st = S.step(0)
        it = fo.implicit_step(fils, 0, _oracle_velocity, max_iter=cap)
        assert it == cap and st["broyden_iterations"] == cap and not st["converged"], (st, it)
        g = S.state()
        assert np.abs(g["x"].reshape(-1, 3) - np.concatenate([f.x for f in fils])).max() < 1e-8
        assert np.abs(g["lam"].reshape(-1, 3) - np.concatenate([f.lam for f in fils])).max() < 1e-6


def test_step_with_fixed_body_matches_oracle():
    xs, xb, c, d, t = geo.sphere_over_wall(n_blob=200, n_fil=6, radius=8.0, gap=50.0,
                                            return_tangents=True)
    P = _params()
    nb = xb.size // 3
    Mbb = oracle.mobility_matrix(xb, 3)
    ns = xs.size // 3
    ids, idb = np.arange(ns), np.arange(ns, ns + nb)

    def vel(xseg, F):
        xseg = xseg.reshape(-1)
        v_b = oracle.apply(xb, 3, xseg, 6, F.reshape(-1), id_tgt=idb, id_src=ids)
        fb = np.linalg.solve(Mbb, -v_b)
        v = oracle.apply(xseg, 6, xseg, 6, F.reshape(-1))
        v += oracle.apply(xseg, 6, xb, 3, fb, id_tgt=ids, id_src=idb)
        return v.reshape(-1, 6)

    with FilamentSystem(xs, t, x_blob=xb, blob_rtol=1e-12) as S:
        fils = _oracle_fils(xs, t, 20, P)
        for nt in range(2):
            st = S.step(nt)
            it = fo.implicit_step(fils, nt, vel, x_blob=xb)
            assert st["converged"] and st["broyden_iterations"] == it, (nt, st, it)
            assert st["max_gmres_rel_residual"] <= 1e-12
            g = S.state()
            assert np.abs(g["x"].reshape(-1, 3) - np.concatenate([f.x for f in fils])).max() < 1e-7


def test_mixed_precision_step_reaches_tolerances():
    x, t = _system(nx=2, ny=2)
    with FilamentSystem(x, t, precision="fp32c", solve_mode="mixed") as S:
        for nt in range(2):
            st = S.step(nt)
            assert st["converged"] and st["final_error"] <= 1.0


def test_projected_guess_reduces_gmres_work():
    xs, xb, c, d, t = geo.sphere_over_wall(n_blob=800, n_fil=30, radius=14.0, gap=50.0,
                                            return_tangents=True)
    out = {}
    for guess in ("previous", "projected"):
        with FilamentSystem(xs, t, x_blob=xb, guess=guess, blob_precond="near") as S:
            its = 0
            for nt in range(3):
                st = S.step(nt)
                assert st["converged"] and st["max_gmres_rel_residual"] <= 1e-10
                its += st["gmres_iterations"]
            out[guess] = (its, S.state()["x"])
    assert out["projected"][0] < out["previous"][0], out
    # both converge to the same state within the Broyden tolerance
    assert np.abs(out["projected"][1] - out["previous"][1]).max() < 1e-3


def test_overlapped_segment_product_gives_the_same_steps():
    """M_ss f_seg on the side stream beside the blob solve: the same step in FP64 (the
    products' source splits differ from the one-launch path, so agreement is to
    rounding)."""
    xs, xb, c, d, t = geo.sphere_over_wall(n_blob=800, n_fil=30, radius=14.0, gap=50.0,
                                            return_tangents=True)
    out = {}
    for overlap in (False, True):
        with FilamentSystem(xs, t, x_blob=xb, guess="projected", blob_precond="dense",
                            overlap=overlap) as S:
            its = []
            for nt in range(3):
                st = S.step(nt)
                assert st["converged"] and st["max_gmres_rel_residual"] <= 1e-10
                its.append(st["broyden_iterations"])
            out[overlap] = (its, S.state()["x"])
    assert out[True][0] == out[False][0]
    assert np.abs(out[True][1] - out[False][1]).max() < 1e-9
