"""CPU pins of the free-swimmer oracle (oracle/swimmer_oracle.py): the Schur-complement
inverse of the approximate Jacobian against a dense solve of the assembled block
matrix, and a converged step against the balances the physics fixes (a force- and
torque-free swimmer exerts no net force or torque on the fluid; the body rows of the
BDF residual vanish)."""
import math

import numpy as np
import pytest

import body_oracle as bo
import filament_oracle as fo
import oracle
import swimmer_oracle as so
from cufcm_fil_amd import geometry as geo


def _swimmer(nseg=10, n_fil=6, nb=100, R=6.0, wall=False, f_ext=None, end_force=100.0):
    P = fo.Params(nseg=nseg, dl=2.2, dimensionless_force=end_force)
    centre = np.array([0.0, 0.0, 40.0])
    ref = geo.fibonacci_sphere(nb, R)
    xs, t = [], []
    dirs = geo.fibonacci_sphere(n_fil, 1.0, offset=0.5)
    for d in dirs:
        base = centre + (R + 2.5) * d      # clear of the blobs' barrier range
        tt = np.tile(d, (nseg, 1))
        xs.append(base + P.dl * np.arange(nseg)[:, None] * d)
        t.append(tt)
    xs, t = np.concatenate(xs), np.concatenate(t)
    q = geo.segment_quaternions(t.reshape(-1), nseg).reshape(-1, nseg, 4)
    pert = 0.01 * P.end_force * np.array([1.0, 0.0, 0.0])
    fils = [fo.Filament(P, xs[f * nseg:(f + 1) * nseg], q[f], pert) for f in range(n_fil)]
    body = so.Body(centre, [1.0, 0.0, 0.0, 0.0], ref)
    K = bo.k_matrix(ref, 1)
    mob = bo.body_mobility(oracle.mobility_matrix(centre + ref, 3, wall=wall), K)
    return P, so.Swimmer(P, body, fils, mob, f_ext), wall


def _velocity(wall):
    def vel(sw, F, fb):
        xs = np.concatenate([f.x for f in sw.fils]).reshape(-1)
        xb = sw.body.blobs().reshape(-1)
        ns, nb = xs.size // 3, xb.size // 3
        ids, idb = np.arange(ns), np.arange(ns, ns + nb)
        v_b = oracle.apply(xb, 3, xs, 6, F.reshape(-1), id_tgt=idb, id_src=ids, wall=wall)
        A = bo.saddle_matrix(oracle.mobility_matrix(xb, 3, wall=wall),
                             bo.k_matrix(sw.body.offsets(), 1))
        z = np.linalg.solve(A, np.concatenate([-v_b, -fb]))
        sw.f_blob = z[:3 * nb]
        V = oracle.apply(xs, 6, xs, 6, F.reshape(-1), wall=wall)
        V += oracle.apply(xs, 6, xb, 3, z[:3 * nb], id_tgt=ids, id_src=idb, wall=wall)
        return V.reshape(-1, 6), z[3 * nb:]
    return vel


def test_schur_inverse_is_the_block_jacobian_inverse():
    P, sw, wall = _swimmer()
    sw.initial_guess(0)
    rng = np.random.default_rng(3)
    for f in sw.fils:                      # a generic state
        f.lam += rng.standard_normal(f.lam.shape)
        f.clamp += rng.standard_normal(3)
    sw.body.u = 0.01 * rng.standard_normal(3)
    Fint = np.concatenate([f.internal_forces(0) for f in sw.fils])
    sw.prepare_jacobian(0, Fint)
    n6 = 6 * P.nseg
    nf = len(sw.fils)
    J = np.zeros((nf * n6 + 6, nf * n6 + 6))
    for i in range(nf):
        J[i * n6:(i + 1) * n6, i * n6:(i + 1) * n6] = sw.Ja[i]
        J[i * n6:(i + 1) * n6, nf * n6:] = sw.A[i]
        J[nf * n6:, i * n6:(i + 1) * n6] = sw.B[i]
    J[nf * n6:, nf * n6:] = sw.S0
    r = rng.standard_normal(J.shape[0])
    got = sw.jinv_mult(r)
    assert np.allclose(got, np.linalg.solve(J, r), rtol=1e-9, atol=1e-9 * np.abs(got).max())


@pytest.mark.parametrize("wall", [False, True])
def test_free_swimmer_steps_balance_force_and_torque(wall):
    P, sw, _ = _swimmer(wall=wall)
    vel = _velocity(wall)
    for nt in range(2):
        it = so.implicit_step(sw, nt, vel, wall=wall)
        assert 0 < it < 60, it
        # converged residual, body rows included
        F_all = np.concatenate([f.internal_forces(nt) for f in sw.fils])
        xs = np.concatenate([f.x for f in sw.fils])
        xb = sw.body.blobs()
        F_all[:, :3] += fo.barrier_forces(xs, P.nseg, P, xb, wall=wall)
        rep = so.blob_barrier_forces(xs, xb, P, xb.shape[0])
        fb = sw.body_force(np.concatenate([f.internal_forces(nt) for f in sw.fils]), rep)
        V, U = vel(sw, F_all, fb)
        e = sw.errors(V, U, nt)
        assert sw.badness(e) <= 1.0
        # no external load: the swimmer exerts no net force or torque on the fluid
        Fb = sw.f_blob.reshape(-1, 3)
        tot_f = F_all[:, :3].sum(0) + Fb.sum(0)
        c = sw.body.x
        tot_t = (F_all[:, 3:] + np.cross(xs - c, F_all[:, :3])).sum(0) + np.cross(xb - c, Fb).sum(0)
        scale = np.abs(F_all[:, :3]).max()
        assert np.abs(tot_f).max() < 1e-9 * scale, tot_f
        assert np.abs(tot_t).max() < 1e-9 * scale * 20.0, tot_t
    # the body moved
    assert np.linalg.norm(sw.body.x - np.array([0.0, 0.0, 40.0])) > 0.0


def test_external_force_sets_the_net_force_on_the_fluid():
    f_ext = np.array([0.0, 0.0, -50.0, 0.0, 0.0, 0.0])
    P, sw, wall = _swimmer(f_ext=f_ext, end_force=0.0)
    vel = _velocity(wall)
    so.implicit_step(sw, 0, vel, wall=wall)
    F_all = np.concatenate([f.internal_forces(0) for f in sw.fils])
    xs = np.concatenate([f.x for f in sw.fils])
    xb = sw.body.blobs()
    F_all[:, :3] += fo.barrier_forces(xs, P.nseg, P, xb, wall=wall)
    fb = sw.body_force(np.concatenate([f.internal_forces(0) for f in sw.fils]),
                       so.blob_barrier_forces(xs, xb, P, xb.shape[0]))
    vel(sw, F_all, fb)
    tot = F_all[:, :3].sum(0) + sw.f_blob.reshape(-1, 3).sum(0)
    assert np.allclose(tot, f_ext[:3], atol=1e-8 * 50.0)
    # it sinks: downward velocity of the body
    assert sw.vel[2] < 0.0


def test_sphere_correction_blocks_are_the_unequal_sphere_rpy_blocks():
    """The blocks of the sphere correction (swimmer.cpp:809-840): translation-
    translation and rotation-rotation are the RPY blocks of the body's (radius R) velocity
    due to a load on a segment (radius a); translation-rotation is the segment's
    translational velocity due to a torque on the body (the orientation the reference
    uses in the body rows)."""
    def rpy_block(xi, ai, xj, aj):
        M = np.zeros((6, 6))
        for k in range(6):
            F = np.zeros(6)
            F[k] = 1.0
            M[:, k] = oracle.free_pair(xi, ai, xj, aj, F, 6, 6, 1.0)
        return M

    rng = np.random.default_rng(4)
    for R, a in ((6.0, 1.0), (20.0, 1.0), (3.0, 0.5)):
        for _ in range(3):
            d = rng.normal(size=3)
            xs = (R + a + rng.uniform(0.5, 30.0)) * d / np.linalg.norm(d)
            Mtt, Mtr, Mrr = so.sphere_blocks(xs, R, a)
            Mbs = rpy_block(np.zeros(3), R, xs, a)
            Msb = rpy_block(xs, a, np.zeros(3), R)
            assert np.allclose(Mtt, Mbs[:3, :3], rtol=1e-12, atol=1e-18)
            assert np.allclose(Mrr, Mbs[3:, 3:], rtol=1e-12, atol=1e-18)
            assert np.allclose(Mtr, Msb[:3, 3:], rtol=1e-12, atol=1e-18)


def test_sphere_correction_changes_convergence_not_the_state():
    out = {}
    for R in (0.0, 6.0):
        P, sw, wall = _swimmer()
        sw.sphere_R = R
        vel = _velocity(wall)
        its = [so.implicit_step(sw, nt, vel, wall=wall) for nt in range(2)]
        out[R] = (its, np.concatenate([f.x for f in sw.fils]), sw.body.x.copy())
    # both converge to the Broyden tolerance (1e-4) around the same state
    assert np.abs(out[0.0][1] - out[6.0][1]).max() < 1e-3
    assert np.abs(out[0.0][2] - out[6.0][2]).max() < 1e-3
    print("broyden iterations without / with the sphere correction", out[0.0][0], out[6.0][0])
