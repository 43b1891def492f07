"""CPU pins of the SOLVER_TYPE 0 oracle (oracle/broyden_only_oracle.py): blob forces (and
a free body's position and orientation) among the Broyden unknowns.

* the blob-force guess is the BDF extrapolation: exact on linear sequences at step 2 and
  on quadratic ones from step 3 (rigid_body.cpp:459-490);
* the body velocity the blob rows use is exact for quadratic motion (BDF2) and for
  rotation at a constant rate about a fixed axis (Lie-algebra BDF2 through dexp);
* jinv_mult is the exact inverse of the assembled block matrix (filament rows, body
  rows, blob rows with the self mobility), and the blob terms of the body's Schur
  complement are exactly the elimination of the blob rows;
* converged steps reach the state of the GMRES-inside-Broyden variant (SOLVER_TYPE 1):
  the two solve the same equations and differ only in how Broyden gets there -- a free
  swimmer (body moving, force- and torque-free) and filaments beside a fixed body."""
import math

import numpy as np
import pytest

import body_oracle as bo
import broyden_only_oracle as bx
import filament_oracle as fo
import oracle
import swimmer_oracle as so
from cufcm_fil_amd import geometry as geo


def test_blob_force_guess_extrapolates_linear_then_quadratic():
    rng = np.random.default_rng(1)
    a, b, c = (rng.standard_normal((5, 3)) for _ in range(3))
    lin = lambda t: a + b * t
    quad = lambda t: a + b * t + c * t * t
    for f, first_exact in ((lin, 2), (quad, 3)):
        fb, m1, m2 = f(0), np.zeros((5, 3)), np.zeros((5, 3))
        for nt in range(1, 7):
            fb, m1, m2 = bx.blob_force_guess(fb, m1, m2, nt)
            if nt == 1:
                assert np.array_equal(fb, f(0)) and np.array_equal(m1, f(0))
            if nt >= first_exact:
                assert np.allclose(fb, f(nt), rtol=0, atol=1e-12 * np.abs(f(nt)).max()), nt
            fb = f(nt)       # the step converges to the sequence's value


def test_body_velocity_is_exact_for_quadratic_motion_and_steady_rotation():
    dt = 0.37
    x0, v0, acc = np.array([1.0, -2.0, 3.0]), np.array([0.3, 0.1, -0.2]), np.array([0.05, -0.02, 0.01])
    x = lambda t: x0 + v0 * t + acc * t * t
    Om0 = np.array([0.2, -0.5, 0.7])
    nt = 3
    t = nt * dt
    u = Om0 * dt
    V, Om = bx.body_velocity(x(t), x(t - dt), x(t - 2 * dt), u, u, dt, nt, 1)
    assert np.allclose(V, v0 + 2 * acc * t, rtol=0, atol=1e-13)
    assert np.allclose(Om, Om0, rtol=0, atol=1e-13)
    # Euler (first step): finite difference and u / dt
    V1, Om1 = bx.body_velocity(x(dt), x(0), x(0), u, u, dt, 0, 1)
    assert np.allclose(V1, (x(dt) - x(0)) / dt) and np.allclose(Om1, Om0)
    # a planted sign error in the BDF2 combination fails the check
    bad = -(x(t) + (4.0 / 3.0) * x(t - dt) + (1.0 / 3.0) * x(t - 2 * dt)) / (-2.0 / 3.0 * dt)
    assert not np.allclose(bad, v0 + 2 * acc * t, atol=1e-6)


def _swimmer(nseg=8, n_fil=4, nb=60, R=5.0, wall=False, end_force=80.0, broyden_only=True):
    P = fo.Params(nseg=nseg, dl=2.2, dimensionless_force=end_force)
    centre = np.array([0.0, 0.0, 40.0])
    ref = geo.fibonacci_sphere(nb, R)
    xs, t = [], []
    for d in geo.fibonacci_sphere(n_fil, 1.0, offset=0.5):
        base = centre + (R + 2.5) * d
        xs.append(base + P.dl * np.arange(nseg)[:, None] * d)
        t.append(np.tile(d, (nseg, 1)))
    xs, t = np.concatenate(xs), np.concatenate(t)
    q = geo.segment_quaternions(t.reshape(-1), nseg).reshape(-1, nseg, 4)
    pert = 0.01 * P.end_force * np.array([1.0, 0.0, 0.0])
    fils = [fo.Filament(P, xs[f * nseg:(f + 1) * nseg], q[f], pert) for f in range(n_fil)]
    body = so.Body(centre, [1.0, 0.0, 0.0, 0.0], ref)
    mob = bo.body_mobility(oracle.mobility_matrix(centre + ref, 3, wall=wall), bo.k_matrix(ref, 1))
    cls = bx.FreeSwimmer if broyden_only else so.Swimmer
    return P, cls(P, body, fils, mob)


def _full_velocity(wall):
    def vel(sw, F, fb):
        xs = np.concatenate([f.x for f in sw.fils])
        xb = sw.body.blobs()
        V, vb = oracle.apply_full(xs, F.reshape(-1), xb, np.asarray(fb).reshape(-1), wall=wall)
        return V.reshape(-1, 6), vb.reshape(-1, 3)
    return vel


def _gmres_velocity(wall):
    def vel(sw, F, fb):
        xs = np.concatenate([f.x for f in sw.fils]).reshape(-1)
        xb = sw.body.blobs().reshape(-1)
        ns, nb = xs.size // 3, xb.size // 3
        ids, idb = np.arange(ns), np.arange(ns, ns + nb)
        v_b = oracle.apply(xb, 3, xs, 6, F.reshape(-1), id_tgt=idb, id_src=ids, wall=wall)
        A = bo.saddle_matrix(oracle.mobility_matrix(xb, 3, wall=wall), bo.k_matrix(sw.body.offsets(), 1))
        z = np.linalg.solve(A, np.concatenate([-v_b, -fb]))
        V = oracle.apply(xs, 6, xs, 6, F.reshape(-1), wall=wall)
        V += oracle.apply(xs, 6, xb, 3, z[:3 * nb], id_tgt=ids, id_src=idb, wall=wall)
        sw.f_blob = z[:3 * nb]
        return V.reshape(-1, 6), z[3 * nb:]
    return vel


def test_jinv_mult_is_the_inverse_of_the_block_matrix():
    P, sw = _swimmer()
    sw.initial_guess(0)
    rng = np.random.default_rng(7)
    for f in sw.fils:
        f.lam += rng.standard_normal(f.lam.shape)
        f.clamp += rng.standard_normal(3)
    sw.body.u = 0.02 * rng.standard_normal(3)
    sw.fb = rng.standard_normal(sw.fb.shape)
    for nt in (0, 3):
        Fint = np.concatenate([f.internal_forces(nt) for f in sw.fils])
        sw.prepare_jacobian(nt, Fint)
        n6, nf, nb = 6 * P.nseg, len(sw.fils), sw.fb.shape[0]
        n0 = nf * n6
        bmf0 = -6.0 * math.pi
        L = n0 + 6 + 3 * nb
        J = np.zeros((L, L))
        d = sw.body.ref @ sw.body.Q_init.T
        X = sw.S.copy()
        for i in range(nf):
            J[i * n6:(i + 1) * n6, i * n6:(i + 1) * n6] = sw.Ja[i]
            J[i * n6:(i + 1) * n6, n0:n0 + 6] = sw.A[i]
            J[n0:n0 + 6, i * n6:(i + 1) * n6] = sw.D[i]
            X += sw.D[i] @ sw.jinv[i] @ sw.A[i]
        for m in range(nb):
            # rcross(v) w = w x v: the torque of blob force g about the centre is
            # d x g = -rcross(d) g; the blob row's Omega x d is rcross(d) Omega
            G = np.vstack([np.eye(3), -fo.rcross(d[m])])         # body rows <- blob force m
            K = np.hstack([np.eye(3), fo.rcross(d[m])])          # blob row m <- body (V, Omega)
            r0 = n0 + 6 + 3 * m
            J[n0:n0 + 6, r0:r0 + 3] = G
            J[r0:r0 + 3, n0:n0 + 6] = sw.dt_fac * K
            J[r0:r0 + 3, r0:r0 + 3] = np.eye(3) / bmf0
            X += bmf0 * sw.dt_fac * G @ K
        J[n0:n0 + 6, n0:n0 + 6] = X
        # what is left of the body block once the filament and blob rows are eliminated
        # is the body's own rotation-rotation block (the blob torques' dependence on the
        # orientation and the filaments' tether and clamp terms): the blob terms of the
        # Schur complement are exactly the elimination of the blob rows
        sc = np.abs(X).max()
        assert np.abs(X[:3, :]).max() < 1e-12 * sc and np.abs(X[3:, :3]).max() < 1e-12 * sc, nt
        r = rng.standard_normal(L)
        got = sw.jinv_mult(r)
        assert np.allclose(got, np.linalg.solve(J, r), rtol=1e-9, atol=1e-9 * np.abs(got).max()), nt


@pytest.mark.parametrize("wall", [False, True])
def test_free_swimmer_reaches_the_gmres_variant_state(wall):
    out = {}
    for broyden_only in (True, False):
        P, sw = _swimmer(wall=wall, broyden_only=broyden_only)
        its = []
        for nt in range(3):
            if broyden_only:
                its.append(bx.swimmer_step(sw, nt, _full_velocity(wall), wall=wall))
            else:
                its.append(so.implicit_step(sw, nt, _gmres_velocity(wall), wall=wall))
        out[broyden_only] = (its, np.concatenate([f.x for f in sw.fils]), sw.body.x.copy(),
                             sw.body.q.copy(), sw)
    its0, x0, b0, q0, sw0 = out[True]
    its1, x1, b1, q1, _ = out[False]
    assert all(0 < i < 400 for i in its0), its0
    # the body moved, and both variants agree to the Broyden tolerance's scale
    assert np.linalg.norm(b0 - np.array([0.0, 0.0, 40.0])) > 1e-4
    assert np.abs(x0 - x1).max() < 2e-3 and np.abs(b0 - b1).max() < 2e-3
    assert np.abs(q0 - q1).max() < 2e-3
    # at convergence the blob rows and the force and torque balance hold (free swimmer:
    # the segment loads plus the blob forces exert no net force or torque)
    V, vb = sw0.last
    assert np.abs(sw0.errors(V, vb, np.zeros(6), 2)[-3 * sw0.fb.shape[0]:]).max() <= sw0.P.tol
    print("Broyden iterations, broyden-only / GMRES-inside", its0, its1)


def test_fixed_blobs_reach_the_gmres_variant_state():
    xs, xb, c, d, t = geo.sphere_over_wall(n_blob=80, n_fil=4, radius=5.0, gap=30.0,
                                            return_tangents=True)
    P = fo.Params(nseg=20, dl=2.2, dimensionless_force=220.0)
    q = geo.segment_quaternions(t, 20).reshape(-1, 20, 4)
    X = xs.reshape(-1, 20, 3)
    nb = xb.size // 3
    ns = xs.size // 3
    ids, idb = np.arange(ns), np.arange(ns, ns + nb)
    Mbb = oracle.mobility_matrix(xb, 3)

    def vel_gmres(xseg, F):
        xseg = xseg.reshape(-1)
        v_b = oracle.apply(xb, 3, xseg, 6, F.reshape(-1), id_tgt=idb, id_src=ids)
        fb = np.linalg.solve(Mbb, -v_b)
        v = oracle.apply(xseg, 6, xseg, 6, F.reshape(-1))
        v += oracle.apply(xseg, 6, xb, 3, fb, id_tgt=ids, id_src=idb)
        return v.reshape(-1, 6)

    def vel_full(xseg, F, fb):
        V, vb = oracle.apply_full(xseg, F.reshape(-1), xb, np.asarray(fb).reshape(-1))
        return V.reshape(-1, 6), vb.reshape(-1, 3)

    fa = [fo.Filament(P, X[f], q[f]) for f in range(X.shape[0])]
    fbx = [fo.Filament(P, X[f], q[f]) for f in range(X.shape[0])]
    S = bx.FixedBlobs(P, fbx, xb)
    for nt in range(3):
        ia = fo.implicit_step(fa, nt, vel_gmres, x_blob=xb)
        ib = bx.fixed_step(S, nt, vel_full)
        assert 0 < ib < 400
        xa = np.concatenate([f.x for f in fa])
        xbx = np.concatenate([f.x for f in fbx])
        assert np.abs(xa - xbx).max() < 2e-3, (nt, np.abs(xa - xbx).max(), ia, ib)
    # the converged blob forces hold the blobs still to the tolerance
    xseg = np.concatenate([f.x for f in fbx])
    F = np.concatenate([f.internal_forces(2) for f in fbx])
    F[:, :3] += fo.barrier_forces(xseg, P.nseg, P, xb)
    _, vb = vel_full(xseg, F, S.fb)
    assert np.abs(vb).max() <= P.tol


def test_body_block_torque_terms_are_derivatives_of_their_torques():
    """The blob-force, tether and end-force terms of the body block are
    d/d(delta) [(exp(delta) a) x f] at delta = 0 with f held (central differences, 1e-8);
    a swapped product rcross(a) rcross(f) fails.  The clamp term is the derivative of the
    moment map dexpinv^T(u) lam through second order in u: the difference shrinks as
    |u|^3, while the first-order part alone (or a sign error in the correction) leaves an
    O(|u|) difference."""
    rng = np.random.default_rng(4)
    h = 1e-6

    def fd(fun, x0):
        J = np.zeros((3, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            J[:, k] = (fun(x0 + e) - fun(x0 - e)) / (2 * h)
        return J

    for _ in range(5):
        a, f = rng.standard_normal(3), rng.standard_normal(3)
        J = bx.torque_rotation_block(f, a)
        ref = fd(lambda d: np.cross(fo.rot_mat(fo.lie_exp(d)) @ a, f), np.zeros(3))
        assert np.abs(J - ref).max() <= 1e-8 * max(1.0, np.abs(ref).max())
        assert np.abs(fo.rcross(a) @ fo.rcross(f) - ref).max() > 1e-3
        lam, dirn = rng.standard_normal(3), rng.standard_normal(3)
        dirn /= np.linalg.norm(dirn)
        errs, first = [], []
        for s in (0.1, 0.05, 0.025):
            u = s * dirn
            ex = fd(lambda v: fo.dexpinv_transpose(v, lam), u)
            errs.append(np.abs(bx.clamp_block(u, lam) - ex).max())
            first.append(np.abs(0.5 * fo.rcross(lam) - ex).max())
            wrong = 0.5 * fo.rcross(lam) - (fo.rcross(np.cross(u, lam)) - fo.rcross(u) @ fo.rcross(lam)) / 12.0
            assert np.abs(wrong - ex).max() > 10 * errs[-1]
        assert errs[0] / errs[2] > 30 and errs[2] < 1e-5 * np.abs(lam).max()   # O(|u|^3)
        assert first[0] / first[2] < 5                                         # O(|u|)
