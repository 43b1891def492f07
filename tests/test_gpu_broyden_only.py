"""Blob forces among the Broyden unknowns (the reference's SOLVER_TYPE 0) on the GPU,
through the C ABI, against the numpy oracle (oracle/broyden_only_oracle.py):

* a free swimmer (body position and orientation and every blob force among the
  unknowns, wall and no wall): the same Broyden iterations as the oracle at every step,
  segment positions within 1e-7, body state within 1e-8, blob forces (themselves Broyden
  unknowns here, carrying the same rounding-path differences as the positions) within
  1e-7 relative;
* filaments beside a body held fixed above the wall (blob forces among the unknowns):
  the same iterations, positions within 1e-7, blob forces within 1e-7 relative;
* FP32C pair sums: the same converged state as FP64 to the Broyden tolerance's scale."""
import numpy as np
import pytest

import broyden_only_oracle as bx
import filament_oracle as fo
import oracle
from cufcm_fil_amd import geometry as geo
from cufcm_fil_amd.filaments import FilamentSystem
from test_broyden_only_oracle import _full_velocity, _swimmer

pytestmark = pytest.mark.gpu


def _gpu_swimmer(sw, P, wall, precision="fp64"):
    xs = np.concatenate([f.x for f in sw.fils]).reshape(-1)
    t = np.concatenate([[fo.tangent(q) for q in f.q] for f in sw.fils]).reshape(-1)
    pert = np.concatenate([f.pert for f in sw.fils])
    S = FilamentSystem(xs, t, x_blob=sw.body.blobs().reshape(-1), wall=wall, nseg=P.nseg,
                       dimensionless_force=P.end_force * P.dl ** 2 * P.nseg ** 2 / P.kb,
                       pert=pert, broyden_only=True, precision=precision)
    S.attach_bodies(sw.body.ref.reshape(-1), sw.body.x, sw.body.q, mobility=sw.mob_ref.reshape(-1))
    return S


@pytest.mark.parametrize("wall", [False, True])
def test_free_swimmer_steps_match_oracle(wall):
    P, sw = _swimmer(wall=wall)
    vel = _full_velocity(wall)
    with _gpu_swimmer(sw, P, wall) as S:
        assert S.n_unknowns == 6 * S.n_seg + 6 + 3 * S.n_blob
        for nt in range(3):
            st = S.step(nt)
            it = bx.swimmer_step(sw, nt, vel, wall=wall)
            assert st["converged"] and st["broyden_iterations"] == it, (nt, st, it)
            assert st["gmres_solves"] == 0
            xs = np.concatenate([f.x for f in sw.fils])
            assert np.abs(S.state()["x"].reshape(-1, 3) - xs).max() < 1e-7, nt
            b = S.body_state()
            assert np.abs(b["x"] - sw.body.x).max() < 1e-8 and np.abs(b["q"] - sw.body.q).max() < 1e-8
            e_fb = float(np.abs(S.blob_forces() - sw.fb).max() / np.abs(sw.fb).max())
            assert e_fb < 1e-7, (nt, e_fb)
    assert np.linalg.norm(sw.body.x - np.array([0.0, 0.0, 40.0])) > 1e-4


def _fixed(n_blob=80, n_fil=4):
    xs, xb, c, d, t = geo.sphere_over_wall(n_blob=n_blob, n_fil=n_fil, radius=5.0, gap=30.0,
                                            return_tangents=True)
    P = fo.Params(nseg=20, dl=2.2, dimensionless_force=220.0)
    q = geo.segment_quaternions(t, 20).reshape(-1, 20, 4)
    X = xs.reshape(-1, 20, 3)
    fils = [fo.Filament(P, X[f], q[f]) for f in range(X.shape[0])]
    return xs, xb, t, P, bx.FixedBlobs(P, fils, xb)


def test_fixed_body_steps_match_oracle():
    xs, xb, t, P, O = _fixed()

    def vel(xseg, F, fb):
        V, vb = oracle.apply_full(xseg, F.reshape(-1), xb, np.asarray(fb).reshape(-1))
        return V.reshape(-1, 6), vb.reshape(-1, 3)

    with FilamentSystem(xs, t, x_blob=xb, broyden_only=True, nseg=20, dl=2.2,
                        dimensionless_force=220.0) as S:
        assert S.n_unknowns == 6 * S.n_seg + 3 * S.n_blob
        for nt in range(3):
            st = S.step(nt)
            it = bx.fixed_step(O, nt, vel)
            assert st["converged"] and st["broyden_iterations"] == it, (nt, st, it)
            xo = np.concatenate([f.x for f in O.fils])
            assert np.abs(S.state()["x"].reshape(-1, 3) - xo).max() < 1e-7, nt
            e_fb = float(np.abs(S.blob_forces() - O.fb).max() / np.abs(O.fb).max())
            assert e_fb < 1e-7, (nt, e_fb)


def test_fp32c_pair_sums_reach_the_fp64_state():
    xs, xb, t, P, _ = _fixed()
    out = {}
    for prec in ("fp64", "fp32c"):
        with FilamentSystem(xs, t, x_blob=xb, broyden_only=True, nseg=20, dl=2.2,
                            dimensionless_force=220.0, precision=prec) as S:
            for nt in range(3):
                st = S.step(nt)
                assert st["converged"], (prec, nt, st)
            out[prec] = (S.state()["x"], S.blob_forces())
    # both meet TOL = 1e-4 on every row; their states agree on that scale
    assert np.abs(out["fp64"][0] - out["fp32c"][0]).max() < 1e-3
    assert np.abs(out["fp64"][1] - out["fp32c"][1]).max() < 1e-2 * np.abs(out["fp64"][1]).max()
