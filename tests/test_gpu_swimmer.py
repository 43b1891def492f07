"""Free swimmers on the GPU (through the C ABI) against the numpy oracle
(oracle/swimmer_oracle.py): a rigid body of blobs carrying elastic filaments, its
position and orientation among the Broyden unknowns.  Same Broyden iterations as the
oracle, positions within 1e-7, body state within 1e-8; the converged swimmer exerts no
net force or torque on the fluid."""
import numpy as np
import pytest

import body_oracle as bo
import filament_oracle as fo
import oracle
import swimmer_oracle as so
from cufcm_fil_amd import geometry as geo
from cufcm_fil_amd.filaments import FilamentSystem
from test_swimmer_oracle import _swimmer, _velocity

pytestmark = pytest.mark.gpu


def _gpu(sw, P, wall, f_ext=None, precond="self", **attach):
    xs = np.concatenate([f.x for f in sw.fils]).reshape(-1)
    t = np.concatenate([[fo.tangent(q) for q in f.q] for f in sw.fils]).reshape(-1)
    pert = np.concatenate([f.pert for f in sw.fils])
    S = FilamentSystem(xs, t, x_blob=sw.body.blobs().reshape(-1), wall=wall, nseg=P.nseg,
                       dimensionless_force=P.end_force * P.dl ** 2 * P.nseg ** 2 / P.kb,
                       blob_rtol=1e-12, blob_precond=precond, pert=pert)
    S.attach_bodies(sw.body.ref.reshape(-1), sw.body.x, sw.body.q, f_ext=f_ext,
                    mobility=sw.mob_ref.reshape(-1), **attach)
    return S


@pytest.mark.parametrize("wall,precond", [(False, "self"), (True, "diag"), (False, "dense"),
                                          (True, "dense")])
def test_free_swimmer_steps_match_oracle(wall, precond):
    """CFM_PC_DENSE is built once and applied in the body frame: exact without the wall
    (a couple of GMRES steps per solve), a preconditioner above it."""
    P, sw, _ = _swimmer(wall=wall)
    vel = _velocity(wall)
    with _gpu(sw, P, wall, precond=precond) as S:
        for nt in range(3):
            st = S.step(nt)
            it = so.implicit_step(sw, nt, vel, wall=wall)
            assert st["converged"] and st["broyden_iterations"] == it, (nt, st, it)
            assert st["max_gmres_rel_residual"] <= 1e-12
            if precond == "dense" and not wall:
                assert st["gmres_iterations"] <= 3 * st["gmres_solves"], st
            g = S.state()
            xs = np.concatenate([f.x for f in sw.fils])
            assert np.abs(g["x"].reshape(-1, 3) - xs).max() < 1e-7, nt
            b = S.body_state()
            assert np.abs(b["x"] - sw.body.x).max() < 1e-8, (b["x"], sw.body.x)
            assert np.abs(b["q"] - sw.body.q).max() < 1e-8
            assert np.abs(b["vel"] - sw.vel).max() < 1e-6 * np.abs(sw.vel).max()
    assert np.linalg.norm(sw.body.x - np.array([0.0, 0.0, 40.0])) > 1e-6


def test_sphere_correction_matches_oracle():
    """The sphere correction of the approximate Jacobian (swimmer.cpp:800-877) on the
    GPU and in the oracle: identical Broyden iterations and states."""
    P, sw, wall = _swimmer()
    sw.sphere_R = 6.0
    vel = _velocity(wall)
    with _gpu(sw, P, wall) as S:
        S.set_body_sphere(6.0)
        for nt in range(3):
            st = S.step(nt)
            it = so.implicit_step(sw, nt, vel, wall=wall)
            assert st["converged"] and st["broyden_iterations"] == it, (nt, st, it)
            xs = np.concatenate([f.x for f in sw.fils])
            assert np.abs(S.state()["x"].reshape(-1, 3) - xs).max() < 1e-7, nt
            assert np.abs(S.body_state()["x"] - sw.body.x).max() < 1e-8


def test_external_load_moves_the_swimmer_as_the_oracle():
    f_ext = np.array([0.0, 0.0, -50.0, 0.0, 5.0, 0.0])
    P, sw, wall = _swimmer(f_ext=f_ext, end_force=0.0)
    vel = _velocity(wall)
    with _gpu(sw, P, wall, f_ext=f_ext) as S:
        for nt in range(2):
            st = S.step(nt)
            it = so.implicit_step(sw, nt, vel, wall=wall)
            assert st["converged"] and st["broyden_iterations"] == it, (nt, st, it)
        b = S.body_state()
    assert np.abs(b["x"] - sw.body.x).max() < 1e-8
    assert b["vel"][2] < 0.0 and b["x"][2] < 40.0


def test_swimmer_writes_reference_state_files(tmp_path):
    from cufcm_fil_amd import simio
    P, sw, wall = _swimmer()
    w = simio.OutputWriter(str(tmp_path) + "/", "swim")
    with _gpu(sw, P, wall) as S:
        for nt in range(2):
            S.step(nt)
            S.write_outputs(w, nt + 1)
        b = S.body_state()
        q = S.state()["q"]
    steps, rows = simio.read_lines(w.path("body_states"))
    assert list(steps) == [1, 2] and rows[-1].size == 7
    assert np.allclose(rows[-1], np.concatenate([b["x"], b["q"]]), rtol=1e-14, atol=1e-14)
    steps, rows = simio.read_lines(w.path("seg_states"))
    assert rows[-1].size == 4 * P.nseg * len(sw.fils) and np.allclose(rows[-1], q, atol=1e-14)
