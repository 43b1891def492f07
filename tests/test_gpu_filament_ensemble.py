"""Ensembles of equal filament systems stepped together (cfm_fil_set_members) on the GPU:
every member's Broyden iteration counts and states equal, bit for bit, those of the same
system stepped alone in its own context -- in FP64 and FP32C pair arithmetic, with
members that converge after different numbers of iterations -- and the refusals."""
import numpy as np
import pytest

from cufcm_fil_amd import geometry as geo
from cufcm_fil_amd import _lib as L
from cufcm_fil_amd.filaments import DEFAULTS, FilamentSystem, end_force

pytestmark = pytest.mark.gpu

NSEG = 20


def _members(E, nx=4, ny=4):
    """E different 4 x 4 lattices of tilted filaments (320 segments each) and tip
    perturbations of different strength, so the members need different iterations."""
    xs, ts, ps = [], [], []
    f_end = end_force(DEFAULTS["dimensionless_force"], DEFAULTS["kb"], DEFAULTS["dl"], NSEG)
    for e in range(E):
        x, t = geo.tilted_lattice(nx=nx, ny=ny, spacing=9.0, nseg=NSEG, seed=30 + e,
                                  return_tangents=True, min_sep=2.5)
        rng = np.random.default_rng(100 + e)
        pert = (0.02 + 0.05 * e) * f_end * rng.standard_normal(3 * nx * ny)
        xs.append(x), ts.append(np.asarray(t).reshape(-1, 3)), ps.append(pert)
    return xs, ts, ps


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_members_equal_lone_runs_bit_for_bit(precision):
    E, steps = 3, 4
    xs, ts, ps = _members(E)
    lone = []
    for e in range(E):
        with FilamentSystem(xs[e], ts[e], pert=ps[e], precision=precision) as S:
            its = [S.step(nt)["broyden_iterations"] for nt in range(steps)]
            lone.append((its, S.state()))
    with FilamentSystem(np.concatenate(xs), np.concatenate(ts), pert=np.concatenate(ps),
                        precision=precision) as S:
        S.set_members(E)
        per = []
        for nt in range(steps):
            st = S.step(nt)
            ms = S.member_stats()
            assert st["converged"] and all(m["converged"] for m in ms)
            assert st["broyden_iterations"] == max(m["broyden_iterations"] for m in ms)
            per.append([m["broyden_iterations"] for m in ms])
        g = S.state()
    n = 320
    for e in range(E):
        its, ref = lone[e]
        assert [p[e] for p in per] == its, (e, per, its)
        for k, dof in (("x", 3), ("q", 4), ("u", 3), ("lam", 3)):
            mine = g[k].reshape(-1, dof)[e * n:(e + 1) * n]
            assert np.array_equal(mine, ref[k].reshape(-1, dof)), (precision, e, k)
        assert np.array_equal(g["clamp"].reshape(-1, 3)[16 * e:16 * (e + 1)], ref["clamp"].reshape(-1, 3))
    # the members really differ in how long they iterate
    assert len({tuple(p) for p in zip(*per)}) > 1, per


def test_members_are_refused_where_they_do_not_apply():
    xs, ts, ps = _members(2)
    x, t = np.concatenate(xs), np.concatenate(ts)
    with FilamentSystem(x, t) as S:
        rc = L.lib().cfm_fil_set_members(S.h, 3)       # 32 filaments: 3 does not divide
        assert rc == L.CFM_EINVAL and b"divide" in L.lib().cfm_last_error(S.h)
        rc = L.lib().cfm_fil_member_stats(S.h, (L.cfm_step_stats * 2)())
        assert rc == L.CFM_ESTATE
    xb = geo.fibonacci_sphere(200, 8.0).reshape(-1) + np.tile([15.0, 15.0, 40.0], 200)
    with FilamentSystem(x, t, x_blob=xb) as S:
        rc = L.lib().cfm_fil_set_members(S.h, 2)
        assert rc == L.CFM_EINVAL and b"wall-bounded" in L.lib().cfm_last_error(S.h)
