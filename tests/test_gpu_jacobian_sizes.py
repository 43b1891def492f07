"""Approximate-Jacobian inverses of filaments of different lengths in one process
(through the C ABI): a context of 8-segment filaments first, then one of 20-segment
filaments, whose 115 KB LDS block needs a larger dynamic-LDS limit than the first.
Run in a fresh interpreter so that no earlier test has raised the limit already; both
inverses come back finite and of their block size (their values are checked against
the oracle's Jacobian in test_gpu_filament.py)."""
import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = textwrap.dedent("""
    import numpy as np
    from cufcm_fil_amd import geometry as geo
    from cufcm_fil_amd.filaments import FilamentSystem
    for nseg in (8, 20):
        x, t = geo.tilted_lattice(nx=2, ny=2, nseg=nseg, return_tangents=True, seed=1)
        with FilamentSystem(x, t, nseg=nseg) as S:
            J = S.jacobian_inverse(0)
            assert J.shape[-1] == 6 * nseg and np.all(np.isfinite(J)), nseg
            print("nseg", nseg, "ok", flush=True)
""")


def test_jacobians_of_growing_size_in_one_process():
    r = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True,
                       timeout=240, cwd=ROOT, env=dict(os.environ, PYTHONPATH=ROOT))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "nseg 8 ok" in r.stdout and "nseg 20 ok" in r.stdout
