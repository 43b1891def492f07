"""GPU checks of the learned near-field correction: the BF16 matrix-core GEMM, the
network forward and gradient against the numpy oracle, and what training buys."""
import numpy as np
import pytest

import gnn_oracle as go
import oracle
from cufcm_fil_amd import RpyMobilitySolver
from cufcm_fil_amd import geometry as geo
from cufcm_fil_amd.filaments import FilamentSystem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K,tA,tB", [(64, 64, 16, False, False), (100, 37, 53, False, False),
                                         (136, 64, 700, True, False), (513, 136, 16, False, True),
                                         (17, 9, 136, True, True)])
def test_gemm_bf16_matches_numpy(M, N, K, tA, tB):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if tA else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tB else (K, N)).astype(np.float32)
    xs = geo.tilted_lattice(nx=1, ny=1)
    with RpyMobilitySolver(20, 0) as m:
        m.set_positions(xs)
        C = m.gemm_bf16(A, B, tA, tB)
    Aq, Bq = go.bf16(A), go.bf16(B)
    ref = (Aq.T if tA else Aq) @ (Bq.T if tB else Bq)
    assert np.abs(C - ref).max() <= 1e-5 * np.abs(ref).max() + 1e-6


def _blob_system(nblob=150, radius=6.0):
    xs, xb, c, d = geo.sphere_over_wall(n_blob=nblob, n_fil=4, radius=radius, gap=50.0)
    return xs, xb


def _blocks_from_dense(W, nbr):
    m = nbr.shape[0]
    jidx = np.concatenate([np.arange(m)[:, None], nbr], axis=1)
    out = np.zeros((m, 17, 3, 3))
    for i in range(m):
        for s, j in enumerate(jidx[i]):
            out[i, s] = W[3 * i:3 * i + 3, 3 * j:3 * j + 3]
    return out


def test_forward_and_gradient_match_oracle():
    xs, xb = _blob_system()
    ns, nb = xs.size // 3, xb.size // 3
    rng = np.random.default_rng(0)
    samples = [rng.standard_normal(3 * nb) for _ in range(3)]
    theta = go.init_params(5).astype(np.float32)
    with RpyMobilitySolver(ns, nb) as m:
        m.set_positions(xs, xb)
        m.gnn_init(seed=1, capacity=8)
        m.gnn_add_samples(np.concatenate(samples))
        m.gnn_set_params(theta)
        C, nbr = m.gnn_correction()
        loss, grad = m.gnn_gradient(0, 3)
        # near-field blocks W from the library's own preconditioner (unit vectors)
        Wd = np.zeros((3 * nb, 3 * nb))
        for col in range(3 * nb):
            e = np.zeros(3 * nb)
            e[col] = 1.0
            Wd[:, col] = m.apply_preconditioner("blob", e, "near")
    net = go.Net(xb, nbr=nbr, use_bf16=True)
    Cr = net.forward(theta.astype(np.float64))
    assert np.abs(C - Cr).max() <= 2e-5 * np.abs(Cr).max()
    Mbb = oracle.mobility_matrix(xb, 3)
    Lr, dC = net.loss_and_dC(Cr, _blocks_from_dense(Wd, nbr), Mbb, samples)
    gr = net.backward(dC)
    assert np.isclose(loss, Lr, rtol=1e-5)
    assert np.linalg.norm(grad - gr) <= 1e-2 * np.linalg.norm(gr)


def test_training_lowers_loss_and_gate_never_costs_iterations():
    """What training buys on held-out right-hand sides: the scale-free one-step
    residual falls; the correction is kept only if it lowers the GMRES iterations of
    validation solves, so with the gate the near-field solves never take more steps
    than without the network, and every solve still meets 1e-10."""
    xs, xb, c, d, t = geo.sphere_over_wall(n_blob=1500, n_fil=40, radius=20.0, gap=50.0,
                                            return_tangents=True)
    with FilamentSystem(xs, t, x_blob=xb, blob_precond="diag", guess="previous") as S:
        S.mob.gnn_init(seed=0, capacity=512)
        S.mob.gnn_record(True)
        for nt in range(3):
            S.step(nt)
        S.mob.gnn_record(False)
        assert S.mob.gnn_samples() >= 30     # one (rhs, solution) pair per blob solve
        rng = np.random.default_rng(3)
        rhs = [-S.mob.evaluate_blob_segment_mobility(rng.standard_normal(6 * S.n_seg))
               for _ in range(4)] + [rng.standard_normal(3 * S.n_blob)]
        val, held = rhs[:2], rhs[2:]
        before = [S.mob.solve("blob", b, rtol=1e-10, precond="near")[1]["iterations"] for b in held]
        gate = S.mob.gnn_train_gated(val, epochs=8, batch=16, lr=1e-3)
        assert gate["loss_after"] < gate["loss_before"]
        assert gate["kept"] == (sum(gate["iterations_with"]) < sum(gate["iterations_without"]))
        after = [S.mob.solve("blob", b, rtol=1e-10, precond="near")[1] for b in held]
        for st in after:
            assert st["converged"] and st["rel_residual"] <= 1e-10
        if not gate["kept"]:
            assert [st["iterations"] for st in after] == before
