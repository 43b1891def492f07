"""CPU pin of oracle/gnn_oracle.py: the hand-written backward pass of the learned
near-field correction against central finite differences of its FP64 forward."""
import numpy as np

import gnn_oracle as go
from cufcm_fil_amd import geometry as geo


def _setup(m=40, seed=0):
    xb = geo.fibonacci_sphere(m, 5.0) + np.array([0.0, 0.0, 20.0])
    net = go.Net(xb.reshape(-1), use_bf16=False)
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((3 * m, 3 * m))
    M = A @ A.T / (3 * m) + np.eye(3 * m)       # symmetric positive definite stand-in
    W = 0.1 * rng.standard_normal((m, go.S, 3, 3))
    samples = [rng.standard_normal(3 * m) for _ in range(3)]
    return net, M, W, samples


def _loss(net, theta, W, M, samples):
    C = net.forward(theta)
    return net.loss_and_dC(C, W, M, samples)


def test_backward_matches_finite_differences():
    net, M, W, samples = _setup()
    theta = go.init_params(3)
    L0, dC = _loss(net, theta, W, M, samples)
    grad = net.backward(dC)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(go.NPARAM, 30, replace=False),
                          [go.OFF["Wo"][0] + 5, go.OFF["bn"][0] + 2, go.OFF["W1_0"][0] + 7]])
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (_loss(net, tp, W, M, samples)[0] - _loss(net, tm, W, M, samples)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grad[k])


def test_zero_output_layer_gives_zero_correction_and_loss_of_W():
    net, M, W, samples = _setup(seed=2)
    theta = go.init_params(4)
    o, n = go.OFF["Wo"]
    theta[o:o + n] = 0.0
    o, n = go.OFF["bo"]
    theta[o:o + n] = 0.0
    C = net.forward(theta)
    assert not C.any()
    L, _ = net.loss_and_dC(C, W, M, samples)
    ref = []
    for b in samples:
        u = M @ net.sparse_apply(W, b).reshape(-1)
        cos2 = (b @ u) ** 2 / ((b @ b) * (u @ u))
        ref.append(1.0 - cos2)                 # the loss is sin^2 of the angle (b, u)
    assert np.isclose(L, np.mean(ref), rtol=1e-10)


def _coarse_setup(m=40, seed=0):
    net, M, W, samples = _setup(m, seed)
    xb = geo.fibonacci_sphere(m, 5.0) + np.array([0.0, 0.0, 20.0])
    Zp = go.linear_fields(xb.reshape(-1))
    rng = np.random.default_rng(seed + 7)
    S = [rng.standard_normal(3 * m) for _ in range(4)]
    return net, M, Zp, S


def _coarse_loss(net, theta, M, Zp, S, delta=1e-6):
    net.forward(theta)
    ZL = net.coarse(theta)
    return go.coarse_loss_and_grad(ZL, Zp, M, S, delta)


def test_coarse_objective_gradient_matches_finite_differences():
    """dL/dZL of the minimal-residual objective (envelope theorem) and its
    back-propagation through the coarse head and the trunk, against central
    differences of the FP64 forward."""
    net, M, Zp, S = _coarse_setup()
    theta = go.init_params(5)
    L0, dZ = _coarse_loss(net, theta, M, Zp, S)
    grad = net.coarse_backward(dZ)
    assert 0.0 < L0 < 1.0
    # first on ZL itself
    ZL = net.coarse(theta)
    rng = np.random.default_rng(3)
    for _ in range(8):
        k, e = rng.integers(go.KC), rng.integers(ZL.shape[1])
        h = 1e-6
        Zp_, Zm_ = ZL.copy(), ZL.copy()
        Zp_[k, e] += h
        Zm_[k, e] -= h
        fd = (go.coarse_loss_and_grad(Zp_, Zp, M, S, 1e-6)[0] -
              go.coarse_loss_and_grad(Zm_, Zp, M, S, 1e-6)[0]) / (2 * h)
        assert abs(fd - dZ[k, e]) <= 1e-5 * max(1e-3, abs(fd)), (k, e, fd, dZ[k, e])
    idx = np.concatenate([rng.choice(go.NPARAM, 25, replace=False),
                          [go.OFF["Wc"][0] + 11, go.OFF["bc"][0] + 4, go.OFF["W2_2"][0] + 9]])
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (_coarse_loss(net, tp, M, Zp, S)[0] - _coarse_loss(net, tm, M, Zp, S)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1e-3, abs(fd)), (k, fd, grad[k])


def test_coarse_objective_closed_forms():
    """With no learned vectors able to help (ZL = 0 rows, delta -> 0) the loss is the
    fraction of ||s||^2 outside span(M Zp); with ZL spanning the samples' own
    directions (M ZL = s) it vanishes."""
    net, M, Zp, S = _coarse_setup(m=30, seed=4)
    MZp = Zp @ M
    Qp, _ = np.linalg.qr(MZp.T)
    ref = np.mean([1.0 - np.linalg.norm(Qp.T @ s) ** 2 / (s @ s) for s in S])
    L, _ = go.coarse_loss_and_grad(np.zeros((go.KC, Zp.shape[1])), Zp, M, S, 1e-12)
    assert np.isclose(L, ref, rtol=1e-6)
    ZL = np.zeros((go.KC, Zp.shape[1]))
    for t, s in enumerate(S):
        ZL[t] = np.linalg.solve(M, s)
    L, _ = go.coarse_loss_and_grad(ZL, Zp, M, S, 1e-12)
    assert L < 1e-12
