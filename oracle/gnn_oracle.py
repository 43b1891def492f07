"""numpy restatement of the learned near-field correction -- TEST INFRASTRUCTURE ONLY.

Mirrors the network of the library's learned correction (16-nearest-neighbour blob
graph, 3 message-passing layers of width 64, 3x3 output block per edge) and its
scale-free training loss L = mean_t ||b_t - alpha_t u_t||^2 / ||b_t||^2 with
u_t = M (W + C) b_t and alpha_t = (b_t . u_t)/(u_t . u_t), with a hand-written
backward pass.  With ``bf16=True`` every GEMM operand is rounded to bfloat16
(round-to-nearest-even), as the matrix-core kernels do, so GPU and oracle agree to
FP32-accumulation rounding.

It also mirrors the network's coarse head (per-blob outputs = KC learned coarse
vectors, from the last hidden state and positional monomials) and the minimal-residual
objective of the two-level preconditioner those vectors extend.

Parity status: both backward passes are pinned by central finite differences of the
FP64 forward (tests/test_gnn_oracle.py).  This component has no counterpart in the
reference; its only oracles are this gradient check and the plain solver's residual.
"""
from __future__ import annotations

import numpy as np

K, S, H, F = 16, 17, 64, 8
XE, XN, O, NL = 2 * H + F, 2 * H, 16, 3
# coarse head: [h_i (64), positional monomials (16)] -> KC vectors x 3 components
QF, KC = 16, 16
XC, OC = H + QF, 3 * KC
# the analytic coarse fields: 3 directions x {1, y_x, y_y, y_z} (y centroid-relative)
KA = 12


def layout():
    off = {}
    o = 0

    def take(name, n):
        nonlocal o
        off[name] = (o, n)
        o += n

    take("Wn", F * H)
    take("bn", H)
    for l in range(NL):
        take(f"W1_{l}", XE * H)
        take(f"b1_{l}", H)
        take(f"W2_{l}", XN * H)
        take(f"b2_{l}", H)
    take("Wo", XE * O)
    take("bo", O)
    take("Wc", XC * OC)
    take("bc", OC)
    return off, o


OFF, NPARAM = layout()


def unpack(theta):
    P = {}
    for k, (o, n) in OFF.items():
        P[k] = theta[o:o + n]
    P["Wn"] = P["Wn"].reshape(F, H)
    for l in range(NL):
        P[f"W1_{l}"] = P[f"W1_{l}"].reshape(XE, H)
        P[f"W2_{l}"] = P[f"W2_{l}"].reshape(XN, H)
    P["Wo"] = P["Wo"].reshape(XE, O)
    P["Wc"] = P["Wc"].reshape(XC, OC)
    return P


def pack(grads):
    g = np.zeros(NPARAM)
    for k, (o, n) in OFF.items():
        if k in grads:
            g[o:o + n] = np.asarray(grads[k]).reshape(-1)
    return g


def bf16(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def knn(x_blob):
    p = np.asarray(x_blob, float).reshape(-1, 3)
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    return np.argsort(d2, axis=1, kind="stable")[:, :K]


def features(x_blob, nbr):
    p = np.asarray(x_blob, float).reshape(-1, 3)
    m = p.shape[0]
    nf = np.zeros((m, F))
    ef = np.zeros((m, S, F))
    for i in range(m):
        ef[i, 0, 5] = 1.0 / p[i, 2]
        ef[i, 0, 7] = 1.0
        d = p[i] - p[nbr[i]]
        r = np.sqrt((d * d).sum(1))
        ef[i, 1:, 0:3] = d / r[:, None]
        ef[i, 1:, 3] = 1.0 / r
        ef[i, 1:, 4] = 1.0 / r ** 2
        ef[i, 1:, 5] = 2.0 / (p[i, 2] + p[nbr[i], 2])
        ef[i, 1:, 6] = 0.25 * r
        ef[i, 1:, 7] = 1.0
        nf[i, 0] = 1.0 / p[i, 2]
        nf[i, 1] = r.mean() / 2.0
        nf[i, 2:5] = d.mean(0) / 2.0
        nf[i, 5] = 1.0
    # the library computes the features in FP64 and stores FP32
    return nf.astype(np.float32).astype(np.float64), ef.astype(np.float32).astype(np.float64)


def positional(x_blob):
    """Positional monomials of the coarse head (degree <= 2 in y = (x - c)/R, c the blob
    centroid, R the largest distance from it), FP32 as the library stores them."""
    p = np.asarray(x_blob, float).reshape(-1, 3)
    c = p.mean(0)
    R = np.sqrt(((p - c) ** 2).sum(1)).max()
    y = (p - c) / R
    q = np.zeros((p.shape[0], QF))
    q[:, 0] = 1.0
    q[:, 1:4] = y
    q[:, 4:7] = y * y
    q[:, 7] = y[:, 0] * y[:, 1]
    q[:, 8] = y[:, 0] * y[:, 2]
    q[:, 9] = y[:, 1] * y[:, 2]
    return q.astype(np.float32).astype(np.float64), y


def linear_fields(x_blob):
    """The KA analytic coarse vectors (system layout [3][m] flattened), column 3 p + d =
    direction d times the monomial p of {1, y_x, y_y, y_z}."""
    _, y = positional(x_blob)
    m = y.shape[0]
    Z = np.zeros((KA, 3, m))
    mon = [np.ones(m), y[:, 0], y[:, 1], y[:, 2]]
    for pidx in range(4):
        for d in range(3):
            Z[3 * pidx + d, d] = mon[pidx]
    return Z.reshape(KA, 3 * m)


def coarse_loss_and_grad(ZL, Zp, M, S_samples, delta):
    """Minimal-residual coarse objective
        L = mean_t min_c (||s_t - M Z c||^2 + delta ||c||^2) / ||s_t||^2
    over Z = [Zp | ZL] (rows are vectors in the system layout): (G + delta I) c = (M Z)^T s,
    G = (M Z)^T (M Z); by the envelope theorem dL/dZL = -2/T M sum_t (s_t - M Z c_t)
    c_t,L^T / ||s_t||^2."""
    Z = np.concatenate([Zp, ZL], axis=0)
    MZ = Z @ M            # rows: (M z)^T (M symmetric)
    G = MZ @ MZ.T
    kk = Z.shape[0]
    tot = 0.0
    D = np.zeros_like(ZL)
    for s in S_samples:
        rhs = MZ @ s
        c = np.linalg.solve(G + delta * np.eye(kk), rhs)
        r = s - MZ.T @ c
        n2 = float(s @ s)
        tot += (float(r @ r) + delta * float(c @ c)) / n2
        D += np.outer(c[Zp.shape[0]:], r) / n2
    T = len(S_samples)
    return tot / T, -2.0 / T * (D @ M)


class Net:
    def __init__(self, x_blob, nbr=None, use_bf16=True):
        self.nbr = knn(x_blob) if nbr is None else np.asarray(nbr)
        self.m = self.nbr.shape[0]
        self.nf, self.ef = features(x_blob, self.nbr)
        self.q = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
        self.jidx = np.concatenate([np.arange(self.m)[:, None], self.nbr], axis=1)  # [m][S]
        self.qf, _ = positional(x_blob)

    def mm(self, A, B):
        return self.q(A) @ self.q(B)

    def edge_input(self, h):
        m = self.m
        X = np.zeros((m, S, XE))
        X[:, :, :H] = h[:, None, :]
        X[:, :, H:2 * H] = h[self.jidx]
        X[:, :, 2 * H:] = self.ef
        return X.reshape(m * S, XE)

    def forward(self, theta):
        P = unpack(np.asarray(theta, dtype=np.float64))
        c = {}
        c["Z0"] = self.mm(self.nf, P["Wn"]) + P["bn"]
        h = np.maximum(c["Z0"], 0.0)
        c["H"] = [h]
        for l in range(NL):
            Xe = self.edge_input(h)
            Ze = self.mm(Xe, P[f"W1_{l}"]) + P[f"b1_{l}"]
            Me = np.maximum(Ze, 0.0).reshape(self.m, S, H)
            Xn = np.concatenate([h, Me[:, 1:, :].mean(1)], axis=1)
            Zn = self.mm(Xn, P[f"W2_{l}"]) + P[f"b2_{l}"]
            h = h + np.maximum(Zn, 0.0)
            c.setdefault("Xe", []).append(Xe)
            c.setdefault("Ze", []).append(Ze)
            c.setdefault("Xn", []).append(Xn)
            c.setdefault("Zn", []).append(Zn)
            c["H"].append(h)
        Xo = self.edge_input(h)
        Co = self.mm(Xo, P["Wo"]) + P["bo"]
        c["Xo"] = Xo
        c["P"] = P
        self.cache = c
        return Co[:, :9].reshape(self.m, S, 3, 3)

    def coarse(self, theta):
        """Learned coarse vectors ZL ([KC][3][m] flattened to KC x 3m) of the last
        trunk (call forward first)."""
        P = unpack(np.asarray(theta, dtype=np.float64))
        Xc = np.concatenate([self.cache["H"][-1], self.qf], axis=1)
        Oc = self.mm(Xc, P["Wc"]) + P["bc"]
        self.cache["Xc"] = Xc
        return Oc.reshape(self.m, KC, 3).transpose(1, 2, 0).reshape(KC, 3 * self.m)

    def coarse_backward(self, dZL):
        """Gradient of the parameters from dL/dZL (same layout as coarse())."""
        c, P = self.cache, self.cache["P"]
        m = self.m
        dOc = np.asarray(dZL).reshape(KC, 3, m).transpose(2, 0, 1).reshape(m, OC)
        g = {"Wc": self.mm(c["Xc"].T, dOc)}
        g["bc"] = dOc.sum(0)
        dH = self.mm(dOc, P["Wc"].T)[:, :H]
        return pack(g) + self._trunk_backward(dH)

    def sparse_apply(self, blocks, v):
        """y_i = sum_s blocks[i, s] v_j(s); v, y AoS (m, 3)."""
        v = np.asarray(v).reshape(-1, 3)
        return np.einsum("ispq,isq->ip", blocks, v[self.jidx])

    def loss_and_dC(self, C, W, M, samples):
        """Loss and dL/dC ([m][S][3][3]); W: near-field blocks, M dense (3m x 3m)."""
        dC = np.zeros_like(C)
        tot = 0.0
        for b in samples:
            b = np.asarray(b).reshape(-1, 3)
            u = M @ self.sparse_apply(W + C, b).reshape(-1)
            alpha = float(b.reshape(-1) @ u) / float(u @ u)
            r = b.reshape(-1) - alpha * u
            nb2 = float((b * b).sum())
            tot += float(r @ r) / nb2
            g = (M.T @ r).reshape(-1, 3)
            dC += -2.0 * alpha / nb2 * np.einsum("ip,isq->ispq", g, b[self.jidx])
        n = len(samples)
        return tot / n, dC / n

    def backward(self, dC):
        c, P = self.cache, self.cache["P"]
        m = self.m
        g = {}
        dCo = np.zeros((m * S, O))
        dCo[:, :9] = dC.reshape(m * S, 9)
        g["Wo"] = self.mm(c["Xo"].T, dCo)
        g["bo"] = dCo.sum(0)
        dXo = self.mm(dCo, P["Wo"].T).reshape(m, S, XE)
        dH = self._scatter(dXo)
        return pack(g) + self._trunk_backward(dH)

    def _trunk_backward(self, dH):
        """Parameter gradient of the message-passing trunk from dL/dh (last layer)."""
        c, P = self.cache, self.cache["P"]
        m = self.m
        g = {}
        for l in reversed(range(NL)):
            dZn = dH * (c["Zn"][l] > 0)
            g[f"W2_{l}"] = self.mm(c["Xn"][l].T, dZn)
            g[f"b2_{l}"] = dZn.sum(0)
            dXn = self.mm(dZn, P[f"W2_{l}"].T)
            dH = dH + dXn[:, :H]
            dMe = np.zeros((m, S, H))
            dMe[:, 1:, :] = dXn[:, None, H:] / K
            dZe = dMe.reshape(m * S, H) * (c["Ze"][l] > 0)
            g[f"W1_{l}"] = self.mm(c["Xe"][l].T, dZe)
            g[f"b1_{l}"] = dZe.sum(0)
            dXe = self.mm(dZe, P[f"W1_{l}"].T).reshape(m, S, XE)
            dH = dH + self._scatter(dXe)
        dZ0 = dH * (c["Z0"] > 0)
        g["Wn"] = self.mm(self.nf.T, dZ0)
        g["bn"] = dZ0.sum(0)
        return pack(g)

    def _scatter(self, dX):
        m = self.m
        dH = dX[:, :, :H].sum(1)
        np.add.at(dH, self.jidx.reshape(-1), dX[:, :, H:2 * H].reshape(m * S, H))
        return dH


def init_params(seed=0):
    """A random parameter vector with non-zero output layers (for checks)."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAM)
    for k, (o, n) in OFF.items():
        if k.startswith("W"):
            fan = {"Wn": F, "Wo": XE, "Wc": XC}.get(k, XE if k.startswith("W1") else XN)
            th[o:o + n] = rng.standard_normal(n) * np.sqrt(2.0 / fan)
        else:
            th[o:o + n] = 0.1 * rng.standard_normal(n)
    return th.astype(np.float32).astype(np.float64)
