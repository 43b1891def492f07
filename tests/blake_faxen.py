"""Independent construction of the wall-bounded RPY grand mobility (test helper).

Starts from Blake's (1971) image solution for a Stokeslet above a no-slip plane,

    G_ij(x, y) = (d_ij/r + r_i r_j/r^3) - (d_ij/R + R_i R_j/R^3)
                 + 2 h P_jk d/dR_k [ h R_i/R^3 - (d_i3/R + R_i R_3/R^3) ],

r = x - y, R = x - y*, y* = (y1, y2, -y3), h = y3, P = diag(1, 1, -1),
u = G F / (8 pi mu), and builds the sphere-sphere blocks with Faxen operators
(Swan & Brady 2007, a = 1 by default):

    tt = (1 + a^2/6 Lap_x)(1 + a^2/6 Lap_y) G / (8 pi)
    rt = 1/2 curl_x  (1 + a^2/6 Lap_y) G / (8 pi)
    tr = torque as an antisymmetric force dipole at y, (1 + a^2/6 Lap_x)
    rr = 1/2 curl_x of the rotlet field

All derivatives are 4th-order central finite differences, so nothing here
repeats the closed-form coefficients of the code under test.  The same
machinery applied to the free-space Oseen tensor reproduces the textbook
Rotne-Prager blocks, and Blake's tensor itself is checked for no slip on the
wall and zero divergence.
"""
from __future__ import annotations

import numpy as np

EPS3 = np.zeros((3, 3, 3))
EPS3[0, 1, 2] = EPS3[1, 2, 0] = EPS3[2, 0, 1] = 1.0
EPS3[0, 2, 1] = EPS3[2, 1, 0] = EPS3[1, 0, 2] = -1.0


def oseen(x, y):
    r = np.asarray(x, float) - np.asarray(y, float)
    s = np.linalg.norm(r)
    return np.eye(3) / s + np.outer(r, r) / s**3


def blake_image(x, y):
    """Image part of Blake's tensor (everything but the Oseen term; smooth at x = y)."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    h = y[2]
    R = x - np.array([y[0], y[1], -y[2]])
    S = np.linalg.norm(R)
    G = -np.eye(3) / S - np.outer(R, R) / S**3
    # D[i, k] = d/dR_k [ h R_i/S^3 - (d_i3/S + R_i R_3/S^3) ]
    D = np.zeros((3, 3))
    for i in range(3):
        for k in range(3):
            dik = 1.0 if i == k else 0.0
            d3k = 1.0 if k == 2 else 0.0
            d_i3 = 1.0 if i == 2 else 0.0
            term_a = h * (dik / S**3 - 3.0 * R[i] * R[k] / S**5)
            term_b = -d_i3 * R[k] / S**3
            term_c = (dik * R[2] + R[i] * d3k) / S**3 - 3.0 * R[i] * R[2] * R[k] / S**5
            D[i, k] = term_a - (term_b + term_c)
    P = np.diag([1.0, 1.0, -1.0])
    return G + 2.0 * h * (D @ P.T)  # G_ij += 2h sum_k D_ik P_jk


def blake(x, y):
    return oseen(x, y) + blake_image(x, y)


_W1 = np.array([1.0, -8.0, 0.0, 8.0, -1.0]) / 12.0           # d/dx, offsets -2..2
_W2 = np.array([-1.0, 16.0, -30.0, 16.0, -1.0]) / 12.0       # d2/dx2
_OFF = np.array([-2, -1, 0, 1, 2])


def _d1(f, p, k, d):
    e = np.zeros(3)
    e[k] = d
    return sum(w * f(p + o * e) for w, o in zip(_W1, _OFF) if w != 0.0) / d


def _lap(f, p, d):
    acc = 0.0
    for k in range(3):
        e = np.zeros(3)
        e[k] = d
        acc = acc + sum(w * f(p + o * e) for w, o in zip(_W2, _OFF)) / (d * d)
    return acc


def blocks(G, x, y, a=1.0, d=2e-2):
    """Return the 6x6 block [[tt, tr], [rt, rr]] for target x, source y."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    c = 1.0 / (8.0 * np.pi)
    a2 = a * a / 6.0

    def Gy(xx):  # (1 + a^2/6 Lap_y) G(xx, y)
        return G(xx, y) + a2 * _lap(lambda yy: G(xx, yy), y, d)

    def Gx(yy):  # (1 + a^2/6 Lap_x) G(x, yy)
        return G(x, yy) + a2 * _lap(lambda xx: G(xx, yy), x, d)

    tt = c * (Gy(x) + a2 * _lap(Gy, x, d))
    # rt[i, j] = 1/2 eps_ikl d/dx_k Gy[l, j]
    dGy = np.stack([_d1(Gy, x, k, d) for k in range(3)])        # [k, l, j]
    rt = 0.5 * c * np.einsum("ikl,klj->ij", EPS3, dGy)
    # tr[i, l] = 1/2 eps_kjl d/dy_k Gx[i, j]
    dGx = np.stack([_d1(Gx, y, k, d) for k in range(3)])        # [k, i, j]
    tr = 0.5 * c * np.einsum("kjl,kij->il", EPS3, dGx)

    # rotlet field u_n(xx) = 1/2 eps_kjl d/dy_k G_nj(xx, y) (per unit torque l)
    def rotlet(xx):
        dG = np.stack([_d1(lambda yy: G(xx, yy), y, k, d) for k in range(3)])  # [k, n, j]
        return 0.5 * np.einsum("kjl,knj->nl", EPS3, dG)

    dR = np.stack([_d1(rotlet, x, m, d) for m in range(3)])     # [m, n, l]
    rr = 0.5 * c * np.einsum("imn,mnl->il", EPS3, dR)
    M = np.zeros((6, 6))
    M[:3, :3] = tt
    M[:3, 3:] = tr
    M[3:, :3] = rt
    M[3:, 3:] = rr
    return M


def divergence_x(G, x, y, d=1e-3):
    """div_x of each column of G(x, y) (should vanish for a Stokes flow)."""
    x = np.asarray(x, float)
    return sum(_d1(lambda xx: G(xx, y)[k, :], x, k, d) for k in range(3))

