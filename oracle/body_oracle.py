"""numpy restatement of the free rigid body made of blobs -- TEST INFRASTRUCTURE ONLY
(imported by tests/ and never by the product).

Follows the reference's rigid-body blob system (USE_BROYDEN_FOR_EVERYTHING 0,
PRESCRIBED_BODY_VELOCITIES 0):
  * blob placement x_b + R(q_b) ref_i           rigid_body.cpp:497-515
  * the saddle-point product
        [ M_bb  -K ] [F]      M_bb F - K U
        [ -K^T   0 ] [U]  =   -K^T F
    with (K U)_i = V_b + Omega_b x r_i, r_i = R ref_i   mobility_solver.cpp:1822-1880
  * the block preconditioner with the scalar blob approximation 6 pi mu a, the
    reference (K^T K)^-1 and body mobility rotated by R      mobility_solver.cpp:1518-1590
  * make_body_reference_matrices: K^T K from the body-frame offsets, and the 6 x 6
    body mobility from six resistance solves with a unit force or torque on the
    body (column 5 - n from right-hand side entry -1 at row 6 n_body - 1 - n)
                                                      mobility_solver.cpp:3049-3146
  * prescribed rotation (BODY_VELOCITY_TYPE 2): the Omega columns and the torque
    rows dropped, [[M, -K_V], [-K_V^T, 0]]      mobility_solver.cpp:1947-1994, 2177-2181
Unknown ordering as the reference: 3 n_blob blob forces (AoS), then 6 per body
(V then Omega).

Parity status: pinned in tests/test_body_oracle.py by what the mathematics fixes --
K^T K against the Gram matrix of an explicitly assembled K, the preconditioner against
the exact inverse of the saddle matrix with M_bb replaced by its scalar
approximation, the body mobility against (K^T M_bb^-1 K)^-1 from a dense solve, its
symmetry, Stokes drag of a blob sphere (1/(6 pi mu R), 1/(8 pi mu R^3)) and the
Faxen/Brenner wall corrections of a sphere's translational mobility.  The
prescribed-rotation solve is pinned against the free system: its velocity under a unit
force equals the clamped mobility ((N^-1)_VV)^-1 of the body mobility N from six free
solves (the Schur complement N_VV - N_VO N_OO^-1 N_OV), and in free space a sphere's
translation is unchanged by holding its rotation; its preconditioner is the exact inverse
of the reduced saddle matrix with M replaced by 6 pi mu a.
"""
from __future__ import annotations

import math

import numpy as np

from filament_oracle import rcross, rot_mat


def place(ref, x, q):
    """Blob positions (n_body nb, 3) and rotated offsets of bodies at centres x
    (n_body, 3) with quaternions q (n_body, 4) (rigid_body.cpp:497-515)."""
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, 3)
    pos, off = [], []
    for xb, qb in zip(np.asarray(x).reshape(-1, 3), np.asarray(q).reshape(-1, 4)):
        r = ref @ rot_mat(qb).T
        off.append(r)
        pos.append(xb + r)
    return np.concatenate(pos), np.concatenate(off)


def k_matrix(offsets, n_body):
    """Dense K (3 n_blob x 6 n_body): column block b maps (V, Omega) of body b to the
    velocities V + Omega x r_i of its blobs."""
    r = np.asarray(offsets).reshape(-1, 3)
    n = r.shape[0]
    nb = n // n_body
    K = np.zeros((3 * n, 6 * n_body))
    for i in range(n):
        b = i // nb
        K[3 * i:3 * i + 3, 6 * b:6 * b + 3] = np.eye(3)
        # Omega x r = rcross(r) Omega   (rcross(v) w = w x v)
        K[3 * i:3 * i + 3, 6 * b + 3:6 * b + 6] = rcross(r[i])
    return K


def ktk_reference(ref):
    """K^T K of one body from its body-frame offsets (mobility_solver.cpp:3052-3070)."""
    ref = np.asarray(ref).reshape(-1, 3)
    A = np.zeros((6, 6))
    A[:3, :3] = ref.shape[0] * np.eye(3)
    for r in ref:
        C = rcross(r)
        A[:3, 3:] += C
        A[3:, :3] -= C
        A[3:, 3:] -= C @ C
    return A


def saddle_matrix(M, K):
    """[[M, -K], [-K^T, 0]] in the reference's unknown order."""
    n, k = K.shape
    A = np.zeros((n + k, n + k))
    A[:n, :n] = M
    A[:n, n:] = -K
    A[n:, :n] = -K.T
    return A


def rotate6(A, R):
    """Each 3 x 3 block of a 6 x 6 matrix conjugated by R (mobility_solver.cpp:1535-1545)."""
    B = np.array(A, dtype=np.float64)
    for i in (0, 3):
        for j in (0, 3):
            B[i:i + 3, j:j + 3] = R @ A[i:i + 3, j:j + 3] @ R.T
    return B


def precondition(v, offsets, q, ktk_inv_ref, mob_ref, mu=1.0, a=1.0):
    """The reference's free-body preconditioner (mobility_solver.cpp:1518-1590)."""
    q = np.asarray(q).reshape(-1, 4)
    nbody = q.shape[0]
    r = np.asarray(offsets).reshape(-1, 3)
    n = r.shape[0]
    nb = n // nbody
    minv = 6.0 * math.pi * mu * a
    out = np.zeros_like(v)
    for b in range(nbody):
        R = rot_mat(q[b])
        Kb = k_matrix(r[b * nb:(b + 1) * nb], 1)
        fb = v[3 * b * nb:3 * (b + 1) * nb]
        ub = v[3 * n + 6 * b:3 * n + 6 * b + 6]
        rhs = -Kb.T @ fb
        vel = rotate6(ktk_inv_ref, R) @ rhs - rotate6(mob_ref, R) @ ub
        out[3 * n + 6 * b:3 * n + 6 * b + 6] = vel
        out[3 * b * nb:3 * (b + 1) * nb] = minv * (fb + Kb @ vel)
    return out


def body_mobility(M, K):
    """(K^T M^-1 K)^-1: the grand body mobility of the bodies whose blob-blob mobility is
    M (what six saddle solves with unit body loads return, column by column)."""
    return np.linalg.inv(K.T @ np.linalg.solve(M, K))


def reference_mobility_by_solves(M, K):
    """make_body_reference_matrices' six solves for a single body: right-hand side zero
    except -1 at row (rows - 1 - n); column 5 - n of the result is the body velocity."""
    A = saddle_matrix(M, K)
    rows = A.shape[0]
    N = np.zeros((6, 6))
    for n in range(6):
        rhs = np.zeros(rows)
        rhs[rows - 1 - n] = -1.0
        z = np.linalg.solve(A, rhs)
        N[:, 5 - n] = z[-6:]
    return N


def faxen_brenner_ratios(R, h):
    """Translational mobility of a sphere of radius R whose centre is h above a no-slip
    wall, relative to 1/(6 pi mu R): parallel (Faxen) and normal (Brenner, leading
    terms of the method-of-reflections series)."""
    s = R / h
    par = 1.0 - 9.0 / 16.0 * s + 1.0 / 8.0 * s ** 3 - 45.0 / 256.0 * s ** 4 - 1.0 / 16.0 * s ** 5
    nor = 1.0 - 9.0 / 8.0 * s + 1.0 / 2.0 * s ** 3 - 57.0 / 100.0 * s ** 4 + 1.0 / 5.0 * s ** 5
    return par, nor


# ---------------------------------------------------------------- squirmer bodies
def legendre(n, x):
    """P_n(x) and P_n'(x) by the three-term recurrences (util.cu:87-100 restates P_n;
    the derivative here is exact where the reference differences P_n numerically)."""
    p0, p1 = 1.0, x
    d0, d1 = 0.0, 1.0
    if n == 0:
        return p0, d0
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        d0, d1 = d1, d0 + (2 * k - 1) * p0
    return p1, d1


def squirmer_slip(ref, A, B):
    """Body-frame slip velocity of each blob (mobility_solver.cpp:470-505 with
    util.cu:69-115): theta, phi the polar and azimuthal angles of the body-frame offset,
    u_r = sum_n A_n P_n(cos theta), u_theta = sum_n B_n V_n(cos theta) with
    V_n(x) = 2 sin(theta) P_n'(x) / (n (n + 1)), mapped to Cartesian components."""
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(ref)
    for i, r in enumerate(ref):
        phi = math.atan2(r[1], r[0])
        th = math.acos(r[2] / np.linalg.norm(r))
        x, s = math.cos(th), math.sin(th)
        ur = ut = 0.0
        for n in range(1, len(A) + 1):
            p, dp = legendre(n, x)
            ur += A[n - 1] * p
            ut += B[n - 1] * 2.0 * s * dp / (n * (n + 1))
        out[i] = [ur * s * math.cos(phi) + ut * x * math.cos(phi),
                  ur * s * math.sin(phi) + ut * x * math.sin(phi),
                  ur * x - ut * s]
    return out


def squirmer_velocity(M, K, slip_lab):
    """Force- and torque-free bodies whose blobs move with the rigid velocity plus the slip:
    M F - K U = u_s, -K^T F = 0 (the body system with the slip as blob right-hand side).
    Returns the body velocities (V, Omega) per body and the blob forces."""
    A = saddle_matrix(M, K)
    n = K.shape[0]
    rhs = np.zeros(A.shape[0])
    rhs[:n] = np.asarray(slip_lab).reshape(-1)
    z = np.linalg.solve(A, rhs)
    return z[n:].reshape(-1, 6), z[:n]


# ---------------------------------------------------------------- prescribed rotation
def translation_columns(n_body):
    """Indices of the V entries among the 6 n_body body unknowns."""
    return np.array([6 * b + k for b in range(n_body) for k in range(3)])


def prescribed_rotation_solve(M, K, rhs_f, rhs_v):
    """BODY_VELOCITY_TYPE 2: Omega columns dropped from the blob rows and the torque rows
    dropped (mobility_solver.cpp:1947-1994, 2177-2181), i.e. the reduced saddle system
    [[M, -K_V], [-K_V^T, 0]] [F; V] = [rhs_f; rhs_v].  Returns F and (V, Omega = 0) per
    body."""
    nbody = K.shape[1] // 6
    Kv = K[:, translation_columns(nbody)]
    A = saddle_matrix(M, Kv)
    n = K.shape[0]
    z = np.linalg.solve(A, np.concatenate([np.asarray(rhs_f, float), np.asarray(rhs_v, float)]))
    U = np.zeros((nbody, 6))
    U[:, :3] = z[n:].reshape(nbody, 3)
    return z[:n], U


def clamped_mobility(N):
    """Translational mobility of a body whose rotation is held: ((N^-1)_VV)^-1."""
    return np.linalg.inv(np.linalg.inv(N)[:3, :3])


def precondition_prescribed_rotation(v, offsets, q, mob_ref, mu=1.0, a=1.0):
    """The block preconditioner of the reduced system: V = -(1/nb) sum f - N'_VV u_V with
    the clamped mobility N' of the rotated reference mobility, Omega = u_Omega, F = 6 pi mu
    a (f + K_V V) (the reference's prescribed-rotation branches
    mobility_solver.cpp:1268-1291, 1506-1512 drop the Omega terms the same way)."""
    q = np.asarray(q).reshape(-1, 4)
    nbody = q.shape[0]
    r = np.asarray(offsets).reshape(-1, 3)
    n = r.shape[0]
    nb = n // nbody
    minv = 6.0 * math.pi * mu * a
    out = np.zeros_like(v)
    for b in range(nbody):
        R = rot_mat(q[b])
        Nc = R @ clamped_mobility(mob_ref) @ R.T
        fb = v[3 * b * nb:3 * (b + 1) * nb].reshape(nb, 3)
        ub = v[3 * n + 6 * b:3 * n + 6 * b + 6]
        V = -fb.sum(axis=0) / nb - Nc @ ub[:3]
        out[3 * n + 6 * b:3 * n + 6 * b + 3] = V
        out[3 * n + 6 * b + 3:3 * n + 6 * b + 6] = ub[3:]
        out[3 * b * nb:3 * (b + 1) * nb] = (minv * (fb + V)).reshape(-1)
    return out


def prescribed_rotation_matrix(M, K):
    """The library's form of the prescribed-rotation product: the saddle matrix with the
    Omega columns of the blob rows zeroed and each torque row the identity on Omega_b,
    in the unknown order (F, then (V, Omega) per body)."""
    A = saddle_matrix(M, K)
    n = K.shape[0]
    for b in range(K.shape[1] // 6):
        for k in range(3, 6):
            j = n + 6 * b + k
            A[:, j] = 0.0
            A[j, :] = 0.0
            A[j, j] = 1.0
    return A
