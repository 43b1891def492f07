"""FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
(``cufcm_fil_amd``) never imports, links or executes anything under
``oracle/``.

Contents
  * ctypes binding of ``rpy_oracle.cpp`` (the FP64 OpenMP restatement of
    ``src/cilia/cuda_functions.cu:9-757``): pair functions and the direct
    all-pairs apply in the reference's AoS layout.
  * ``mobility_matrix``: the explicit grand mobility matrix for small systems,
    assembled column by column from the pair functions.
  * ``gmres_reference``: restatement of the reference's right-preconditioned
    GMRES (``src/cilia/mobility/mobility_solver.cpp:2187-2669``: MGS Arnoldi
    :2453-2465, Givens :2467-2506, relative residual :2510, back substitution
    :2520-2559), with an optional restart length.

Parity status: the pair functions are pinned by the independent checks in
``tests/test_oracle_pins.py``; ``gmres_reference`` is pinned against dense
``numpy.linalg.solve``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librpy_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_wall_self.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.oracle_wall_pair.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oracle_free_pair.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _dp,
                                       ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.oracle_apply.argtypes = [ctypes.c_int, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, _dp, _i64p, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, _dp, _i64p, ctypes.c_double,
                                   _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_wtime.restype = ctypes.c_double
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def wall_self(z, F, vd=6, fd=6):
    F6 = np.zeros(6)
    F6[:fd] = F[:fd]
    V = np.zeros(6)
    lib().oracle_wall_self(vd, fd, float(z), _p(F6), _p(V))
    return V[:vd]


def wall_pair(xi, xj, F, vd=6, fd=6):
    F6 = np.zeros(6)
    F6[:fd] = F[:fd]
    V = np.zeros(6)
    lib().oracle_wall_pair(vd, fd, _p(_c(xi)), _p(_c(xj)), _p(F6), _p(V))
    return V[:vd]


def free_pair(xi, ai, xj, aj, F, vd=6, fd=6, mu=1.0):
    F6 = np.zeros(6)
    F6[:fd] = F[:fd]
    V = np.zeros(6)
    lib().oracle_free_pair(vd, fd, _p(_c(xi)), float(ai), _p(_c(xj)), float(aj), float(mu),
                           _p(F6), _p(V))
    return V[:vd]


def apply(x_tgt, vd, x_src, fd, f_src, *, wall=True, id_tgt=None, id_src=None,
          a_tgt=1.0, a_src=1.0, mu=1.0, out=None, accumulate=False, nthreads=0):
    """V_tgt = M(tgt <- src) F_src in AoS layout (Mss/Mbb/Msb/Mbs_mult)."""
    x_tgt = _c(x_tgt).reshape(-1)
    x_src = _c(x_src).reshape(-1)
    f_src = _c(f_src).reshape(-1)
    nt = x_tgt.size // 3
    ns = x_src.size // 3
    assert f_src.size == fd * ns
    if id_tgt is None:
        id_tgt = np.arange(nt, dtype=np.int64)
    if id_src is None:
        id_src = np.arange(ns, dtype=np.int64)
    id_tgt = _c(id_tgt, np.int64)
    id_src = _c(id_src, np.int64)
    if out is None:
        out = np.zeros(vd * nt)
    assert out.size == vd * nt and out.dtype == np.float64 and out.flags.c_contiguous
    lib().oracle_apply(int(wall), float(mu), nt, vd, _p(x_tgt), id_tgt.ctypes.data_as(_i64p),
                       float(a_tgt), ns, fd, _p(x_src), id_src.ctypes.data_as(_i64p),
                       float(a_src), _p(f_src), _p(out), int(accumulate), int(nthreads))
    return out


def apply_full(x_seg, f_seg, x_blob=None, f_blob=None, *, seg_dof=6, wall=True, mu=1.0,
               nthreads=0):
    """Velocities of every particle: V_seg = Mss F_seg + Msb F_blob,
    V_blob = Mbs F_seg + Mbb F_blob (evaluate_full_mobility,
    rpy_mobility_solver.cu:330-338).  Global ids: segments first, then blobs."""
    x_seg = _c(x_seg).reshape(-1)
    ns = x_seg.size // 3
    id_s = np.arange(ns, dtype=np.int64)
    v_seg = apply(x_seg, seg_dof, x_seg, seg_dof, f_seg, wall=wall, id_tgt=id_s, id_src=id_s,
                  mu=mu, nthreads=nthreads)
    if x_blob is None or len(x_blob) == 0:
        return v_seg, np.zeros(0)
    x_blob = _c(x_blob).reshape(-1)
    nb = x_blob.size // 3
    id_b = np.arange(ns, ns + nb, dtype=np.int64)
    apply(x_seg, seg_dof, x_blob, 3, f_blob, wall=wall, id_tgt=id_s, id_src=id_b, mu=mu,
          out=v_seg, accumulate=True, nthreads=nthreads)
    v_blob = apply(x_blob, 3, x_blob, 3, f_blob, wall=wall, id_tgt=id_b, id_src=id_b, mu=mu,
                   nthreads=nthreads)
    apply(x_blob, 3, x_seg, seg_dof, f_seg, wall=wall, id_tgt=id_b, id_src=id_s, mu=mu,
          out=v_blob, accumulate=True, nthreads=nthreads)
    return v_seg, v_blob


def mobility_matrix(x, dof=6, wall=True, a=1.0, mu=1.0):
    """Dense grand mobility (dof*N x dof*N) of one species, one column per unit load."""
    x = _c(x).reshape(-1, 3)
    n = x.shape[0]
    M = np.zeros((dof * n, dof * n))
    e = np.zeros(dof * n)
    for c in range(dof * n):
        e[:] = 0.0
        e[c] = 1.0
        M[:, c] = apply(x, dof, x, dof, e, wall=wall, a_tgt=a, a_src=a, mu=mu)
    return M


def gmres_reference(matvec, precond, rhs, x0=None, tol=1e-8, restart=None, max_iter=350):
    """Right-preconditioned GMRES as in mobility_solver.cpp:2187-2669.

    Solves A x = rhs with A = matvec, preconditioner application precond(v) = P^-1 v.
    Returns (x, iterations, final relative residual estimate).  ``restart=None``
    reproduces the reference's unrestarted loop (max_iter = min(n, 350)).
    """
    rhs = np.asarray(rhs, dtype=np.float64)
    n = rhs.size
    nrm_b = np.linalg.norm(rhs)
    if nrm_b == 0.0:
        return np.zeros(n), 0, 0.0
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    total = 0
    m = min(n, max_iter) if restart is None else min(restart, n)
    while True:
        r = rhs - matvec(x)
        beta0 = np.linalg.norm(r)
        if beta0 <= tol * nrm_b:
            return x, total, beta0 / nrm_b
        if total == 0 and beta0 > nrm_b:  # reject a poor initial guess (:2406-2417)
            x[:] = 0.0
            r = rhs.copy()
            beta0 = nrm_b
        Q = np.zeros((n, m + 1))
        H = np.zeros((m + 1, m + 1))
        cs = np.zeros(m)
        sn = np.zeros(m)
        beta = np.zeros(m + 1)
        beta[0] = beta0
        Q[:, 0] = r / beta0
        for it in range(1, m + 1):
            w = matvec(precond(Q[:, it - 1]))
            for i in range(it):  # modified Gram-Schmidt
                H[i, it - 1] = np.dot(w, Q[:, i])
                w = w - H[i, it - 1] * Q[:, i]
            H[it, it - 1] = np.linalg.norm(w)
            Q[:, it] = w / H[it, it - 1] if H[it, it - 1] != 0.0 else 0.0
            for i in range(1, it):  # previous rotations
                t = cs[i - 1] * H[i - 1, it - 1] + sn[i - 1] * H[i, it - 1]
                H[i, it - 1] = cs[i - 1] * H[i, it - 1] - sn[i - 1] * H[i - 1, it - 1]
                H[i - 1, it - 1] = t
            a, b = H[it - 1, it - 1], H[it, it - 1]
            if b == 0.0:
                cs[it - 1], sn[it - 1] = 1.0, 0.0
            elif abs(b) > abs(a):
                t = a / b
                sn[it - 1] = (2.0 * (b > 0) - 1.0) / np.sqrt(1.0 + t * t)
                cs[it - 1] = t * sn[it - 1]
            else:
                t = b / a
                cs[it - 1] = (2.0 * (a > 0) - 1.0) / np.sqrt(1.0 + t * t)
                sn[it - 1] = t * cs[it - 1]
            H[it - 1, it - 1] = cs[it - 1] * a + sn[it - 1] * b
            H[it, it - 1] = 0.0
            beta[it] = -sn[it - 1] * beta[it - 1]
            beta[it - 1] = cs[it - 1] * beta[it - 1]
            total += 1
            rel = abs(beta[it]) / nrm_b
            if rel <= tol or it == m or total >= max_iter:
                y = np.zeros(it)
                for i in range(it - 1, -1, -1):
                    y[i] = (beta[i] - np.dot(H[i, i + 1:it], y[i + 1:it])) / H[i, i]
                x = x + precond(Q[:, :it] @ y)
                break
        if rel <= tol or total >= max_iter or restart is None:
            r = rhs - matvec(x)
            return x, total, np.linalg.norm(r) / nrm_b


def wtime() -> float:
    return lib().oracle_wtime()


def max_threads() -> int:
    return lib().oracle_max_threads()
