"""Host-side mirror of the reference's mobility-solver plugin interface.

``RpyMobilitySolver`` keeps the method names of the reference's abstract
``mobility_solver`` / ``rpy_mobility_solver`` (src/cilia/mobility/mobility_solver.hpp:26-46,
rpy_mobility_solver.cu:14-338) and holds the same host buffers (x/f/v of
segments and blobs, reference AoS layout), while every product runs in the HIP
library through the C ABI of include/cufcm_fil.h.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


class RpyMobilitySolver:
    """Wall-corrected (or unbounded) RPY mobility of segments + blobs on one GPU."""

    def __init__(self, n_seg: int, n_blob: int = 0, *, seg_dof: int = 6, wall: bool = True,
                 a_seg: float = 1.0, a_blob: float = 1.0, mu: float = 1.0,
                 precision: str = "fp64", device: int = 0, tile_split: int = 0):
        lib = L.lib()
        cfg = L.cfm_config(n_seg=n_seg, n_blob=n_blob, seg_dof=seg_dof, wall=int(wall),
                           a_seg=a_seg, a_blob=a_blob, mu=mu,
                           precision=self._prec(precision), device=device,
                           tile_split=tile_split, reserved=0)
        h = ctypes.c_void_p()
        code = lib.cfm_create(ctypes.byref(cfg), ctypes.byref(h))
        if code != L.CFM_OK:
            raise L.CfmError(code, lib.cfm_last_error(None).decode())
        self._h = h
        self.n_seg, self.n_blob, self.seg_dof = n_seg, n_blob, seg_dof
        self.wall, self.mu = wall, mu
        # host buffers of the reference interface (mobility_solver.hpp:57-63)
        self.x_segs_host = np.zeros(3 * n_seg)
        self.x_blobs_host = np.zeros(3 * n_blob)
        self.f_segs_host = np.zeros(seg_dof * n_seg)
        self.f_blobs_host = np.zeros(3 * n_blob)
        self.v_segs_host = np.zeros(seg_dof * n_seg)
        self.v_blobs_host = np.zeros(3 * n_blob)

    @staticmethod
    def _prec(p):
        if isinstance(p, int):
            return p
        return {"fp64": L.CFM_FP64, "fp32c": L.CFM_FP32C, "fp32": L.CFM_FP32C}[p.lower()]

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            L.lib().cfm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_precision(self, precision):
        L.check(self._h, L.lib().cfm_set_precision(self._h, self._prec(precision)))

    # -- transfers (copy_*_positions_to_device) -----------------------------
    def set_positions(self, x_seg=None, x_blob=None):
        if x_seg is not None:
            self.x_segs_host[:] = L.f64(x_seg, 3 * self.n_seg)
        if x_blob is not None:
            self.x_blobs_host[:] = L.f64(x_blob, 3 * self.n_blob)
        L.check(self._h, L.lib().cfm_set_positions(
            self._h, L.ptr(self.x_segs_host), L.ptr(self.x_blobs_host)))

    copy_segment_positions_to_device = set_positions

    # -- products -------------------------------------------------------------
    def apply(self, which: int, f_seg=None, f_blob=None):
        if f_seg is not None:
            self.f_segs_host[:] = L.f64(f_seg, self.seg_dof * self.n_seg)
        if f_blob is not None:
            self.f_blobs_host[:] = L.f64(f_blob, 3 * self.n_blob)
        L.check(self._h, L.lib().cfm_apply(self._h, int(which), L.ptr(self.f_segs_host),
                                           L.ptr(self.f_blobs_host), L.ptr(self.v_segs_host),
                                           L.ptr(self.v_blobs_host)))
        return self.v_segs_host, self.v_blobs_host

    def evaluate_segment_segment_mobility(self, f_seg=None):
        return self.apply(L.CFM_SS, f_seg=f_seg)[0].copy()

    def evaluate_segment_blob_mobility(self, f_blob=None):
        return self.apply(L.CFM_SB, f_blob=f_blob)[0].copy()

    def evaluate_blob_blob_mobility(self, f_blob=None):
        return self.apply(L.CFM_BB, f_blob=f_blob)[1].copy()

    def evaluate_blob_segment_mobility(self, f_seg=None):
        return self.apply(L.CFM_BS, f_seg=f_seg)[1].copy()

    def evaluate_full_mobility(self, f_seg=None, f_blob=None):
        v_s, v_b = self.apply(L.CFM_FULL, f_seg=f_seg, f_blob=f_blob)
        return v_s.copy(), v_b.copy()

    def time_apply(self, which: int, reps: int = 10) -> float:
        ms = ctypes.c_double()
        L.check(self._h, L.lib().cfm_time_apply(self._h, int(which), int(reps), ctypes.byref(ms)))
        return ms.value

    # -- linear solves (solve_linear_system / apply_preconditioner_new) -------
    _SYS = {"seg": L.CFM_SYS_SEG, "blob": L.CFM_SYS_BLOB, "body": L.CFM_SYS_BODY}
    _PC = {"none": L.CFM_PC_NONE, "diag": L.CFM_PC_DIAG, "self": L.CFM_PC_SELF,
           "near": L.CFM_PC_NEAR, "dense": L.CFM_PC_DENSE}
    _MODE = {"fp64": L.CFM_SOLVE_FP64, "mixed": L.CFM_SOLVE_MIXED}

    def _sys_size(self, system):
        k = self._SYS[system]
        if k == L.CFM_SYS_SEG:
            return self.seg_dof * self.n_seg
        if k == L.CFM_SYS_BODY:
            return 3 * self.n_blob + 6 * getattr(self, "_n_body", 0)
        return 3 * self.n_blob

    def solve(self, system: str, rhs, x0=None, *, rtol=1e-8, restart=50, max_iter=1000,
              precond="self", mode="fp64", raise_on_noconv=True):
        """GMRES solve of M x = rhs for one species; returns (x, stats dict)."""
        n = self._sys_size(system)
        b = L.f64(rhs, n)
        x = np.zeros(n) if x0 is None else L.f64(x0, n).copy()
        o = L.cfm_solve_opts(system=self._SYS[system], precond=self._PC[precond],
                             mode=self._MODE[mode], restart=restart, max_iter=max_iter,
                             use_guess=int(x0 is not None), rtol=rtol)
        st = L.cfm_solve_stats()
        code = L.lib().cfm_gmres(self._h, ctypes.byref(o), L.ptr(b), L.ptr(x), ctypes.byref(st))
        if code != L.CFM_OK and (raise_on_noconv or code != L.CFM_ENOCONV):
            L.check(self._h, code)
        return x, st.as_dict()

    solve_linear_system = solve

    def apply_preconditioner(self, system: str, vec, precond="self"):
        n = self._sys_size(system)
        v = L.f64(vec, n)
        out = np.zeros(n)
        L.check(self._h, L.lib().cfm_apply_preconditioner(self._h, self._SYS[system],
                                                           self._PC[precond], L.ptr(v), L.ptr(out)))
        return out

    def system_apply(self, system: str, x, precision="fp64"):
        """y = A x of a whole system (system_matrix_mult, mobility_solver.cpp:1596-1881)."""
        n = self._sys_size(system)
        v = L.f64(x, n)
        out = np.zeros(n)
        L.check(self._h, L.lib().cfm_system_apply(self._h, self._SYS[system],
                                                   self._prec(precision), L.ptr(v), L.ptr(out)))
        return out

    # -- free rigid bodies made of the blobs (rigid_body.cpp, mobility_solver.cpp) --
    def body_init(self, ref, n_body=1):
        """n_body bodies sharing the body-frame blob offsets ref (3 n_blob / n_body)."""
        r = L.f64(ref, 3 * (self.n_blob // max(int(n_body), 1)))
        L.check(self._h, L.lib().cfm_body_init(self._h, int(n_body), L.ptr(r)))
        self._n_body = int(n_body)

    def body_set_state(self, x, q):
        """Body centres (3 n_body) and unit quaternions (w, x, y, z) (4 n_body)."""
        nb = self._n_body
        xs, qs = L.f64(x, 3 * nb), L.f64(q, 4 * nb)
        L.check(self._h, L.lib().cfm_body_set_state(self._h, L.ptr(xs), L.ptr(qs)))

    def make_body_reference_matrices(self, body=0, *, rtol=1e-10, precond="dense", mode="fp64",
                                     restart=50, max_iter=1000):
        """(K^T K)^-1 and the 6 x 6 body mobility, body frame (mobility_solver.cpp:3049-3146)."""
        o = L.cfm_solve_opts(system=L.CFM_SYS_BODY, precond=self._PC[precond],
                             mode=self._MODE[mode], restart=restart, max_iter=max_iter,
                             use_guess=0, rtol=rtol)
        ktk, mob = np.zeros(36), np.zeros(36)
        st = L.cfm_solve_stats()
        L.check(self._h, L.lib().cfm_body_reference_matrices(
            self._h, int(body), ctypes.byref(o), L.ptr(ktk), L.ptr(mob), ctypes.byref(st)))
        return ktk.reshape(6, 6), mob.reshape(6, 6), st.as_dict()

    def body_set_reference_matrices(self, ktk_inv=None, mobility=None):
        a = None if ktk_inv is None else L.f64(ktk_inv, 36)
        b = None if mobility is None else L.f64(mobility, 36)
        L.check(self._h, L.lib().cfm_body_set_reference_matrices(
            self._h, None if a is None else L.ptr(a), None if b is None else L.ptr(b)))

    def body_reference_matrices(self):
        ktk, mob = np.zeros(36), np.zeros(36)
        L.check(self._h, L.lib().cfm_body_get_reference_matrices(self._h, L.ptr(ktk), L.ptr(mob)))
        return ktk.reshape(6, 6), mob.reshape(6, 6)

    _MOTION = {"free": L.CFM_BODY_FREE, "prescribed_velocity": L.CFM_BODY_PRESCRIBED_VELOCITY,
               "prescribed_rotation": L.CFM_BODY_PRESCRIBED_ROTATION}

    def body_set_motion(self, motion):
        """The reference's BODY_VELOCITY_TYPE (config.hpp:94-101): "free" (0) or
        "prescribed_rotation" (2: Omega held at zero, the torque rows of CFM_SYS_BODY read
        Omega; mobility_solver.cpp:1947-1994, 2177-2181).  "prescribed_velocity" (1) is
        the blob system of bodies held fixed and is refused."""
        L.check(self._h, L.lib().cfm_body_set_motion(self._h, int(self._MOTION[motion])))

    def body_motion(self):
        t = ctypes.c_int32(-1)
        L.check(self._h, L.lib().cfm_body_get_motion(self._h, ctypes.byref(t)))
        return {v: k for k, v in self._MOTION.items()}[t.value]

    # -- measurement --------------------------------------------------------------
    def profile(self, enable=True):
        L.check(self._h, L.lib().cfm_profile(self._h, int(bool(enable))))

    def profile_read(self):
        st = L.cfm_profile_stats()
        L.check(self._h, L.lib().cfm_profile_read(self._h, ctypes.byref(st)))
        return st.as_dict()

    def set_pair_symmetry(self, on=True):
        """FP32C wall segment products evaluate each unordered pair once (default) or,
        off, with the one-directional kernel."""
        L.check(self._h, L.lib().cfm_set_pair_symmetry(self._h, int(bool(on))))

    # -- learned near-field correction of the blob preconditioner ----------------
    def gnn_init(self, seed=0, capacity=256):
        L.check(self._h, L.lib().cfm_gnn_init(self._h, int(seed), int(capacity)))
        return L.lib().cfm_gnn_num_params(self._h)

    def gnn_record(self, on=True):
        L.check(self._h, L.lib().cfm_gnn_record(self._h, int(bool(on))))

    def gnn_samples(self):
        return L.lib().cfm_gnn_samples(self._h)

    def gnn_add_samples(self, rhs, sol=None):
        rhs = L.f64(rhs)
        n = rhs.size // (3 * self.n_blob)
        sol = None if sol is None else L.f64(sol, rhs.size)
        L.check(self._h, L.lib().cfm_gnn_add_samples(self._h, n, L.ptr(rhs), L.ptr(sol)))

    def gnn_train(self, epochs=20, batch=8, lr=1e-3):
        l0, l1 = ctypes.c_double(), ctypes.c_double()
        L.check(self._h, L.lib().cfm_gnn_train(self._h, int(epochs), int(batch), float(lr),
                                               ctypes.byref(l0), ctypes.byref(l1)))
        return l0.value, l1.value

    def gnn_enable(self, on=True):
        L.check(self._h, L.lib().cfm_gnn_enable(self._h, int(bool(on))))

    def gnn_train_gated(self, validation_rhs, epochs=8, batch=16, lr=1e-3, rtol=1e-10,
                        max_iter=3000):
        """Train on the recorded samples, then keep the correction only if it lowers
        the total GMRES iterations of the validation blob solves (near-field
        preconditioner, FP64, rtol); otherwise the analytic preconditioner alone stays
        in use.  Every validation solve must reach rtol either way."""
        def iterations():
            its = []
            for b in validation_rhs:
                _, st = self.solve("blob", b, rtol=rtol, precond="near", mode="fp64",
                                   max_iter=max_iter)
                if not st["converged"]:
                    raise L.CfmError(L.CFM_ENOCONV, "validation solve did not converge")
                its.append(st["iterations"])
            return its

        self.gnn_enable(False)
        its0 = iterations()
        l0, l1 = self.gnn_train(epochs=epochs, batch=batch, lr=lr)
        self.gnn_enable(True)
        its1 = iterations()
        keep = sum(its1) < sum(its0)
        if not keep:
            self.gnn_enable(False)
        return {"loss_before": l0, "loss_after": l1, "iterations_without": its0,
                "iterations_with": its1, "kept": keep}

    def gnn_params(self):
        th = np.zeros(L.lib().cfm_gnn_num_params(self._h), dtype=np.float32)
        L.check(self._h, L.lib().cfm_gnn_get_params(self._h, L.fptr(th)))
        return th

    def gnn_set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        L.check(self._h, L.lib().cfm_gnn_set_params(self._h, L.fptr(th)))

    def gnn_correction(self):
        corr = np.zeros(self.n_blob * 17 * 9)
        nbr = np.zeros(self.n_blob * 16, dtype=np.int32)
        L.check(self._h, L.lib().cfm_gnn_correction(self._h, L.ptr(corr), L.iptr(nbr)))
        return corr.reshape(self.n_blob, 17, 3, 3), nbr.reshape(self.n_blob, 16)

    def gnn_gradient(self, s0, s1):
        g = np.zeros(L.lib().cfm_gnn_num_params(self._h), dtype=np.float32)
        loss = ctypes.c_double()
        L.check(self._h, L.lib().cfm_gnn_gradient(self._h, int(s0), int(s1), L.fptr(g),
                                                  ctypes.byref(loss)))
        return loss.value, g

    # -- squirmer bodies (blobs only; bodies from body_init / body_set_state) ---------
    def squirmer_init(self, B, A=None, *, dt=1e-3, precond="dense", mode="fp64", rtol=1e-10,
                      restart=50, max_iter=2000):
        """Slip modes B_n (tangential) and A_n (radial), n = 1.. (at most 8)."""
        B = list(np.atleast_1d(np.asarray(B, dtype=np.float64)))
        A = [0.0] * len(B) if A is None else list(np.atleast_1d(np.asarray(A, dtype=np.float64)))
        n = max(len(A), len(B))
        P = L.cfm_squirmer_params()
        P.n_modes = n
        for k in range(n):
            P.A[k] = A[k] if k < len(A) else 0.0
            P.B[k] = B[k] if k < len(B) else 0.0
        P.dt = dt
        P.solve = L.cfm_solve_opts(system=L.CFM_SYS_BODY, precond=self._PC[precond],
                                   mode=self._MODE[mode], restart=restart, max_iter=max_iter,
                                   use_guess=1, rtol=rtol)
        L.check(self._h, L.lib().cfm_squirmer_init(self._h, ctypes.byref(P)))

    def squirmer_step(self, nt):
        """One step; returns ((V, Omega) per body as an (n_body, 6) array, solve stats)."""
        nb = self._n_body_cached()
        vel = np.zeros(6 * nb)
        st = L.cfm_solve_stats()
        L.check(self._h, L.lib().cfm_squirmer_step(self._h, int(nt), L.ptr(vel), ctypes.byref(st)))
        return vel.reshape(nb, 6), st.as_dict()

    def _n_body_cached(self):
        return getattr(self, "_n_body", 1)

    # -- coarse correction of the near-field preconditioner --------------------------
    _COARSE = {"none": L.CFM_COARSE_NONE, "linear": L.CFM_COARSE_LINEAR,
               "learned": L.CFM_COARSE_LEARNED}

    def nearfield_coarse(self, mode="linear"):
        """Coarse correction added to the 'near' blob preconditioner: 'none', 'linear'
        (the 12 linear fields of the blob positions) or 'learned' (those plus the
        network's 16 coarse vectors)."""
        L.check(self._h, L.lib().cfm_nearfield_coarse(self._h, self._COARSE[mode]))

    def gnn_train_coarse(self, epochs=10, batch=32, lr=1e-3):
        l0, l1 = ctypes.c_double(), ctypes.c_double()
        L.check(self._h, L.lib().cfm_gnn_train_coarse(self._h, int(epochs), int(batch), float(lr),
                                                      ctypes.byref(l0), ctypes.byref(l1)))
        return l0.value, l1.value

    def gnn_coarse_gradient(self, s0, s1):
        g = np.zeros(L.lib().cfm_gnn_num_params(self._h), dtype=np.float32)
        loss, delta = ctypes.c_double(), ctypes.c_double()
        L.check(self._h, L.lib().cfm_gnn_coarse_gradient(self._h, int(s0), int(s1), L.fptr(g),
                                                         ctypes.byref(loss), ctypes.byref(delta)))
        return loss.value, g, delta.value

    def gnn_coarse_vectors(self):
        z = np.zeros(16 * 3 * self.n_blob)
        L.check(self._h, L.lib().cfm_gnn_coarse_vectors(self._h, L.ptr(z)))
        return z.reshape(16, 3 * self.n_blob)

    def gnn_train_coarse_gated(self, validation_rhs, epochs=10, batch=32, lr=1e-3, rtol=1e-10,
                               max_iter=3000):
        """Train the coarse head on the recorded right-hand sides, then keep the learned
        coarse space only if it lowers the total GMRES iterations of the validation blob
        solves against the linear fields alone (near-field preconditioner, FP64, rtol);
        otherwise the linear fields stay in use.  Every validation solve must reach rtol."""
        def iterations(mode):
            self.nearfield_coarse(mode)
            its = []
            for b in validation_rhs:
                _, st = self.solve("blob", b, rtol=rtol, precond="near", mode="fp64",
                                   max_iter=max_iter)
                if not st["converged"]:
                    raise L.CfmError(L.CFM_ENOCONV, "validation solve did not converge")
                its.append(st["iterations"])
            return its

        its0 = iterations("linear")
        l0, l1 = self.gnn_train_coarse(epochs=epochs, batch=batch, lr=lr)
        its1 = iterations("learned")
        keep = sum(its1) < sum(its0)
        self.nearfield_coarse("learned" if keep else "linear")
        return {"loss_before": l0, "loss_after": l1, "iterations_linear": its0,
                "iterations_learned": its1, "kept": keep}

    def gemm_bf16(self, A, B, transA=False, transB=False):
        A = np.ascontiguousarray(A, dtype=np.float32)
        B = np.ascontiguousarray(B, dtype=np.float32)
        M, K = (A.shape[1], A.shape[0]) if transA else A.shape
        N = B.shape[0] if transB else B.shape[1]
        C = np.zeros((M, N), dtype=np.float32)
        L.check(self._h, L.lib().cfm_gemm_bf16(self._h, M, N, K, L.fptr(A), int(transA), L.fptr(B),
                                               int(transB), L.fptr(C)))
        return C
