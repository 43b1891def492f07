"""Batched ensembles: many independent wall systems of uneven size in one context.

The reference runs every member of an ensemble as its own process with its own
mobility solver (pyfile/driver/func.py:544 launches one executable per parameter
set); ``BatchedMobility`` holds all members together on one GPU.  Arrays are the
members' reference-layout (AoS) arrays concatenated in member order.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .mobility import RpyMobilitySolver


class BatchedMobility:
    """Independent segment systems sharing one device context.

    sizes   particles (segments) of each system
    """

    def __init__(self, sizes, seg_dof=6, wall=True, a=1.0, mu=1.0, device=0):
        self.sizes = np.asarray(sizes, dtype=np.int32).reshape(-1)
        self.n_sys = int(self.sizes.size)
        self.dof = int(seg_dof)
        self.N = int(self.sizes.sum())
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        cfg = L.cfm_batch_config(n_sys=self.n_sys, seg_dof=self.dof, wall=int(bool(wall)),
                                 device=device, a_seg=a, mu=mu)
        h = ctypes.c_void_p()
        L.check_batch(None, L.lib().cfm_batch_create(ctypes.byref(cfg), L.iptr(self.sizes),
                                                     ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            L.lib().cfm_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()

    def _chk(self, code):
        L.check_batch(self.h, code)

    def set_positions(self, x):
        x = L.f64(x, 3 * self.N)
        self._chk(L.lib().cfm_batch_set_positions(self.h, L.ptr(x)))

    def apply(self, f, precision="fp64"):
        f = L.f64(f, self.dof * self.N)
        v = np.empty(self.dof * self.N)
        self._chk(L.lib().cfm_batch_apply(self.h, RpyMobilitySolver._prec(precision), L.ptr(f),
                                          L.ptr(v)))
        return v

    def tuning(self):
        t = L.cfm_batch_tuning()
        self._chk(L.lib().cfm_batch_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, max_sources, grid):
        t = L.cfm_batch_tuning(max_sources=max_sources, grid=grid)
        self._chk(L.lib().cfm_batch_set_tuning(self.h, ctypes.byref(t)))

    def tune(self, precision="fp32c"):
        """Autotune the persistent pair kernel's decomposition; returns (chosen, tried)."""
        best = L.cfm_batch_tuning()
        tried = (L.cfm_batch_tuning * 64)()
        n = ctypes.c_int32(0)
        self._chk(L.lib().cfm_batch_tune(self.h, RpyMobilitySolver._prec(precision),
                                         ctypes.byref(best), tried, 64, ctypes.byref(n)))
        return best.as_dict(), [tried[i].as_dict() for i in range(n.value)]

    def solve(self, rhs, x0=None, rtol=1e-10, restart=50, max_iter=1000, precond="self",
              mode="fp64"):
        """All systems' resistance problems M_s x_s = rhs_s; returns (x, per-system
        stats list, total milliseconds)."""
        rhs = L.f64(rhs, self.dof * self.N)
        x = np.zeros(self.dof * self.N) if x0 is None else L.f64(x0, self.dof * self.N).copy()
        o = L.cfm_solve_opts(system=L.CFM_SYS_SEG, precond=RpyMobilitySolver._PC[precond],
                             mode=RpyMobilitySolver._MODE[mode], restart=restart,
                             max_iter=max_iter, use_guess=int(x0 is not None), rtol=rtol)
        st = (L.cfm_batch_sys_stats * self.n_sys)()
        ms = ctypes.c_double(0.0)
        self._chk(L.lib().cfm_batch_gmres(self.h, L.ptr(rhs), L.ptr(x), ctypes.byref(o), st,
                                          ctypes.byref(ms)))
        stats = [{"iterations": s.iterations, "converged": bool(s.converged),
                  "rel_residual": s.rel_residual, "latency_ms": s.latency_ms} for s in st]
        return x, stats, ms.value

    def split(self, v, per=None):
        """Per-system views of a concatenated vector."""
        per = self.dof if per is None else per
        return [v[per * o:per * (o + n)] for o, n in zip(self.offsets, self.sizes)]
