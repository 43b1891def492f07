"""The allocation-free entry point cfm_apply_noalloc (through the C ABI) against the
existing cfm_apply: every block of a segment + blob context, from device arrays, on two
calls in a row on one GPU -- the second from a smaller context, so the shared static
scratch array still holds the first call's larger partial sums.  Tolerances as for
every product (FP64 <= 1e-12 relative; FP32C <= 1e-6 absolute) against cfm_apply's
default, which forms single-species wall products with the symmetric kernels; the
allocation-free entry point runs the one-directional kernels (their partial sums fit
its scratch array), so with a fixed split it is bitwise equal to cfm_apply with pair
symmetry off, which runs the same launches.  Device memory for the test comes from the
HIP runtime the library itself links."""
import ctypes

import numpy as np
import pytest

from cufcm_fil_amd import CFM_BB, CFM_BS, CFM_FULL, CFM_SB, CFM_SS, RpyMobilitySolver
from cufcm_fil_amd import _lib as L
from cufcm_fil_amd import geometry as geo

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        L.lib()  # the library's own HIP runtime (rpath /opt/rocm/lib)
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class Dev:
    """A device copy of a host float64 array."""

    def __init__(self, a):
        self.n = a.size
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(8, 8 * self.n)) == 0
        assert hip().hipMemcpy(self.p, a.ctypes.data, 8 * self.n, 1) == 0   # host -> device

    def get(self):
        out = np.empty(self.n)
        assert hip().hipMemcpy(out.ctypes.data, self.p, 8 * self.n, 2) == 0  # device -> host
        return out

    def free(self):
        hip().hipFree(self.p)


def _system(nx, ny, nblob, seed):
    xs = geo.tilted_lattice(nx=nx, ny=ny, seed=seed)
    rng = np.random.default_rng(seed + 1)
    xb = np.column_stack([rng.uniform(-5, 10 * nx, nblob), rng.uniform(-5, 10 * ny, nblob),
                          rng.uniform(50.0, 70.0, nblob)]).reshape(-1)
    return xs, xb, geo.loads(xs.size // 3, 6, seed + 2), geo.loads(nblob, 3, seed + 3)


def _err(got, ref, prec):
    if prec == "fp64":
        return np.abs(got - ref).max() / np.abs(ref).max(), 1e-12
    return np.abs(got - ref).max(), 1e-6


def _compare(xs, xb, fs, fb, prec, split):
    ns, nb = xs.size // 3, xb.size // 3
    lib = L.lib()
    with RpyMobilitySolver(ns, nb, precision=prec, tile_split=split) as m:
        m.set_positions(xs, xb)
        dfs, dfb = Dev(fs), Dev(fb)
        v0s = np.random.default_rng(9).standard_normal(6 * ns)  # accumulated into by SB
        try:
            for which in (CFM_SS, CFM_SB, CFM_BB, CFM_BS, CFM_FULL):
                refs = []
                for sym in (1, 0):   # cfm_apply's default, then the one-directional kernels
                    L.check(m.handle, lib.cfm_set_pair_symmetry(m.handle, sym))
                    r_s = v0s.copy() if which == CFM_SB else np.zeros(6 * ns)
                    r_b = np.zeros(3 * nb)
                    L.check(m.handle, lib.cfm_apply(m.handle, which, L.ptr(fs), L.ptr(fb),
                                                    L.ptr(r_s), L.ptr(r_b)))
                    refs.append((r_s, r_b))
                L.check(m.handle, lib.cfm_set_pair_symmetry(m.handle, 1))
                (ref_s, ref_b), (dir_s, dir_b) = refs
                dvs = Dev(v0s.copy() if which == CFM_SB else np.zeros(6 * ns))
                dvb = Dev(np.zeros(3 * nb))
                try:
                    L.check(m.handle, lib.cfm_apply_noalloc(m.handle, which, dfs.p, dfb.p, dvs.p, dvb.p))
                    got_s, got_b = dvs.get(), dvb.get()
                finally:
                    dvs.free()
                    dvb.free()
                if which in (CFM_SS, CFM_SB, CFM_FULL):
                    if split > 0:
                        assert np.array_equal(got_s, dir_s), which
                    e, t = _err(got_s, ref_s, prec)
                    assert e < t, (which, e)
                if which in (CFM_BB, CFM_BS, CFM_FULL):
                    if split > 0:
                        assert np.array_equal(got_b, dir_b), which
                    e, t = _err(got_b, ref_b, prec)
                    assert e < t, (which, e)
        finally:
            dfs.free()
            dfb.free()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("split", [0, 5])
def test_noalloc_apply_matches_apply_on_two_calls_in_a_row(prec, split):
    _compare(*_system(8, 6, 900, seed=1), prec, split)     # 960 segments + 900 blobs
    _compare(*_system(3, 2, 70, seed=4), prec, split)      # then 120 segments + 70 blobs


def test_noalloc_apply_allocates_nothing():
    xs, xb, fs, fb = _system(6, 6, 500, seed=2)
    ns, nb = xs.size // 3, xb.size // 3
    lib = L.lib()
    with RpyMobilitySolver(ns, nb, precision="fp32c") as m:
        m.set_positions(xs, xb)
        dfs, dfb, dvs, dvb = Dev(fs), Dev(fb), Dev(np.zeros(6 * ns)), Dev(np.zeros(3 * nb))
        try:
            L.check(m.handle, lib.cfm_apply_noalloc(m.handle, CFM_FULL, dfs.p, dfb.p, dvs.p, dvb.p))
            free0, free1, tot = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
            assert hip().hipMemGetInfo(ctypes.byref(free0), ctypes.byref(tot)) == 0
            for _ in range(3):
                L.check(m.handle, lib.cfm_apply_noalloc(m.handle, CFM_FULL, dfs.p, dfb.p, dvs.p, dvb.p))
            assert hip().hipMemGetInfo(ctypes.byref(free1), ctypes.byref(tot)) == 0
            assert free1.value == free0.value
        finally:
            for d in (dfs, dfb, dvs, dvb):
                d.free()


def test_noalloc_rejects_partials_beyond_the_scratch():
    """A fixed split whose partial sums exceed CFM_NOALLOC_SCRATCH_BYTES is refused
    before anything is launched (64 x 6 x 12,800 doubles = 39 MB > 32 MiB)."""
    xs = geo.wave_lattice()
    n = xs.size // 3
    with RpyMobilitySolver(n, 0, tile_split=64) as m:
        m.set_positions(xs)
        d = Dev(np.zeros(6 * n))
        try:
            code = L.lib().cfm_apply_noalloc(m.handle, CFM_SS, d.p, None, d.p, None)
            assert code == L.CFM_EINVAL
            assert b"scratch" in L.lib().cfm_last_error(m.handle)
        finally:
            d.free()
