"""CPU checks of the C-ABI library: it loads, exports every declared symbol and
rejects bad configurations before touching a device."""
import ctypes
import os
import subprocess

import pytest

from cufcm_fil_amd import _lib as L


def test_library_exports_every_header_symbol():
    syms = L.header_symbols()
    assert len(syms) >= 8
    lib = L.lib()
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(syms) <= exported
    assert lib.cfm_abi_version() == 1


def test_probe_library_exports_its_header_and_stays_out_of_the_product():
    psyms = L.header_symbols(L.PROBE_HEADER)
    assert psyms == ["cfm_probe_valu_rate"]
    assert hasattr(L.probe_lib(), "cfm_probe_valu_rate")
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    assert "cfm_probe_valu_rate" not in out
    assert not set(psyms) & set(L.header_symbols())


def _cfg(**kw):
    base = dict(n_seg=40, n_blob=0, seg_dof=6, wall=1, a_seg=1.0, a_blob=1.0, mu=1.0,
                precision=0, device=0, tile_split=0, reserved=0)
    base.update(kw)
    return L.cfm_config(**base)


@pytest.mark.parametrize("bad", [dict(seg_dof=4), dict(n_seg=-1), dict(n_seg=0, n_blob=0),
                                 dict(a_seg=2.0), dict(mu=0.0), dict(precision=7),
                                 dict(tile_split=-2), dict(tile_split=65)])
def test_create_rejects_bad_config(bad):
    h = ctypes.c_void_p()
    code = L.lib().cfm_create(ctypes.byref(_cfg(**bad)), ctypes.byref(h))
    assert code == L.CFM_EINVAL
    assert h.value is None
    assert L.lib().cfm_last_error(None)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_reports_hip_error():
    h = ctypes.c_void_p()
    code = L.lib().cfm_create(ctypes.byref(_cfg()), ctypes.byref(h))
    assert code == L.CFM_EHIP
    assert b"device" in L.lib().cfm_last_error(None).lower()


def test_null_context_is_rejected():
    assert L.lib().cfm_set_positions(None, None, None) == L.CFM_EINVAL
    assert L.lib().cfm_destroy(None) == L.CFM_OK


def _bcfg(**kw):
    base = dict(n_sys=3, seg_dof=6, wall=1, device=0, a_seg=1.0, mu=1.0)
    base.update(kw)
    return L.cfm_batch_config(**base)


@pytest.mark.parametrize("bad,sizes", [(dict(n_sys=0), [20]), (dict(seg_dof=5), [20, 20, 20]),
                                       (dict(a_seg=2.0), [20, 20, 20]), (dict(mu=-1.0), [20, 20, 20]),
                                       (dict(), [20, 0, 20])])
def test_batch_create_rejects_bad_config(bad, sizes):
    h = ctypes.c_void_p()
    n = (ctypes.c_int32 * len(sizes))(*sizes)
    code = L.lib().cfm_batch_create(ctypes.byref(_bcfg(**bad)), n, ctypes.byref(h))
    assert code == L.CFM_EINVAL and h.value is None
    assert L.lib().cfm_batch_last_error(None)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_batch_create_without_gpu_reports_hip_error():
    h = ctypes.c_void_p()
    n = (ctypes.c_int32 * 3)(20, 40, 60)
    assert L.lib().cfm_batch_create(ctypes.byref(_bcfg()), n, ctypes.byref(h)) == L.CFM_EHIP
    assert b"device" in L.lib().cfm_batch_last_error(None).lower()
    assert L.lib().cfm_batch_destroy(None) == L.CFM_OK
