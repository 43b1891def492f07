"""The FP32 issue-rate probe (cfm_probe_valu_rate, libcfm_probe.so) through its C ABI: every form reports
a rate between 1e13 and 2e14 lane-instructions per second on one MI355X (256 CUs at
2.4 GHz: 16 lanes per SIMD per clock is 3.9e13), and a bad form is refused."""
import ctypes

import pytest

from cufcm_fil_amd import _lib as L

pytestmark = pytest.mark.gpu


def test_probe_rates_are_plausible_and_bad_forms_refused():
    lib = L.probe_lib()
    for form in range(5):
        r = ctypes.c_double(0.0)
        assert lib.cfm_probe_valu_rate(0, form, ctypes.byref(r)) == L.CFM_OK
        assert 1e13 < r.value < 2e14, (form, r.value)
    r = ctypes.c_double(0.0)
    assert lib.cfm_probe_valu_rate(0, 5, ctypes.byref(r)) == L.CFM_EINVAL
    assert lib.cfm_probe_valu_rate(0, 0, None) == L.CFM_EINVAL
