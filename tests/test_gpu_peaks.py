"""pfem_device_peaks: the STREAM-style copy, the read-only stream and the FP64 FMA loop
return rates inside the physical range of the device (below the published 8 TB/s and
78.6 TFLOP/s, and not absurdly low), and bad arguments are rejected."""
import ctypes as C

import pytest

import plasticity_amd._lib as L

pytestmark = pytest.mark.gpu


def test_device_peaks_in_physical_range(need_gpu):
    gbs, rd, tf = C.c_double(), C.c_double(), C.c_double()
    L.check(L.lib().pfem_device_peaks(1 << 30, 5, C.byref(gbs), C.byref(rd), C.byref(tf)), "pfem_device_peaks")
    print(f"[peaks] copy {gbs.value:.0f} GB/s, read {rd.value:.0f} GB/s, FP64 FMA {tf.value:.1f} TFLOP/s", flush=True)
    assert 500.0 < gbs.value < 8000.0 * 1.02
    assert 500.0 < rd.value < 8000.0 * 1.02
    assert 5.0 < tf.value < 78.6 * 1.02


def test_device_peaks_rejects_bad_arguments(need_gpu):
    gbs, tf = C.c_double(), C.c_double()
    assert L.lib().pfem_device_peaks(1 << 10, 5, C.byref(gbs), None, C.byref(tf)) == L.PFEM_ERR_BAD_ARG
    assert L.lib().pfem_device_peaks(1 << 30, 0, C.byref(gbs), None, C.byref(tf)) == L.PFEM_ERR_BAD_ARG
