"""The VALU issue probe the instruction-bound filters are measured against."""
import ctypes as C

import pytest

from zim_amd import _capi


@pytest.mark.gpu
def test_valu_probe_rates_are_plausible(ctx):
    """FP32 and FP64 FMA issue rates of the device: positive and within the
    single-lane issue limit (one wave-instruction per SIMD every four cycles at
    no more than 3 GHz), FP64 not faster than FP32."""
    a, b = C.c_double(), C.c_double()
    _capi.check(ctx.lib.psgc_probe_valu_rate(ctx.h, C.byref(a), C.byref(b)))
    props = __import__("torch").cuda.get_device_properties(0)
    limit = props.multi_processor_count * 4 * 3.0e9 / 4 / 1e9
    assert 0 < b.value <= a.value * 1.05
    assert 0.25 * limit < a.value <= limit
