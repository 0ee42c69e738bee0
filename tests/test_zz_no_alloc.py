"""sfx_total_volume_noalloc (the allocation-free variant of sfx_total_volume, with its
partial sums in a static device array) through the C ABI: its result equals
sfx_total_volume's on two calls in a row, the second on a smaller mesh, and stays the
same bits when called again."""
import ctypes as C

import numpy as np
import pytest

from helpers import gpu_for
from sfincs_amd import _lib
from sfincs_amd import workloads as W

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself uses (the test, not the entry point, allocates)."""
    L = _lib.lib()
    del L
    for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            h = None
    assert h is not None, "libamdhip64 not found"
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc(inst, hip, dev_out):
    L = _lib.lib()
    f = L.sfx_total_volume_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p]
    rc = f(inst.h, dev_out)
    assert rc == 0, _lib.last_error()
    assert hip.hipDeviceSynchronize() == 0
    v = C.c_double()
    assert hip.hipMemcpy(C.byref(v), dev_out, 8, 2) == 0          # device to host
    return v.value


def test_noalloc_volume_equals_total_volume():
    hip = _hip()
    dev_out = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev_out), 8) == 0
    big = W.subgrid_case(300, 260, dx=25.0, seed=91, t_end=600.0, nsrc=2, rain_mmh=20.0)
    small = W.regular_case(40, 52, dx=50.0, seed=92, t_end=600.0)
    for k, sc in enumerate((big, small)):
        print(f"case noalloc call {k + 1}: {sc.name}", flush=True)
        inst = gpu_for(sc.build_model())
        inst.update(-1000.0)
        ref = inst.total_volume()
        got = _noalloc(inst, hip, dev_out)
        print(sc.name, ref, got, flush=True)
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (got, ref)
        assert _noalloc(inst, hip, dev_out) == got          # summed in a fixed order: the same bits again
        inst.close()
    assert hip.hipFree(dev_out) == 0
