"""x**(1/3), x**(5/3) and x**(7/3) with the C library's exact powf results (sfx_libm.hpp).

The reference evaluates hu**expo, hu**(5.0/3.0) and hgate**(7.0/3.0) with the C
library's powf; the
emulation must return the same float for every input the tables cover. Checked here on
the host implementation over every table key and a dense random sample, and on the
device (GPU test) over the same inputs."""
import ctypes as C

import numpy as np
import pytest

from sfincs_amd import _lib

libm = C.CDLL("libm.so.6")
libm.powf.restype = C.c_float
libm.powf.argtypes = [C.c_float, C.c_float]
EXP = [np.float32(1.0) / np.float32(3.0), np.float32(5.0) / np.float32(3.0), np.float32(7.0) / np.float32(3.0)]


def glibc_powf(x, which):
    y = float(EXP[which])
    return np.array([libm.powf(float(v), y) for v in x], np.float32)


def sample(n, seed):
    r = np.random.default_rng(seed)
    # log-uniform over the covered range plus hand-picked mismatching values
    x = np.exp(r.uniform(np.log(2.0 ** -40), np.log(2.0 ** 14), n)).astype(np.float32)
    return np.concatenate([x, np.array([1.6265767e-05, 8.68599415, 0.05, 1.0, 2.0, 10.0], np.float32)])


def emulate(x, which, device=False):
    L = _lib.lib()
    f = L.sfx_libm_pow_device if device else L.sfx_libm_pow_host
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]
    f.restype = C.c_int
    out = np.empty_like(x)
    _lib.check(f(x.ctypes.data, len(x), which, out.ctypes.data), "libm_pow")
    return out


@pytest.mark.parametrize("which", [0, 1, 2])
def test_host_emulation_equals_glibc(which):
    x = sample(200000, 7 + which)
    assert np.array_equal(emulate(x, which).view(np.uint32), glibc_powf(x, which).view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1, 2])
def test_device_emulation_equals_glibc(which):
    x = sample(2000000, 17 + which)
    want = emulate(x, which)            # host emulation, equal to glibc (test above)
    got = emulate(x, which, device=True)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, (bad.size, x[bad[:5]], got[bad[:5]], want[bad[:5]])
