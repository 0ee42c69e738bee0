"""x**(1/3), x**(5/3), x**(7/3) and x**1.5 with the C library's exact powf results, and
cos and sin with its exact cosf and sinf results (sfx_libm.hpp).

The reference evaluates hu**expo, hu**(5.0/3.0), hgate**(7.0/3.0) and the weir's h1**1.5
with the C library's powf, and the uniform wind's cos(vdir), sin(vdir) with its cosf and
sinf (vdir in [0, 4 pi]); the
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
for _f in (libm.cosf, libm.sinf):
    _f.restype = C.c_float
    _f.argtypes = [C.c_float]
EXP = [np.float32(1.0) / np.float32(3.0), np.float32(5.0) / np.float32(3.0), np.float32(7.0) / np.float32(3.0),
       np.float32(1.5)]


def glibc_powf(x, which):
    if which == 4:
        return np.array([libm.cosf(float(v)) for v in x], np.float32)
    if which == 5:
        return np.array([libm.sinf(float(v)) for v in x], np.float32)
    y = float(EXP[which])
    return np.array([libm.powf(float(v), y) for v in x], np.float32)


def angles(n, seed):
    """Directions as the wind update makes them: uniform on [0, 4 pi], plus values next to
    multiples of pi/2 (where cos or sin is small and the argument reduction matters),
    zeros of both signs and the largest float below 4 pi + 0.03."""
    r = np.random.default_rng(seed)
    x = r.uniform(0.0, 4 * np.pi, n).astype(np.float32)
    k = (np.arange(1, 9) * np.float64(np.pi) / 2).astype(np.float32)
    near = np.concatenate([np.nextafter(k, np.float32(np.inf)), np.nextafter(k, np.float32(0)), k])
    return np.concatenate([x, near, np.array([0.0, -0.0, 1e-30, 1e-8, np.nextafter(np.float32(12.6), np.float32(0))],
                                             np.float32)])


def exact_midpoints():
    """x = a^2 4^k with a odd in [257, 321]: x**1.5 = a^3 8^k has 25 significant bits and
    its last one set, an exact midpoint between two floats, which powf must round (k <= -2
    keeps x inside the tables' range, below 2^14)."""
    a = np.arange(257, 322, 2, dtype=np.float64)
    k = np.arange(-12, -1, dtype=np.float64)
    x = (a[:, None] ** 2 * 4.0 ** k[None, :]).reshape(-1)
    assert np.all((a ** 3 >= 2 ** 24) & (a ** 3 < 2 ** 25))
    return x.astype(np.float32)


def sample(n, seed):
    r = np.random.default_rng(seed)
    # log-uniform over the covered range plus hand-picked mismatching values
    x = np.exp(r.uniform(np.log(2.0 ** -40), np.log(2.0 ** 14), n)).astype(np.float32)
    squares = (np.arange(1, 1000, dtype=np.float32) ** 2 / np.float32(64.0)).astype(np.float32)
    return np.concatenate([x, np.array([1.6265767e-05, 8.68599415, 0.05, 1.0, 2.0, 10.0], np.float32),
                           exact_midpoints(), squares])


def emulate(x, which, device=False):
    L = _lib.lib()
    f = L.sfx_libm_pow_device if device else L.sfx_libm_pow_host
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]
    f.restype = C.c_int
    out = np.empty_like(x)
    _lib.check(f(x.ctypes.data, len(x), which, out.ctypes.data), "libm_pow")
    return out


@pytest.mark.parametrize("which", [0, 1, 2, 3, 4, 5])
def test_host_emulation_equals_glibc(which):
    x = sample(200000, 7 + which) if which < 4 else angles(200000, 7 + which)
    assert np.array_equal(emulate(x, which).view(np.uint32), glibc_powf(x, which).view(np.uint32))


def test_midpoints_are_decided_by_the_table():
    """Every exact midpoint of x**1.5 is a table entry, and powf rounds them to even."""
    x = exact_midpoints()
    g = glibc_powf(x, 3)
    exact = (x.astype(np.float64) ** 1.5)
    lo = exact.astype(np.float32)
    assert np.array_equal(emulate(x, 3).view(np.uint32), g.view(np.uint32))
    assert np.all(np.abs(g.astype(np.float64) - exact) == np.abs(lo.astype(np.float64) - exact))


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1, 2, 3, 4, 5])
def test_device_emulation_equals_glibc(which):
    x = sample(2000000, 17 + which) if which < 4 else angles(2000000, 17 + which)
    want = emulate(x, which)            # host emulation, equal to glibc (test above)
    got = emulate(x, which, device=True)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, (bad.size, x[bad[:5]], got[bad[:5]], want[bad[:5]])


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1, 2])
def test_device_estimate_is_within_its_bound_for_every_table_input(which):
    """The kernels' powf estimate (fused multiply-adds, hardware log2 of the mantissa) against
    pow_estimate, which the table build verifies within 2^-46 (2^-44 for 7/3) of pow for every
    float of [2^-40, 2^14): within 2^-44 of it for every one of those floats, so within 2^-43
    of the exact power, far inside the 2^-40 the rounding decisions need."""
    L = _lib.lib()
    f = L.sfx_libm_estimate_check
    f.argtypes = [C.c_int32, C.c_void_p]
    f.restype = C.c_int
    out = np.zeros(2, np.uint64)
    _lib.check(f(which, out.ctypes.data), "libm_estimate_check")
    worst = out[1:2].view(np.float64)[0]
    print("estimate", which, "beyond 2^-44:", int(out[0]), "largest relative difference", worst, flush=True)
    assert out[0] == 0 and worst < 2.0 ** -44
