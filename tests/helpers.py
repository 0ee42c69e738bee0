"""Shared helpers of the parity tests: run the GPU path and the oracle on one scenario."""
import numpy as np

import oracle as O

# Tolerances of the parity gate (after the simulated period, FP32 arithmetic):
ZS_TOL = 1e-4       # m
Q_TOL = 1e-5        # m2/s


def oracle_for(sc, model=None):
    m = model or sc.build_model()
    o = O.OracleRun(m).configure(sc.params, m)
    return m, o


def gpu_for(model, fp64=False):
    from sfincs_amd import _lib
    if fp64:
        import ctypes as C
        L = _lib.lib()
        out = C.c_void_p()
        f = L.sfx_create_ex
        f.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
        f.restype = C.c_int
        _lib.check(f(model.h, 0, 1, C.byref(out)), "sfx_create_ex")
        inst = _lib.Instance.__new__(_lib.Instance)
        inst.h = out
        inst.sizes = model.sizes()
        return inst
    return _lib.Instance(model)


def maxdiff(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


def compare_state(inst, orc, fields=("zs", "q", "uv", "zsmax")):
    out = {}
    for f in fields:
        out[f] = maxdiff(inst.get(f), orc.state(f))
    return out
