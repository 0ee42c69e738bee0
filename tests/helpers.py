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
    return _lib.Instance(model, fp64=fp64)


def maxdiff(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


def compare_state(inst, orc, fields=("zs", "q", "uv", "zsmax")):
    out = {}
    for f in fields:
        out[f] = maxdiff(inst.get(f), orc.state(f))
    return out


SUBGRID_FIELDS = ("zs0", "zsderv")


def compare_subgrid_state(inst, orc):
    """The subgrid cell state beside zs: the volume (largest difference relative to
    1 + |V|) and the wiggle-suppression levels zs0 / zsderv."""
    va, vb = inst.get("z_volume"), orc.state("z_volume")
    out = {"z_volume_rel": float(np.max(np.abs(va - vb) / (1.0 + np.abs(vb)))) if va.size else 0.0}
    out.update(compare_state(inst, orc, SUBGRID_FIELDS))
    return out
