"""Continuous batching on the GPU: requests that arrive while others run join the batch,
and every result equals the same scenario run alone."""
import numpy as np
import pytest

from helpers import ZS_TOL, gpu_for
from sfincs_amd import ensemble as E
from sfincs_amd import workloads as W

pytestmark = pytest.mark.gpu


def _requests():
    return [(0.0, W.regular_case(30, 36, dx=50.0, seed=51, t_end=1800.0)),
            (0.0, W.subgrid_case(24, 28, dx=25.0, seed=52, t_end=1200.0, nsrc=1, rain_mmh=40.0)),
            (0.02, W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=53, t_end=1500.0)),
            (0.04, W.regular_case(20, 44, dx=50.0, seed=54, t_end=900.0, tide_amp=2.0))]


def test_served_results_equal_alone():
    print("case serving mix of 4", flush=True)
    alone = []
    for _, sc in _requests():
        inst = gpu_for(sc.build_model())
        inst.update(-1000.0)
        alone.append((inst.get("zs"), inst.time()[2]))
        inst.close()
    srv = E.Server(hbm_budget=1 << 34, round_steps=8, workers=2, keep_state=True)
    for a, sc in _requests():
        srv.submit(a, sc)
    out = srv.run()
    assert out["requests"] == 4
    assert out["max_concurrent"] >= 2
    by_seq = {r["seq"]: r for r in out["results"]}
    for k, (zs, nt) in enumerate(alone):
        assert by_seq[k]["steps"] == nt
        assert np.max(np.abs(by_seq[k]["zs"] - zs)) <= ZS_TOL


def test_ensemble_runners_agree():
    """run_batched and run_sequential give the same member states and step counts."""
    print("case ensemble runners", flush=True)
    scs = W.ensemble_members(5, t_end=900.0, nmin=24, nmax=96)
    models = [sc.build_model() for sc in scs]
    t_ends = [sc.t_end for sc in scs]
    a = E.run_sequential(models, t_ends, keep_state=True)
    b = E.run_batched(models, t_ends, round_steps=16, keep_state=True)
    assert a["steps"] == b["steps"]
    for za, zb in zip(a["zs"], b["zs"]):
        assert np.max(np.abs(za - zb)) <= ZS_TOL
