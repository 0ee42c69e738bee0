"""The FP64 volume balance on the GPU: between two reads, the interior volume change plus
the storage used equals the volume that entered through boundary faces, point sources,
rain and qext (sfx_get_balance), to 1e-6 of the water volume, on regular, subgrid and
quadtree meshes with every inflow kind; the configuration-1 recipe over its whole 6 h."""
import numpy as np
import pytest

from helpers import gpu_for
from sfincs_amd import workloads as W

pytestmark = pytest.mark.gpu

REL = 1e-6


def _water_volume(model, inst):
    """Water volume of the interior cells from the state: V (subgrid) or (zs - zb) A."""
    kcs = model.table("kcs")
    inner = kcs == 1
    if model.sizes()["subgrid"]:
        return float(np.sum(inst.get("z_volume")[inner]))
    area = model.table("cell_area")[model.table("z_flags_iref") - 1].astype(np.float64)
    h = np.maximum(inst.get("zs") - model.table("zb"), 0.0)
    return float(np.sum((h * area)[inner]))


def _check(sc, qext=None, steps=None):
    m = sc.build_model()
    inst = gpu_for(m)
    b0 = inst.balance()
    if qext is not None:
        inst.set("qext", qext)
    if steps is None:
        inst.update(-1000.0)
    else:
        inst.update(float(steps))
    b1 = inst.balance()
    dv = b1["volume"] - b0["volume"] + (b0["storage"] - b1["storage"])
    inflow = sum(b1[k] - b0[k] for k in ("boundary", "sources", "rain", "qext"))
    vw = _water_volume(m, inst)
    err = abs(dv - inflow) / vw
    print(sc.name, {k: b1[k] - b0[k] for k in b1}, "water", vw, "rel err", err, flush=True)
    assert err <= REL, (dv, inflow, vw)
    return b1, b0


def test_regular_tide_and_rain():
    b1, b0 = _check(W.regular_case(48, 56, dx=50.0, seed=81, t_end=3600.0, tide_amp=2.0, rain_mmh=40.0))
    assert abs(b1["boundary"]) > 0 and b1["rain"] > 0


def test_regular_sources_with_bed_clamp():
    sc = W.regular_sources(W.regular_case(40, 52, dx=50.0, seed=28, t_end=3600.0, tide_amp=1.5, rain_mmh=0.0))
    b1, _ = _check(sc)
    assert b1["sources"] != 0.0


def test_subgrid_sources_rain_and_storage():
    sc = W.subgrid_case(40, 48, dx=25.0, seed=82, t_end=3600.0, nsrc=3, rain_mmh=30.0)
    sc.storage = lambda model: np.full(model.sizes()["np"], 50.0, np.float32)
    b1, b0 = _check(sc)
    assert b1["sources"] > 0 and b1["rain"] > 0 and b1["storage"] < b0["storage"]


def test_quadtree_surge_and_neumann():
    sc = W.quadtree_case(12, 14, nlev=3, dx0=100.0, seed=83, t_end=3600.0, surge_amp=3.0, wind=False)
    b1, _ = _check(sc)
    assert b1["boundary"] > 0


def test_neumann_side_boundary():
    sc = W.regular_case(36, 44, dx=50.0, seed=63, t_end=1800.0, tide_amp=2.0)
    msk = sc.msk.reshape(44, 36).copy()
    msk[1:, 35] = 6
    sc.msk = msk.reshape(-1)
    _check(sc)


def test_qext_inflow():
    sc = W.subgrid_case(30, 34, dx=25.0, seed=84, t_end=1800.0, nsrc=0, use_qext=1)
    m = sc.build_model()
    qx = np.zeros(m.sizes()["np"], np.float32)
    qx[::37] = 2.0e-5
    b1, _ = _check(sc, qext=qx)
    assert b1["qext"] > 0


def test_config1_recipe_whole_period():
    """The configuration-1 recipe (fBm subgrid DEM, 8 rivers, 1.5 m tide) at 512 x 512,
    the whole 6 h."""
    sc = W.subgrid_case(512, 512, dx=25.0, refi=5, nl=10, seed=1001, t_end=21600.0, tide_amp=1.5, nsrc=8)
    b1, _ = _check(sc)
    assert b1["boundary"] != 0 and b1["sources"] > 0
