"""Downstream river boundaries (msk=5 cells forced from a bdr file): the host tables
(sfincs_boundaries.f90:217-285, 587-613), the oracle's boundary level
(update_boundary_conditions, :878-901) and the Fortran host's bdr reader.

The reference's boundary module needs the NetCDF library, so it is not compiled here;
the oracle's river branch is pinned by brute-force lookups and by the cases the formula
reduces to: zero slope (the interior level), a large slope (the bed), and the level drop
of slope x distance in between."""
import numpy as np
import pytest

from helpers import oracle_for
from sfincs_amd import _lib
from sfincs_amd import workloads as W


def river_case(slope=1.0e-3, kind="regular"):
    if kind == "regular":
        sc = W.regular_case(30, 40, dx=50.0, seed=91, t_end=1800.0, tide_amp=1.5)
    else:
        sc = W.subgrid_case(30, 40, dx=25.0, seed=92, t_end=1800.0, nsrc=0, tide_amp=1.5)
    return W.with_river(sc, slope=slope)


@pytest.mark.parametrize("kind", ["regular", "subgrid"])
def test_river_tables(kind):
    sc = river_case(kind=kind)
    m = sc.build_model()
    kcs, nmind = m.table("kcs"), m.table("nmindbnd") - 1
    gb, zsi, dzs = m.table("index_bdr_gbp"), m.table("index_zsi_bdr"), m.table("dzs_bdr")
    x, y, xi, yi, slope, dist = sc.bdr
    xz, yz = m.table("z_xz"), m.table("z_yz")
    riv = kcs[nmind] == 5
    assert riv.sum() == 8 and np.all(gb[~riv] == 0)
    for ib in np.flatnonzero(riv):   # nearest bdr point, the first of equals
        c = nmind[ib]
        d = np.sqrt((x - xz[c]) ** 2 + (y - yz[c]) ** 2).astype(np.float32)
        assert gb[ib] == int(np.argmin(d)) + 1
    for j in range(len(x)):          # the cell holding the interior point, an ordinary cell
        c = zsi[j] - 1
        assert kcs[c] == 1 and abs(xz[c] - xi[j]) <= sc.dx / 2 and abs(yz[c] - yi[j]) <= sc.dy / 2
    hyp = np.sqrt((xi - x) ** 2 + (yi - y) ** 2).astype(np.float32)
    assert np.allclose(dzs, -slope * hyp, rtol=1e-6) and np.all(dzs < 0)


def test_river_distance_given():
    sc = river_case()
    x, y, xi, yi, slope, _ = sc.bdr
    sc.bdr = (x, y, xi, yi, slope, np.array([400.0, 1000.0], np.float32))
    m = sc.build_model()
    assert np.array_equal(m.table("dzs_bdr"), (-slope * np.array([400.0, 1000.0], np.float32)).astype(np.float32))


def river_levels(slope, kind="regular"):
    """Oracle state after 600 s, then one forcing call: (river level, interior level, bed)
    per river cell."""
    sc = river_case(slope, kind)
    m, orc = oracle_for(sc)
    orc.update(600.0)
    t, dt = orc.get("t") + 0.5, np.float32(0.5)
    orc.update_forcing(t, dt)
    kcs, nmind = m.table("kcs"), m.table("nmindbnd") - 1
    gb, zsi = m.table("index_bdr_gbp"), m.table("index_zsi_bdr")
    bed = m.table("z_zmin") if m.sizes()["subgrid"] else m.table("zb")
    zs = orc.state("zs")
    out = [(zs[c], zs[zsi[gb[ib] - 1] - 1], bed[c]) for ib, c in enumerate(nmind) if kcs[c] == 5]
    return np.array(out, np.float64)


@pytest.mark.parametrize("kind", ["regular", "subgrid"])
def test_oracle_river_levels(kind):
    flat = river_levels(0.0, kind)                 # no slope: the interior level (as f32)
    wet = flat[:, 1] > flat[:, 2]
    assert wet.any()
    assert np.array_equal(flat[wet, 0], flat[wet, 1].astype(np.float32).astype(np.float64))
    drop = river_levels(1.0e-3, kind)              # 1e-3 over the 150 m (regular) / 75 m between the points
    d = 0.15 if kind == "regular" else 0.075
    wet = drop[:, 1] + (-d) > drop[:, 2]
    assert wet.any()
    assert np.allclose(drop[wet, 1] - drop[wet, 0], d, atol=2e-6)
    steep = river_levels(10.0, kind)               # far below the bed: the bed
    assert np.array_equal(steep[:, 0], steep[:, 2].astype(np.float64))


def test_river_drains_the_domain():
    """With a slope the river cells lie below the interior, so water leaves through them."""
    sc = river_case(1.0e-3)
    m, orc = oracle_for(sc)
    orc.update(900.0)
    k2, nmb = m.table("index_kcuv2") - 1, m.table("nmbkcuv2") - 1
    kcs = m.table("kcs")
    q = orc.state("q")[k2[kcs[nmb] == 5]]
    assert np.all(q < 0.0)                         # southward, out of the domain


def test_river_errors():
    sc = W.regular_case(30, 40, dx=50.0, seed=91, t_end=1800.0)
    msk = sc.msk.reshape(40, 30).copy()
    msk[3:6, 0] = 5
    sc.msk = msk.reshape(-1)
    with pytest.raises(_lib.SfxError):             # msk=5 without a bdr file
        sc.build_model()
    for xi, yi in ((-500.0, 100.0), (25.0, 25.0)):  # outside the mesh; inside a boundary cell
        sc = river_case()
        x, y, _, _, s, dd = sc.bdr
        sc.bdr = (x, y, np.full(2, xi, np.float32), np.full(2, yi, np.float32), s, dd)
        with pytest.raises(_lib.SfxError):
            sc.build_model()
