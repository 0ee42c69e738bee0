"""Incoming infragravity levels at water-level boundaries (bzi file; the zig term of
update_boundary_conditions, sfincs_boundaries.f90:748-753, 829-866).

The reference's boundary module needs the NetCDF library, so it is not compiled here;
the oracle's zig is pinned by the cases it reduces to: zero levels change nothing, a
constant level c separates the total boundary level from the still level by c, and a
series equal to the water-level series doubles the boundary level (the same time and
space interpolation, summed exactly in f32)."""
import numpy as np
import pytest

from helpers import oracle_for
from sfincs_amd import workloads as W


def case(nbnd=1):
    sc = W.regular_case(24, 30, dx=50.0, seed=95, t_end=1800.0, tide_amp=1.0, rain_mmh=0.0)
    if nbnd > 1:   # three boundary points along the west edge, different tides
        x, y, t, zs = sc.bnd
        sc.bnd = (np.full(3, x[0], np.float32), np.array([0.0, 600.0, 1200.0], np.float32), t,
                  np.concatenate([zs, 0.8 * zs, 1.2 * zs], 1).astype(np.float32))
    return sc


def boundary_levels(sc, t_stop=900.0):
    m, orc = oracle_for(sc)
    orc.update(t_stop)
    return m, orc, orc.state("zsb").copy(), orc.state("zsb0").copy()


@pytest.mark.parametrize("nbnd", [1, 3])
def test_zero_ig_changes_nothing(nbnd):
    a = case(nbnd)
    b = case(nbnd)
    b.bzi = np.zeros_like(b.bnd[3])
    ma, oa, _, _ = boundary_levels(a)
    mb, ob, _, _ = boundary_levels(b)
    assert len(mb.table("zsi_bnd")) == b.bzi.size
    for n in ("zs", "q", "uv"):
        assert np.array_equal(oa.state(n), ob.state(n)), n


def test_constant_ig_offsets_the_total_level():
    sc = case(1)
    sc.bzi = np.full_like(sc.bnd[3], 0.25)
    m, orc, zsb, zsb0 = boundary_levels(sc)
    kcs, nmind = m.table("kcs"), m.table("nmindbnd") - 1
    wl = kcs[nmind] == 2
    assert wl.any()
    above = zsb0[wl] > m.table("zb")[nmind[wl]]        # not clamped to the bed
    assert above.any()
    d = (zsb[wl] - zsb0[wl])[above]
    assert np.allclose(d, 0.25, atol=1e-6)


@pytest.mark.parametrize("nbnd", [1, 3])
def test_ig_equal_to_the_tide_doubles_the_level(nbnd):
    sc = case(nbnd)
    sc.bzi = sc.bnd[3].copy()
    m, orc, zsb, zsb0 = boundary_levels(sc)
    kcs, nmind = m.table("kcs"), m.table("nmindbnd") - 1
    wl = kcs[nmind] == 2
    above = (zsb0[wl] > m.table("zb")[nmind[wl]]) & (2 * zsb0[wl] > m.table("zb")[nmind[wl]])
    assert above.any()
    assert np.array_equal(zsb[wl][above], (np.float32(2) * zsb0[wl][above]).astype(np.float32))


def test_ig_waves_reach_the_domain():
    """An IG wave at the boundary moves water through the boundary points."""
    a = case(1)
    b = W.with_ig_waves(case(1), amp=0.3, period=100.0)
    _, oa, _, _ = boundary_levels(a, 600.0)
    _, ob, _, _ = boundary_levels(b, 600.0)
    assert not np.array_equal(oa.state("q"), ob.state("q"))
