"""Drainage structures (drnfile; sfincs_discharges.f90:176-290, 371-655).

The reference's discharges module needs NetCDF and cannot be compiled here, so the
oracle's structure update is pinned by what the formulas fix, driving
oracle_update_discharges with the levels held:
* a pump delivers Q through the relaxation q_n = a Q + (1 - a) q_(n-1), a = dt / relax,
  i.e. Q (1 - (1 - a)^n), and cannot take more than its intake cell holds;
* a culvert tends to c sqrt(dz) in the direction of the fall; a check valve only forward;
* a gate whose intake level leaves [zmin, zmax] closes by dt / tclose per step, after which
  its flow decays as (1 - a)^n;
* an open gate at held levels tends to Manning's flow through its width,
  w h^(5/3) sqrt(|s|) / n, the fixed point of the Bates et al. (2010) formula;
* a time-controlled gate is closed between t_close and t_open.
The intake loses what the outfall gains. Continuity treats the structures' points as
point sources: the oracle against the reference's own compute_water_levels, bit for bit.
"""
import math

import numpy as np
import pytest

import oracle as O
from sfincs_amd import workloads as W

G = 9.81


def _basin(structures, depth=4.0, nmax=10, mmax=12, dx=50.0, zb=None, **params):
    zb_grid = np.full((mmax, nmax), -depth, np.float32) if zb is None else zb
    p = dict(t0=0.0, t1=3600.0)
    p.update(params)
    sc = W.Scenario(name="drn_basin", params=p, nmax=nmax, mmax=mmax, dx=dx, dy=dx,
                    msk=np.ones(nmax * mmax, np.int8), zb_grid=zb_grid, t_end=3600.0)
    return W.with_drainage(sc, structures)


def _run(sc):
    m = sc.build_model()
    return m, O.OracleRun(m).configure(sc.params, m)


def _cells(m):
    k = m.table("nmindsrc") - 1
    return k[0::2], k[1::2]


def _step(orc, n, dt, t0=100.0, levels=None):
    zs = orc.state("zs")
    for i in range(n):
        if levels is not None:
            for c, v in levels.items():
                zs[c] = v
        orc.L.oracle_update_discharges(orc.m, t0 + i * dt, np.float32(dt))
    return orc.state("qtsrc")


def test_pump_relaxes_to_its_capacity():
    sc = _basin([(75.0, 75.0, 425.0, 425.0, 1, [3.0])])
    m, orc = _run(sc)
    dt, n = 2.0, 40
    q = _step(orc, n, dt)
    a = dt / 10.0
    expect = 3.0 * (1.0 - (1.0 - a) ** n)
    assert abs(q[1] - expect) < 1e-5 and q[0] == -q[1]
    assert m.sizes()["ndrn"] == 1 and len(m.table("nmindsrc")) == 2
    assert abs(m.table("drainage_distance")[0] - math.hypot(350.0, 350.0)) < 1e-3


def test_pump_cannot_empty_its_intake_below_the_bed():
    zb = np.full((12, 10), -4.0, np.float32)
    zb[1, 1] = -0.0002                    # the intake cell holds 0.2 mm of water
    sc = _basin([(75.0, 75.0, 425.0, 425.0, 1, [3.0])], zb=zb)
    m, orc = _run(sc)
    cin, cout = _cells(m)
    q = _step(orc, 5, 1.0)
    cap = 0.0002 * 2500.0 / 1.0           # (zs - zb) A / dt, below the relaxed pump rate
    assert abs(q[1] - cap) < 1e-4 * cap


@pytest.mark.parametrize("kind", [2, 3])
def test_culvert_and_check_valve(kind):
    sc = _basin([(75.0, 75.0, 425.0, 425.0, kind, [2.0])])
    m, orc = _run(sc)
    cin, cout = _cells(m)
    q = _step(orc, 300, 1.0, levels={cin[0]: 0.9, cout[0]: 0.1})
    assert abs(q[1] - 2.0 * math.sqrt(0.8)) < 1e-4
    q = _step(orc, 300, 1.0, levels={cin[0]: 0.1, cout[0]: 0.5})
    if kind == 2:
        assert abs(q[1] + 2.0 * math.sqrt(0.4)) < 1e-4          # backwards
    else:
        assert abs(q[1]) < 1e-5                                  # the valve does not open backwards


def test_open_gate_tends_to_manning_flow():
    w, n, sill = 3.0, 0.03, -3.0
    sc = _basin([(75.0, 75.0, 425.0, 425.0, 4, [w, sill, n, -5.0, 5.0, 300.0])])
    m, orc = _run(sc)
    cin, cout = _cells(m)
    L = float(m.table("drainage_distance")[0])
    zi, zo = 0.4, 0.1
    q = _step(orc, 30000, 1.0, levels={cin[0]: zi, cout[0]: zo})
    h = max(zi, zo) - sill
    s = (zo - zi) / L
    manning = w * h ** (5.0 / 3.0) * math.sqrt(abs(s)) / n
    assert abs(q[1] - manning) < 2e-4 * manning, (q[1], manning)


def test_gate_closes_when_the_intake_leaves_its_range():
    tcl = 100.0
    sc = _basin([(75.0, 75.0, 425.0, 425.0, 4, [3.0, -3.0, 0.03, -1.0, 0.3, tcl])])
    m, orc = _run(sc)
    cin, cout = _cells(m)
    q_open = _step(orc, 400, 1.0, levels={cin[0]: 0.25, cout[0]: 0.0})[1]
    assert q_open > 0.1 and orc.state("drainage_status")[0] == 1
    _step(orc, 50, 1.0, levels={cin[0]: 0.5, cout[0]: 0.0})   # above zmax: closing
    assert orc.state("drainage_status")[0] == 2
    assert abs(orc.state("drainage_fraction_open")[0] - (1.0 - 50 * 1.0 / tcl)) < 1e-5
    _step(orc, 60, 1.0, levels={cin[0]: 0.5, cout[0]: 0.0})
    assert orc.state("drainage_status")[0] == 0 and orc.state("drainage_fraction_open")[0] == 0.0
    q0 = float(orc.state("qtsrc")[1])
    q = _step(orc, 20, 1.0, levels={cin[0]: 0.5, cout[0]: 0.0})
    assert abs(q[1] - q0 * (1.0 - 0.1) ** 20) < 1e-5 * max(1.0, abs(q0))


def test_timed_gate_is_closed_between_its_times():
    sc = _basin([(75.0, 75.0, 425.0, 425.0, 5, [3.0, -3.0, 0.03, 200.0, 400.0, 10.0])])
    m, orc = _run(sc)
    cin, cout = _cells(m)
    zs = orc.state("zs")
    st = []
    for i in range(500):
        zs[cin[0]], zs[cout[0]] = 0.3, 0.0
        orc.L.oracle_update_discharges(orc.m, float(i), np.float32(1.0))
        st.append(int(orc.state("drainage_status")[0]))
    assert st[150] == 1 and st[205] == 2 and st[250] == 0 and st[350] == 0 and st[402] == 3 and st[450] == 1
    assert abs(orc.state("qtsrc")[1]) > 0.1


def subgrid_drainage_case():
    """A subgrid tidal domain with a pump, a culvert and a gate from wet cells (lowest
    level below -2 m) to cells inland."""
    sc = W.subgrid_case(30, 44, dx=25.0, seed=72, t_end=1800.0, nsrc=1, tide_amp=2.0)
    m = sc.build_model()
    zmin = m.table("z_zmin")
    xz, yz = m.table("z_xz"), m.table("z_yz")
    wet = np.nonzero(zmin < -2.0)[0]
    dry = np.nonzero(zmin > 1.0)[0]
    pick = lambda cells, k: (float(xz[cells[k]]), float(yz[cells[k]]))
    st = [(*pick(wet, 5), *pick(dry, 7), 1, [1.5]),
          (*pick(wet, len(wet) // 2), *pick(dry, len(dry) // 2), 2, [3.0]),
          (*pick(wet, -3), *pick(dry, -5), 4, [3.0, -2.0, 0.03, -5.0, 0.8, 200.0])]
    return W.with_drainage(sc, st)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("kind", ["regular", "subgrid"])
def test_drainage_in_continuity_bit_exact_with_reference(kind):
    """The structures' intake and outfall points in compute_water_levels (nsrcdrn,
    sfincs_continuity.f90:87-106, 311-325): oracle and reference after every step."""
    if kind == "regular":
        sc = W.beach_drainage(W.regular_case(40, 48, dx=50.0, seed=71, t_end=1800.0, tide_amp=2.5, rain_mmh=0.0))
    else:
        sc = subgrid_drainage_case()
    m = sc.build_model()
    a = O.OracleRun(m).configure(sc.params, m)
    b = O.OracleRun(m).configure(sc.params, m)
    r = O.RefRun(b)
    fields = ["zs", "q", "uv", "kfuv", "zsmax"] + (["z_volume"] if kind == "subgrid" else [])
    for i in range(250):
        a.update(1e30, 1)
        r.step()
        for f in fields:
            assert np.array_equal(a.state(f), b.state(f)), (i, f)
    assert np.count_nonzero(a.state("qtsrc")[-6:]) >= 2
