"""Gridded and spiderweb meteo: the amu/amv/amp/ampr equidistant grids and the spw
spiderweb (read_meteo_data, update_meteo_fields, update_meteo_forcing,
sfincs_meteo.f90:5-245, 450-1419; readers sfincs_spiderweb.f90).

Pins: the ASCII readers against the reference's own readers (sfincs_spiderweb.f90,
compiled into oracle/_ref); the oracle's interpolation to the cells against closed forms
(a uniform wind gives the drag law, a linear pressure field is reproduced exactly while
the reference's bound check leaves its northern band outside, frames blend in time and
block rain holds, a spiderweb with linear radial profiles gives the profiles back with
the merge ring and the 6-hour fade); the per-step forcing by the reference's own
momentum and continuity (RefRun, bit for bit), a pressure gradient's sign and the
boundary pressure correction; the library's fields at every update time against the
oracle's bit for bit; and the GPU against the oracle.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from helpers import compare_state, compare_subgrid_state, gpu_for, oracle_for
from sfincs_amd import deck
from sfincs_amd import workloads as W

F = np.float32
G = 9.81
RHOA, RHOW = 1.25, 1024.0
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs /root/reference)")


def cd_curve(w):
    if w >= 50.0:
        return 0.0015
    if w >= 28.0:
        return 0.0025 + (w - 28.0) * (0.0015 - 0.0025) / 22.0
    return 0.0010 + w * (0.0025 - 0.0010) / 28.0


def cd_bins(v):
    """The drag coefficient of the table bin above v, and of the bins either side (a float
    speed within rounding of a bin edge may land in either)."""
    i = int(v * 10)
    return [cd_curve((k + 1) * 0.1) for k in (i - 1, i, i + 1)]


def basin(nmax=20, mmax=24, dx=50.0, t1=1800.0, rotation=0.0, depth=3.0, **params):
    p = dict(t0=0.0, t1=t1, zsini=0.0)
    p.update(params)
    sc = W.Scenario(name="metbasin", params=p, nmax=nmax, mmax=mmax, dx=dx, dy=dx,
                    msk=np.ones(nmax * mmax, np.int8), zb_grid=np.full((mmax, nmax), -depth, F), t_end=t1)
    sc.rotation = rotation
    return sc


def fields(sc, T):
    """The oracle's fields on the cells at meteo time T (update_meteo_fields)."""
    m = sc.build_model()
    o = O.OracleRun(m).configure(sc.params, m)
    o.set("meteo_t0", T)
    o.set("meteo_t1", T + 1800.0)
    o.L.oracle_meteo_fields(o.m)
    return m, {k: o.keep[k + "0"].copy() for k in ("tauwu", "tauwv", "patm", "prcp")}


def grid_over(sc, nx, ny, margin=2.0):
    """An equidistant grid `margin` nodes wider than the basin on each side (0: its nodes
    from half a spacing inside the basin's west and south edges to half a spacing beyond
    its east and north edges, so that the northern band lies inside the basin)."""
    lx, ly = sc.mmax * sc.dx, sc.nmax * sc.dy
    gdx, gdy = lx / (nx - 2 * margin - 1), ly / (ny - 2 * margin - 1)
    return dict(x_llcorner=-margin * gdx, y_llcorner=-margin * gdy, dx=gdx, dy=gdy)


def node_xy(g, nx, ny):
    xc = g["x_llcorner"] + (np.arange(nx) + 0.5) * g["dx"]
    yc = g["y_llcorner"] + (ny - np.arange(ny) - 0.5) * g["dy"]      # row 0 northernmost
    return np.meshgrid(xc, yc)


def inside(m, g, nx, ny):
    """Cells the reference interpolates (update_amuv_data :898: rows 2 .. nrows-1 of the
    cell's lower-left node, columns 1 .. ncols-1, the indices truncated towards zero as
    Fortran's int does), from the cell centres in double."""
    x, y = m.table("z_xz").astype(np.float64), m.table("z_yz").astype(np.float64)
    yul = g["y_llcorner"] + (ny - 1) * g["dy"] + 0.5 * g["dy"]
    xll = g["x_llcorner"] + 0.5 * g["dx"]
    iy = np.trunc((yul - y) / g["dy"]) + 1
    ix = np.trunc((x - xll) / g["dx"]) + 1
    return (iy > 1) & (iy < ny) & (ix > 0) & (ix < nx)


# ------------------------------------------------------------------------- readers
@needs_ref
def test_readers_equal_the_reference_readers(tmp_path):
    """amu, amp, ampr and spw files as the deck writer writes them (minutes since noon the
    day before tref; hours since tref), read by the reference's read_amuv_file /
    read_spw_file and by sfx_model_read_meteo_file."""
    from sfincs_amd import _lib
    sc = W.with_spiderweb(W.with_gridded_meteo(W.regular_case(24, 30, dx=50.0, seed=5, t_end=1800.0)))
    deck.write(sc, str(tmp_path))
    p = _lib.default_params(**sc.params)
    m = _lib.Model.regular(p, sc.nmax, sc.mmax, sc.dx, sc.dy, sc.msk)
    for kind in ("amu", "amv", "amp", "ampr", "spw"):
        path = os.path.join(str(tmp_path), f"sfincs.{kind}")
        m.read_meteo_file(kind, path)
        ref = O.ref_read_meteo(path, kind)
        g = sc.meteo[kind]
        assert np.array_equal(ref["times"], g["times"]) and np.array_equal(m.table(f"{kind}_times"), g["times"])
        if kind == "spw":
            for k, mine in (("vmag", "vmag"), ("vdir", "vdir"), ("pdrop", "pdrop"), ("prcp", "rain")):
                want = np.asarray(g[mine], F)
                assert np.array_equal(ref[k], want), k
                assert np.array_equal(m.table(f"spw_{'prcp' if k == 'prcp' else k}").reshape(want.shape), want), k
            assert np.array_equal(ref["xe"], g["xe"]) and np.array_equal(m.table("spw_xe"), g["xe"])
            assert np.array_equal(ref["ye"], g["ye"]) and np.array_equal(m.table("spw_ye"), g["ye"])
            assert ref["geo"][4] == F(g["radius"]) == m.table("spw_radius")[0]
        else:
            assert np.array_equal(ref["values"], g["values"])
            assert np.array_equal(m.table(f"{kind}_values").reshape(g["values"].shape), g["values"])
            geo = np.array([g["x_llcorner"], g["y_llcorner"], g["dx"], g["dy"]], F)
            assert np.array_equal(ref["geo"][:4], geo) and np.array_equal(m.table(f"{kind}_geo"), geo)


# -------------------------------------------------------------------- closed forms
@pytest.mark.parametrize("rotation", [0.0, 30.0])
def test_uniform_wind_grid_gives_the_drag_law(rotation):
    """A uniform (u, v) = (12, -5) m/s: inside the grid tau = rhoa Cd(|V|) |V| (u, v) / rhow,
    turned into the grid's axes; outside (its northern band and beyond) no wind."""
    sc = basin(rotation=rotation)
    g = grid_over(sc, 9, 8, margin=0.0)
    t = np.array([0.0, 7200.0], F)
    sc.meteo = {"amu": dict(times=t, values=np.full((2, 8, 9), 12.0, F), **g),
                "amv": dict(times=t, values=np.full((2, 8, 9), -5.0, F), **g)}
    m, f = fields(sc, 900.0)
    ins = inside(m, g, 9, 8)
    assert ins.sum() > 50 and (~ins).sum() > 0
    th = math.radians(rotation)
    u, v = 12.0 * math.cos(th) - 5.0 * math.sin(th), -12.0 * math.sin(th) - 5.0 * math.cos(th)
    for c in np.flatnonzero(ins):
        ok = False
        for cd in cd_bins(13.0):
            s = RHOA * cd * 13.0 / RHOW
            ok |= abs(f["tauwu"][c] - s * u) <= 2e-6 * abs(s * 13) and abs(f["tauwv"][c] - s * v) <= 2e-6 * abs(s * 13)
        assert ok, (c, f["tauwu"][c], f["tauwv"][c])
    assert np.all(f["tauwu"][~ins] == 0.0) and np.all(f["tauwv"][~ins] == 0.0)


def test_linear_pressure_field_is_reproduced_and_its_northern_band_is_outside():
    sc = basin(baro=1)
    nx, ny = 10, 9
    g = grid_over(sc, nx, ny, margin=0.0)
    X, Y = node_xy(g, nx, ny)
    p = (100000.0 + 0.8 * X - 0.5 * Y).astype(F)
    sc.meteo = {"amp": dict(times=np.array([0.0, 7200.0], F), values=np.stack([p, p]), **g)}
    m, f = fields(sc, 600.0)
    ins = inside(m, g, nx, ny)
    x, y = m.table("z_xz").astype(np.float64), m.table("z_yz").astype(np.float64)
    want = 100000.0 + 0.8 * x - 0.5 * y
    assert np.max(np.abs(f["patm"][ins] - want[ins])) < 0.05
    north = (~ins) & (y > y.max() - 2 * sc.dy)
    assert north.sum() > 0 and np.all(f["patm"][~ins] == F(101200.0))


def test_frames_blend_in_time_and_rain_is_held_per_interval():
    sc = basin(baro=1)
    g = grid_over(sc, 8, 8)
    t = np.array([0.0, 3600.0], F)
    sc.meteo = {"amp": dict(times=t, values=np.stack([np.full((8, 8), 100000.0, F), np.full((8, 8), 102000.0, F)]), **g),
                "ampr": dict(times=t, values=np.stack([np.full((8, 8), 10.0, F), np.full((8, 8), 30.0, F)]), **g)}
    m, f = fields(sc, 1800.0)
    ins = inside(m, g, 8, 8)
    rel = lambda a, b: np.max(np.abs(a - b) / abs(b))
    assert rel(f["patm"][ins], 101000.0) < 1e-6
    assert rel(f["prcp"][ins], 10.0 / 3.6e6) < 1e-6                  # block: the interval's first frame
    m, f = fields(sc, 9000.0)                                        # past the last frame: the last
    assert rel(f["patm"][ins], 102000.0) < 1e-6
    sc.params["amprblock"] = 0
    m, f = fields(sc, 900.0)
    assert rel(f["prcp"][ins], 15.0 / 3.6e6) < 1e-6
    assert rel(f["patm"][ins], 100500.0) < 1e-6


def _spw_case(xe, ye, times, V0=30.0, P0=3000.0, R=900.0, nrows=20, ncols=36):
    sc = basin(nmax=30, mmax=30, dx=50.0, baro=1)
    nt = len(times)
    r = (np.arange(1, nrows + 1) * R / nrows)[:, None] + np.zeros((1, ncols))
    vm = np.repeat((V0 * r / R)[None].astype(F), nt, axis=0)
    pd = np.repeat((P0 * (1.0 - r / R))[None].astype(F), nt, axis=0)
    sc.meteo = {"spw": dict(times=np.asarray(times, F), xe=np.asarray(xe, F), ye=np.asarray(ye, F), radius=R,
                            vmag=vm, vdir=np.full_like(vm, 90.0), pdrop=pd, rain=None)}
    return sc


def _check_spw(m, f, xe, ye, fade, V0=30.0, P0=3000.0, R=900.0, nrows=20):
    x, y = m.table("z_xz").astype(np.float64), m.table("z_yz").astype(np.float64)
    d = np.hypot(x - xe, y - ye)
    dr = R / nrows
    ring = (d >= dr * 1.01) & (d <= R - dr * 1.01)
    merge = np.where(d > 0.5 * R, (R - d) / (0.5 * R), 1.0) * fade
    assert ring.sum() > 100
    for c in np.flatnonzero(ring):
        v = V0 * d[c] / R
        ok = any(abs(f["tauwu"][c] + merge[c] * RHOA * cd * v * v / RHOW) <= 1e-4 * RHOA * cd * v * v / RHOW + 1e-9
                 for cd in cd_bins(v))
        assert ok, (c, d[c], f["tauwu"][c])
        assert abs(f["tauwv"][c]) <= 1e-5 * abs(f["tauwu"][c]) + 1e-12
        want_p = (1 - merge[c]) * 101200.0 + merge[c] * (101200.0 - P0 * (1.0 - d[c] / R))
        assert abs(f["patm"][c] - want_p) < 0.1, (c, f["patm"][c], want_p)
    out = d > R * 1.001
    assert out.sum() > 0 and np.all(f["tauwu"][out] == 0.0) and np.all(f["patm"][out] == F(101200.0))


def test_spiderweb_linear_profiles_merge_ring_and_moving_eye():
    """Speed V0 r / R from the east (nautical 90) and pressure drop P0 (1 - r / R): inside
    the spiderweb the stress is -rhoa Cd V^2 / rhow along x and the pressure gapres - drop,
    faded linearly to the background from half the radius outwards; the eye moves linearly
    between its positions."""
    sc = _spw_case([500.0, 1100.0], [700.0, 800.0], [0.0, 7200.0])
    m, f = fields(sc, 3600.0)
    _check_spw(m, f, 800.0, 750.0, 1.0)


def test_spiderweb_fades_in_over_six_hours_before_its_first_time():
    sc = _spw_case([750.0, 750.0], [750.0, 750.0], [5400.0, 12600.0])
    m, f = fields(sc, 0.0)
    _check_spw(m, f, 750.0, 750.0, 1.0 - 5400.0 / 21600.0)


def test_pressure_gradient_drives_flow_towards_low_pressure():
    sc = basin(baro=1, t1=600.0)
    nx, ny = 10, 9
    g = grid_over(sc, nx, ny)
    X, _ = node_xy(g, nx, ny)
    p = (101200.0 + 0.5 * X).astype(F)                      # higher in the east
    sc.meteo = {"amp": dict(times=np.array([0.0, 7200.0], F), values=np.stack([p, p]), **g)}
    m, o = oracle_for(sc)
    o.update(120.0)
    q = o.state("q")[: m.sizes()["npuv"]]
    u = m.table("uv_flags_dir")[: m.sizes()["npuv"]] == 0
    assert np.mean(q[u]) < -1e-4 and abs(np.mean(q[~u])) < 1e-3 * abs(np.mean(q[u]))


def test_boundary_levels_carry_the_pressure_correction():
    """pavbnd > 1: a water-level boundary takes (pavbnd - patm) / (rhow g) on top of its
    series (sfincs_boundaries.f90:820-826)."""
    def zsb(with_low):
        sc = W.regular_case(16, 20, dx=50.0, seed=3, t_end=600.0, rain_mmh=0.0, tide_amp=0.0, pavbnd=101200.0)
        sc.rain = None
        if with_low:
            g = grid_over(sc, 8, 8, margin=2.0)
            sc.meteo = {"amp": dict(times=np.array([0.0, 7200.0], F), values=np.full((2, 8, 8), 100200.0, F), **g)}
        m, o = oracle_for(sc)
        o.update(1e30, 1)
        return m, o.state("zsb").copy()
    m, a = zsb(False)
    _, b = zsb(True)
    d = F(1000.0) / F(F(1024.0) * F(9.81))
    assert a.size > 0 and np.allclose(b - a, d, atol=2e-6), (b - a)[:5]


def test_gridded_and_time_series_wind_together_are_refused():
    sc = basin()
    g = grid_over(sc, 8, 8)
    t = np.array([0.0, 7200.0], F)
    sc.meteo = {"amu": dict(times=t, values=np.ones((2, 8, 8), F), **g),
                "amv": dict(times=t, values=np.ones((2, 8, 8), F), **g)}
    sc.wind = (t, np.array([5.0, 5.0], F), np.array([90.0, 90.0], F))
    with pytest.raises(Exception, match="time series and as gridded"):
        sc.build_model()


# ------------------------------------------------------------ library against oracle
def meteo_case(kind):
    if kind == "regular":
        sc = W.regular_case(36, 44, dx=50.0, seed=95, t_end=1800.0, rain_mmh=0.0, tide_amp=1.0, pavbnd=101200.0,
                            dtwnd=600.0)
        sc.rain = None
        return W.with_gridded_meteo(sc, t_step=900.0)
    if kind == "subgrid":
        sc = W.subgrid_case(30, 36, dx=25.0, seed=96, t_end=1800.0, nsrc=1, spinup_meteo=1, tspinup=600.0)
        W.with_gridded_meteo(sc, wind=False, pressure=True, rain=True, rain_mmh=80.0)
        return W.with_infiltration(sc, qinf=25.0, qinf_zmin=-100.0)
    sc = W.quadtree_case(10, 14, nlev=3, dx0=100.0, seed=97, t_end=1800.0, surge_amp=2.0, wind=False,
                         factor_wind=1.2, spw_merge_frac=0.6)
    return W.with_spiderweb(sc, t_step=1200.0)


def test_library_fields_equal_the_oracle_at_every_update_time():
    for kind in ("regular", "subgrid", "quadtree"):
        sc = meteo_case(kind)
        m = sc.build_model()
        o = O.OracleRun(m).configure(sc.params, m)
        T = m.table("meteo_times")
        n = m.sizes()["np"]
        frames = {k: m.table("meteo_" + k) for k in ("tauwu", "tauwv", "patm", "prcp")}
        assert len(T) >= 4
        for k in range(len(T) - 1):
            o.set("meteo_t0", T[k])
            o.set("meteo_t1", T[k + 1])
            o.L.oracle_meteo_fields(o.m)
            for name, fr in frames.items():
                if fr.size == 0:
                    continue
                fr = fr.reshape(len(T), n)
                assert np.array_equal(o.keep[name + "0"], fr[k]) and np.array_equal(o.keep[name + "1"], fr[k + 1]), \
                    (kind, k, name)


@needs_ref
@pytest.mark.parametrize("kind", ["regular", "subgrid", "quadtree"])
def test_meteo_oracle_bit_exact_with_reference(kind):
    """The oracle's steps with per-cell wind stress, pressure (its gradient in momentum) and
    rain against the reference's own momentum and continuity, 250 steps."""
    sc = meteo_case(kind)
    m = sc.build_model()
    a = O.OracleRun(m).configure(sc.params, m)
    b = O.OracleRun(m).configure(sc.params, m)
    assert a.vals["meteo3d"] == 1
    r = O.RefRun(b)
    fields_ = ["zs", "q", "uv", "kfuv", "zsmax", "netprcp", "tauwu", "patm"] + (["z_volume"] if kind == "subgrid" else [])
    for i in range(250):
        a.update(1e30, 1)
        r.step()
        for f in fields_:
            assert np.array_equal(a.state(f), b.state(f)), (i, f)
    if a.vals["patmos"]:
        assert np.ptp(a.state("patm")) > 100.0


# ----------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["regular", "subgrid", "quadtree"])
def test_gridded_meteo_matches_oracle(kind):
    """Gridded wind, pressure and rain (regular, with the boundary pressure correction and
    windows every 600 s), pressure and rain with meteo spin-up and infiltration (subgrid),
    a spiderweb (quadtree): GPU and oracle bit for bit over 30 minutes, and the volume
    balance closes with the gridded rain."""
    print("case gridded meteo", kind, flush=True)
    sc = meteo_case(kind)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    b0 = inst.balance()
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc)
    if m.sizes()["subgrid"]:
        d.update(compare_subgrid_state(inst, orc))
    print(sc.name, st["nt"], int(orc.get("nt")), d, flush=True)
    assert rc == 0 and st["nt"] == int(orc.get("nt")) and abs(st["t"] - orc.get("t")) < 1e-6
    for f in ("zs", "q", "uv", "zsmax") + (("z_volume",) if m.sizes()["subgrid"] else ()):
        assert np.array_equal(inst.get(f), orc.state(f)), f
    b1 = inst.balance()
    dv = b1["volume"] - b0["volume"] + (b0["storage"] - b1["storage"])
    inflow = sum(b1[k] - b0[k] for k in ("boundary", "sources", "rain", "qext"))
    kcs = m.table("kcs")
    if m.sizes()["subgrid"]:
        water = float(np.sum(inst.get("z_volume")[kcs == 1]))
    else:
        area = m.table("cell_area")[m.table("z_flags_iref") - 1].astype(np.float64)
        water = float(np.sum((np.maximum(inst.get("zs") - m.table("zb"), 0.0) * area)[kcs == 1]))
    assert abs(dv - inflow) <= 1e-6 * water, (dv, inflow, water)
    if kind != "quadtree":
        assert b1["rain"] - b0["rain"] != 0.0
