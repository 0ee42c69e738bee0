"""Thin dams and weirs (thdfile, weirfile; sfincs_structures.f90).

The structures module needs the quadtree module, which needs NetCDF, so it cannot be
compiled here. What it does is pinned by geometry and closed forms instead:
* the uv-points a polyline closes or crosses (find_uv_points_intersected_by_polyline,
  sfincs_quadtree.F90:853-1010) equal a brute-force search over every uv-point with exact
  rational orientation tests, on regular, rotated and quadtree meshes (the inputs keep
  every orientation far from the 1e-6 dead band of geometry.f90:124-139);
* a weir's crest is the polyline's crest interpolated at the uv-point, its length the
  uv-point's face projected onto the polyline (a straight north-south weir: 1 grid
  spacing; at 45 degrees: cos 45); a crest at or below the bed makes no structure;
* the weir discharge (:592-697) against the broad-crested weir formula in double:
  Cd 1.7049 h1^1.5 L free, Cd h2 sqrt(2 g (h1 - h2)) L submerged, 0 below the crest,
  signed by the direction of the fall;
* a thin dam across a closed basin lets no water through;
* the oracle's step with weirs and thin dams against the reference's own
  compute_fluxes / compute_water_levels (the weir routine restated between them, as the
  reference's loop calls it, sfincs_lib.f90:598-602), bit for bit after every step, on
  regular, subgrid and quadtree meshes (where combined uv-points next to a weir average
  its last step's discharge, sfincs_momentum.f90:757-779).
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from sfincs_amd import _lib
from sfincs_amd import workloads as W

F = np.float32


def _orient(a, b, c):
    ax, ay = map(Fraction, map(float, a))
    bx, by = map(Fraction, map(float, b))
    cx, cy = map(Fraction, map(float, c))
    return (by - ay) * (cx - bx) - (bx - ax) * (cy - by)


def brute_force_crossed(m, polylines):
    """Every uv-point whose cell-centre connection properly crosses a polyline segment,
    with the smallest |orientation| met (to show the case is well conditioned)."""
    nm = m.table("uv_index_z_nm") - 1
    nmu = m.table("uv_index_z_nmu") - 1
    xz, yz = m.table("z_xz"), m.table("z_yz")
    out, worst = set(), np.inf
    for pl in polylines:
        pl = np.asarray(pl, F)
        for k in range(len(pl) - 1):
            a, b = pl[k, :2], pl[k + 1, :2]
            lo, hi = np.minimum(a, b), np.maximum(a, b)
            for ip in range(len(nm)):
                c, d = (xz[nm[ip]], yz[nm[ip]]), (xz[nmu[ip]], yz[nmu[ip]])
                if max(c[0], d[0]) < lo[0] or min(c[0], d[0]) > hi[0] or \
                        max(c[1], d[1]) < lo[1] or min(c[1], d[1]) > hi[1]:
                    continue
                o = [_orient(a, b, c), _orient(a, b, d), _orient(c, d, a), _orient(c, d, b)]
                worst = min(worst, min(abs(float(v)) for v in o))
                if (o[0] > 0) != (o[1] > 0) and (o[2] > 0) != (o[3] > 0):
                    out.add(ip + 1)
    return out, worst


def _regular_basin(nmax=16, mmax=20, dx=10.0, rotation=0.0, zb=-2.0):
    p = _lib.default_params(t0=0.0, t1=600.0)
    m = _lib.Model.regular(p, nmax, mmax, dx, dx, np.ones(nmax * mmax, np.int8), x0=0.0, y0=0.0,
                           rotation=rotation)
    return m, np.full(nmax * mmax, zb, F)


DAMS = [[(23.7, 11.3), (61.2, 97.9), (143.1, 121.7)], [(171.3, 3.9), (118.6, 52.1)]]


@pytest.mark.parametrize("rotation", [0.0, 30.0])
def test_thin_dam_uv_points_equal_brute_force(rotation):
    m, zb = _regular_basin(rotation=rotation)
    if rotation:
        # rotate the dams with the grid about its origin
        c, s = np.cos(np.radians(rotation)), np.sin(np.radians(rotation))
        dams = [[(c * x - s * y, s * x + c * y) for x, y in d] for d in DAMS]
    else:
        dams = DAMS
    m.set_bathymetry(zb)
    m.set_thin_dams([np.array(d, F) for d in dams])
    m.initialize()
    want, worst = brute_force_crossed(m, dams)
    assert worst > 1.0
    got = set(m.table("thindam_uv_index").tolist())
    assert got == want and len(got) > 10
    kc = m.table("kcuv")
    assert np.all(kc[np.array(sorted(got)) - 1] == 0)
    assert m.sizes()["nthd"] == len(got)


def test_thin_dam_uv_points_on_a_quadtree_mesh_equal_brute_force():
    sc = W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=29, t_end=600.0, surge_amp=1.0)
    dams = [[(131.7, 23.9), (412.3, 377.1), (871.9, 612.3)], [(950.3, 80.7), (620.1, 540.9)]]
    W.with_thin_dams(sc, dams)
    m = sc.build_model()
    assert m.sizes()["nref"] == 3
    want, worst = brute_force_crossed(m, dams)
    assert worst > 1.0
    got = set(m.table("thindam_uv_index").tolist())
    assert got == want and len(got) > 10
    iref = m.table("uv_flags_iref")[np.array(sorted(got)) - 1]
    assert len(set(iref.tolist())) >= 2            # dams cross more than one level


def test_weir_tables_straight_and_diagonal():
    m, zb = _regular_basin(nmax=16, mmax=20, dx=10.0)
    zb = zb.reshape(20, 16)
    zb[15:, 4:] = 1.5                                # x > 150 m, y > 40 m: above the diagonal's crest
    m.set_bathymetry(zb.reshape(-1))
    ly = 160.0
    straight = np.array([(73.1, -5.0, 0.5, 0.6), (73.1, ly + 5.0, 2.5, 0.8)], F)
    diag = np.array([(123.3, 2.7, 1.0, 0.7), (193.3, 72.7, 1.0, 0.7)], F)
    m.set_weirs([straight, diag])
    m.initialize()
    ip = m.table("structure_uv_index") - 1
    want, _ = brute_force_crossed(m, [straight, diag])
    nm = m.table("uv_index_z_nm")[ip] - 1
    xz, yz = m.table("z_xz")[nm], m.table("z_yz")[nm]
    d = m.table("uv_flags_dir")[ip]
    h, cd, ln = m.table("structure_height"), m.table("structure_cd"), m.table("structure_length")
    on_straight = np.abs(xz - 65.0) < 1e-3           # u-points from column 6 to 7 (x = 65 to 75)
    assert np.all(d[on_straight] == 0)
    ys = yz[on_straight].astype(np.float64)
    # crest and Cd weighted by the distances of the face centre (x = 70) to the two
    # vertices: w = d1 / (d1 + d2)
    d1 = np.hypot(70.0 - 73.1, ys + 5.0)
    d2 = np.hypot(70.0 - 73.1, ly + 5.0 - ys)
    t = d1 / (d1 + d2)
    assert np.allclose(h[on_straight], 0.5 + 2.0 * t, atol=2e-5)
    assert np.allclose(cd[on_straight], 0.6 + 0.2 * t, atol=2e-6)
    assert np.allclose(ln[on_straight], 1.0, atol=1e-4)
    # crest above the bed everywhere on the straight weir: every crossed u-point is one
    assert on_straight.sum() == 16
    # the diagonal: faces projected onto a 45 degree line, where its crest (1.0) is above
    # the bed; crossed points on the raised corner (bed 1.5) are no structure
    diag_pts = ~on_straight
    assert np.allclose(ln[diag_pts], np.cos(np.pi / 4), atol=1e-4)
    crossed_diag = {i for i in want if abs(m.table("z_xz")[m.table("uv_index_z_nm")[i - 1] - 1] - 65.0) > 1e-3}
    high = {i for i in crossed_diag if m.table("zbuvmx")[i - 1] >= 1.0}
    assert set((ip[diag_pts] + 1).tolist()) == crossed_diag - high and len(high) > 0
    kc = m.table("kcuv")
    assert np.all(kc[ip] == 3)
    assert np.all(kc[np.array(sorted(high)) - 1] == 1)


def _weir_row_model(crest=0.0, cd=0.55):
    m, zb = _regular_basin(nmax=4, mmax=6, dx=10.0, zb=-3.0)
    m.set_bathymetry(zb)
    m.set_weirs([np.array([(30.0 + 1.7, -3.0, crest, cd), (30.0 + 1.7, 43.0, crest, cd)], F)])
    m.initialize()
    orc = O.OracleRun(m).configure(dict(t0=0.0, t1=600.0), m)
    return m, orc


@pytest.mark.parametrize("left,right", [(0.5, -0.2), (0.5, 0.4), (-0.1, 0.35), (0.2, 0.5), (-0.3, -0.1),
                                        (0.7, 0.7)])
def test_weir_discharge_closed_forms(left, right):
    cd, crest = 0.55, 0.0
    m, orc = _weir_row_model(crest, cd)
    ip = m.table("structure_uv_index") - 1
    assert len(ip) == 4
    zs = orc.state("zs")
    xz = m.table("z_xz")
    zs[:] = np.where(xz < 31.7, left, right)
    orc.state("uv")[:] = 1.0
    orc.L.oracle_compute_fluxes_over_structures(orc.m)
    q = orc.state("q")[ip]
    ln = float(m.table("structure_length")[0])
    up, lo = max(left, right), min(left, right)
    h1, h2 = up - crest, lo - crest
    if up < crest:
        expect = 0.0
    elif h2 > 2.0 / 3.0 * h1:
        expect = cd * h2 * np.sqrt(2 * 9.81 * (h1 - h2)) * ln
    else:
        expect = cd * 1.7049 * h1 ** 1.5 * ln
    expect = expect if left > right else -expect
    assert np.allclose(q, expect, rtol=2e-6, atol=1e-9), (q, expect)
    assert np.all(orc.state("uv")[ip] == 0.0)


def test_thin_dam_across_a_closed_basin_holds_the_water():
    nmax, mmax, dx = 12, 16, 10.0
    levels = np.where(np.arange(mmax) < 8, 0.8, 0.1)[:, None] + np.zeros((1, nmax))

    def run(with_dam):
        p = dict(t0=0.0, t1=600.0)
        sc = W.Scenario(name="basin", params=p, nmax=nmax, mmax=mmax, dx=dx, dy=dx,
                        msk=np.ones(nmax * mmax, np.int8), zb_grid=np.full((mmax, nmax), -1.0, F), t_end=600.0)
        if with_dam:   # between columns 7 and 8, ends outside the mesh
            W.with_thin_dams(sc, [[(80.3, -7.1), (79.6, 127.3)]])
        m = sc.build_model()
        orc = O.OracleRun(m).configure(p, m)
        orc.state("zs")[:] = levels.reshape(-1)
        east = m.table("z_xz") > 80.0
        v0 = float(orc.state("zs")[east].sum())
        orc.update(600.0)
        return m, float(orc.state("zs")[east].sum()) - v0
    m, gained = run(True)
    assert m.sizes()["nthd"] == nmax and gained == 0.0
    _, gained_open = run(False)
    assert gained_open > 1.0


def _structure_cases():
    reg = W.beach_structures(W.regular_case(40, 48, dx=50.0, seed=71, t_end=1800.0, tide_amp=2.5, rain_mmh=0.0),
                             crest=(-0.3, 0.4))
    sub = W.subgrid_case(30, 44, dx=25.0, seed=73, t_end=1800.0, nsrc=1, tide_amp=2.5)
    W.beach_structures(sub, crest=(-0.5, 0.4), weir_x=0.22, dam_x=(0.25, 0.3, 0.35))
    qt = W.quadtree_case(10, 14, nlev=3, dx0=100.0, seed=29, t_end=1800.0, surge_amp=3.0)
    W.beach_structures(qt, crest=(-0.5, 0.5), extent=(1400.0, 1000.0), cell=25.0, weir_x=0.09,
                       dam_x=(0.75, 0.85, 0.95))
    return {"regular": reg, "subgrid": sub, "quadtree": qt}


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("kind", ["regular", "subgrid", "quadtree"])
def test_structures_oracle_bit_exact_with_reference(kind):
    sc = _structure_cases()[kind]
    m = sc.build_model()
    s = m.sizes()
    assert s["nstruc"] > 10 and s["nthd"] > 10
    a = O.OracleRun(m).configure(sc.params, m)
    b = O.OracleRun(m).configure(sc.params, m)
    r = O.RefRun(b)
    fields = ["zs", "q", "uv", "kfuv", "zsmax"] + (["z_volume"] if kind == "subgrid" else [])
    for i in range(300):
        a.update(1e30, 1)
        r.step()
        for f in fields:
            assert np.array_equal(a.state(f), b.state(f)), (i, f)
    assert np.count_nonzero(a.state("q")[m.table("structure_uv_index") - 1]) > 0
