"""GPU parity: the HIP path (through the C ABI) against the oracle on identical inputs.

Tolerances are the parity gate: max|dzs| <= 1e-4 m, max|dq| <= 1e-5 m2/s; index tables
and masks are compared bit-exactly in test_tables.py."""
import numpy as np
import pytest

from helpers import Q_TOL, ZS_TOL, compare_state, compare_subgrid_state, gpu_for, oracle_for
from sfincs_amd import workloads as W

pytestmark = pytest.mark.gpu


def _cases():
    return [
        ("regular", W.regular_case(48, 64, dx=50.0, seed=11, t_end=1800.0)),
        ("subgrid", W.subgrid_case(40, 56, dx=25.0, seed=12, t_end=1800.0, nsrc=3, rain_mmh=30.0)),
    ]


@pytest.mark.parametrize("name,sc", _cases(), ids=[c[0] for c in _cases()])
def test_single_steps_match_oracle(name, sc):
    print(f"case {name}", flush=True)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    worst = {}
    for step in range(60):
        rc, st = inst.update(0.0)           # one step (sfincs_update with dtrange <= 0)
        orc.update(1e30, 1)
        d = compare_state(inst, orc)
        if m.sizes()["subgrid"]:
            d.update(compare_subgrid_state(inst, orc))
        for k, v in d.items():
            worst[k] = max(worst.get(k, 0.0), v)
        assert abs(st["t"] - orc.get("t")) < 1e-6, (step, st["t"], orc.get("t"))
    print(name, worst, flush=True)
    assert worst["zs"] <= ZS_TOL and worst["q"] <= Q_TOL
    if m.sizes()["subgrid"]:   # volumes to 1e-7 relative; zs0 / zsderv as zs
        assert worst["z_volume_rel"] <= 1e-7 and worst["zs0"] <= ZS_TOL and worst["zsderv"] <= ZS_TOL


@pytest.mark.parametrize("name,sc", _cases(), ids=[c[0] for c in _cases()])
def test_full_period_matches_oracle(name, sc):
    print(f"case {name}", flush=True)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    rc, st = inst.update(-1000.0)           # run to t1 (graph-captured loop)
    orc.update(sc.t_end)
    d = compare_state(inst, orc)
    if m.sizes()["subgrid"]:
        d.update(compare_subgrid_state(inst, orc))
    print(name, st["nt"], int(orc.get("nt")), d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL
    if m.sizes()["subgrid"]:
        assert d["z_volume_rel"] <= 1e-7 and d["zs0"] <= ZS_TOL and d["zsderv"] <= ZS_TOL


def _masked_case():
    sc = W.regular_case(40, 52, dx=50.0, seed=13, t_end=1800.0, tide_amp=2.0)
    r = np.random.default_rng(13)
    msk = sc.msk.reshape(52, 40).copy()
    holes = r.random(msk.shape) < 0.08
    holes[0, :] = False
    msk[holes] = 0
    sc.msk = msk.reshape(-1)
    return sc


INDEXED = [
    ("quadtree", lambda: W.quadtree_case(10, 14, nlev=3, dx0=100.0, seed=5, t_end=1800.0, surge_amp=4.0)),
    ("masked_regular", _masked_case),
]


@pytest.mark.parametrize("name,make", INDEXED, ids=[c[0] for c in INDEXED])
def test_indexed_path_matches_oracle(name, make):
    """Quadtree and masked grids run the table-gather kernels (and the combined-uv pass)."""
    print(f"case {name}", flush=True)
    sc = make()
    m, orc = oracle_for(sc)
    assert m.sizes()["dense"] == 0
    inst = gpu_for(m)
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc)
    print(name, st["nt"], int(orc.get("nt")), d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL


def test_standalone_entry_points_match_oracle():
    """sfx_update_forcing / sfx_compute_fluxes / sfx_compute_water_levels called one by one
    (the reference's call sequence, sfincs_lib.f90:530-618) against the oracle's pieces."""
    sc = W.subgrid_case(24, 30, dx=25.0, seed=14, t_end=1800.0, nsrc=2, rain_mmh=50.0)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    t = 0.0
    min_dt = np.float32(1.0e-6)
    for step in range(40):
        dt = np.float32(np.float32(0.5) * min_dt)
        t = t + float(dt)
        inst.update_forcing(t, float(dt))
        orc.update_forcing(t, dt)
        g = inst.compute_fluxes(float(dt))
        o = orc.compute_fluxes(dt)
        inst.compute_water_levels(t, float(dt))
        orc.compute_water_levels(t, dt)
        assert abs(g - o) <= 1e-6 * abs(o), (step, g, o)
        min_dt = np.float32(o)
    d = compare_state(inst, orc, ("zs", "q", "z_volume", "zsderv"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL, d


def test_batch_member_equals_alone():
    """A scenario advanced inside a batch with other scenarios gives the same result as
    the same scenario run alone (within the parity tolerance; in practice identical)."""
    from sfincs_amd import _lib
    scs = [W.regular_case(30, 38, dx=50.0, seed=31, t_end=900.0),
           W.subgrid_case(26, 34, dx=25.0, seed=32, t_end=900.0, nsrc=1, rain_mmh=20.0),
           W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=33, t_end=900.0)]
    alone = []
    for sc in scs:
        inst = gpu_for(sc.build_model())
        inst.update(-1000.0)
        alone.append((inst.get("zs"), inst.get("q"), inst.time()[2]))
    b = _lib.Batch()
    insts = [gpu_for(sc.build_model()) for sc in scs]
    for sc, inst in zip(scs, insts):
        b.add(inst, sc.t_end)
    n_active = len(insts)
    rounds = 0
    while n_active and rounds < 10000:
        n_active = b.advance(16)
        rounds += 1
    for (zs, q, nt), inst in zip(alone, insts):
        assert inst.time()[2] == nt
        assert np.max(np.abs(inst.get("zs") - zs)) <= ZS_TOL
        assert np.max(np.abs(inst.get("q") - q)) <= Q_TOL


def test_fp64_twin_all_wet():
    """The FP64 instantiation of the same kernels on an all-wet basin (no wet/dry or table
    switches): the FP32 path's rounding error over 1800 s of tide stays below 1e-4 m."""
    sc = W.regular_case(32, 40, dx=50.0, seed=15, t_end=1800.0, tide_amp=0.5, rain_mmh=0.0)
    sc.zb_grid = np.full_like(sc.zb_grid, -8.0) + 0.05 * sc.zb_grid
    m = sc.build_model()
    a = gpu_for(m)
    b = gpu_for(m, fp64=True)
    a.update(-1000.0)
    b.update(-1000.0)
    dzs = float(np.max(np.abs(a.get("zs") - b.get("zs"))))
    dq = float(np.max(np.abs(a.get("q").astype(np.float64) - b.get("q").astype(np.float64))))
    print("all wet: fp32 vs fp64 max|dzs|", dzs, "max|dq|", dq, flush=True)
    assert dzs <= 1e-4
    assert dq <= 1e-4


def _error_stats(a, b):
    d = np.abs(a.get("zs") - b.get("zs"))
    va, vb = a.total_volume(), b.total_volume()
    return dict(max=float(d.max()), p99=float(np.percentile(d, 99)), p90=float(np.percentile(d, 90)),
                mean=float(d.mean()), frac_gt_1mm=float(np.mean(d > 1e-3)), dvol_rel=abs(va - vb) / abs(vb))


def test_fp64_twin_wetting_statistics():
    """FP32 against the FP64 twin on a wetting/drying subgrid case, each with its own
    adaptive dt. Cells at a wet/dry front can take the other branch (cm-scale differences
    there), and the two runs take different dt sequences, so the error is stated as
    statistics over cells and as the volume difference (measured: mean 1.3 mm, p90 2.7 mm,
    p99 1.7 cm, volume 1.4e-6 relative)."""
    sc = W.subgrid_case(32, 40, dx=25.0, seed=15, t_end=1800.0, nsrc=2)
    m = sc.build_model()
    a = gpu_for(m)
    b = gpu_for(m, fp64=True)
    a.update(-1000.0)
    b.update(-1000.0)
    stats = _error_stats(a, b)
    print("wetting, own dt: fp32 vs fp64", stats, flush=True)
    assert stats["mean"] <= 3e-3 and stats["p90"] <= 6e-3 and stats["p99"] <= 5e-2
    assert stats["dvol_rel"] <= 1e-5


def test_fp64_twin_same_dt_sequence():
    """The same comparison with both runs driven through the standalone entry points on
    the FP32 run's dt sequence, so that only the arithmetic differs."""
    sc = W.subgrid_case(32, 40, dx=25.0, seed=15, t_end=1800.0, nsrc=2)
    m = sc.build_model()
    a = gpu_for(m)
    b = gpu_for(m, fp64=True)
    t = 0.0
    min_dt = np.float32(1.0e-6)
    alpha = np.float32(0.5)
    steps = 0
    while t < 1800.0:
        dt = np.float32(alpha * min_dt)
        t = t + float(dt)
        for x in (a, b):
            x.update_forcing(t, float(dt))
        min_dt = np.float32(a.compute_fluxes(float(dt)))
        b.compute_fluxes(float(dt))
        for x in (a, b):
            x.compute_water_levels(t, float(dt))
        steps += 1
    stats = _error_stats(a, b)
    print(f"wetting, same dt ({steps} steps): fp32 vs fp64", stats, flush=True)
    assert stats["mean"] <= 3e-3 and stats["p90"] <= 6e-3 and stats["p99"] <= 5e-2
    assert stats["dvol_rel"] <= 1e-5


def test_lake_at_rest_and_volume():
    """Flat water over a closed basin stays at rest (q == 0, zs unchanged) and the FP64
    volume accumulator is conserved; with uniform rain it grows by rain * area * t."""
    sc = W.regular_case(20, 24, dx=50.0, seed=16, t_end=600.0, tide_amp=0.0, rain_mmh=0.0, zsini=1.0)
    msk = np.ones(20 * 24, np.int8)
    sc.msk = msk
    sc.bnd = None
    m = sc.build_model()
    inst = gpu_for(m)
    v0 = inst.total_volume()
    zs0 = inst.get("zs")
    inst.update(-1000.0)
    assert np.max(np.abs(inst.get("q"))) == 0.0
    wet = zs0 > m.table("zb") + 1e-9
    assert np.max(np.abs(inst.get("zs")[wet] - zs0[wet])) == 0.0
    assert abs(inst.total_volume() - v0) <= 1e-9 * abs(v0)


def test_branch_counters_cover_every_point():
    """The per-tile branch counters add up to every uv-point slot of every sweep (dense:
    two slots per cell; indexed: npuv), and the CFL fold gives the same dt sequence with
    counting on or off."""
    for sc in (W.subgrid_case(24, 30, dx=25.0, seed=17, t_end=600.0, nsrc=1),
               W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=18, t_end=600.0)):
        m = sc.build_model()
        sz = m.sizes()
        a = gpu_for(m)
        b = gpu_for(m)
        a.set_counting(True)
        for _ in range(25):
            a.update(0.0)
            b.update(0.0)
        per_sweep = 2 * sz["np"] if sz["dense"] else sz["npuv"]
        assert sum(a.counters()) == 25 * per_sweep
        st = a.store_counters()        # continuity stores only where a value changed
        assert all(0 <= x <= 25 * sz["np"] for x in st) and st[0] > 0
        assert a.time() == b.time()
        assert np.array_equal(a.get("zs"), b.get("zs"))


SHAPES = [(0, (64, 4), (64, 4)), (0, (256, 1), (128, 2)), (0, (32, 8), (256, 1)), (4, (64, 4), (32, 8)),
          (8, (64, 4), (64, 4)), (16, (64, 4), (64, 4)), (-1, (64, 4), (64, 4)), (-1, (32, 8), (128, 2)),
          (-1, (256, 1), (64, 4)), (-1, (16, 16), (16, 16)), (0, (16, 16), (32, 8)),
          (-2, (64, 4), (64, 4)), (-2, (256, 1), (128, 2)), (-2, (16, 16), (64, 4)), (-4, (64, 4), (32, 8)),
          (-4, (32, 8), (64, 4)), (-4, (128, 2), (256, 1))]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"kernel{a}_mom{b[0]}x{b[1]}_cont{c[0]}x{c[1]}" for a, b, c in SHAPES])
def test_every_dense_kernel_shape_matches_oracle(shape):
    """Each dense momentum variant (per-cell tiles, per-point tiles with 1, 2 or 4 points per
    thread, strip marches) and continuity tile the autotuner can pick gives the oracle's
    result; the grid (70 x 83) is not a multiple of any tile, so partial tiles and strips
    are exercised."""
    strip, mt, ct = shape
    print(f"case shape {shape}", flush=True)
    sc = W.subgrid_case(70, 83, dx=25.0, seed=19, t_end=900.0, nsrc=2, rain_mmh=20.0)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    inst.set_tuning(strip, mt, ct)
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc)
    print(shape, st["nt"], d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL


REG_POINT = [(-1, (64, 4)), (-2, (64, 4)), (-2, (16, 16)), (-4, (32, 8)), (-4, (256, 1))]


@pytest.mark.parametrize("shape", REG_POINT, ids=[f"kernel{a}_mom{b[0]}x{b[1]}" for a, b in REG_POINT])
def test_regular_point_kernels_match_oracle(shape):
    """The per-point momentum kernels with 1, 2 and 4 points per thread on a regular
    (non-subgrid) mesh of 53 x 77 cells, which no tile divides."""
    kind, mt = shape
    print(f"case regular point kernel {shape}", flush=True)
    sc = W.regular_case(53, 77, dx=50.0, seed=23, t_end=1200.0, rain_mmh=30.0)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    inst.set_tuning(kind, mt, (64, 4))
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc)
    print(shape, st["nt"], d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL



@pytest.mark.parametrize("mesh", ["dense", "quadtree", "subgrid"])
def test_state_write_over_long_dry_points_matches_oracle(mesh):
    """Momentum stops storing zeros at points that stayed dry for two steps (both ping-pong
    buffers hold them) and continuity skips cells that stayed unchanged without inflow.
    A host write of state must end both: after 900 s of a beach with a tide, q + 1 mm2/s
    and zs + 1 mm are written into both runs (the oracle's state, so the two start equal;
    a skipped cell would miss the zsmax rise of a regular cell and the level a subgrid cell
    takes back from its volume) and the run continues to its end on both."""
    print(f"case state write over dry points and steady cells, {mesh}", flush=True)
    if mesh == "dense":
        sc = W.regular_case(40, 56, dx=50.0, seed=37, t_end=1800.0, rain_mmh=0.0, tide_amp=1.5)
    elif mesh == "quadtree":
        sc = W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=38, t_end=1800.0)
    else:
        sc = W.subgrid_case(36, 44, dx=25.0, seed=39, t_end=1800.0, nsrc=1, rain_mmh=0.0)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    if mesh != "quadtree":
        inst.set_tuning(-2, (64, 4), (64, 4))
    inst.update(900.0)
    orc.update(900.0)
    assert inst.time()[2] == int(orc.get("nt"))
    npuv = inst.sizes["npuv"]
    q = orc.state("q")
    q[:npuv] += np.float32(1e-3)          # in place: the oracle's own state
    inst.set("q", q)
    zs = orc.state("zs")
    zs += 1e-3
    inst.set("zs", zs)
    inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc)
    if m.sizes()["subgrid"]:
        d.update(compare_subgrid_state(inst, orc))
    print(d, flush=True)
    assert inst.time()[2] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL
    if m.sizes()["subgrid"]:
        assert d["z_volume_rel"] <= 1e-7 and d["zs0"] <= ZS_TOL and d["zsderv"] <= ZS_TOL


QUIET = [(-1, (16, 16)), (-2, (64, 4)), (-4, (32, 8))]


@pytest.mark.parametrize("shape", QUIET, ids=[f"kernel{a}_mom{b[0]}x{b[1]}" for a, b in QUIET])
def test_quiescent_tiles_wake_with_the_tide(shape):
    """Without rain most per-point tiles go quiescent (all points dry with their zeros held,
    no level they read changing) and are skipped; a rising tide must wake them as the water
    reaches their cells. Checked against the oracle after 2 h of a 2 m tide, with the
    skip counted so that the path is known to run."""
    kind, mt = shape
    print(f"case quiescent tiles {shape}", flush=True)
    sc = W.subgrid_case(70, 83, dx=25.0, seed=41, t_end=7200.0, tide_amp=2.0, nsrc=0, rain_mmh=0.0)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    inst.set_tuning(kind, mt, (64, 4))
    inst.set_counting(True)
    rc, st = inst.update(-1000.0)
    skipped = inst.counters()[5]
    orc.update(sc.t_end)
    d = compare_state(inst, orc)
    print(shape, st["nt"], skipped, d, flush=True)
    assert skipped > 0
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL


STORE = dict(store_velmax=1, store_fluxmax=1, store_twet=1)
STORE_CASES = [
    ("regular", lambda: W.regular_case(40, 52, dx=50.0, seed=71, t_end=3600.0, tide_amp=2.0, **STORE)),
    ("subgrid", lambda: W.subgrid_case(36, 44, dx=25.0, seed=72, t_end=3600.0, nsrc=2, rain_mmh=30.0, tide_amp=2.0,
                                       twet_threshold=0.05, **STORE)),
    ("quadtree", lambda: W.quadtree_case(10, 14, nlev=3, dx0=100.0, seed=73, t_end=3600.0, surge_amp=4.0, **STORE)),
]


@pytest.mark.parametrize("name,make", STORE_CASES, ids=[c[0] for c in STORE_CASES])
def test_store_variables_match_oracle(name, make):
    """vmax, qmax and twet (compute_store_variables, fused into continuity) against the
    oracle over a whole run, then a reset through set_state (as after a maximum output)
    and another stretch. Tolerances: vmax 1e-5 m/s and qmax 1e-5 m2/s, as for q; twet
    1e-3 s (one step is about 1 s, so a cell counted wet in a different step fails)."""
    print(f"case store variables {name}", flush=True)
    sc = make()
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    inst.update(1800.0)
    orc.update(1800.0)
    assert inst.time()[2] == int(orc.get("nt"))
    d = compare_state(inst, orc, ("zs", "q", "vmax", "qmax", "twet"))
    print(name, inst.time()[2], d, flush=True)
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL
    assert d["vmax"] <= 1e-5 and d["qmax"] <= 1e-5 and d["twet"] <= 1e-3
    assert np.max(inst.get("vmax")) > 0.0 and np.max(inst.get("twet")) > 0.0
    n = m.sizes()["np"]
    for f, v in (("vmax", -999.0), ("qmax", -999.0), ("twet", 0.0)):
        inst.set(f, np.full(n, v, np.float32))
        orc.state(f)[:] = v
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc, ("zs", "q", "vmax", "qmax", "twet"))
    print(name, st["nt"], d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["vmax"] <= 1e-5 and d["qmax"] <= 1e-5 and d["twet"] <= 1e-3


def test_store_variables_off_are_not_stored():
    """With the switches off (the default) the fields do not exist on the device."""
    m, _ = oracle_for(W.regular_case(16, 20, dx=50.0, seed=74, t_end=600.0))
    inst = gpu_for(m)
    for f in ("vmax", "qmax", "twet"):
        with pytest.raises(Exception):
            inst.get(f)


def _records(raw):
    out, i = [], 0
    while i < len(raw):
        n = int(np.frombuffer(raw[i:i + 4], np.int32)[0])
        out.append(raw[i + 4:i + 4 + n])
        assert int(np.frombuffer(raw[i + 4 + n:i + 8 + n], np.int32)[0]) == n
        i += n + 8
    return out


RESTART_CASES = [
    ("subgrid", lambda: W.subgrid_case(32, 40, dx=25.0, seed=76, t_end=1800.0, nsrc=2, tide_amp=2.0)),
    ("quadtree", lambda: W.quadtree_case(10, 14, nlev=3, dx0=100.0, seed=77, t_end=1800.0, surge_amp=4.0)),
]


@pytest.mark.parametrize("name,make", RESTART_CASES, ids=[c[0] for c in RESTART_CASES])
def test_restart_round_trip_matches_oracle(name, make, tmp_path):
    """sfx_write_restart after 900 s holds the instance's state in write_rst_file's layout
    (type 1; zs as f32, q, uvmean; exact); a model initialised from that file then runs
    900 s on the GPU as on the oracle (parity tolerances)."""
    print(f"case restart round trip {name}", flush=True)
    sc = make()
    m = sc.build_model()
    inst = gpu_for(m)
    inst.update(900.0)
    path = tmp_path / "sfincs.rst"
    inst.write_restart(path)
    rec = _records(path.read_bytes())
    assert len(rec) == 4 and int(np.frombuffer(rec[0], np.int32)[0]) == 1
    assert np.array_equal(np.frombuffer(rec[1], np.float32), inst.get("zs").astype(np.float32))
    assert np.array_equal(np.frombuffer(rec[2], np.float32), inst.get("q"))
    assert np.array_equal(np.frombuffer(rec[3], np.float32), inst.get("uvmean"))
    assert np.any(inst.get("q"))
    sc2 = make()
    sc2.rstfile = str(path)
    sc2.t_end = 900.0
    sc2.params["t1"] = 900.0
    m2, orc = oracle_for(sc2)
    assert np.array_equal(m2.table("zs"), np.maximum(m2.table("z_zmin" if m2.sizes()["subgrid"] else "zb"),
                                                     inst.get("zs").astype(np.float32)).astype(np.float64))
    inst2 = gpu_for(m2)
    rc, st = inst2.update(-1000.0)
    orc.update(sc2.t_end)
    d = compare_state(inst2, orc)
    print(name, st["nt"], d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL


RIVER_CASES = [
    ("regular", lambda: W.with_river(W.regular_case(40, 52, dx=50.0, seed=78, t_end=1800.0, tide_amp=1.5))),
    ("subgrid", lambda: W.with_river(W.subgrid_case(32, 40, dx=25.0, seed=79, t_end=1800.0, nsrc=1, tide_amp=1.5),
                                     slope=2.0e-3)),
]


@pytest.mark.parametrize("name,make", RIVER_CASES, ids=[c[0] for c in RIVER_CASES])
def test_river_boundaries_match_oracle(name, make):
    """Downstream river cells (msk=5) take their level from an interior cell in the boundary
    pass of k_forcing; whole run against the oracle (subgrid volumes included)."""
    print(f"case river boundaries {name}", flush=True)
    sc = make()
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc, ("zs", "q", "uv", "zsmax") + (("z_volume",) if m.sizes()["subgrid"] else ()))
    print(name, st["nt"], int(orc.get("nt")), d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL


def test_source_into_dry_cell_keeps_its_volume():
    """A point source filling a dry subgrid cell with no face flux: continuity stores a field
    only when it changes, and the source's volume must count as a change (the volume is
    compared against its loaded value, not against the value after the sources). Volumes
    are compared step by step with the oracle (relative 1e-7; the source adds 4.4e-5 m3
    to a dry cell in the first step)."""
    print("case source into a dry cell", flush=True)
    sc = W.subgrid_case(32, 40, dx=25.0, seed=79, t_end=1800.0, nsrc=1, tide_amp=1.5)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    c = int(m.table("nmindsrc")[0]) - 1
    for step in range(40):
        inst.update(0.0)
        orc.update(1e30, 1)
        va, vb = inst.get("z_volume"), orc.state("z_volume")
        assert np.all(np.abs(va - vb) <= 1e-7 * (1.0 + np.abs(vb))), (step, va[c], vb[c])
    assert orc.state("z_volume")[c] > 0.0


IG_CASES = [
    ("regular", lambda: W.with_ig_waves(W.regular_case(40, 52, dx=50.0, seed=96, t_end=1800.0, tide_amp=1.0))),
    ("subgrid", lambda: W.with_ig_waves(W.subgrid_case(32, 40, dx=25.0, seed=97, t_end=1800.0, nsrc=1, tide_amp=1.5),
                                        amp=0.3, period=90.0)),
]


@pytest.mark.parametrize("name,make", IG_CASES, ids=[c[0] for c in IG_CASES])
def test_ig_waves_match_oracle(name, make):
    """Incoming infragravity levels (bzi) added to the water-level boundary in the boundary
    pass of k_forcing; whole run against the oracle."""
    print(f"case ig waves {name}", flush=True)
    sc = make()
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc)
    if m.sizes()["subgrid"]:
        d.update(compare_subgrid_state(inst, orc))
    print(name, st["nt"], int(orc.get("nt")), d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL


TSA_CASES = [
    ("regular", lambda: W.regular_case(40, 52, dx=50.0, seed=98, t_end=1800.0, tide_amp=2.0, timestep_analysis=1)),
    ("subgrid", lambda: W.subgrid_case(32, 40, dx=25.0, seed=99, t_end=1800.0, nsrc=2, timestep_analysis=1)),
    ("quadtree", lambda: W.quadtree_case(10, 14, nlev=3, dx0=100.0, seed=100, t_end=1800.0, surge_amp=4.0,
                                         timestep_analysis=1)),
]


@pytest.mark.parametrize("name,make", TSA_CASES, ids=[c[0] for c in TSA_CASES])
def test_timestep_analysis_matches_oracle(name, make):
    """timestep_analysis accumulators (k_tsa_min + k_tsa_update after each momentum sweep)
    against the oracle over a whole run: wet and limiting counts exact, the sums of
    required time steps to 1e-6 relative. Autotuning before the run must not count its
    timed sweeps."""
    print(f"case timestep analysis {name}", flush=True)
    sc = make()
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    inst.autotune(2)
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    assert st["nt"] == int(orc.get("nt"))
    n = m.sizes()["npuv"]
    wa, wb = inst.get("tsa_wet"), orc.state("tsa_wet")[:n]
    la, lb = inst.get("tsa_lim"), orc.state("tsa_lim")[:n]
    sa, sb = inst.get("tsa_sum"), orc.state("tsa_sum")[:n]
    print(name, st["nt"], int(wb.max()), int(lb.sum()), float(np.max(np.abs(sa - sb) / (1.0 + np.abs(sb)))), flush=True)
    assert np.array_equal(wa, wb) and np.array_equal(la, lb)
    assert np.all(np.abs(sa - sb) <= 1e-6 * (1.0 + np.abs(sb)))
    assert lb.sum() >= st["nt"]


@pytest.mark.parametrize("kind", ["regular", "subgrid", "quadtree"])
def test_qext_matches_oracle(kind):
    """External sources per cell (use_qext, written with set_state "qext"), changed halfway
    as a coupled model would; whole run against the oracle."""
    print(f"case qext {kind}", flush=True)
    if kind == "regular":
        sc = W.regular_case(40, 52, dx=50.0, seed=101, t_end=1800.0, tide_amp=1.5, use_qext=1)
    elif kind == "subgrid":
        sc = W.subgrid_case(32, 40, dx=25.0, seed=102, t_end=1800.0, nsrc=1, use_qext=1)
    else:
        sc = W.quadtree_case(10, 14, nlev=3, dx0=100.0, seed=103, t_end=1800.0, surge_amp=4.0, use_qext=1)
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    rng = np.random.default_rng(104)
    n = m.sizes()["np"]
    for t_stop in (900.0, sc.t_end):
        field = rng.uniform(-2e-5, 5e-5, n).astype(np.float32)
        inst.set("qext", field)
        orc.state("qext")[:] = field
        inst.update(t_stop - inst.time()[0])
        orc.update(t_stop)
        assert inst.time()[2] == int(orc.get("nt"))
    d = compare_state(inst, orc)
    if m.sizes()["subgrid"]:
        d.update(compare_subgrid_state(inst, orc))
    print(kind, inst.time()[2], d, flush=True)
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL


def test_storage_volume_matches_oracle():
    """Storage volume per subgrid cell (use_storage_volume) with rain filling it; whole run
    against the oracle, the storage left included."""
    print("case storage volume", flush=True)
    def storage(m):
        r = np.random.default_rng(105)
        v = r.uniform(0.0, 40.0, m.sizes()["np"]).astype(np.float32)
        v[r.random(v.size) < 0.3] = 0.0
        return v
    sc = W.subgrid_case(32, 40, dx=25.0, seed=106, t_end=1800.0, nsrc=1, rain_mmh=60.0)
    sc.storage = storage
    m, orc = oracle_for(sc)
    inst = gpu_for(m)
    rc, st = inst.update(-1000.0)
    orc.update(sc.t_end)
    d = compare_state(inst, orc, ("zs", "q", "uv", "zsmax", "storage_volume"))
    d.update(compare_subgrid_state(inst, orc))
    print(st["nt"], d, flush=True)
    assert st["nt"] == int(orc.get("nt"))
    assert d["zs"] <= ZS_TOL and d["q"] <= Q_TOL and d["z_volume_rel"] <= 1e-7
    assert d["storage_volume"] <= 1e-3   # m3
