"""The Fortran host (sfincs_lib entry points + BMI names over the C ABI).

CPU: the host library exports the reference's driver and BMI names. GPU: the host
executable, run on an input deck written in the reference's formats, reproduces the
Python-driven run of the same scenario exactly (same kernels, same inputs)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from sfincs_amd import deck
from sfincs_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "sfincs_amd", "fortran")
HOSTLIB = os.path.join(FDIR, "libsfincs_host.so")
EXE = os.path.join(FDIR, "sfincs_mi355x")

needs_host = pytest.mark.skipif(not os.path.exists(HOSTLIB), reason="Fortran host not built (amdflang)")


def records(path, dtype):
    """The records of a Fortran unformatted sequential file (4-byte length markers)."""
    raw = open(path, "rb").read()
    out, k = [], 0
    while k < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, k)[0])
        out.append(np.frombuffer(raw, dtype, n // np.dtype(dtype).itemsize, k + 4).copy())
        assert int(np.frombuffer(raw, np.int32, 1, k + 4 + n)[0]) == n
        k += n + 8
    return out


@needs_host
def test_host_exports_driver_and_bmi_names():
    lib = ctypes.CDLL(HOSTLIB)
    for name in ("initialize", "update", "update_until", "finalize", "get_value_ptr", "get_current_time",
                 "get_start_time", "get_end_time", "get_time_step", "get_var_shape", "get_var_type",
                 "get_var_rank", "get_sfincs_cell_index", "get_sfincs_cell_indices", "get_sfincs_cell_area",
                 "get_last_bmi_error", "set_logical", "update_zbuv", "update_apparent_roughness"):
        assert hasattr(lib, name), name
    syms = subprocess.run(["nm", "-D", HOSTLIB], capture_output=True, text=True).stdout
    for name in ("sfincs_initialize", "sfincs_update", "sfincs_finalize"):
        assert name in syms, name
    # the BMI string lengths (sfincs_bmi.f90:41-52)
    for name, v in (("BMI_LENVARADDRESS", 64), ("BMI_LENVARTYPE", 64), ("BMI_LENGRIDTYPE", 64),
                    ("BMI_LENCOMPONENTNAME", 64), ("BMI_LENVERSION", 256), ("BMI_LENERRMESSAGE", 1024)):
        assert ctypes.c_int.in_dll(lib, name).value == v, name


def test_deck_round_trip_of_regular_mesh(tmp_path):
    sc = W.regular_case(12, 16, dx=50.0, seed=4, t_end=600.0)
    d = deck.write(sc, str(tmp_path))
    ind = np.fromfile(os.path.join(d, "sfincs.ind"), np.int32)
    assert ind[0] == 12 * 16 and np.array_equal(ind[1:], np.arange(1, 12 * 16 + 1))
    dep = np.fromfile(os.path.join(d, "sfincs.dep"), np.float32)
    assert np.array_equal(dep, sc.zb_grid.reshape(-1))


TABLES = ("params", "kcs", "zb", "z_flags_iref", "uv_index_z_nm", "uv_index_z_nmu", "uv_index_u_nmd",
          "uv_index_u_nmu", "uv_index_u_ndm", "uv_index_u_num", "uv_index_v_ndm", "uv_index_v_nm",
          "uv_index_v_ndmu", "uv_index_v_nmu", "uv_flags_iref", "uv_flags_type", "uv_flags_dir", "kcuv",
          "mask_adv", "cuv_index_uv", "zbuv", "zbuvmx", "gn2uv", "dxrinv", "cell_area", "zs", "zsb", "zsb0",
          "nmindbnd", "index_kcuv2", "nmbkcuv2", "nmikcuv2", "ibuvdir", "ind1_bnd_gbp", "ind2_bnd_gbp",
          "fac_bnd_gbp", "nmi_gbp", "nmindsrc", "t_bnd", "zs_bnd", "tsrc", "qsrc", "tprcpt", "tprcpv",
          "twnd", "wndmag", "wnddir", "cdval", "index_bdr_gbp", "index_zsi_bdr", "dzs_bdr", "zsi_bnd")


def _table(lib, m, name):
    n = ctypes.c_size_t()
    if lib.sfx_model_table_bytes(ctypes.c_void_p(m), name.encode(), ctypes.byref(n)) != 0:
        return None
    buf = (ctypes.c_char * max(1, n.value))()
    assert lib.sfx_model_get_table(ctypes.c_void_p(m), name.encode(), buf, n) == 0
    return bytes(buf)[:n.value]


SUBGRID_TABLES = ("z_zmin", "z_zmax", "z_volmax", "z_dep", "uv_zmin", "uv_zmax", "uv_havg", "uv_nrep", "uv_pwet",
                  "uv_fnfit", "uv_havg_zmax", "uv_nrep_zmax", "uv_navg_w", "z_volume")


@needs_host
@pytest.mark.parametrize("kind", ["regular", "quadtree", "wind_sources", "store_switches", "river", "ig_waves",
                                  "subgrid_binary", "drainage", "structures", "infiltration", "meteo",
                                  "quadtree_netcdf", "subgrid_netcdf", "curve_number"])
def test_deck_reader_builds_the_same_model(kind, tmp_path):
    """The Fortran deck reader (no GPU) builds, table for table and byte for byte, the
    model that the Python builder makes from the same scenario (parameters included)."""
    from sfincs_amd import _lib
    if kind == "regular":
        sc = W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0)
    elif kind == "ig_waves":
        sc = W.with_ig_waves(W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0))
    elif kind == "river":
        sc = W.with_river(W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0), slope=2.0e-3)
    elif kind == "store_switches":
        sc = W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0, store_velmax=1, store_fluxmax=1, store_twet=1,
                            twet_threshold=0.02, timestep_analysis=1)
    elif kind in ("quadtree", "quadtree_netcdf"):
        sc = W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=42, t_end=900.0)
    elif kind == "subgrid_binary":
        sc = W.with_subgrid_binary(W.subgrid_case(22, 26, dx=25.0, seed=40, t_end=900.0, nsrc=2, rain_mmh=10.0),
                                   tmp_path / "gen.sbg")
    elif kind == "subgrid_netcdf":   # sbgfile as NetCDF
        sc = W.subgrid_case(22, 26, dx=25.0, seed=40, t_end=900.0, nsrc=2, rain_mmh=10.0)
    elif kind == "curve_number":     # scsfile and sfacinf
        sc = W.with_curve_number(W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0, rain_mmh=20.0),
                                 lambda m: (m.table("z_xz") / 3000.0).astype(np.float32), sfacinf=0.1)
    elif kind == "drainage":
        sc = W.beach_drainage(W.regular_case(40, 48, dx=50.0, seed=71, t_end=900.0, tide_amp=2.5,
                                             structure_relax=7.5))
    elif kind == "structures":
        sc = W.beach_structures(W.regular_case(40, 48, dx=50.0, seed=71, t_end=900.0, tide_amp=2.5),
                                crest=(-0.3, 0.4))
    elif kind == "infiltration":     # qinffile map, with qinf and qinf_zmin keys that it overrides
        sc = W.with_infiltration(W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0), qinf_zmin=-3.0,
                                 qinf_map=lambda m: (m.table("z_xz") / 100.0).astype(np.float32))
    elif kind == "meteo":            # amu/amv/amp/ampr and spw files, meteo keys
        sc = W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0, dtwnd=450.0, pavbnd=101000.0, spinup_meteo=1,
                            tspinup=300.0, factor_pres=1.5, spw_merge_frac=0.7)
        sc.rain = None
        W.with_spiderweb(W.with_gridded_meteo(sc, t_step=600.0))
    else:
        sc = W.subgrid_case(20, 24, dx=25.0, seed=43, t_end=900.0, nsrc=2, rain_mmh=10.0)
        sc.subgrid = None
        sc.zb_grid = W.beach_with_bumps(W.rng(43), 20, 24, 25.0, 25.0)
        sc.wind = (np.array([0.0, 2000.0], np.float32), np.array([12.0, 18.0], np.float32),
                   np.array([200.0, 250.0], np.float32))
    d = deck.write(sc, str(tmp_path), netcdf=kind.endswith("netcdf"))
    host = ctypes.CDLL(HOSTLIB)
    host.sfincs_host_model.restype = ctypes.c_void_p
    cwd = os.getcwd()
    os.chdir(d)
    try:
        assert host.sfincs_host_read_model() == 0
    finally:
        os.chdir(cwd)
    mf = host.sfincs_host_model()
    mp = sc.build_model()
    lib = _lib.lib()
    extra = SUBGRID_TABLES if kind.startswith("subgrid") else ()
    if kind == "drainage":
        extra = ("drainage_type", "drainage_params", "drainage_distance")
        assert mp.sizes()["ndrn"] == 5 and mp.params().structure_relax == 7.5
    if kind == "structures":
        extra = ("thindam_uv_index", "structure_uv_index", "structure_height", "structure_cd", "structure_length")
        assert mp.sizes()["nstruc"] > 10 and mp.sizes()["nthd"] > 10
    if kind == "curve_number":
        extra = ("scs_m",)
        assert np.count_nonzero(mp.table("scs_m")) > 100 and mp.params().sfacinf == np.float32(0.1)
    if kind == "infiltration":
        extra = ("qinffield",)
        assert np.count_nonzero(mp.table("qinffield")) > 100
    if kind == "meteo":
        extra = ("meteo_times", "meteo_tauwu", "meteo_tauwv", "meteo_patm", "meteo_prcp") + tuple(
            f"{k}_{s}" for k in ("amu", "amv", "amp", "ampr") for s in ("times", "values", "geo")) + (
            "spw_times", "spw_xe", "spw_ye", "spw_vmag", "spw_vdir", "spw_pdrop", "spw_prcp", "spw_radius")
        assert len(mp.table("meteo_times")) >= 4 and np.ptp(mp.table("meteo_patm")) > 100.0
    for name in TABLES + extra:
        a, b = _table(lib, mf, name), _table(lib, mp.h.value, name)
        assert (a is None) == (b is None), name
        assert a == b, name
    if kind.startswith("subgrid"):
        assert mp.sizes()["subgrid"] == 1 and len(_table(lib, mf, "uv_havg")) > 0
    if kind.endswith("netcdf"):
        assert any(f.endswith(".nc") for f in os.listdir(d))


@needs_host
@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["regular", "quadtree", "subgrid_binary", "drainage", "structures",
                                  "infiltration", "meteo", "quadtree_netcdf", "subgrid_netcdf", "curve_number"])
def test_fortran_host_matches_python_path(kind, tmp_path):
    print(f"case fortran host {kind}", flush=True)
    from sfincs_amd import _lib
    if kind == "regular":
        sc = W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0)
    elif kind == "subgrid_binary":   # a subgrid deck through sfincs_initialize (sbgfile)
        sc = W.with_subgrid_binary(W.subgrid_case(30, 34, dx=25.0, seed=40, t_end=900.0, nsrc=2, rain_mmh=10.0),
                                   tmp_path / "gen.sbg")
    elif kind == "drainage":         # drnfile
        sc = W.beach_drainage(W.regular_case(40, 48, dx=50.0, seed=71, t_end=900.0, tide_amp=2.5))
    elif kind == "structures":       # thdfile and weirfile
        sc = W.beach_structures(W.regular_case(40, 48, dx=50.0, seed=71, t_end=900.0, tide_amp=2.5),
                                crest=(-0.3, 0.4))
    elif kind == "infiltration":     # qinffile
        sc = W.with_infiltration(W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0, rain_mmh=30.0),
                                 qinf_map=lambda m: (m.table("z_xz") / 40.0).astype(np.float32))
    elif kind == "meteo":            # amu/amv/amp/ampr and spw files
        sc = W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0, dtwnd=450.0, pavbnd=101000.0)
        sc.rain = None
        W.with_spiderweb(W.with_gridded_meteo(sc, t_step=600.0))
    elif kind == "subgrid_netcdf":   # NetCDF sbgfile
        sc = W.subgrid_case(30, 34, dx=25.0, seed=40, t_end=900.0, nsrc=2, rain_mmh=10.0)
    elif kind == "curve_number":     # scsfile
        sc = W.with_curve_number(W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0, rain_mmh=40.0),
                                 lambda m: (m.table("z_xz") / 300.0).astype(np.float32))
    else:                            # legacy binary or NetCDF qtrfile
        sc = W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=42, t_end=900.0)
    d = deck.write(sc, str(tmp_path), netcdf=kind.endswith("netcdf"))
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    zs_f = np.fromfile(os.path.join(d, "zs_final.dat"), np.float64)
    inst = _lib.Instance(sc.build_model())
    inst.update(-1000.0)
    zs_p = inst.get("zs")
    assert zs_f.shape == zs_p.shape
    assert np.array_equal(zs_f, zs_p)


@needs_host
def test_deck_reader_applies_the_restart_file(tmp_path):
    """`rstfile` in sfincs.inp: the Fortran host reads the binary restart file into the same
    initial state as the Python builder does (no GPU)."""
    from sfincs_amd import _lib
    from test_restart import write_rst
    sc = W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=44, t_end=900.0)
    d = deck.write(sc, str(tmp_path))
    base = sc.build_model()
    s = base.sizes()
    r = np.random.default_rng(5)
    rst = os.path.join(d, "sfincs.rst")
    write_rst(rst, 1, base.table("zb") + r.uniform(-0.5, 1.0, s["np"]),
              r.uniform(-0.2, 0.2, s["npuv"] + s["ncuv"] + 1), r.uniform(-0.1, 0.1, s["nkcuv2"]))
    with open(os.path.join(d, "sfincs.inp"), "a") as f:
        f.write("rstfile          = sfincs.rst\n")
    host = ctypes.CDLL(HOSTLIB)
    host.sfincs_host_model.restype = ctypes.c_void_p
    cwd = os.getcwd()
    os.chdir(d)
    try:
        assert host.sfincs_host_read_model() == 0
    finally:
        os.chdir(cwd)
    mf = host.sfincs_host_model()
    sc.rstfile = rst
    mp = sc.build_model()
    lib = _lib.lib()
    for name in ("zs", "q", "uv", "zs0", "uvmean"):
        a, b = _table(lib, mf, name), _table(lib, mp.h.value, name)
        assert a is not None and a == b, name
    assert _table(lib, mf, "zs") != _table(lib, base.h.value, "zs")


@needs_host
@pytest.mark.gpu
def test_fortran_host_writes_store_fields_and_timestep_analysis(tmp_path):
    """storevelmax / storefluxmax / storetwet and timestep_analysis in sfincs.inp: the host
    executable writes vmax, qmax, twet and the per-cell time step analysis at finalize,
    equal to the Python-driven run of the same scenario."""
    print("case fortran host store fields", flush=True)
    from sfincs_amd import _lib
    sc = W.regular_case(28, 36, dx=50.0, seed=45, t_end=900.0, tide_amp=2.0, store_velmax=1, store_fluxmax=1,
                        store_twet=1, timestep_analysis=1)
    d = deck.write(sc, str(tmp_path))
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = sc.build_model()
    inst = _lib.Instance(m)
    inst.update(-1000.0)
    # the maximum output at the end of the run (finalize_output, sfincs_output.f90:262-310)
    for f, name in (("vmax", "velmax.dat"), ("qmax", "qmax.dat"), ("twet", "tmax.dat")):
        rec = records(os.path.join(d, name), np.float32)
        assert len(rec) == 1 and np.array_equal(rec[0], inst.get(f)), f
    dtc, pct = m.timestep_analysis_per_cell(inst.get("tsa_sum"), inst.get("tsa_wet"), inst.get("tsa_lim"), inst.time()[2])
    assert np.array_equal(np.fromfile(os.path.join(d, "timestep_analysis_dt.dat"), np.float32), dtc)
    assert np.array_equal(np.fromfile(os.path.join(d, "timestep_analysis_pct.dat"), np.float32), pct)
    assert np.any(pct > 0)


@needs_host
@pytest.mark.parametrize("kind", ["regular", "quadtree"])
def test_bmi_mesh_queries(kind, tmp_path):
    """The BMI's mesh queries after a host-only deck read (no GPU): cell lookup by point
    (sfincs_bmi.f90:299-349), cell areas, and the mesh fields with their shape, type and
    rank (:95-224); switches of processes outside this build answer -1."""
    if kind == "regular":
        sc = W.regular_case(12, 16, dx=50.0, seed=46, t_end=600.0)
    else:
        sc = W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=47, t_end=600.0)
    d = deck.write(sc, str(tmp_path))
    host = ctypes.CDLL(HOSTLIB)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        assert host.sfincs_host_read_model() == 0
    finally:
        os.chdir(cwd)
    m = sc.build_model()
    np_ = m.sizes()["np"]
    xz, yz, iref = m.table("z_xz"), m.table("z_yz"), m.table("z_flags_iref")
    area = m.table("cell_area")
    pick = np.arange(0, np_, max(1, np_ // 17))
    idx = (ctypes.c_int * len(pick))()
    xs = (ctypes.c_double * len(pick))(*xz[pick].astype(np.float64))
    ys = (ctypes.c_double * len(pick))(*yz[pick].astype(np.float64))
    assert host.get_sfincs_cell_indices(xs, ys, idx, ctypes.c_int(len(pick))) == 0
    assert list(idx) == list(pick + 1)
    one = ctypes.c_int()
    assert host.get_sfincs_cell_index(ctypes.byref(ctypes.c_double(-1.0e6)), ctypes.byref(ctypes.c_double(0.0)),
                                      ctypes.byref(one)) == 0 and one.value == 0
    a = ctypes.c_double()
    c = ctypes.c_int(int(pick[-1]) + 1)
    assert host.get_sfincs_cell_area(ctypes.byref(c), ctypes.byref(a)) == 0
    assert a.value == float(area[iref[pick[-1]] - 1])
    shape = (ctypes.c_int * 6)()
    assert host.get_var_shape(b"z_xz", shape) == 0 and shape[0] == np_
    tbuf = ctypes.create_string_buffer(64)
    assert host.get_var_type(b"zs", tbuf) == 0 and tbuf.value == b"double"
    rank = ctypes.c_int()
    assert host.get_var_rank(b"zb", ctypes.byref(rank)) == 0 and rank.value == 1
    for name in (b"z_index_z_n", b"z_index_z_m"):   # shape and type only, as the reference (:150-176)
        assert host.get_var_shape(name, shape) == 0 and shape[0] == np_
        assert host.get_var_type(name, tbuf) == 0 and tbuf.value == b"integer"
    ptr = ctypes.c_void_p()
    assert host.get_value_ptr(b"z_yz", ctypes.byref(ptr)) == 0
    got = np.ctypeslib.as_array((ctypes.c_float * np_).from_address(ptr.value))
    assert np.array_equal(got, yz)
    assert host.get_value_ptr(b"qext", ctypes.byref(ptr)) == -1
    assert host.update_zbuv() == -1


@needs_host
@pytest.mark.gpu
def test_bmi_qext_coupling(tmp_path):
    """A coupler's external sources through the BMI (sfincs_bmi.f90:60, 95-129): initialize
    carries a qext field, get_value_ptr("qext") hands it out, update_until steps with what
    was written there; equal to the Python-driven run given the same field."""
    print("case bmi qext coupling", flush=True)
    from sfincs_amd import _lib
    sc = W.regular_case(28, 36, dx=50.0, seed=48, t_end=1800.0, tide_amp=1.5)
    d = deck.write(sc, str(tmp_path))
    host = ctypes.CDLL(HOSTLIB)
    host.update_until.argtypes = [ctypes.c_double]
    cwd = os.getcwd()
    os.chdir(d)
    try:
        assert host.initialize() == 0
        np_ = sc.build_model().sizes()["np"]
        ptr = ctypes.c_void_p()
        assert host.get_value_ptr(b"qext", ctypes.byref(ptr)) == 0
        qext = np.ctypeslib.as_array((ctypes.c_float * np_).from_address(ptr.value))
        field = np.random.default_rng(49).uniform(-2e-5, 5e-5, np_).astype(np.float32)
        qext[:] = field
        assert host.update_until(ctypes.c_double(900.0)) == 0
        assert host.get_value_ptr(b"zs", ctypes.byref(ptr)) == 0
        zs_f = np.ctypeslib.as_array((ctypes.c_double * np_).from_address(ptr.value)).copy()
        assert host.finalize() == 0
    finally:
        os.chdir(cwd)
    sc.params["use_qext"] = 1
    inst = _lib.Instance(sc.build_model())
    inst.set("qext", field)
    inst.update(900.0)
    assert np.array_equal(zs_f, inst.get("zs"))


@needs_host
@pytest.mark.gpu
def test_bmi_bed_and_level_writes(tmp_path):
    """A coupler writes the bed through get_value_ptr("zb") and calls update_zbuv
    (sfincs_bmi.f90:277-284), later raises the water level through get_value_ptr("zs"):
    the host uploads both before stepping, and the run equals the Python-driven run given
    the same writes (sfx_set_state "zb" + sfx_update_zbuvmx, "zs")."""
    print("case bmi zb / zs writes", flush=True)
    from sfincs_amd import _lib
    sc = W.regular_case(28, 36, dx=50.0, seed=50, t_end=1800.0, tide_amp=1.5)
    d = deck.write(sc, str(tmp_path))
    host = ctypes.CDLL(HOSTLIB)
    host.update_until.argtypes = [ctypes.c_double]
    np_ = sc.build_model().sizes()["np"]
    r = np.random.default_rng(51)
    bump = np.zeros(np_, np.float32)
    bump[r.choice(np_, 60, replace=False)] = 2.0
    cwd = os.getcwd()
    os.chdir(d)
    try:
        assert host.initialize() == 0
        ptr = ctypes.c_void_p()
        assert host.update_until(ctypes.c_double(300.0)) == 0
        assert host.get_value_ptr(b"zb", ctypes.byref(ptr)) == 0
        zb = np.ctypeslib.as_array((ctypes.c_float * np_).from_address(ptr.value))
        zb_new = (zb + bump).astype(np.float32)
        zb[:] = zb_new
        assert host.update_zbuv() == 0
        assert host.update_until(ctypes.c_double(900.0)) == 0
        assert host.get_value_ptr(b"zs", ctypes.byref(ptr)) == 0
        zs = np.ctypeslib.as_array((ctypes.c_double * np_).from_address(ptr.value))
        zs_mid = zs.copy()
        zs_new = zs_mid + 0.05
        zs[:] = zs_new
        assert host.update_until(ctypes.c_double(1500.0)) == 0
        zs_f = zs.copy()                     # the live copy shows the model's levels again
        assert host.finalize() == 0
    finally:
        os.chdir(cwd)
    inst = _lib.Instance(sc.build_model())
    inst.update_to(300.0)
    inst.set("zb", zb_new)
    inst.update_zbuvmx()
    inst.update_to(900.0)
    assert np.array_equal(inst.get("zs"), zs_mid)
    inst.set("zs", zs_new)
    inst.update_to(1500.0)
    assert np.array_equal(zs_f, inst.get("zs"))
    assert np.array_equal(inst.get("zb"), zb_new)


@needs_host
@pytest.mark.gpu
def test_fortran_host_scheduled_output(tmp_path):
    """dtout, dtmaxout and dtrstout in sfincs.inp (sfincs_lib.f90:419-489,
    sfincs_output.f90:80-258, 457-519, 657-729): zs.dat holds zs at the first step reaching
    t0out, t0out + dtout, ...; zsmax.dat the maxima of each dtmaxout interval with dry
    cells at -999, after which the maxima start again from -999; sfincs.<date>.rst the
    restart at dtrstout. Equal to the Python-driven run stopped at the same times."""
    print("case fortran host scheduled output", flush=True)
    from sfincs_amd import _lib
    sc = W.regular_case(28, 36, dx=50.0, seed=52, t_end=1800.0, tide_amp=2.0)
    d = deck.write(sc, str(tmp_path))
    with open(os.path.join(d, "sfincs.inp"), "a") as f:
        f.write("dtout = 300\ndtmaxout = 600\ndtrstout = 900\noutputformat = bin\n")
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = sc.build_model()
    zb = m.table("zb")
    inst = _lib.Instance(m)
    zs_rec = records(os.path.join(d, "zs.dat"), np.float64)
    mx_rec = records(os.path.join(d, "zsmax.dat"), np.float32)
    assert len(zs_rec) == 7 and len(mx_rec) == 3
    inst.update(0.0)                                   # t0out = t0: the first step
    assert np.array_equal(zs_rec[0], inst.get("zs"))
    k_max = 0
    for k in range(1, 7):
        inst.update_to(300.0 * k)
        assert np.array_equal(zs_rec[k], inst.get("zs")), k
        if k == 3:
            rst = tmp_path / "py.rst"
            inst.write_restart(rst)
            assert open(rst, "rb").read() == open(os.path.join(d, "sfincs.20000101.001500.rst"), "rb").read()
        if k % 2 == 0:
            zmx = inst.get("zsmax")
            zmx = np.where(zmx < zb + np.float32(0.05), np.float32(-999.0), zmx).astype(np.float32)
            assert np.array_equal(mx_rec[k_max], zmx), k
            assert np.count_nonzero(zmx > -999.0) > 0
            k_max += 1
            inst.set("zsmax", np.full(m.sizes()["np"], -999.0, np.float32))
    assert np.array_equal(np.fromfile(os.path.join(d, "zs_final.dat"), np.float64), inst.get("zs"))


def test_deck_reader_refuses_an_infiltration_map_on_a_quadtree_mesh(tmp_path):
    """As the reference (sfincs_infiltration.f90:181-185): a quadtree mesh takes its
    infiltration map from the NetCDF infiltrationfile only."""
    sc = W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=42, t_end=900.0)
    sc.rain = (np.array([0.0, 1.0e4], np.float32), np.array([10.0, 10.0], np.float32))
    W.with_infiltration(sc, qinf_map=lambda m: np.ones(m.sizes()["np"], np.float32))
    d = deck.write(sc, str(tmp_path))
    host = ctypes.CDLL(HOSTLIB)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        assert host.sfincs_host_read_model() != 0
    finally:
        os.chdir(cwd)


@needs_host
@pytest.mark.gpu
def test_fortran_host_writes_history_at_observation_points(tmp_path):
    """obsfile and dthisout (read_obs_points, sfincs_obspoints.f90:5-135; write_his_output,
    sfincs_output.f90:593-655; trigger sfincs_lib.f90:470-477): zst.txt holds t and the
    level at each observation point (-999 for a point outside the mesh) at the first step
    reaching t0, t0 + dthisout, ...; prcp.txt the rain of the step's series in mm/h;
    qdrain.txt the drainage structures' intake discharges. Equal, to the files' three
    decimals, to the Python-driven run stopped at the same times."""
    print("case fortran host history output", flush=True)
    from sfincs_amd import _lib
    sc = W.beach_drainage(W.regular_case(40, 48, dx=50.0, seed=71, t_end=900.0, tide_amp=2.5, rain_mmh=30.0))
    d = deck.write(sc, str(tmp_path))
    pts = [(610.0, 420.0), (1210.0, 1630.0), (2230.0, 1010.0), (9.0e5, 9.0e5)]
    with open(os.path.join(d, "sfincs.obs"), "w") as f:
        for k, (x, y) in enumerate(pts):
            f.write(f"{x} {y} 'gauge{k}'\n")
    with open(os.path.join(d, "sfincs.inp"), "a") as f:
        f.write("obsfile = sfincs.obs\ndthisout = 300\n")
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    zst = np.loadtxt(os.path.join(d, "zst.txt"))
    prcp = np.loadtxt(os.path.join(d, "prcp.txt"))
    qdr = np.loadtxt(os.path.join(d, "qdrain.txt"))
    assert zst.shape == (4, 1 + len(pts)) and prcp.shape == zst.shape and qdr.shape[0] == 4
    m = sc.build_model()
    nm = m.find_cells([p[0] for p in pts], [p[1] for p in pts])
    assert nm[-1] == 0 and np.all(nm[:-1] > 0)
    s = m.sizes()
    inst = _lib.Instance(m)
    tp, vp = (np.asarray(a, np.float32) for a in sc.rain)
    for k in range(4):
        if k == 0:
            rc, st = inst.update(0.0)
        else:
            rc, st = inst.update_to(300.0 * k)
        t, dt = st["t"], st["dt"]
        assert zst[k, 0] == round(max(300.0 * k, t - dt), 1)
        zs = inst.get("zs").astype(np.float32)
        want = np.where(nm > 0, zs[np.maximum(nm, 1) - 1], np.float32(-999.0))
        assert np.max(np.abs(zst[k, 1:] - want)) <= 5.1e-4, (k, zst[k], want)
        i = int(np.searchsorted(tp, t, side="right"))
        i = max(min(i, len(tp) - 1), 1)
        tw = np.float32((t - float(tp[i - 1])) / float(tp[i] - tp[i - 1]))
        rain = np.float32(np.float32(vp[i - 1] * np.float32(1 - tw) + vp[i] * tw) / np.float32(3.6e6)) * np.float32(3.6e6)
        assert np.all(np.abs(prcp[k, 1:-1] - rain) <= 5.1e-4) and prcp[k, -1] == -999.0
        q = inst.get("qtsrc")[s["nsrc"]::2]
        assert np.max(np.abs(qdr[k, 1:] - q)) <= 5.1e-4, (k, qdr[k], q)
