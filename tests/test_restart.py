"""Binary restart files as initial conditions (sfx_model_set_restart): the reader of
read_binary_restart_file (sfincs_initial_conditions.F90:183-260) and the initial state the
reference then builds from it (:98-174, sfincs_domain.f90:2269-2311, 2396).

The reference holds no restart fixture and its reader sits in a module that needs the
NetCDF library, so these are known-answer tests on files written here in the documented
layout (Fortran sequential records: length, payload, length; type, zs f32, q f32 of
npuv+ncuv+1, uvmean f32 of nkcuv2), plus the special case that pins the rest: a restart
holding one uniform level equals initialising with zsini at that level."""
import numpy as np
import pytest

from sfincs_amd import _lib
from sfincs_amd import workloads as W


def record(a):
    b = np.ascontiguousarray(a).tobytes()
    n = np.array([len(b)], np.int32).tobytes()
    return n + b + n


def write_rst(path, rsttype, zs4=None, q=None, uvmean=None):
    with open(path, "wb") as f:
        f.write(record(np.array([rsttype], np.int32)))
        if zs4 is not None:
            f.write(record(np.asarray(zs4, np.float32)))
        if q is not None:
            f.write(record(np.asarray(q, np.float32)))
        if uvmean is not None:
            f.write(record(np.asarray(uvmean, np.float32)))


def cases():
    return {
        "regular": lambda: W.regular_case(20, 24, dx=50.0, seed=81, t_end=600.0),
        "subgrid": lambda: W.subgrid_case(16, 20, dx=25.0, seed=82, t_end=600.0, nsrc=1),
        "quadtree": lambda: W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=83, t_end=600.0),
    }


@pytest.mark.parametrize("kind", list(cases()))
def test_type1_restart_sets_the_initial_state(kind, tmp_path):
    sc = cases()[kind]()
    base = sc.build_model()
    s = base.sizes()
    np_, npuv, ncuv, nk = s["np"], s["npuv"], s["ncuv"], s["nkcuv2"]
    r = np.random.default_rng(7)
    bed = base.table("z_zmin") if s["subgrid"] else base.table("zb")
    zs4 = (bed + r.uniform(-0.5, 1.5, np_)).astype(np.float32)     # some below the bed
    q = r.uniform(-0.3, 0.3, npuv + ncuv + 1).astype(np.float32)
    uvm = r.uniform(-0.1, 0.1, nk).astype(np.float32)
    path = tmp_path / "sfincs.rst"
    write_rst(path, 1, zs4, q, uvm)
    sc.rstfile = str(path)
    m = sc.build_model()
    zs = m.table("zs")
    assert np.array_equal(zs, np.maximum(bed.astype(np.float64), zs4.astype(np.float64)))
    mq = m.table("q")
    assert np.array_equal(mq[:npuv], q[:npuv])
    if ncuv:
        c = m.table("cuv_index_uv") - 1
        a, b = m.table("cuv_index_uv1") - 1, m.table("cuv_index_uv2") - 1
        assert np.array_equal(mq[c], (np.float32(0.5) * (mq[a] + mq[b])).astype(np.float32))
    assert mq[-1] == 0.0                                  # the dummy point
    assert not np.any(m.table("uv"))                      # uv = 0 after the restart (:2311)
    assert np.array_equal(m.table("uvmean"), uvm)
    assert np.array_equal(m.table("zs0"), zs.astype(np.float32))
    if s["subgrid"]:                                      # volumes follow the new levels (monotone)
        dz, dv = zs - base.table("zs"), m.table("z_volume") - base.table("z_volume")
        assert np.all(dv[dz > 0] >= 0) and np.all(dv[dz < 0] <= 0) and np.all(dv[dz == 0] == 0)


@pytest.mark.parametrize("kind", list(cases()))
def test_uniform_restart_equals_zsini(kind, tmp_path):
    """A restart at one level L (q = 0, uvmean = 0) gives the state of zsini = L, volumes
    included."""
    lev = 0.75
    sc = cases()[kind]()
    sc.params["zsini"] = lev
    ref = sc.build_model()
    s = ref.sizes()
    sc2 = cases()[kind]()
    path = tmp_path / "uniform.rst"
    write_rst(path, 1, np.full(s["np"], lev, np.float32), np.zeros(s["npuv"] + s["ncuv"] + 1, np.float32),
              np.zeros(s["nkcuv2"], np.float32))
    sc2.rstfile = str(path)
    m = sc2.build_model()
    names = ["zs", "q", "uv", "zs0", "uvmean"] + (["z_volume"] if s["subgrid"] else [])
    for n in names:
        assert np.array_equal(m.table(n), ref.table(n)), n


def test_type3_sets_levels_only(tmp_path):
    sc = cases()["regular"]()
    base = sc.build_model()
    s = base.sizes()
    zs4 = (base.table("zb") + 1.0).astype(np.float32)
    path = tmp_path / "t3.rst"
    write_rst(path, 3, zs4)
    sc.rstfile = str(path)
    m = sc.build_model()
    assert np.array_equal(m.table("zs"), zs4.astype(np.float64))
    assert not np.any(m.table("q")) and m.table("uvmean").size == s["nkcuv2"]


def test_unknown_type_is_skipped(tmp_path):
    """As the reference (a warning, then the normal initial conditions)."""
    sc = cases()["regular"]()
    base = sc.build_model()
    path = tmp_path / "t9.rst"
    write_rst(path, 9, np.ones(base.sizes()["np"], np.float32))
    sc.rstfile = str(path)
    m = sc.build_model()
    assert np.array_equal(m.table("zs"), base.table("zs"))


def test_bad_restart_files_are_errors(tmp_path):
    sc = cases()["quadtree"]()
    s = sc.build_model().sizes()
    bad = {
        "infiltration": (4, np.zeros(s["np"]), np.zeros(s["npuv"] + s["ncuv"] + 1), np.zeros(s["nkcuv2"])),
        "short_zs": (1, np.zeros(s["np"] - 1), np.zeros(s["npuv"] + s["ncuv"] + 1), np.zeros(s["nkcuv2"])),
        "old_q_length": (1, np.zeros(s["np"]), np.zeros(s["npuv"] + s["ncuv"]), np.zeros(s["nkcuv2"])),
        "short_uvmean": (1, np.zeros(s["np"]), np.zeros(s["npuv"] + s["ncuv"] + 1), np.zeros(s["nkcuv2"] + 1)),
    }
    for name, (t, zs4, q, uvm) in bad.items():
        path = tmp_path / f"{name}.rst"
        write_rst(path, t, zs4, q, uvm)
        sc.rstfile = str(path)
        with pytest.raises(_lib.SfxError):
            sc.build_model()
    path = tmp_path / "truncated.rst"
    write_rst(path, 1, np.zeros(s["np"]), np.zeros(s["npuv"] + s["ncuv"] + 1), np.zeros(s["nkcuv2"]))
    data = path.read_bytes()
    path.write_bytes(data[:-6])
    sc.rstfile = str(path)
    with pytest.raises(_lib.SfxError):
        sc.build_model()
    sc.rstfile = str(tmp_path / "missing.rst")
    with pytest.raises(_lib.SfxError):
        sc.build_model()
