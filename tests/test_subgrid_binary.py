"""The legacy binary subgrid file (read_subgrid_file_original, sfincs_subgrid.F90:458-903).

* Known answer: a hand-written file for a 3 x 2 regular mesh, and every derived table
  checked against the reader's rules: level 1 of the cell levels max(zmin, -20), level 1
  of havg = huthresh (kept at its input value on this path), nrep = g max(navg, 1e-4)^2
  with level 1 a copy of level 2, pwet = 1, no fit, uv zmin raised to both cells' zmin,
  +0.001 when zmax and zmin lie closer than 0.001, havg_zmax = havg(nl) - zmax and
  navg_w = nrep_zmax (:873-900).
* A file whose size does not match the mesh is refused.
* The oracle against the reference's own hot path on a legacy-file subgrid mesh (huthresh
  0.05, which the NetCDF path never has), bit for bit after every step.
* The writer and reader agree on a generated configuration-1 style mesh: the tables read
  back equal the legacy rules applied to the generated tables.
"""
import numpy as np
import pytest

import oracle as O
from sfincs_amd import _lib
from sfincs_amd import workloads as W

G = np.float32(9.81)


def _mesh(nmax=3, mmax=2, **over):
    p = _lib.default_params(**over)
    return _lib.Model.regular(p, nmax, mmax, 10.0, 10.0, np.ones(nmax * mmax, np.int8))


def _write(path, np_, nbins, z, u, v, pad=0):
    with open(path, "wb") as f:
        np.array([np_, 0, nbins], np.int32).tofile(f)
        for a in z:
            np.asarray(a, np.float32).tofile(f)
        for rec in (u, v):
            for a in rec:
                np.asarray(a, np.float32).tofile(f)
        if pad:
            np.zeros(pad, np.float32).tofile(f)


def test_known_answer_tables(tmp_path):
    m = _mesh()
    npc = 6
    nb = 2                                    # 3 levels
    zmin = np.array([-1.0, -30.0, 0.5, 2.0, 3.0, 4.0], np.float32)
    zmax = np.array([1.0, -25.0, 0.5004, 4.0, 5.0, 6.0], np.float32)
    volmax = np.array([100.0, 200.0, 0.01, 300.0, 400.0, 500.0], np.float32)
    dep2 = zmin + 0.5
    dep3 = zmax
    # per cell: its right u-point / upper v-point
    uzmin = np.array([-2.0, 0.0, 1.0, 2.5, 0.0, 0.0], np.float32)
    uzmax = np.array([3.0, 0.0, 1.0005, 6.0, 0.0, 0.0], np.float32)
    uh = [np.full(npc, 0.3, np.float32), np.full(npc, 0.8, np.float32)]
    un = [np.array([0.02, 0, 0.00001, 0.05, 0, 0], np.float32), np.array([0.03, 0, 0.04, 0.06, 0, 0], np.float32)]
    vzmin = np.array([-0.5, -40.0, 0.0, 2.0, 3.5, 0.0], np.float32)
    vzmax = np.array([0.5, -20.0, 2.0, 2.0005, 5.5, 0.0], np.float32)
    vh = [np.full(npc, 0.2, np.float32), np.full(npc, 0.9, np.float32)]
    vn = [np.full(npc, 0.025, np.float32), np.full(npc, 0.035, np.float32)]
    path = tmp_path / "sfincs.sbg"
    _write(path, npc, nb, [zmin, zmax, volmax, dep2, dep3],
           [uzmin, uzmax, np.zeros(npc), uh[0], uh[1], un[0], un[1]],
           [vzmin, vzmax, np.zeros(npc), vh[0], vh[1], vn[0], vn[1]])
    m.read_subgrid_binary(path)
    m.initialize()
    assert abs(m.params().huthresh - 0.05) < 1e-9          # kept on this path
    nl = 3
    zdep = m.table("z_dep").reshape(npc, nl)
    assert np.array_equal(zdep[:, 0], np.maximum(zmin, np.float32(-20.0)))
    assert np.array_equal(zdep[:, 1], dep2) and np.array_equal(zdep[:, 2], dep3)
    zmx = m.table("z_zmax")
    expect_zmax = np.where(zmax - zmin < np.float32(0.001), zmax + np.float32(0.001), zmax).astype(np.float32)
    assert np.array_equal(zmx, expect_zmax)
    # uv-points: which cell and direction each belongs to
    nm = m.table("uv_index_z_nm") - 1
    nmu = m.table("uv_index_z_nmu") - 1
    d = m.table("uv_flags_dir")
    npuv = len(nm)
    havg = m.table("uv_havg").reshape(npuv, nl)
    nrep = m.table("uv_nrep").reshape(npuv, nl)
    assert np.all(havg[:, 0] == np.float32(0.05))
    assert np.all(m.table("uv_pwet") == 1.0) and np.all(m.table("uv_fnfit") == 0.0)
    for i in range(npuv):
        c = nm[i]
        src = (uzmin, uzmax, uh, un) if d[i] == 0 else (vzmin, vzmax, vh, vn)
        zmn = max(src[0][c], zmin[c], zmin[nmu[i]])
        zmx_uv = src[1][c]
        if zmx_uv - zmn < np.float32(0.001):
            zmx_uv = np.float32(zmx_uv + np.float32(0.001))
        assert m.table("uv_zmin")[i] == np.float32(zmn), i
        assert m.table("uv_zmax")[i] == np.float32(zmx_uv), i
        assert havg[i, 1] == src[2][0][c] and havg[i, 2] == src[2][1][c]
        for lv in (1, 2):
            n = np.float32(max(src[3][lv - 1][c], np.float32(0.0001)))
            assert nrep[i, lv] == np.float32(G * np.float32(n * n)), (i, lv)
        assert nrep[i, 0] == nrep[i, 1]
        assert m.table("uv_havg_zmax")[i] == np.float32(havg[i, 2] - m.table("uv_zmax")[i])
        assert m.table("uv_nrep_zmax")[i] == nrep[i, 2] and m.table("uv_navg_w")[i] == nrep[i, 2]


def test_wrong_size_is_refused(tmp_path):
    m = _mesh()
    path = tmp_path / "bad.sbg"
    z = [np.zeros(6)] * 5
    rec = [np.zeros(6)] * 7
    _write(path, 6, 2, z, rec, rec, pad=3)
    with pytest.raises(_lib.SfxError, match="bytes"):
        m.read_subgrid_binary(path)
    _write(path, 6, 2, z[:4], rec, rec)
    with pytest.raises(_lib.SfxError):
        _mesh().read_subgrid_binary(path)


def _legacy_case(tmp_path, seed=41, **kw):
    sc = W.subgrid_case(30, 26, dx=25.0, seed=seed, t_end=1800.0, nsrc=2, rain_mmh=25.0, **kw)
    return W.with_subgrid_binary(sc, tmp_path / f"legacy{seed}.sbg")


def test_tables_read_back_follow_the_legacy_rules(tmp_path):
    sc = W.subgrid_case(24, 20, dx=25.0, seed=43, t_end=600.0, nsrc=0)
    m0 = sc.build_model()                           # NetCDF path, for the generated tables
    sg = sc.subgrid(m0)
    m = W.with_subgrid_binary(sc, tmp_path / "rb.sbg").build_model()
    nl = int(sg["nlevels"])
    npuv = m.sizes()["npuv"]
    havg = m.table("uv_havg").reshape(npuv, nl)
    assert np.array_equal(havg[:, 1:], np.asarray(sg["uv_havg"], np.float32).reshape(npuv, nl)[:, 1:])
    n = np.maximum(np.asarray(sg["uv_nrep"], np.float32).reshape(npuv, nl)[:, 1:], np.float32(0.0001))
    assert np.array_equal(m.table("uv_nrep").reshape(npuv, nl)[:, 1:], (G * (n * n)).astype(np.float32))
    zl = np.asarray(sg["z_level"], np.float32).reshape(-1, nl)
    assert np.array_equal(m.table("z_dep").reshape(-1, nl)[:, 1:], zl[:, 1:])
    assert m.sizes()["nlevels"] == nl


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs /root/reference)")
def test_legacy_subgrid_oracle_bit_exact_with_reference(tmp_path):
    sc = _legacy_case(tmp_path)
    m = sc.build_model()
    assert abs(m.params().huthresh - 0.05) < 1e-9
    a = O.OracleRun(m).configure(sc.params, m)
    b = O.OracleRun(m).configure(sc.params, m)
    assert abs(a.vals["huthresh"] - 0.05) < 1e-7
    r = O.RefRun(b)
    for i in range(300):
        a.update(1e30, 1)
        r.step()
        for n in ("zs", "q", "uv", "kfuv", "zsmax", "z_volume", "zsderv", "zs0"):
            assert np.array_equal(a.state(n), b.state(n)), (i, n)
        assert a.get("min_dt") == b.get("min_dt")
    assert np.count_nonzero(a.state("kfuv")) > 10
