"""Quadtree meshes: the generator's output and the host builder's tables, pinned by
geometry (the reference's quadtree builder needs netCDF and cannot run here):
leaves tile the domain, faces are 2:1 balanced, every uv-point is a face between two
face-adjacent cells and every such face has exactly one uv-point; combined points
average the two fine faces of a coarse side; one level reduces to the regular mesh."""
import numpy as np
import pytest

from sfincs_amd import _lib
from sfincs_amd import quadtree as Q
from sfincs_amd import workloads as W


def small_case(seed=5, nlev=3):
    return W.quadtree_case(10, 14, nlev=nlev, dx0=100.0, seed=seed, t_end=3600.0, surge_amp=4.0)


def cell_rect(level, n, m, dx0):
    s = dx0 / 2 ** (level - 1)
    return (m - 1) * s, m * s, (n - 1) * s, n * s     # x0, x1, y0, y1


def test_leaves_tile_domain_and_are_balanced():
    sc = small_case()
    qt = sc.qt
    s = 100.0 / 2 ** (qt["level"].astype(float) - 1)
    assert np.isclose(np.sum(s * s), 10 * 14 * 100.0 * 100.0)
    for side in ("md", "mu", "nd", "nu"):
        for k in range(len(qt["n"])):
            j = qt[side + "1"][k]
            if j:
                assert abs(int(qt["level"][j - 1]) - int(qt["level"][k])) <= 1


def test_neighbour_arrays_share_faces():
    sc = small_case()
    qt = sc.qt
    for k in range(len(qt["n"])):
        x0, x1, y0, y1 = cell_rect(qt["level"][k], qt["n"][k], qt["m"][k], 100.0)
        for side in ("mu", "md", "nu", "nd"):
            for j in (qt[side + "1"][k], qt[side + "2"][k]):
                if not j:
                    continue
                a0, a1, b0, b1 = cell_rect(qt["level"][j - 1], qt["n"][j - 1], qt["m"][j - 1], 100.0)
                if side == "mu":
                    assert np.isclose(a0, x1) and max(b0, y0) < min(b1, y1)
                if side == "md":
                    assert np.isclose(a1, x0) and max(b0, y0) < min(b1, y1)
                if side == "nu":
                    assert np.isclose(b0, y1) and max(a0, x0) < min(a1, x1)
                if side == "nd":
                    assert np.isclose(b1, y0) and max(a0, x0) < min(a1, x1)


def test_uv_points_are_the_faces():
    sc = small_case()
    m = sc.build_model()
    s = m.sizes()
    lev, n, mm = m.table("z_flags_iref"), m.table("z_index_z_n"), m.table("z_index_z_m")
    rect = [cell_rect(lev[c], n[c], mm[c], 100.0) for c in range(s["np"])]
    nm_t, nmu_t, dirs = m.table("uv_index_z_nm"), m.table("uv_index_z_nmu"), m.table("uv_flags_dir")
    faces = set()
    for i in range(s["npuv"]):
        a, b = nm_t[i] - 1, nmu_t[i] - 1
        ra, rb = rect[a], rect[b]
        if dirs[i] == 0:
            assert np.isclose(rb[0], ra[1]) and max(ra[2], rb[2]) < min(ra[3], rb[3])
        else:
            assert np.isclose(rb[2], ra[3]) and max(ra[0], rb[0]) < min(ra[1], rb[1])
        faces.add((a, b))
    assert len(faces) == s["npuv"]
    # every geometric face between two cells has a uv-point (vectorised pair test)
    R = np.array(rect)
    ax1, bx0 = R[:, 1][:, None], R[:, 0][None, :]
    ay1, by0 = R[:, 3][:, None], R[:, 2][None, :]
    ovy = np.minimum(R[:, 3][:, None], R[:, 3][None, :]) - np.maximum(R[:, 2][:, None], R[:, 2][None, :])
    ovx = np.minimum(R[:, 1][:, None], R[:, 1][None, :]) - np.maximum(R[:, 0][:, None], R[:, 0][None, :])
    xf = np.isclose(bx0, ax1) & (ovy > 1e-9)
    yf = np.isclose(by0, ay1) & (ovx > 1e-9)
    pairs = set(map(tuple, np.argwhere(xf | yf)))
    assert pairs == faces

    # combined points: the two fine faces along one coarse side, and the coarse cell's
    # z_index_uv on that side points at the combined point
    ncuv = s["ncuv"]
    assert ncuv > 0
    c, u1, u2 = m.table("cuv_index_uv"), m.table("cuv_index_uv1"), m.table("cuv_index_uv2")
    assert np.array_equal(c, np.arange(s["npuv"] + 1, s["npuv"] + ncuv + 1))
    types = m.table("uv_flags_type")
    for j in range(ncuv):
        f1, f2 = u1[j] - 1, u2[j] - 1
        assert types[f1] != 0 and types[f2] != 0 and dirs[f1] == dirs[f2]
        shared = {nm_t[f1], nmu_t[f1]} & {nm_t[f2], nmu_t[f2]}
        assert len(shared) == 1


def test_single_level_quadtree_reduces_to_regular():
    r = np.random.default_rng(3)
    nmax, mmax = 7, 9
    dem = r.uniform(-10, 10, (nmax, mmax)).astype(np.float32)
    _, leaf, _, _ = Q.build(dem, 1)
    qt = Q.quadtree_arrays(leaf, 1, 0.0, 0.0, 10.0, 10.0)
    p = _lib.default_params(t0=0.0, t1=60.0)
    mq = _lib.Model.quadtree(p, qt)
    mr = _lib.Model.regular(p, nmax, mmax, 10.0, 10.0, np.ones(nmax * mmax, np.int8))
    for name in ("uv_index_z_nm", "uv_index_z_nmu", "uv_index_u_nmd", "uv_index_u_nmu", "uv_index_u_ndm",
                 "uv_index_u_num", "uv_index_v_ndm", "uv_index_v_nm", "uv_index_v_ndmu", "uv_index_v_nmu",
                 "z_index_uv_md", "z_index_uv_mu", "z_index_uv_nd", "z_index_uv_nu", "uv_flags_dir"):
        assert np.array_equal(mq.table(name), mr.table(name)), name
    assert mq.sizes()["quadtree"] == 0      # one level reverts to the regular path
