"""Index tables and masks of the host model builder, pinned without running the reference
(its mesh builder needs netCDF and cannot be built here):

* full regular grids: every table against its closed form (a different formulation
  from the builder's loops over cells): cell nm = (m-1)*nmax + n, uv-points per cell
  right-u then above-v, neighbour rules of sfincs_domain.f90:1005-1261;
* grids with inactive cells: geometric invariants that any correct table must satisfy;
* kcuv / mask_adv: closed forms for a west water-level boundary column.
"""
import numpy as np
import pytest

from sfincs_amd import _lib
from sfincs_amd import workloads as W


def regular_model(nmax, mmax, msk=None, **over):
    p = _lib.default_params(t0=0.0, t1=60.0, **over)
    if msk is None:
        msk = np.ones((mmax, nmax), np.int8)
    m = _lib.Model.regular(p, nmax, mmax, 10.0, 10.0, msk.reshape(-1))
    return m


def closed_form(nmax, mmax):
    """Reference uv numbering of a full nmax x mmax grid (1-based)."""
    per_col = 2 * nmax - 1

    def ip_u(n, m):            # u right of (n, m), m < mmax
        return (m - 1) * per_col + 2 * (n - 1) + 1

    def ip_v(n, m):            # v above (n, m), n < nmax
        if m < mmax:
            return (m - 1) * per_col + 2 * (n - 1) + 2
        return (mmax - 1) * per_col + (n - 1) + 1
    return ip_u, ip_v


@pytest.mark.parametrize("nmax,mmax", [(5, 7), (9, 4), (1, 6), (6, 1), (17, 13)])
def test_full_grid_tables_closed_form(nmax, mmax):
    m = regular_model(nmax, mmax)
    s = m.sizes()
    npuv = nmax * (mmax - 1) + (nmax - 1) * mmax
    assert s["np"] == nmax * mmax and s["npuv"] == npuv and s["ncuv"] == 0 and s["dense"] == 1
    dummy = npuv + 1
    ip_u, ip_v = closed_form(nmax, mmax)
    nm_t, nmu_t = m.table("uv_index_z_nm"), m.table("uv_index_z_nmu")
    dirs = m.table("uv_flags_dir")
    md, mu, nd, nu = (m.table(k) for k in ("z_index_uv_md", "z_index_uv_mu", "z_index_uv_nd", "z_index_uv_nu"))
    nb = {k: m.table("uv_index_" + k) for k in ("u_nmd", "u_nmu", "u_ndm", "u_num", "v_ndm", "v_nm", "v_ndmu", "v_nmu")}
    cell = lambda n, mm: (mm - 1) * nmax + n
    U = lambda n, mm: ip_u(n, mm) if (1 <= n <= nmax and 1 <= mm < mmax) else dummy
    Vv = lambda n, mm: ip_v(n, mm) if (1 <= n < nmax and 1 <= mm <= mmax) else dummy
    for mm in range(1, mmax + 1):
        for n in range(1, nmax + 1):
            c = cell(n, mm) - 1
            assert md[c] == U(n, mm - 1) and mu[c] == U(n, mm)
            assert nd[c] == Vv(n - 1, mm) and nu[c] == Vv(n, mm)
            if mm < mmax:
                i = ip_u(n, mm) - 1
                assert (nm_t[i], nmu_t[i], dirs[i]) == (cell(n, mm), cell(n, mm + 1), 0)
                assert nb["u_nmd"][i] == U(n, mm - 1) and nb["u_nmu"][i] == U(n, mm + 1)
                assert nb["u_ndm"][i] == (U(n - 1, mm) if n > 1 else i + 1)
                assert nb["u_num"][i] == (U(n + 1, mm) if n < nmax else i + 1)
                assert nb["v_nm"][i] == Vv(n, mm) and nb["v_ndm"][i] == Vv(n - 1, mm)
                assert nb["v_nmu"][i] == Vv(n, mm + 1) and nb["v_ndmu"][i] == Vv(n - 1, mm + 1)
            if n < nmax:
                i = ip_v(n, mm) - 1
                assert (nm_t[i], nmu_t[i], dirs[i]) == (cell(n, mm), cell(n + 1, mm), 1)
                assert nb["u_nmd"][i] == Vv(n - 1, mm) and nb["u_nmu"][i] == Vv(n + 1, mm)
                assert nb["u_ndm"][i] == (Vv(n, mm - 1) if mm > 1 else i + 1)
                assert nb["u_num"][i] == (Vv(n, mm + 1) if mm < mmax else i + 1)
                assert nb["v_nm"][i] == U(n, mm) and nb["v_ndm"][i] == U(n, mm - 1)
                assert nb["v_nmu"][i] == U(n + 1, mm) and nb["v_ndmu"][i] == U(n + 1, mm - 1)


def test_kcuv_and_mask_adv_closed_form():
    nmax, mmax = 6, 8
    msk = np.ones((mmax, nmax), np.int8)
    msk[0, :] = 2                                  # west water-level boundary column
    m = regular_model(nmax, mmax, msk)
    m.set_bathymetry(np.zeros(nmax * mmax, np.float32))
    m.set_boundary(np.array([-5.0], np.float32), np.array([30.0], np.float32))
    m.initialize()
    ip_u, ip_v = closed_form(nmax, mmax)
    kcuv = m.table("kcuv")
    ma = m.table("mask_adv")
    assert len(kcuv) == m.sizes()["npuv"] + 1 and kcuv[-1] == 0
    for mm in range(1, mmax + 1):
        for n in range(1, nmax + 1):
            if mm < mmax:
                i = ip_u(n, mm) - 1
                assert kcuv[i] == (2 if mm == 1 else 1)
                interior = (mm >= 3 and mm + 1 < mmax and 1 < n < nmax)
                assert ma[i] == (1 if interior else 0), (n, mm)
            if n < nmax:
                i = ip_v(n, mm) - 1
                assert kcuv[i] == (0 if mm == 1 else 1)
                interior = (mm >= 3 and mm < mmax and 1 < n < nmax - 1)
                assert ma[i] == (1 if interior else 0), (n, mm)
    # boundary uv-points: the west faces, boundary cell on the left (ibuvdir = 1)
    assert np.array_equal(m.table("ibuvdir"), np.ones(nmax, np.int8))
    assert np.array_equal(m.table("index_kcuv2"), np.array([ip_u(n, 1) for n in range(1, nmax + 1)], np.int32))


def test_masked_grid_invariants():
    r = np.random.default_rng(5)
    nmax, mmax = 23, 31
    msk = (r.random((mmax, nmax)) > 0.25).astype(np.int8)
    m = regular_model(nmax, mmax, msk)
    s = m.sizes()
    assert s["dense"] == 0
    act = np.argwhere(msk.T > 0)                   # (n, m) zero-based, column-major order below
    order = sorted(((mm, n) for n, mm in act))
    cells = {(n, mm): k + 1 for k, (mm, n) in enumerate(order)}
    assert s["np"] == len(cells)
    nm_t, nmu_t, dirs = m.table("uv_index_z_nm"), m.table("uv_index_z_nmu"), m.table("uv_flags_dir")
    faces = set()
    nn = m.table("z_index_z_n"); mmv = m.table("z_index_z_m")
    for i in range(s["npuv"]):
        a, b = nm_t[i] - 1, nmu_t[i] - 1
        da = (nn[b] - nn[a], mmv[b] - mmv[a])
        assert da == ((0, 1) if dirs[i] == 0 else (1, 0))
        faces.add((a, b))
    expect = set()
    for (n, mm), k in cells.items():
        if (n, mm + 1) in cells:
            expect.add((k - 1, cells[(n, mm + 1)] - 1))
        if (n + 1, mm) in cells:
            expect.add((k - 1, cells[(n + 1, mm)] - 1))
    assert faces == expect and len(faces) == s["npuv"]
    # uv-points are numbered in cell order, right before above
    key = [(nm_t[i], dirs[i]) for i in range(s["npuv"])]
    assert key == sorted(key)


def test_dense_recipes_are_dense():
    assert W.regular_case(12, 9, t_end=60.0).build_model().sizes()["dense"] == 1
