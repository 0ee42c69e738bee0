"""The uv-point neighbour tables of quadtree and masked meshes, derived again from
geometry alone and compared bit for bit with the host builder's tables.

The reference builds them from the quadtree's neighbour-index arrays
(initialize_mesh, sfincs_domain.f90:727-1003 for the cell faces and :1007-1261 for the
eight uv neighbours, with the self / dummy fallbacks at :1236-1250). Its builder needs
netCDF and cannot run here, so the tables are pinned by an independent construction:
cells are squares, a uv-point is the segment its two cells share, a cell side is
represented by its one face, by the combined point of its two finer faces, or by none;
and each neighbour is found by locating a point just across a cell side. A swapped
fine cell (1 / 2), a wrong side or a wrong fallback changes the tables."""
import numpy as np
import pytest

from sfincs_amd import workloads as W

EPS = 1e-3


class Geometry:
    """Cells as squares (x0, x1, y0, y1) from their level and grid indices, and uv-points as
    the shared segment of their two cells, from a host model (rotation 0). `leaves` lists
    every leaf of the mesh file as (level, n, m, active), inactive ones included."""

    def __init__(self, m, dx0, leaves, x0=0.0, y0=0.0):
        s = m.sizes()
        self.m, self.s = m, s
        self.lev = m.table("z_flags_iref").astype(int)
        n, mm = m.table("z_index_z_n"), m.table("z_index_z_m")
        size = dx0 / 2.0 ** (self.lev - 1)
        self.rect = np.stack([x0 + (mm - 1) * size, x0 + mm * size, y0 + (n - 1) * size, y0 + n * size], 1)
        self.dx0, self.x0, self.y0 = dx0, x0, y0
        cell_of = {(int(self.lev[c]), int(n[c]), int(mm[c])): c + 1 for c in range(s["np"])}
        self.leaf = {}
        for L, ln, lm, act in leaves:
            c = cell_of.get((int(L), int(ln), int(lm)), 0)
            assert bool(c) == bool(act)
            self.leaf[(int(L), int(ln), int(lm))] = c
        self.nlev = max(k[0] for k in self.leaf)
        self.dummy = s["npuv"] + s["ncuv"] + 1
        nm, nmu, d = m.table("uv_index_z_nm"), m.table("uv_index_z_nmu"), m.table("uv_flags_dir")
        # faces on each side of each cell
        self.side = {k: [[] for _ in range(s["np"] + 1)] for k in ("md", "mu", "nd", "nu")}
        for i in range(s["npuv"]):
            a, b = int(nm[i]), int(nmu[i])
            if d[i] == 0:
                self.side["mu"][a].append(i + 1)
                self.side["md"][b].append(i + 1)
            else:
                self.side["nu"][a].append(i + 1)
                self.side["nd"][b].append(i + 1)
        self.nm, self.nmu, self.dir = nm, nmu, d
        # combined point of each fine face (a coarse side keeps its combined point when one of
        # its two finer neighbours is inactive; the missing half is the dummy)
        c1, c2 = m.table("cuv_index_uv1"), m.table("cuv_index_uv2")
        self.combined = {}
        for j, pair in enumerate(zip(c1, c2)):
            for f in pair:
                if f != self.dummy:
                    self.combined[int(f)] = s["npuv"] + j + 1

    def leaf_at(self, x, y):
        """(level, active cell 1-based or 0) of the leaf holding point (x, y); (0, 0) outside."""
        for L in range(1, self.nlev + 1):
            size = self.dx0 / 2.0 ** (L - 1)
            k = (L, int(np.floor((y - self.y0) / size)) + 1, int(np.floor((x - self.x0) / size)) + 1)
            if k in self.leaf:
                return L, self.leaf[k]
        return 0, 0

    def face_mid(self, ip):
        """Midpoint of uv-point ip's segment, along the side (y for u, x for v)."""
        a, b = self.rect[self.nm[ip - 1] - 1], self.rect[self.nmu[ip - 1] - 1]
        if self.dir[ip - 1] == 0:
            return 0.5 * (max(a[2], b[2]) + min(a[3], b[3]))
        return 0.5 * (max(a[0], b[0]) + min(a[1], b[1]))

    def finer_across(self, c, ip):
        """Whether face ip joins cell c to a finer cell."""
        other = self.nmu[ip - 1] if self.nm[ip - 1] == c else self.nm[ip - 1]
        return self.lev[other - 1] > self.lev[c - 1]

    def face(self, c, side):
        """A cell side as one uv index: its face when the cell across is as coarse or coarser,
        the combined point of the side's finer faces, or the dummy point when it has none."""
        f = self.side[side][c]
        if not f:
            return self.dummy
        if self.finer_across(c, f[0]):
            pts = {self.combined[i] for i in f}
            assert len(pts) == 1
            return pts.pop()
        assert len(f) == 1
        return f[0]

    def fine(self, c, side):
        """Slots 1 and 2 of a side: a face to a finer cell in the slot of the half of the side
        it lies on (lower / left first), any other face in slot 1; 0 where none."""
        out = [0, 0]
        r = self.rect[c - 1]
        mid = 0.5 * (r[2] + r[3]) if side in ("md", "mu") else 0.5 * (r[0] + r[1])
        for i in self.side[side][c]:
            k = (1 if self.face_mid(i) > mid else 0) if self.finer_across(c, i) else 0
            assert out[k] == 0
            out[k] = i
        return out

    def across(self, a, x, y, side):
        """The neighbour rule of the cross-direction u neighbours: the leaf holding (x, y),
        just across one side of cell a at that side's lower / left end. A leaf of a's level
        or finer gives its face on `side`; a coarser leaf, an inactive leaf or none gives 0
        (the reference reads the second neighbour index there, which the mesh file leaves 0
        for a coarser neighbour, sfincs_domain.f90:1031-1062)."""
        L, c = self.leaf_at(x, y)
        if c == 0 or L < self.lev[a - 1]:
            return 0
        return self.face(c, side)


def expected_tables(G):
    s = G.s
    npuv, dummy = s["npuv"], G.dummy
    out = {k: np.zeros(npuv, np.int64) for k in ("uv_index_u_nmd", "uv_index_u_nmu", "uv_index_u_ndm",
                                                  "uv_index_u_num", "uv_index_v_ndm", "uv_index_v_nm",
                                                  "uv_index_v_ndmu", "uv_index_v_nmu")}
    R = G.rect
    for ip in range(1, npuv + 1):
        a, b = int(G.nm[ip - 1]), int(G.nmu[ip - 1])
        ra = R[a - 1]
        if G.dir[ip - 1] == 0:
            # u-point: a left of the face, b right of it
            out["uv_index_u_nmd"][ip - 1] = G.face(a, "md")
            out["uv_index_u_nmu"][ip - 1] = G.face(b, "mu")
            out["uv_index_u_ndm"][ip - 1] = G.across(a, ra[0] + EPS, ra[2] - EPS, "mu")   # below a
            out["uv_index_u_num"][ip - 1] = G.across(a, ra[0] + EPS, ra[3] + EPS, "mu")   # above a
            out["uv_index_v_nm"][ip - 1] = G.face(a, "nu")
            out["uv_index_v_ndm"][ip - 1] = G.face(a, "nd")
            out["uv_index_v_nmu"][ip - 1] = G.face(b, "nu")
            out["uv_index_v_ndmu"][ip - 1] = G.face(b, "nd")
        else:
            # v-point: a below the face, b above it (the mirror image)
            out["uv_index_u_nmd"][ip - 1] = G.face(a, "nd")
            out["uv_index_u_nmu"][ip - 1] = G.face(b, "nu")
            out["uv_index_u_ndm"][ip - 1] = G.across(a, ra[0] - EPS, ra[2] + EPS, "nu")   # left of a
            out["uv_index_u_num"][ip - 1] = G.across(a, ra[1] + EPS, ra[2] + EPS, "nu")   # right of a
            out["uv_index_v_nm"][ip - 1] = G.face(a, "mu")
            out["uv_index_v_ndm"][ip - 1] = G.face(a, "md")
            out["uv_index_v_nmu"][ip - 1] = G.face(b, "mu")
            out["uv_index_v_ndmu"][ip - 1] = G.face(b, "md")
        for k in ("uv_index_u_nmd", "uv_index_u_nmu", "uv_index_u_ndm", "uv_index_u_num"):
            if out[k][ip - 1] == 0:
                out[k][ip - 1] = ip                # missing u neighbours: the point itself
        for k in ("uv_index_v_ndm", "uv_index_v_nm", "uv_index_v_ndmu", "uv_index_v_nmu"):
            if out[k][ip - 1] == 0:
                out[k][ip - 1] = dummy             # missing v neighbours: the dummy point
    return out


def _masked_regular():
    sc = W.regular_case(23, 31, dx=50.0, seed=13, t_end=600.0)
    r = np.random.default_rng(13)
    msk = sc.msk.reshape(31, 23).copy()
    msk[r.random(msk.shape) < 0.15] = 0
    msk[0, :] = 2
    sc.msk = msk.reshape(-1)
    mk = msk                                          # (m, n)
    leaves = [(1, n + 1, m + 1, mk[m, n] > 0) for m in range(mk.shape[0]) for n in range(mk.shape[1])]
    return sc, 50.0, leaves


def _quadtree(nb, mb, nlev, seed, holes=0.0):
    sc = W.quadtree_case(nb, mb, nlev=nlev, dx0=100.0, seed=seed, t_end=600.0)
    if holes:
        r = np.random.default_rng(seed)
        mask = sc.qt["mask"].copy()
        mask[(r.random(mask.size) < holes) & (sc.qt["m"] > 1)] = 0
        sc.qt["mask"] = mask
    q = sc.qt
    return sc, 100.0, list(zip(q["level"], q["n"], q["m"], q["mask"] > 0))


MESHES = {
    "quadtree_3_levels": lambda: _quadtree(10, 14, 3, 5),
    "quadtree_4_levels": lambda: _quadtree(12, 12, 4, 1002),
    "quadtree_4_levels_masked": lambda: _quadtree(10, 12, 4, 61, holes=0.1),
    "masked_regular": _masked_regular,
}


@pytest.mark.parametrize("name", list(MESHES))
def test_uv_neighbour_tables_from_geometry(name):
    sc, dx0, leaves = MESHES[name]()
    m = sc.build_model()
    s = m.sizes()
    assert s["dense"] == 0
    G = Geometry(m, dx0, leaves)
    if name.startswith("quadtree"):
        assert s["ncuv"] > 0 and G.nlev >= 3
    exp = expected_tables(G)
    for k, v in exp.items():
        got = m.table(k)
        bad = np.nonzero(got != v)[0]
        assert bad.size == 0, f"{k}: {bad.size} of {v.size} differ, first uv {bad[0] + 1}: {got[bad[0]]} vs {v[bad[0]]}"


@pytest.mark.parametrize("name", list(MESHES))
def test_cell_face_tables_and_flags_from_geometry(name):
    """z_index_uv_md/mu/nd/nu (one face, the combined point of two finer faces, or 0), the
    fine faces md1/md2 ... nu1/nu2, and per uv-point its refinement level (the finer of its
    two cells), direction and type (-1: finer to coarser, 0: same level, 1: coarser to
    finer, seen from its first cell)."""
    sc, dx0, leaves = MESHES[name]()
    m = sc.build_model()
    G = Geometry(m, dx0, leaves)
    n = G.s["np"]
    for side in ("md", "mu", "nd", "nu"):
        exp = np.array([G.face(c, side) for c in range(1, n + 1)])
        assert np.array_equal(m.table("z_index_uv_" + side), exp), side
        f = np.array([G.fine(c, side) for c in range(1, n + 1)])
        # a side with one face lists it as its first fine face
        assert np.array_equal(m.table(f"z_index_uv_{side}1"), f[:, 0]), side + "1"
        assert np.array_equal(m.table(f"z_index_uv_{side}2"), f[:, 1]), side + "2"
    la, lb = G.lev[G.nm - 1], G.lev[G.nmu - 1]
    assert np.array_equal(m.table("uv_flags_iref"), np.maximum(la, lb))
    assert np.array_equal(m.table("uv_flags_type"), np.sign(lb - la))
    ra, rb = G.rect[G.nm - 1], G.rect[G.nmu - 1]
    is_u = np.isclose(rb[:, 0], ra[:, 1])
    assert np.array_equal(m.table("uv_flags_dir"), np.where(is_u, 0, 1))


@pytest.mark.parametrize("name", [k for k in MESHES if k.startswith("quadtree")])
def test_combined_points_from_geometry(name):
    """Each combined point covers one coarse side: cuv_index_uv1 is the face on the lower /
    left half, cuv_index_uv2 the one on the upper / right half (the dummy where that finer
    cell is inactive), and the coarse cell's side table points at it."""
    sc, dx0, leaves = MESHES[name]()
    m = sc.build_model()
    G = Geometry(m, dx0, leaves)
    s = G.s
    u1, u2 = m.table("cuv_index_uv1"), m.table("cuv_index_uv2")
    seen = set()
    for c in range(1, s["np"] + 1):
        for side in ("md", "mu", "nd", "nu"):
            f = G.side[side][c]
            if not f or not G.finer_across(c, f[0]):
                continue
            j = G.face(c, side) - s["npuv"] - 1
            slots = [x if x else G.dummy for x in G.fine(c, side)]
            assert (u1[j], u2[j]) == tuple(slots), (c, side)
            seen.add(j)
    assert seen == set(range(s["ncuv"]))
