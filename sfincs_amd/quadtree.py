"""Seeded quadtree meshes in the reference's quadtree-file semantics.

The reference reads quadtree meshes built by an external tool (HydroMT-SFINCS) as arrays
n, m, level, md/mu/nd/nu (-1 coarser, 0 same, 1 finer) and md1..nu2 (1-based point
indices, 0 = none), points sorted level-major then by nm = (m-1)*nmx + n
(sfincs_quadtree.F90:165-186, 295-442). This module builds such meshes from a
refinement criterion on a fine-resolution bed level, with 2:1 balance across faces
(the reference handles jumps of one level only, sfincs_domain.f90:731-1001).
"""
import numpy as np


def _pool(a, op):
    n, m = a.shape
    b = a.reshape(n // 2, 2, m // 2, 2)
    return op(op(b, axis=3), axis=1)


def _dilate4(x):
    y = x.copy()
    y[1:, :] |= x[:-1, :]
    y[:-1, :] |= x[1:, :]
    y[:, 1:] |= x[:, :-1]
    y[:, :-1] |= x[:, 1:]
    return y


def build(dem, nlev, refine_range=(-5.0, 5.0)):
    """dem: (nb*F, mb*F) bed level at the finest resolution, F = 2**(nlev-1), rows = n.
    A cell is refined while its bed level range intersects refine_range.

    Returns (levels, leaf) where leaf[L] marks the level-L cells that are mesh cells."""
    F = 2 ** (nlev - 1)
    ny, nx = dem.shape
    assert ny % F == 0 and nx % F == 0
    mins, maxs = [None] * (nlev + 1), [None] * (nlev + 1)
    mins[nlev], maxs[nlev] = dem, dem
    for L in range(nlev - 1, 0, -1):
        mins[L] = _pool(mins[L + 1], np.min)
        maxs[L] = _pool(maxs[L + 1], np.max)
    lo, hi = refine_range
    refined = [None] * (nlev + 1)
    for L in range(1, nlev):
        refined[L] = (mins[L] <= hi) & (maxs[L] >= lo)
    refined[nlev] = np.zeros_like(dem, dtype=bool)
    # 2:1 balance: every face neighbour of a refined level-L node must exist, i.e. its
    # parent must be refined; processed from fine to coarse.
    for L in range(nlev - 1, 1, -1):
        need = _dilate4(refined[L])
        par = _pool(need, np.max)
        refined[L - 1] |= par
    exists = [None] * (nlev + 1)
    exists[1] = np.ones_like(refined[1], dtype=bool)
    for L in range(1, nlev):
        exists[L + 1] = np.repeat(np.repeat(exists[L] & refined[L], 2, axis=0), 2, axis=1)
    leaf = [None] * (nlev + 1)
    for L in range(1, nlev + 1):
        leaf[L] = exists[L] & ~refined[L]
    return exists, leaf, mins, maxs


def quadtree_arrays(leaf, nlev, x0, y0, dx, dy, zb_of_leaf=None, mask_of_leaf=None):
    """Quadtree-file arrays of the leaves (1-based neighbour indices, level-major order)."""
    idx = [None] * (nlev + 2)
    pts = []
    count = 0
    for L in range(1, nlev + 1):
        nrows = leaf[L].shape[0]
        n0, m0 = np.nonzero(leaf[L])
        order = np.lexsort((n0, m0))          # by m, then n  (nm = (m-1)*nmx + n)
        n0, m0 = n0[order], m0[order]
        k = np.arange(count + 1, count + 1 + len(n0), dtype=np.int32)
        a = np.zeros(leaf[L].shape, np.int32)
        a[n0, m0] = k
        idx[L] = a
        pts.append((L, n0, m0))
        count += len(n0)
    total = count

    def get(L, n, m):
        """Point index of level-L cell (n, m) (zero-based), 0 when absent/outside."""
        if L < 1 or L > nlev:
            return np.zeros(len(n), np.int32)
        a = idx[L]
        ok = (n >= 0) & (m >= 0) & (n < a.shape[0]) & (m < a.shape[1])
        out = np.zeros(len(n), np.int32)
        out[ok] = a[n[ok], m[ok]]
        return out

    def inside(L, n, m):
        a = idx[L]
        return (n >= 0) & (m >= 0) & (n < a.shape[0]) & (m < a.shape[1])

    out = {k: np.zeros(total, np.int32) for k in ("n", "m", "md1", "md2", "mu1", "mu2", "nd1", "nd2", "nu1", "nu2")}
    for k in ("level", "md", "mu", "nd", "nu"):
        out[k] = np.zeros(total, np.int8)
    s = 0
    for L, n0, m0 in pts:
        e = s + len(n0)
        out["level"][s:e] = L
        out["n"][s:e] = n0 + 1
        out["m"][s:e] = m0 + 1
        for side, dn, dm in (("mu", 0, 1), ("md", 0, -1), ("nu", 1, 0), ("nd", -1, 0)):
            nn, mm = n0 + dn, m0 + dm
            same = get(L, nn, mm)
            ins = inside(L, nn, mm)
            flag = np.zeros(len(n0), np.int8)
            i1 = same.copy()
            i2 = np.zeros(len(n0), np.int32)
            # finer neighbours: the two level-(L+1) cells along the shared face
            if L < nlev:
                if side == "mu":
                    a1 = get(L + 1, 2 * n0, 2 * mm); a2 = get(L + 1, 2 * n0 + 1, 2 * mm)
                elif side == "md":
                    a1 = get(L + 1, 2 * n0, 2 * mm + 1); a2 = get(L + 1, 2 * n0 + 1, 2 * mm + 1)
                elif side == "nu":
                    a1 = get(L + 1, 2 * nn, 2 * m0); a2 = get(L + 1, 2 * nn, 2 * m0 + 1)
                else:
                    a1 = get(L + 1, 2 * nn + 1, 2 * m0); a2 = get(L + 1, 2 * nn + 1, 2 * m0 + 1)
                fin = ins & (same == 0) & ((a1 > 0) | (a2 > 0))
                flag[fin] = 1
                i1[fin] = a1[fin]
                i2[fin] = a2[fin]
            # coarser neighbour: the level-(L-1) cell containing (nn, mm)
            if L > 1:
                c = get(L - 1, nn // 2, mm // 2)
                crs = ins & (same == 0) & (flag == 0) & (c > 0)
                flag[crs] = -1
                i1[crs] = c[crs]
            out[side][s:e] = flag
            out[side + "1"][s:e] = i1
            out[side + "2"][s:e] = i2
        s = e
    out["nr_levels"] = nlev
    out["x0"], out["y0"], out["dx"], out["dy"] = float(x0), float(y0), float(dx), float(dy)
    out["mask"] = np.ones(total, np.int8) if mask_of_leaf is None else mask_of_leaf(out)
    return out


def leaf_bed_level(qt, dem, nlev):
    """Block mean of the finest-resolution bed level over every leaf."""
    F = 2 ** (nlev - 1)
    means = {}
    zb = np.zeros(len(qt["n"]), np.float32)
    for L in range(1, nlev + 1):
        s = F // 2 ** (L - 1)
        ny, nx = dem.shape
        means[L] = dem.reshape(ny // s, s, nx // s, s).mean(axis=(1, 3))
        sel = qt["level"] == L
        zb[sel] = means[L][qt["n"][sel] - 1, qt["m"][sel] - 1]
    return zb
