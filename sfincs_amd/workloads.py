"""Seeded synthetic scenarios for the benchmark configurations and the parity tests.

Seed convention: numpy Generator(PCG64(seed)); configuration k uses seed 1000+k and
member/request j of configuration k uses (1000+k)*100000 + j.

Every scenario is returned as a `Scenario` (plain numpy arrays + parameters) that can be
turned into a host model with `build_model()`. The same scenario object feeds the
GPU path and the CPU oracle, so both see identical inputs.
"""
import dataclasses
import math

import numpy as np

from . import _lib


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


@dataclasses.dataclass
class Scenario:
    name: str
    params: dict
    kind: str = "regular"                 # "regular" | "quadtree"
    nmax: int = 0
    mmax: int = 0
    dx: float = 0.0
    dy: float = 0.0
    msk: np.ndarray = None                # (nmax*mmax,) int8, n fastest
    zb_grid: np.ndarray = None            # (mmax, nmax) bed level per grid cell (regular)
    qt: dict = None                       # quadtree-file arrays
    zb_qt: np.ndarray = None              # per quadtree point
    subgrid: object = None                # callable(model) -> subgrid dict, or None
    bnd: tuple = None                     # (x, y, t, zs[nt, nbnd])
    src: tuple = None                     # (x, y, t, q[nt, nsrc])
    rain: tuple = None                    # (t, mm/h)
    wind: tuple = None                    # (t, mag, dir)
    t_end: float = 3600.0

    def build_model(self):
        p = _lib.default_params(**self.params)
        if self.kind == "regular":
            m = _lib.Model.regular(p, self.nmax, self.mmax, self.dx, self.dy, self.msk)
        else:
            m = _lib.Model.quadtree(p, self.qt)
        if self.subgrid is not None:
            m.set_subgrid(self.subgrid(m))
        else:
            m.set_bathymetry(self.active_zb(m))
        if self.bnd is not None:
            x, y, t, zs = self.bnd
            m.set_boundary(x, y, t, zs)
        if self.src is not None:
            x, y, t, q = self.src
            m.set_sources(x, y, t, q)
        if self.rain is not None:
            m.set_rain(*self.rain)
        if self.wind is not None:
            m.set_wind(*self.wind)
        m.initialize()
        return m

    def active_zb(self, m):
        if self.kind == "regular":
            n = m.table("z_index_z_n") - 1
            mm = m.table("z_index_z_m") - 1
            return self.zb_grid[mm, n].astype(np.float32)
        active = self.qt["mask"] > 0
        return self.zb_qt[active].astype(np.float32)


# ---------------------------------------------------------------- bathymetry recipes
def beach_with_bumps(r, nmax, mmax, dx, dy, nbumps=20):
    """Linear beach 1:200 from -10 m to +5 m (flat land above) plus Gaussian bumps."""
    x = (np.arange(mmax) + 0.5) * dx
    y = (np.arange(nmax) + 0.5) * dy
    zb = np.clip(-10.0 + x / 200.0, -10.0, 5.0)[:, None] + np.zeros((1, nmax))
    xs, ys = mmax * dx, nmax * dy
    amp = r.uniform(-2.0, 2.0, nbumps)
    sig = r.uniform(100.0, 500.0, nbumps)
    cx = r.uniform(0.0, xs, nbumps)
    cy = r.uniform(0.0, ys, nbumps)
    X = x[:, None]
    Y = y[None, :]
    for a, s, px, py in zip(amp, sig, cx, cy):
        zb = zb + a * np.exp(-((X - px) ** 2 + (Y - py) ** 2) / (2.0 * s * s))
    return zb.astype(np.float32)          # (mmax, nmax)


def fbm(r, n, m, hurst=0.7, lo=-15.0, hi=20.0):
    """2-D fractional Brownian surface by spectral synthesis, P(k) ~ k^-(2H+2), rescaled."""
    from scipy import fft as sfft
    ky = np.fft.fftfreq(n).astype(np.float32)
    kx = np.fft.rfftfreq(m).astype(np.float32)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    k2[0, 0] = 1.0
    amp = k2 ** (-(2.0 * hurst + 2.0) / 4.0)
    amp[0, 0] = 0.0
    phase = r.uniform(0.0, 2.0 * np.pi, size=amp.shape).astype(np.float32)
    spec = (amp * (np.cos(phase) + 1j * np.sin(phase))).astype(np.complex64)
    del phase, k2
    z = sfft.irfft2(spec, s=(n, m), workers=8).astype(np.float32)
    del spec
    zmin, zmax = float(z.min()), float(z.max())
    z = lo + (z - zmin) * ((hi - lo) / (zmax - zmin))
    return z.astype(np.float32)           # (n rows, m cols)


def landuse_manning(r, n, m, values=(0.02, 0.035, 0.06)):
    """Seeded land-use: smoothed noise split into three quantile classes."""
    from scipy import ndimage
    noise = r.standard_normal((n, m)).astype(np.float32)
    sm = ndimage.gaussian_filter(noise, sigma=max(2.0, min(n, m) / 64.0))
    q1, q2 = np.quantile(sm, [1.0 / 3.0, 2.0 / 3.0])
    out = np.where(sm < q1, values[0], np.where(sm < q2, values[1], values[2])).astype(np.float32)
    return out


# ---------------------------------------------------------------- subgrid tables
def _volume_levels(zsorted, nl, pix_area):
    """Cell volume table: z at nl equally spaced volumes between 0 and volmax."""
    npx = zsorted.shape[1]
    zmin = zsorted[:, 0]
    zmax = zsorted[:, -1]
    # cumulative volume at each sorted pixel elevation: V(z_k) = sum_{j<k} (z_k - z_j) * A
    csum = np.cumsum(zsorted, axis=1, dtype=np.float64)
    j = np.arange(npx)
    vk = (zsorted.astype(np.float64) * j[None, :] - np.concatenate(
        [np.zeros((zsorted.shape[0], 1)), csum[:, :-1]], axis=1)) * pix_area
    volmax = vk[:, -1]
    zl = np.empty((zsorted.shape[0], nl), dtype=np.float64)
    for lev in range(nl):
        target = volmax * lev / (nl - 1)
        # find segment k with vk[k] <= target < vk[k+1]
        k = np.clip((vk <= target[:, None]).sum(axis=1) - 1, 0, npx - 1)
        kk = np.minimum(k + 1, npx - 1)
        v0 = np.take_along_axis(vk, k[:, None], 1)[:, 0]
        z0 = np.take_along_axis(zsorted, k[:, None], 1)[:, 0].astype(np.float64)
        z1 = np.take_along_axis(zsorted, kk[:, None], 1)[:, 0].astype(np.float64)
        wet = (k + 1).astype(np.float64) * pix_area          # wetted area in segment
        dz = (target - v0) / wet
        zl[:, lev] = np.where(kk > k, np.minimum(z0 + dz, z1), z0 + dz)
    zl[:, 0] = zmin
    zl[:, -1] = zmax
    return zmin, zmax, volmax, zl


def _uv_tables(zpix, npix, nl):
    """Face tables from the pixels around a uv-point: zmin/zmax, havg/nrep/pwet per level."""
    zmin = zpix.min(axis=1)
    zmax = zpix.max(axis=1)
    zmax = np.maximum(zmax, zmin + 0.01)
    levels = zmin[:, None] + (zmax - zmin)[:, None] * (np.arange(nl)[None, :] / (nl - 1))
    havg = np.empty((zpix.shape[0], nl), np.float32)
    nrep = np.empty_like(havg)
    pwet = np.empty_like(havg)
    navg = npix.mean(axis=1)
    for lev in range(nl):
        h = np.maximum(levels[:, lev][:, None] - zpix, 0.0)
        ha = h.mean(axis=1)
        havg[:, lev] = ha
        pwet[:, lev] = (h > 0).mean(axis=1)
        conv = (h ** (5.0 / 3.0) / npix).mean(axis=1)
        with np.errstate(divide="ignore", invalid="ignore"):
            nr = np.where(conv > 0, ha ** (5.0 / 3.0) / conv, navg)
        nrep[:, lev] = np.where(np.isfinite(nr) & (nr > 0), nr, navg)
    # fit parameter of the above-zmax roughness transition (build-owned recipe)
    ffit = (1.0 / np.maximum(zmax - zmin, 0.1)).astype(np.float32)
    return zmin.astype(np.float32), zmax.astype(np.float32), havg, nrep, pwet, navg.astype(np.float32), ffit


def regular_subgrid_tables(dem, nman, refi, dx, nl, model):
    """Tables for a dense regular grid from a DEM with refi x refi pixels per cell.

    dem/nman: (nmax*refi, mmax*refi) arrays, rows = n (y), cols = m (x)."""
    ny, nx = dem.shape
    nmax, mmax = ny // refi, nx // refi
    pix_area = (dx / refi) ** 2
    # cells in sfincs order (column-major: n fastest)
    blocks = dem.reshape(nmax, refi, mmax, refi).transpose(2, 0, 1, 3).reshape(nmax * mmax, refi * refi)
    zs_sorted = np.sort(blocks, axis=1)
    zmin, zmax, volmax, zl = _volume_levels(zs_sorted, nl, pix_area)
    del blocks, zs_sorted
    nm = model.table("uv_index_z_nm") - 1
    nmu = model.table("uv_index_z_nmu") - 1
    dirs = model.table("uv_flags_dir")
    nsrc_n = model.table("z_index_z_n") - 1
    nsrc_m = model.table("z_index_z_m") - 1
    npuv = len(nm)
    half = refi // 2 + 1
    out = {k: np.empty(npuv, np.float32) for k in ("uv_zmin", "uv_zmax", "uv_navg", "uv_ffit")}
    for k in ("uv_havg", "uv_nrep", "uv_pwet"):
        out[k] = np.empty((npuv, nl), np.float32)
    # window of 2*half pixel columns (u) or rows (v) centred on the face, refi wide across
    chunk = 1 << 20
    for d in (0, 1):
        sel = np.nonzero(dirs == d)[0]
        for s0 in range(0, len(sel), chunk):
            idx = sel[s0:s0 + chunk]
            n0 = nsrc_n[nm[idx]]
            m0 = nsrc_m[nm[idx]]
            if d == 0:
                rows = n0[:, None] * refi + np.arange(refi)[None, :]
                cols = (m0[:, None] + 1) * refi - half + np.arange(2 * half)[None, :]
                cols = np.clip(cols, 0, nx - 1)
                zp = dem[rows[:, :, None], cols[:, None, :]].reshape(len(idx), -1)
                npx = nman[rows[:, :, None], cols[:, None, :]].reshape(len(idx), -1)
            else:
                rows = (n0[:, None] + 1) * refi - half + np.arange(2 * half)[None, :]
                rows = np.clip(rows, 0, ny - 1)
                cols = m0[:, None] * refi + np.arange(refi)[None, :]
                zp = dem[rows[:, :, None], cols[:, None, :]].reshape(len(idx), -1)
                npx = nman[rows[:, :, None], cols[:, None, :]].reshape(len(idx), -1)
            zmn, zmx, ha, nr, pw, na, ff = _uv_tables(zp, npx, nl)
            out["uv_zmin"][idx] = zmn
            out["uv_zmax"][idx] = zmx
            out["uv_havg"][idx] = ha
            out["uv_nrep"][idx] = nr
            out["uv_pwet"][idx] = pw
            out["uv_navg"][idx] = na
            out["uv_ffit"][idx] = ff
    sg = dict(nlevels=nl, z_zmin=zmin.astype(np.float32), z_zmax=zmax.astype(np.float32),
              z_volmax=volmax.astype(np.float32), z_level=zl.astype(np.float32))
    sg.update(out)
    return sg


# ---------------------------------------------------------------- scenario recipes
def tide(amplitude, t_end, period=43200.0, step=10.0, phase=0.0):
    t = np.arange(0.0, t_end + 2 * step, step, dtype=np.float64)
    return t.astype(np.float32), (amplitude * np.sin(2 * np.pi * t / period + phase)).astype(np.float32)[:, None]


def regular_case(nmax, mmax, dx=50.0, seed=1000, t_end=3600.0, rain_mmh=20.0, rain_until=1800.0,
                 tide_amp=1.0, manning=0.02, wind=None, nbumps=20, **params):
    """Configuration-0 recipe at any size (regular mesh, beach + bumps, tide + rain)."""
    r = rng(seed)
    zb = beach_with_bumps(r, nmax, mmax, dx, dx, nbumps)
    msk = np.ones((mmax, nmax), np.int8)
    msk[0, :] = 2                                   # west column is a water-level boundary
    t, zs = tide(tide_amp, t_end)
    p = dict(t0=0.0, t1=float(t_end), manning=manning)
    p.update(params)
    rain = None
    if rain_mmh and rain_mmh > 0:
        rain = (np.array([0.0, rain_until, rain_until + 1.0, t_end + 3600.0], np.float32),
                np.array([rain_mmh, rain_mmh, 0.0, 0.0], np.float32))
    return Scenario(name=f"regular_{nmax}x{mmax}", params=p, kind="regular", nmax=nmax, mmax=mmax,
                    dx=dx, dy=dx, msk=msk.reshape(-1), zb_grid=zb,
                    bnd=(np.array([-dx / 2], np.float32), np.array([nmax * dx / 2], np.float32), t, zs),
                    rain=rain, wind=wind, t_end=float(t_end))


def subgrid_case(nmax, mmax, dx=25.0, refi=5, nl=10, seed=1001, t_end=21600.0, tide_amp=1.5,
                 nsrc=8, rain_mmh=0.0, **params):
    """Configuration-1 recipe at any size (regular mesh with subgrid tables from an fBm DEM)."""
    r = rng(seed)
    dem = fbm(r, nmax * refi, mmax * refi)
    nman = landuse_manning(r, nmax * refi, mmax * refi)
    zcell_min = dem.reshape(nmax, refi, mmax, refi).min(axis=(1, 3))     # (nmax, mmax)
    msk = np.ones((mmax, nmax), np.int8)
    msk[0, :] = np.where(zcell_min[:, 0] < 0.0, 2, 1)
    t, zs = tide(tide_amp, t_end)
    # point sources at cells whose lowest pixel lies in [0, 5] m
    cand = np.argwhere((zcell_min >= 0.0) & (zcell_min <= 5.0) & (np.arange(mmax)[None, :] > 0))
    if len(cand) < nsrc:
        cand = np.argwhere(np.arange(mmax)[None, :].repeat(nmax, 0) > 0)
    pick = cand[r.choice(len(cand), size=nsrc, replace=False)] if nsrc > 0 else np.zeros((0, 2), int)
    xs = ((pick[:, 1] + 0.5) * dx).astype(np.float32)
    ys = ((pick[:, 0] + 0.5) * dx).astype(np.float32)
    qv = r.uniform(50.0, 500.0, nsrc).astype(np.float32)
    ts = np.array([0.0, t_end + 3600.0], np.float32)
    src = (xs, ys, ts, np.stack([qv, qv])) if nsrc > 0 else None
    p = dict(t0=0.0, t1=float(t_end))
    p.update(params)
    rain = None
    if rain_mmh and rain_mmh > 0:
        rain = (np.array([0.0, 1800.0, 1801.0, t_end + 3600.0], np.float32),
                np.array([rain_mmh, rain_mmh, 0.0, 0.0], np.float32))

    def tables(model, _dem=dem, _nman=nman):
        return regular_subgrid_tables(_dem, _nman, refi, dx, nl, model)

    sc = Scenario(name=f"subgrid_{nmax}x{mmax}", params=p, kind="regular", nmax=nmax, mmax=mmax,
                  dx=dx, dy=dx, msk=msk.reshape(-1), subgrid=tables,
                  bnd=(np.array([-dx / 2], np.float32), np.array([nmax * dx / 2], np.float32), t, zs),
                  src=src, rain=rain, t_end=float(t_end))
    return sc


def config0():
    return regular_case(512, 512, dx=50.0, seed=1000, t_end=3600.0)


def config1(n=2048):
    return subgrid_case(n, n, dx=25.0, refi=5, nl=10, seed=1001, t_end=21600.0, tide_amp=1.5, nsrc=8)


def ensemble_members(count=64, seed_base=1003, t_end=10800.0, nmin=256, nmax=2048, scale=1.0):
    """Configuration-3 members: log-uniform sizes, regular or subgrid by a seeded coin flip."""
    out = []
    for j in range(count):
        r = rng(seed_base * 100000 + j)
        n = int(round(math.exp(r.uniform(math.log(nmin), math.log(nmax))) * scale))
        n = max(16, n)
        amp = float(r.uniform(0.5, 4.0))
        rain = float(r.uniform(0.0, 100.0))
        if r.uniform() < 0.5:
            sc = regular_case(n, n, dx=50.0, seed=seed_base * 100000 + j, t_end=t_end, tide_amp=amp,
                              rain_mmh=rain, rain_until=1800.0)
        else:
            sc = subgrid_case(n, n, dx=25.0, seed=seed_base * 100000 + j, t_end=t_end, tide_amp=amp,
                              nsrc=0, rain_mmh=rain)
        sc.name = f"member{j}_{sc.name}"
        out.append(sc)
    return out
