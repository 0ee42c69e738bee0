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
    rstfile: str = None                   # binary restart file as initial condition
    bdr: tuple = None                     # river boundary points (x, y, x_in, y_in, slope, distance)
    bzi: np.ndarray = None                # infragravity levels [nt, nbnd] at the bnd series times
    storage: object = None                # callable(model) -> storage volume per cell (m3), or None
    drn: tuple = None                     # drainage structures (x_in, y_in, x_out, y_out, type, params[n, 6])
    thd: list = None                      # thin dams: polylines, each (n, 2) x, y
    weirs: list = None                    # weirs: polylines, each (n, 4) x, y, crest level, Cd
    qinf_map: object = None               # callable(model) -> infiltration per cell (mm/h), or None
    scs_map: object = None                # callable(model) -> curve-number retention S per cell (inches)
    meteo: dict = None                    # gridded meteo: "amu", "amv", "amp", "ampr" grids, "spw" spiderweb
    rotation: float = 0.0                 # regular mesh rotation (degrees)

    def build_model(self):
        p = _lib.default_params(**self.params)
        if self.kind == "regular":
            m = _lib.Model.regular(p, self.nmax, self.mmax, self.dx, self.dy, self.msk, rotation=self.rotation)
        else:
            m = _lib.Model.quadtree(p, self.qt)
        if self.subgrid is not None:
            sg = self.subgrid(m)
            if isinstance(sg, str):
                m.read_subgrid_binary(sg)      # legacy binary subgrid file
            else:
                m.set_subgrid(sg)
        else:
            m.set_bathymetry(self.active_zb(m))
        if self.bnd is not None:
            x, y, t, zs = self.bnd
            m.set_boundary(x, y, t, zs)
            if self.bzi is not None:
                m.set_boundary_ig(self.bzi)
        if self.src is not None:
            x, y, t, q = self.src
            m.set_sources(x, y, t, q)
        if self.rain is not None:
            m.set_rain(*self.rain)
        if self.wind is not None:
            m.set_wind(*self.wind)
        if self.bdr is not None:
            m.set_river_boundary(*self.bdr)
        if self.drn is not None:
            m.set_drainage(*self.drn)
        if self.thd is not None:
            m.set_thin_dams(self.thd)
        if self.weirs is not None:
            m.set_weirs(self.weirs)
        if self.storage is not None:
            m.set_storage_volume(self.storage(m))
        if self.qinf_map is not None:
            m.set_infiltration(self.qinf_map(m))
        if self.scs_map is not None:
            m.set_curve_number(self.scs_map(m))
        for kind, g in (self.meteo or {}).items():
            if kind == "spw":
                m.set_spiderweb(**g)
            else:
                m.set_meteo_grid(kind, **g)
        if self.rstfile is not None:
            m.set_restart(self.rstfile)
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


def _device(npix):
    """Torch device for input generation: the GPU for large inputs when present."""
    import torch
    if npix >= (1 << 20) and torch.cuda.is_available():
        return torch.device("cuda")
    return torch.device("cpu")


def fbm(r, n, m, hurst=0.7, lo=-15.0, hi=20.0):
    """2-D fractional Brownian surface by spectral synthesis, P(k) ~ k^-(2H+2), rescaled
    affinely to [lo, hi]. Random phases come from the numpy generator `r`."""
    import torch
    dev = _device(n * m)
    phase = torch.from_numpy(r.uniform(0.0, 2.0 * np.pi, size=(n, m // 2 + 1)).astype(np.float32)).to(dev)
    ky = torch.fft.fftfreq(n, device=dev, dtype=torch.float32)
    kx = torch.fft.rfftfreq(m, device=dev, dtype=torch.float32)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    k2[0, 0] = 1.0
    amp = k2 ** (-(2.0 * hurst + 2.0) / 4.0)
    amp[0, 0] = 0.0
    spec = torch.polar(amp, phase)
    del phase, k2, amp
    z = torch.fft.irfft2(spec, s=(n, m))
    del spec
    zmin, zmax = z.min(), z.max()
    z = lo + (z - zmin) * ((hi - lo) / (zmax - zmin))
    return z.float().cpu().numpy()          # (n rows, m cols)


def landuse_manning(r, n, m, values=(0.02, 0.035, 0.06), refi=1):
    """Seeded land-use: smoothed noise on the cell grid split into three quantile classes,
    repeated onto the refi x refi pixels of each cell."""
    import torch
    nc, mc = n // refi, m // refi
    noise = torch.from_numpy(r.standard_normal((nc, mc)).astype(np.float32))
    sig = max(2.0, min(nc, mc) / 64.0)
    ky = torch.fft.fftfreq(nc)[:, None]
    kx = torch.fft.rfftfreq(mc)[None, :]
    filt = torch.exp(-2.0 * (np.pi ** 2) * sig * sig * (ky ** 2 + kx ** 2))
    sm = torch.fft.irfft2(torch.fft.rfft2(noise) * filt, s=(nc, mc)).numpy()
    q1, q2 = np.quantile(sm, [1.0 / 3.0, 2.0 / 3.0])
    cls = np.where(sm < q1, values[0], np.where(sm < q2, values[1], values[2])).astype(np.float32)
    return np.repeat(np.repeat(cls, refi, axis=0), refi, axis=1)


# ---------------------------------------------------------------- subgrid tables
def _volume_levels(zsorted, nl, pix_area):
    """Cell volume table: z at nl equally spaced volumes between 0 and volmax.
    zsorted: (ncell, npx) torch tensor of sorted pixel elevations."""
    import torch
    z = zsorted.double()
    npx = z.shape[1]
    csum = torch.cumsum(z, dim=1)
    j = torch.arange(npx, device=z.device, dtype=torch.float64)
    prev = torch.cat([torch.zeros_like(csum[:, :1]), csum[:, :-1]], dim=1)
    vk = (z * j[None, :] - prev) * pix_area          # volume below pixel k's elevation
    volmax = vk[:, -1]
    zl = torch.empty((z.shape[0], nl), dtype=torch.float64, device=z.device)
    for lev in range(nl):
        target = volmax * lev / (nl - 1)
        k = ((vk <= target[:, None]).sum(dim=1) - 1).clamp(0, npx - 1)
        kk = (k + 1).clamp(max=npx - 1)
        v0 = vk.gather(1, k[:, None])[:, 0]
        z0 = z.gather(1, k[:, None])[:, 0]
        z1 = z.gather(1, kk[:, None])[:, 0]
        wet = (k + 1).double() * pix_area
        dz = (target - v0) / wet
        zl[:, lev] = torch.where(kk > k, torch.minimum(z0 + dz, z1), z0 + dz)
    zl[:, 0] = z[:, 0]
    zl[:, -1] = z[:, -1]
    return z[:, 0], z[:, -1], volmax, zl


def _uv_tables(zpix, npix, nl):
    """Face tables from the pixels around a uv-point: zmin/zmax, havg/nrep/pwet per level
    (build-owned recipe; torch tensors in, float32 numpy arrays out)."""
    import torch
    zmin = zpix.min(dim=1).values
    zmax = torch.maximum(zpix.max(dim=1).values, zmin + 0.01)
    navg = npix.mean(dim=1)
    P = zpix.shape[0]
    havg = torch.empty((P, nl), dtype=torch.float32, device=zpix.device)
    nrep = torch.empty_like(havg)
    pwet = torch.empty_like(havg)
    for lev in range(nl):
        level = zmin + (zmax - zmin) * (lev / (nl - 1))
        h = (level[:, None] - zpix).clamp(min=0.0)
        ha = h.mean(dim=1)
        havg[:, lev] = ha
        pwet[:, lev] = (h > 0).float().mean(dim=1)
        conv = (h.pow(5.0 / 3.0) / npix).mean(dim=1)
        nr = torch.where(conv > 0, ha.pow(5.0 / 3.0) / conv.clamp(min=1e-30), navg)
        nrep[:, lev] = torch.where(torch.isfinite(nr) & (nr > 0), nr, navg)
    ffit = 1.0 / (zmax - zmin).clamp(min=0.1)
    f = lambda t: t.float().cpu().numpy()
    return f(zmin), f(zmax), f(havg), f(nrep), f(pwet), f(navg), f(ffit)


def regular_subgrid_tables(dem, nman, refi, dx, nl, model):
    """Tables for a dense regular grid from a DEM with refi x refi pixels per cell.

    dem/nman: (nmax*refi, mmax*refi) arrays, rows = n (y), cols = m (x)."""
    import torch
    ny, nx = dem.shape
    nmax, mmax = ny // refi, nx // refi
    dev = _device(ny * nx)
    pix_area = (dx / refi) ** 2
    D = torch.from_numpy(np.ascontiguousarray(dem)).to(dev)
    N = torch.from_numpy(np.ascontiguousarray(nman)).to(dev)
    # cells in sfincs order (column-major: n fastest)
    blocks = D.reshape(nmax, refi, mmax, refi).permute(2, 0, 1, 3).reshape(nmax * mmax, refi * refi)
    zmin, zmax, volmax, zl = _volume_levels(torch.sort(blocks, dim=1).values, nl, pix_area)
    del blocks
    nm = torch.from_numpy(model.table("uv_index_z_nm").astype(np.int64) - 1).to(dev)
    dirs = torch.from_numpy(model.table("uv_flags_dir").astype(np.int64)).to(dev)
    zn = torch.from_numpy(model.table("z_index_z_n").astype(np.int64) - 1).to(dev)
    zm = torch.from_numpy(model.table("z_index_z_m").astype(np.int64) - 1).to(dev)
    npuv = nm.shape[0]
    half = refi // 2 + 1
    out = {k: np.empty(npuv, np.float32) for k in ("uv_zmin", "uv_zmax", "uv_navg", "uv_ffit")}
    for k in ("uv_havg", "uv_nrep", "uv_pwet"):
        out[k] = np.empty((npuv, nl), np.float32)
    ar_r = torch.arange(refi, device=dev)
    ar_h = torch.arange(2 * half, device=dev)
    chunk = 1 << 20
    for d in (0, 1):
        sel = torch.nonzero(dirs == d)[:, 0]
        for s0 in range(0, sel.shape[0], chunk):
            idx = sel[s0:s0 + chunk]
            n0 = zn[nm[idx]]
            m0 = zm[nm[idx]]
            if d == 0:      # window across the face between columns m0 and m0+1
                rows = n0[:, None] * refi + ar_r[None, :]
                cols = ((m0[:, None] + 1) * refi - half + ar_h[None, :]).clamp(0, nx - 1)
            else:           # window across the face between rows n0 and n0+1
                rows = ((n0[:, None] + 1) * refi - half + ar_h[None, :]).clamp(0, ny - 1)
                cols = m0[:, None] * refi + ar_r[None, :]
            zp = D[rows[:, :, None], cols[:, None, :]].reshape(idx.shape[0], -1)
            npx = N[rows[:, :, None], cols[:, None, :]].reshape(idx.shape[0], -1)
            zmn, zmx, ha, nr, pw, na, ff = _uv_tables(zp, npx, nl)
            ii = idx.cpu().numpy()
            out["uv_zmin"][ii] = zmn
            out["uv_zmax"][ii] = zmx
            out["uv_havg"][ii] = ha
            out["uv_nrep"][ii] = nr
            out["uv_pwet"][ii] = pw
            out["uv_navg"][ii] = na
            out["uv_ffit"][ii] = ff
    f = lambda t: t.float().cpu().numpy()
    sg = dict(nlevels=nl, z_zmin=f(zmin), z_zmax=f(zmax), z_volmax=f(volmax), z_level=f(zl))
    sg.update(out)
    return sg


def write_subgrid_binary(path, sg, model):
    """Writes NetCDF-style tables `sg` of a regular mesh (as regular_subgrid_tables returns
    them) as a legacy binary subgrid file, the layout read_subgrid_file_original reads
    (sfincs_subgrid.F90:671-866): int32 cells, option, bins (= levels - 1); per cell zmin,
    zmax, volmax and the levels 2..nl; then for the u-points and the v-points, per cell
    through its right / upper point, zmin, zmax, dhdz (unused, 0), hrep (havg) and navg
    (Manning's n) of levels 2..nl. Cells without such a point write 0."""
    nl = int(sg["nlevels"])
    np_ = len(sg["z_zmin"])
    zl = np.asarray(sg["z_level"], np.float32).reshape(np_, nl)
    havg = np.asarray(sg["uv_havg"], np.float32).reshape(-1, nl)
    nrep = np.asarray(sg["uv_nrep"], np.float32).reshape(-1, nl)
    with open(path, "wb") as f:
        np.array([np_, 0, nl - 1], np.int32).tofile(f)
        for k in ("z_zmin", "z_zmax", "z_volmax"):
            np.asarray(sg[k], np.float32).tofile(f)
        for lv in range(1, nl):
            np.ascontiguousarray(zl[:, lv]).tofile(f)
        for name in ("z_index_uv_mu1", "z_index_uv_nu1"):
            ip = model.table(name).astype(np.int64)
            has = ip > 0
            sel = np.where(has, ip - 1, 0)

            def per_cell(a):
                return np.where(has, np.asarray(a, np.float32)[sel], np.float32(0.0)).astype(np.float32)
            per_cell(sg["uv_zmin"]).tofile(f)
            per_cell(sg["uv_zmax"]).tofile(f)
            np.zeros(np_, np.float32).tofile(f)
            for lv in range(1, nl):
                per_cell(havg[:, lv]).tofile(f)
            for lv in range(1, nl):
                per_cell(nrep[:, lv]).tofile(f)


def with_subgrid_binary(sc, path):
    """The scenario with its subgrid tables read from a legacy binary file written at `path`
    (huthresh keeps its input value on that path)."""
    make = sc.subgrid

    def tables_via_file(model, _make=make, _path=str(path)):
        write_subgrid_binary(_path, _make(model), model)
        return _path
    sc.subgrid = tables_via_file
    sc.name += "_sbgbin"
    return sc


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


def with_ig_waves(sc, amp=0.2, period=120.0):
    """Adds incoming infragravity levels (a bzi series) to a scenario's water-level
    boundary: a sine of `amp` metres and `period` seconds, sampled at the series times
    refined to 10 s so that the wave is resolved."""
    x, y, t, zs = sc.bnd
    tt = np.arange(float(t[0]), float(t[-1]) + 1.0, 10.0, dtype=np.float64)
    zs2 = np.stack([np.interp(tt, t.astype(np.float64), zs[:, k].astype(np.float64)) for k in range(zs.shape[1])], 1)
    sc.bnd = (x, y, tt.astype(np.float32), zs2.astype(np.float32))
    ig = amp * np.sin(2.0 * np.pi * tt / period)
    sc.bzi = np.repeat(ig[:, None], zs.shape[1], 1).astype(np.float32)
    sc.name += "_ig"
    return sc


def with_sources(sc, x, y, q):
    """Adds point discharges (src/dis files) at points (x, y) with constant rates q in m3/s
    (negative rates drain the cell) to a scenario without sources."""
    q = np.asarray(q, np.float32)
    ts = np.array([0.0, sc.t_end + 3600.0], np.float32)
    sc.src = (np.asarray(x, np.float32), np.asarray(y, np.float32), ts, np.stack([q, q]))
    sc.name += "_src"
    return sc


def regular_sources(sc):
    """Four sources on a configuration-0 style beach (the west half lies below 0 m): a
    filling and a draining one on dry land (the draining one is held at the bed by the
    clamp of the regular branch), one draining a cell the tide wets, and one draining deep
    water."""
    zb = sc.zb_grid                                   # (mmax, nmax)
    inner = np.zeros_like(zb, bool)
    inner[2:-2, 2:-2] = True

    def pick(lo, hi, k):
        cand = np.argwhere(inner & (zb > lo) & (zb < hi))
        m, n = cand[min(k, len(cand) - 1)]
        return (m + 0.5) * sc.dx, (n + 0.5) * sc.dy
    pts = [pick(2.0, 99.0, 3), pick(2.0, 99.0, 40), pick(0.2, 1.0, 7), pick(-99.0, -5.0, 11)]
    return with_sources(sc, [p[0] for p in pts], [p[1] for p in pts], [2.0, -1.0, -3.0, -5.0])


def with_drainage(sc, structures):
    """Adds drainage structures (drnfile): a list of (x_in, y_in, x_out, y_out, type,
    params) with up to 6 params (type 1 pump [Q]; 2 culvert, 3 check valve [coefficient];
    4 gate [width, sill, n, zmin, zmax, closing time]; 5 gate [width, sill, n, t_close,
    t_open, closing time])."""
    n = len(structures)
    par = np.zeros((n, 6), np.float32)
    for k, st in enumerate(structures):
        par[k, :len(st[5])] = st[5]
    col = lambda j: np.array([st[j] for st in structures], np.float32)
    sc.drn = (col(0), col(1), col(2), col(3), np.array([st[4] for st in structures], np.int32), par)
    sc.name += "_drn"
    return sc


def beach_drainage(sc):
    """Five structures on a configuration-0 style beach (sea in the west, land rising to the
    east): a pump lifting sea water onto land, a culvert and a check valve between the sea
    and a coastal cell, a gate that closes when the sea rises above 0.5 m, and a gate closed
    from 600 s to 1500 s."""
    dx = sc.dx
    ny = sc.nmax * dx

    def pt(col, row):
        return (col + 0.5) * dx, (row + 0.5) * dx
    rows = [int(sc.nmax * f) for f in (0.15, 0.3, 0.45, 0.6, 0.8)]
    sea, coast, land = 3, 18, 40
    out = []
    out.append((*pt(sea, rows[0]), *pt(land, rows[0]), 1, [2.0]))
    out.append((*pt(sea, rows[1]), *pt(coast, rows[1]), 2, [5.0]))
    out.append((*pt(sea, rows[2]), *pt(coast, rows[2]), 3, [5.0]))
    out.append((*pt(sea, rows[3]), *pt(coast, rows[3]), 4, [4.0, -3.0, 0.03, -1.0, 0.5, 300.0]))
    out.append((*pt(sea, rows[4]), *pt(coast, rows[4]), 5, [4.0, -3.0, 0.03, 600.0, 1500.0, 200.0]))
    del ny
    return with_drainage(sc, out)


def with_thin_dams(sc, polylines):
    """Adds thin dams (thdfile): polylines of (x, y) vertices."""
    sc.thd = [np.asarray(p, np.float32).reshape(-1, 2) for p in polylines]
    sc.name += "_thd"
    return sc


def with_weirs(sc, polylines):
    """Adds weirs (weirfile): polylines of (x, y, crest level, Cd) vertices."""
    sc.weirs = [np.asarray(p, np.float32).reshape(-1, 4) for p in polylines]
    sc.name += "_weir"
    return sc


def with_infiltration(sc, qinf=None, qinf_zmin=None, qinf_map=None):
    """Adds constant infiltration: uniform `qinf` mm/h on cells whose bed lies above
    `qinf_zmin` (sfincs.inp qinf, qinf_zmin), or a map (qinffile): `qinf_map(model)` gives mm/h
    per active cell. It acts only where the scenario has rain."""
    if qinf is not None:
        sc.params = dict(sc.params, qinf=float(qinf))
    if qinf_zmin is not None:
        sc.params = dict(sc.params, qinf_zmin=float(qinf_zmin))
    sc.qinf_map = qinf_map
    sc.name += "_inf"
    return sc


def with_curve_number(sc, scs_map, sfacinf=None):
    """Adds curve-number infiltration, method A (scsfile): `scs_map(model)` gives the maximum
    retention S in inches per active cell; `sfacinf` the initial abstraction fraction. It
    acts only where the scenario has rain."""
    if sfacinf is not None:
        sc.params = dict(sc.params, sfacinf=float(sfacinf))
    sc.scs_map = scs_map
    sc.name += "_cn"
    return sc


def domain_extent(sc):
    """(x0, y0, x1, y1) of a scenario's mesh (regular: from the grid; quadtree: level 1)."""
    if sc.kind == "regular":
        return 0.0, 0.0, sc.mmax * sc.dx, sc.nmax * sc.dy
    q = sc.qt
    f = 2.0 ** (q["level"].astype(np.float64) - 1)
    return (float(q["x0"]), float(q["y0"]), float(q["x0"] + np.max(q["m"] * q["dx"] / f)),
            float(q["y0"] + np.max(q["n"] * q["dy"] / f)))


def with_gridded_meteo(sc, nx=14, ny=12, t_step=1800.0, nt=None, wind=True, pressure=True, rain=True,
                       dp=2500.0, u0=8.0, v0=-3.0, rain_mmh=40.0, seed=11):
    """Gridded meteo on an equidistant grid one cell wider than the domain on every side (amu,
    amv, amp and ampr contents): a pressure low of depth dp crossing the domain from south-west
    to north-east, with a wind circulating around it over a uniform (u0, v0) and rain near its
    centre, at times 0, t_step, ... covering the run."""
    x0, y0, x1, y1 = domain_extent(sc)
    gdx, gdy = (x1 - x0) / (nx - 3), (y1 - y0) / (ny - 3)
    xll, yll = x0 - 1.5 * gdx, y0 - 1.5 * gdy
    nt = nt or int(np.ceil(sc.t_end / t_step)) + 2
    t = (np.arange(nt) * t_step).astype(np.float32)
    xc = xll + (np.arange(nx) + 0.5) * gdx
    yc = (yll + (np.arange(ny) + 0.5) * gdy)[::-1]          # row 0 northernmost
    X, Y = np.meshgrid(xc, yc)
    L = 0.35 * max(x1 - x0, y1 - y0)
    r = rng(seed)
    jitter = r.uniform(-0.5, 0.5, size=(nt, ny, nx))
    P, U, V, R = (np.zeros((nt, ny, nx), np.float32) for _ in range(4))
    for k in range(nt):
        f = k / max(nt - 1, 1)
        cx, cy = x0 + (x1 - x0) * (0.1 + 0.8 * f), y0 + (y1 - y0) * (0.2 + 0.6 * f)
        g = np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / L ** 2)
        P[k] = 101200.0 - dp * g + 20.0 * jitter[k]
        U[k] = u0 - 15.0 * g * (Y - cy) / L + jitter[k]
        V[k] = v0 + 15.0 * g * (X - cx) / L - jitter[k]
        R[k] = rain_mmh * g
    geo = dict(x_llcorner=float(xll), y_llcorner=float(yll), dx=float(gdx), dy=float(gdy))
    sc.meteo = dict(sc.meteo or {})
    if wind:
        sc.meteo["amu"] = dict(times=t, values=U, **geo)
        sc.meteo["amv"] = dict(times=t, values=V, **geo)
    if pressure:
        sc.meteo["amp"] = dict(times=t, values=P, **geo)
    if rain:
        sc.meteo["ampr"] = dict(times=t, values=R, **geo)
    sc.name += "_met"
    return sc


def with_spiderweb(sc, nrows=16, ncols=24, radius_frac=0.8, vmax=35.0, rmax_frac=0.2, dp=4000.0, rain_mmh=60.0,
                   t_step=1800.0, nt=None, t_first=0.0):
    """A spiderweb (spw contents): a tropical cyclone whose eye crosses the domain from west
    to east, with a Rankine-like wind speed profile peaking vmax at rmax, the wind turning
    anticlockwise round the eye with 20 degrees of inflow, a pressure drop dp exp(-rmax / r)
    and rain rain_mmh exp(-r / (2 rmax)); times t_first, t_first + t_step, ..."""
    x0, y0, x1, y1 = domain_extent(sc)
    ext = max(x1 - x0, y1 - y0)
    radius = radius_frac * ext
    rmax = rmax_frac * ext
    nt = nt or int(np.ceil(sc.t_end / t_step)) + 2
    t = (t_first + np.arange(nt) * t_step).astype(np.float32)
    f = np.arange(nt) / max(nt - 1, 1)
    xe = np.round(x0 - 0.2 * (x1 - x0) + 1.4 * (x1 - x0) * f, 1).astype(np.float32)
    ye = np.round(y0 + (y1 - y0) * (0.35 + 0.3 * f), 1).astype(np.float32)
    rr = (np.arange(1, nrows + 1) * radius / nrows)[:, None]              # row i at radius i * dr
    bearing = (np.arange(ncols) * 360.0 / ncols)[None, :]
    speed = np.where(rr < rmax, vmax * rr / rmax, vmax * (rmax / rr) ** 0.6) + 0.0 * bearing
    vdir = np.mod(bearing + 90.0 + 20.0, 360.0) + 0.0 * rr
    pdrop = dp * np.exp(-rmax / rr) + 0.0 * bearing
    rain = rain_mmh * np.exp(-rr / (2 * rmax)) + 0.0 * bearing
    rep = lambda a: np.repeat(a[None].astype(np.float32), nt, axis=0)
    decay = np.linspace(1.0, 0.8, nt, dtype=np.float32)[:, None, None]
    sc.meteo = dict(sc.meteo or {})
    sc.meteo["spw"] = dict(times=t, xe=xe, ye=ye, radius=float(radius), vmag=rep(speed) * decay, vdir=rep(vdir),
                           pdrop=rep(pdrop) * decay, rain=rep(rain))
    sc.name += "_spw"
    return sc


def beach_structures(sc, crest=(0.6, 1.2), cd=0.6, extent=None, cell=None, weir_x=0.4, dam_x=(0.55, 0.8, 0.9)):
    """A weir along a north-south line at `weir_x` of the way inland on a beach with the sea
    in the west, its crest rising from crest[0] in the south to crest[1] in the north, so
    the tide overtops it first in the south; and a thin dam running diagonally across the
    land behind it (vertices at `dam_x` of the length). The lines are offset from cell
    faces and centres by fractions of a cell (`cell`, default the grid spacing; `extent` =
    (x length, y length), default the grid's)."""
    dx = sc.dx if cell is None else cell
    lx, ly = (sc.mmax * sc.dx, sc.nmax * sc.dy) if extent is None else extent
    xw = weir_x * lx + 0.31 * dx
    weir = [(xw, -0.37 * dx, crest[0], cd), (xw + 0.23 * dx, 0.5 * ly, 0.5 * (crest[0] + crest[1]), cd),
            (xw - 0.11 * dx, ly + 0.41 * dx, crest[1], cd)]
    dam = [(dam_x[0] * lx + 0.13 * dx, 0.2 * ly + 0.29 * dx), (dam_x[1] * lx + 0.37 * dx, 0.55 * ly + 0.17 * dx),
           (dam_x[2] * lx + 0.21 * dx, 0.6 * ly + 0.43 * dx)]
    with_weirs(sc, [weir])
    return with_thin_dams(sc, [dam])


def with_river(sc, first=1, ncells=8, slope=1.0e-3):
    """Turns cells `first`..`first+ncells-1` of the south row (n = 0) of a regular-kind
    scenario into downstream river boundary cells (msk=5), forced by two bdr points whose
    interior points lie two cells inside (distance < 0: taken from the two points)."""
    msk = sc.msk.reshape(sc.mmax, sc.nmax).copy()
    msk[first:first + ncells, 0] = 5
    sc.msk = msk.reshape(-1)
    xs = np.array([(first + 1.5) * sc.dx, (first + ncells - 1.5) * sc.dx], np.float32)
    ys = np.full(2, -0.5 * sc.dy, np.float32)
    sc.bdr = (xs, ys, xs, np.full(2, 2.5 * sc.dy, np.float32), np.full(2, slope, np.float32),
              np.full(2, -1.0, np.float32))
    sc.name += "_river"
    return sc


def face_low_edge_west(dem, refi):
    """The DEM turned so that its edge with the most cells whose lowest pixel lies below
    0 m is the west edge (m = 0), where the tidal boundary is. Candidates: as generated,
    mirrored east-west, and on square grids the two transposes; the first of the best
    is kept. A seeded fBm can rise above 0 m along its whole original west edge
    (configuration 1's surface does: 2.1 .. 18.1 m there), which would keep the tide out
    of the domain altogether. dem: (nmax*refi, mmax*refi), rows = n (y), cols = m (x)."""
    cands = [dem, dem[:, ::-1]]
    if dem.shape[0] == dem.shape[1]:
        cands += [dem.T, dem.T[:, ::-1]]
    best, best_n = cands[0], -1
    for c in cands:
        west = c[:, :refi].reshape(-1, refi, refi).min(axis=(1, 2))
        n_low = int(np.count_nonzero(west < 0.0))
        if n_low > best_n:
            best, best_n = c, n_low
    return np.ascontiguousarray(best)


def subgrid_case(nmax, mmax, dx=25.0, refi=5, nl=10, seed=1001, t_end=21600.0, tide_amp=1.5,
                 nsrc=8, rain_mmh=0.0, **params):
    """Configuration-1 recipe at any size (regular mesh with subgrid tables from an fBm DEM)."""
    r = rng(seed)
    dem = face_low_edge_west(fbm(r, nmax * refi, mmax * refi), refi)
    nman = landuse_manning(r, nmax * refi, mmax * refi, refi=refi)
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


def member(seed, t_end=10800.0, nmin=256, nmax=2048, scale=1.0, name="member"):
    """One configuration-3 scenario: log-uniform size, regular or subgrid by a seeded coin
    flip, surge amplitude U(0.5, 4) m, rain U(0, 100) mm/h."""
    r = rng(seed)
    n = int(round(math.exp(r.uniform(math.log(nmin), math.log(nmax))) * scale))
    n = max(16, n)
    amp = float(r.uniform(0.5, 4.0))
    rain = float(r.uniform(0.0, 100.0))
    if r.uniform() < 0.5:
        sc = regular_case(n, n, dx=50.0, seed=seed, t_end=t_end, tide_amp=amp, rain_mmh=rain, rain_until=1800.0)
    else:
        sc = subgrid_case(n, n, dx=25.0, seed=seed, t_end=t_end, tide_amp=amp, nsrc=0, rain_mmh=rain)
    sc.name = f"{name}_{sc.name}"
    return sc


def ensemble_members(count=64, seed_base=1003, t_end=10800.0, nmin=256, nmax=2048, scale=1.0):
    """Configuration-3 members j = 0..count-1, seeded PCG64(seed_base*100000 + j)."""
    return [member(seed_base * 100000 + j, t_end, nmin, nmax, scale, name=f"member{j}") for j in range(count)]


def quadtree_case(nb, mb, nlev=4, dx0=100.0, seed=1002, t_end=43200.0, refine=(-5.0, 5.0), surge_amp=3.0,
                  wind=True, **params):
    """Configuration-2 recipe at any size: quadtree mesh refined where the seeded fBm bed
    level lies within `refine`, 2:1 balanced; Gaussian surge (peak at 6 h, sigma 3 h) on
    the west edge, uniform 20 m/s wind from the south-west."""
    from . import quadtree as Q
    r = rng(seed)
    F = 2 ** (nlev - 1)
    dem = fbm(r, nb * F, mb * F)
    _, leaf, _, _ = Q.build(dem, nlev, refine)
    qt = Q.quadtree_arrays(leaf, nlev, 0.0, 0.0, dx0, dx0)
    zb = Q.leaf_bed_level(qt, dem, nlev)
    mask = np.ones(len(zb), np.int8)
    mask[(qt["m"] == 1) & (zb < 0.0)] = 2
    qt["mask"] = mask
    t = np.arange(0.0, t_end + 1200.0, 600.0, dtype=np.float64)
    zs = (surge_amp * np.exp(-((t - 21600.0) ** 2) / (2.0 * 10800.0 ** 2))).astype(np.float32)[:, None]
    p = dict(t0=0.0, t1=float(t_end))
    p.update(params)
    wnd = None
    if wind:
        tw = np.array([0.0, t_end + 3600.0], np.float32)
        wnd = (tw, np.array([20.0, 20.0], np.float32), np.array([225.0, 225.0], np.float32))
    return Scenario(name=f"quadtree_{nb}x{mb}_L{nlev}", params=p, kind="quadtree", qt=qt, zb_qt=zb,
                    bnd=(np.array([-dx0 / 2], np.float32), np.array([nb * dx0 / 2], np.float32),
                         t.astype(np.float32), zs),
                    wind=wnd, t_end=float(t_end))


def config2(nb=512, mb=512):
    """Configuration 2: quadtree, 4 levels 100 m .. 12.5 m, about 7.7M active cells on a
    512 x 512 base grid (51.2 km square)."""
    return quadtree_case(nb, mb, nlev=4, dx0=100.0, seed=1002, t_end=43200.0)
