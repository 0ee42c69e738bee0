"""Write a scenario as a SFINCS input deck in the reference's file formats.

sfincs.inp (key = value), inputformat = bin: indexfile (int32 np, then the column-major
indices (m-1)*nmax+n of the active cells), mskfile (int8 per index), depfile (float32 per
index) -- sfincs_domain.f90:244-260, 1670-1676; quadtree meshes as the legacy binary
qtrfile (sfincs_quadtree.F90:210-292) with a binary mskfile, or (netcdf=True) as the
NetCDF qtrfile (:295-442); subgrid scenarios as the legacy binary sbgfile
(sfincs_subgrid.F90:458-903) or (netcdf=True) the NetCDF sbgfile (:57-455), both NetCDF
files in the classic 64-bit-offset format (scipy writes it); ASCII bnd/bzs, src/dis,
prc and wnd files (sfincs_boundaries.f90:8-140, sfincs_discharges.f90:36-126,
sfincs_meteo.f90:247-312). Values are written with 9 significant digits so that
float32 round-trips exactly.
"""
import datetime
import os

import numpy as np

_DEFAULTS_SKIP = {"t0", "t1"}


def _fmt(v):
    return f"{np.float32(v):.9g}"


def _date(seconds):
    d = datetime.datetime(2000, 1, 1) + datetime.timedelta(seconds=float(seconds))
    return d.strftime("%Y%m%d %H%M%S")


def _write_subgrid(sc, folder):
    """The scenario's legacy binary subgrid file as sfincs.sbg. Only scenarios whose tables
    come from such a file (workloads.with_subgrid_binary) can be written this way: the two
    formats do not describe the same tables (huthresh, pwet and the fit differ), so tables
    in the NetCDF semantics go to a NetCDF file (write(netcdf=True))."""
    import shutil
    from . import _lib
    p = _lib.default_params(**sc.params)
    if sc.kind == "regular":
        mesh = _lib.Model.regular(p, sc.nmax, sc.mmax, sc.dx, sc.dy, sc.msk)
    else:
        mesh = _lib.Model.quadtree(p, sc.qt)
    sg = sc.subgrid(mesh)
    if not isinstance(sg, str):
        raise ValueError("deck.write: subgrid tables must come from a legacy binary file (with_subgrid_binary)")
    shutil.copyfile(sg, os.path.join(folder, "sfincs.sbg"))


def _model_of(sc):
    from . import _lib
    p = _lib.default_params(**sc.params)
    if sc.kind == "regular":
        return _lib.Model.regular(p, sc.nmax, sc.mmax, sc.dx, sc.dy, sc.msk)
    return _lib.Model.quadtree(p, sc.qt)


def write_quadtree_netcdf(path, qt, z, version=2):
    """Quadtree arrays (the dict Model.quadtree takes) and the bed level z of every
    quadtree point as the NetCDF quadtree file: dimension mesh2d_nFaces, variables n, m,
    level, md/mu/nd/nu (+1, +2), z, mask, global attributes x0, y0, dx, dy, rotation, nmax,
    mmax, nr_levels (sfincs_quadtree.F90:295-442)."""
    from scipy.io import netcdf_file
    npt = len(qt["n"])
    with netcdf_file(path, "w", version=version) as f:
        f.createDimension("mesh2d_nFaces", npt)
        for k in ("n", "m", "md1", "md2", "mu1", "mu2", "nd1", "nd2", "nu1", "nu2"):
            f.createVariable(k, "i4", ("mesh2d_nFaces",))[:] = np.asarray(qt[k], np.int32)
        for k in ("level", "md", "mu", "nd", "nu", "mask"):
            f.createVariable(k, "b", ("mesh2d_nFaces",))[:] = np.asarray(qt[k], np.int8)
        f.createVariable("z", "f4", ("mesh2d_nFaces",))[:] = np.asarray(z, np.float32)
        lev = np.asarray(qt["level"], np.int64)
        f.x0, f.y0 = np.float32(qt["x0"]), np.float32(qt["y0"])
        f.dx, f.dy = np.float32(qt["dx"]), np.float32(qt["dy"])
        f.rotation = np.float32(qt.get("rotation", 0.0))
        f.nmax = np.int32(np.max((np.asarray(qt["n"]) - 1) // 2 ** (lev - 1)) + 1)
        f.mmax = np.int32(np.max((np.asarray(qt["m"]) - 1) // 2 ** (lev - 1)) + 1)
        f.nr_levels = np.int32(qt["nr_levels"])


def subgrid_file_order(model, qt=None):
    """The file record of each active cell and uv-point in a NetCDF subgrid file: sfincs
    order on regular meshes; on quadtree meshes the quadtree point, and per quadtree point
    its right then upper uv-points (sfincs_subgrid.F90:156-240). Returns (zsrc, uvsrc, np
    records, npuv records)."""
    s = model.sizes()
    if qt is None:
        return np.arange(s["np"]), np.arange(s["npuv"]), s["np"], s["npuv"]
    mask = np.asarray(qt["mask"])
    sidx = np.where(mask > 0, np.cumsum(mask > 0), 0)
    zsrc = np.nonzero(mask > 0)[0]
    t = {k: model.table(f"z_index_uv_{k}") for k in ("mu1", "mu2", "nu1", "nu2")}
    uvsrc, nq = [], 0
    for ip in range(len(mask)):
        nm = int(sidx[ip])
        for side, j in (("mu", "1"), ("mu", "2"), ("nu", "1"), ("nu", "2")):
            if int(qt[side + j][ip]) > 0 and (j == "1" or int(qt[side][ip]) >= 1):
                nq += 1
                if nm > 0 and t[side + j][nm - 1] > 0:
                    uvsrc.append(nq - 1)
    return zsrc, np.asarray(uvsrc), len(mask), nq


def write_subgrid_netcdf(path, sg, model, qt=None, version=2, fill=0.0):
    """NetCDF-style tables `sg` (per active cell / uv-point, as Model.set_subgrid takes them)
    as the NetCDF subgrid file, records in the file order of subgrid_file_order; records of
    inactive points hold `fill`."""
    from scipy.io import netcdf_file
    nl = int(sg["nlevels"])
    zsrc, uvsrc, npf, nuvf = subgrid_file_order(model, qt)
    with netcdf_file(path, "w", version=version) as f:
        f.createDimension("np", npf)
        f.createDimension("npuv", nuvf)
        f.createDimension("levels", nl)
        for k, src, n in (("z_zmin", zsrc, npf), ("z_zmax", zsrc, npf), ("z_volmax", zsrc, npf),
                          ("uv_zmin", uvsrc, nuvf), ("uv_zmax", uvsrc, nuvf), ("uv_navg", uvsrc, nuvf),
                          ("uv_ffit", uvsrc, nuvf)):
            a = np.full(n, fill, np.float32)
            a[src] = np.asarray(sg[k], np.float32)
            f.createVariable(k, "f4", ("npuv" if k.startswith("uv") else "np",))[:] = a
        for k, src, n, d in (("z_level", zsrc, npf, "np"), ("uv_havg", uvsrc, nuvf, "npuv"),
                             ("uv_nrep", uvsrc, nuvf, "npuv"), ("uv_pwet", uvsrc, nuvf, "npuv")):
            a = np.full((n, nl), fill, np.float32)
            a[src] = np.asarray(sg[k], np.float32).reshape(-1, nl)
            f.createVariable(k, "f4", (d, "levels"))[:] = a


_QUANT = {"amu": ("x_wind", "m s-1"), "amv": ("y_wind", "m s-1"), "amp": ("air_pressure", "Pa"),
          "ampr": ("precipitation", "mm/hr")}


def _rows(f, a):
    for row in a:
        f.write(" ".join(_fmt(v) for v in row) + "\n")


def _write_amuv(path, kind, g):
    """An equidistant-grid meteo file (read_amuv_file, sfincs_spiderweb.f90:135-229): header
    keys, then per time a TIME line -- here in minutes since 12:00 on the day before tref, so
    that the reader's date offset is exercised -- and the rows from the north."""
    v = np.asarray(g["values"], np.float32)
    nt, nr, nc = v.shape
    q, u = _QUANT[kind]
    with open(path, "w") as f:
        f.write(f"### {kind} file written by sfincs_amd.deck\n")
        for k, val in (("FileVersion", "1.03"), ("filetype", "meteo_on_equidistant_grid"), ("NODATA_value", "-999"),
                       ("n_cols", str(nc)), ("n_rows", str(nr)), ("grid_unit", "m"),
                       ("x_llcorner", _fmt(g["x_llcorner"])), ("y_llcorner", _fmt(g["y_llcorner"])),
                       ("dx", _fmt(g["dx"])), ("dy", _fmt(g["dy"])), ("n_quantity", "1"), ("quantity1", q),
                       ("unit1", u)):
            f.write(f"{k:<17s}=   {val}\n")
        for it in range(nt):
            f.write(f"TIME = {_fmt((float(g['times'][it]) + 43200.0) / 60.0)} minutes since "
                    f"1999-12-31 12:00:00 +00:00\n")
            _rows(f, v[it])


def _write_spw(path, kind, g):
    """A spiderweb file (read_spw_file, sfincs_spiderweb.f90:8-133): header, then per time a
    TIME line (hours since tref), the eye (its values within the 12 characters after '='),
    and the rows of speed, direction, pressure drop and, with rain, precipitation."""
    vm = np.asarray(g["vmag"], np.float32)
    nt, nr, nc = vm.shape
    rain = g.get("rain")
    nq = 4 if rain is not None else 3
    with open(path, "w") as f:
        f.write("### spiderweb written by sfincs_amd.deck\n")
        for k, val in (("FileVersion", "1.03"), ("filetype", "meteo_on_spiderweb_grid"), ("NODATA_value", "-999"),
                       ("n_cols", str(nc)), ("n_rows", str(nr)), ("grid_unit", "m"), ("spw_radius", _fmt(g["radius"])),
                       ("spw_rad_unit", "m"), ("n_quantity", str(nq)), ("quantity1", "wind_speed"),
                       ("quantity2", "wind_from_direction"), ("quantity3", "p_drop")) + \
                ((("quantity4", "precipitation"),) if nq == 4 else ()):
            f.write(f"{k:<17s}=   {val}\n")
        for it in range(nt):
            f.write(f"TIME = {_fmt(float(g['times'][it]) / 3600.0)} hours since 2000-01-01 00:00:00 +00:00\n")
            for key, val in (("x_spw_eye", g["xe"][it]), ("y_spw_eye", g["ye"][it])):
                txt = _fmt(val)
                if len(txt) > 11:
                    raise ValueError(f"deck.write: spiderweb eye {key} = {txt} needs more than 11 characters")
                f.write(f"{key:<17s}= {txt}\n")
            f.write(f"{'p_drop_spw_eye':<17s}= 0\n")
            _rows(f, vm[it])
            _rows(f, np.asarray(g["vdir"], np.float32)[it])
            _rows(f, np.asarray(g["pdrop"], np.float32)[it])
            if nq == 4:
                _rows(f, np.asarray(rain, np.float32)[it])


def write(sc, folder, netcdf=False):
    """Writes the deck; netcdf=True writes quadtree meshes and subgrid tables as NetCDF
    files (qtrfile sfincs_qtr.nc, sbgfile sfincs_subgrid.nc)."""
    os.makedirs(folder, exist_ok=True)
    p = dict(sc.params)
    lines = []
    t0 = float(p.get("t0", 0.0))
    t1 = float(p.get("t1", sc.t_end))
    lines += [f"tref = {_date(0)}", f"tstart = {_date(t0)}", f"tstop = {_date(t1)}"]
    for k, v in p.items():
        if k in _DEFAULTS_SKIP:
            continue
        if k == "zsini":
            lines.append(f"zsini = {_fmt(v)}")
        elif isinstance(v, (int, np.integer)) and k in ("viscosity", "coriolis", "friction2d", "advection_mask",
                                                         "wiggle_suppression", "h73table"):
            lines.append(f"{k} = {int(v)}")
        elif k == "advection_scheme":
            lines.append(f"advection_scheme = {'original' if int(v) == 0 else 'upw1'}")
        elif k == "advection":
            lines.append(f"advection = {int(v)}")
        elif k in ("store_velmax", "store_fluxmax", "store_twet"):   # the input names of the switches
            lines.append(f"{'store' + k[6:]} = {int(v)}")
        elif k == "timestep_analysis":
            lines.append(f"timestep_analysis = {int(v)}")
        elif k == "structure_relax":
            lines.append(f"structure_relax = {_fmt(v)}")
        elif k == "spw_merge_frac":
            lines.append(f"spwmergefrac = {_fmt(v)}")
        elif k in ("baro", "spinup_meteo", "amprblock", "usespwprecip"):
            lines.append(f"{k} = {int(v)}")
        else:
            lines.append(f"{k} = {_fmt(v)}")
    if sc.subgrid is not None and netcdf:
        mesh = _model_of(sc)
        sg = sc.subgrid(mesh)
        if isinstance(sg, str):
            raise ValueError("deck.write: a legacy binary subgrid file cannot be written as NetCDF")
        write_subgrid_netcdf(os.path.join(folder, "sfincs_subgrid.nc"), sg, mesh,
                             qt=None if sc.kind == "regular" else sc.qt)
        lines.append("sbgfile = sfincs_subgrid.nc")
    elif sc.subgrid is not None:
        _write_subgrid(sc, folder)
        lines.append("sbgfile = sfincs.sbg")
    if sc.kind == "regular":
        lines += [f"nmax = {sc.nmax}", f"mmax = {sc.mmax}", f"dx = {_fmt(sc.dx)}", f"dy = {_fmt(sc.dy)}",
                  "x0 = 0", "y0 = 0", "rotation = 0", "inputformat = bin", "indexfile = sfincs.ind",
                  "mskfile = sfincs.msk"]
        if sc.subgrid is None:
            lines.append("depfile = sfincs.dep")
        msk = sc.msk.reshape(sc.mmax, sc.nmax)          # [m][n]
        mm, nn = np.nonzero(msk > 0)
        order = np.lexsort((nn, mm))
        mm, nn = mm[order], nn[order]
        ind = (mm * sc.nmax + nn + 1).astype(np.int32)
        with open(os.path.join(folder, "sfincs.ind"), "wb") as f:
            f.write(np.int32(len(ind)).tobytes())
            f.write(ind.tobytes())
        with open(os.path.join(folder, "sfincs.msk"), "wb") as f:
            f.write(msk[mm, nn].astype(np.int8).tobytes())
        if sc.subgrid is None:
            with open(os.path.join(folder, "sfincs.dep"), "wb") as f:
                f.write(sc.zb_grid[mm, nn].astype(np.float32).tobytes())
    elif netcdf:
        lines.append("qtrfile = sfincs_qtr.nc")
        write_quadtree_netcdf(os.path.join(folder, "sfincs_qtr.nc"), sc.qt, sc.zb_qt)
    else:
        qt = sc.qt
        lines += ["qtrfile = sfincs.qtr", "mskfile = sfincs.msk"]
        with open(os.path.join(folder, "sfincs.qtr"), "wb") as f:
            f.write(np.int8(1).tobytes())
            f.write(np.int32(0).tobytes())
            f.write(np.int32(len(qt["n"])).tobytes())
            f.write(np.int8(qt["nr_levels"]).tobytes())
            for k in ("x0", "y0", "dx", "dy"):
                f.write(np.float32(qt[k]).tobytes())
            f.write(np.float32(qt.get("rotation", 0.0)).tobytes())
            f.write(qt["level"].astype(np.int8).tobytes())
            f.write(qt["n"].astype(np.int32).tobytes())
            f.write(qt["m"].astype(np.int32).tobytes())
            for side in ("nu", "mu", "nd", "md"):
                f.write(qt[side].astype(np.int8).tobytes())
                f.write(qt[side + "1"].astype(np.int32).tobytes())
                f.write(qt[side + "2"].astype(np.int32).tobytes())
            f.write(sc.zb_qt.astype(np.float32).tobytes())
        with open(os.path.join(folder, "sfincs.msk"), "wb") as f:
            f.write(qt["mask"].astype(np.int8).tobytes())
    if sc.bnd is not None:
        x, y, t, zs = sc.bnd
        lines += ["bndfile = sfincs.bnd", "bzsfile = sfincs.bzs"]
        with open(os.path.join(folder, "sfincs.bnd"), "w") as f:
            for a, b in zip(x, y):
                f.write(f"{_fmt(a)} {_fmt(b)}\n")
        with open(os.path.join(folder, "sfincs.bzs"), "w") as f:
            for it in range(len(t)):
                f.write(_fmt(t[it]) + " " + " ".join(_fmt(v) for v in zs[it]) + "\n")
        if getattr(sc, "bzi", None) is not None:
            lines.append("bzifile = sfincs.bzi")
            with open(os.path.join(folder, "sfincs.bzi"), "w") as f:
                for it in range(len(t)):
                    f.write(_fmt(t[it]) + " " + " ".join(_fmt(v) for v in sc.bzi[it]) + "\n")
    if getattr(sc, "bdr", None) is not None:
        lines.append("bdrfile = sfincs.bdr")
        with open(os.path.join(folder, "sfincs.bdr"), "w") as f:
            for row in zip(*sc.bdr):
                f.write(" ".join(_fmt(v) for v in row) + "\n")
    if getattr(sc, "drn", None) is not None:   # x_in y_in x_out y_out type params (sfincs_discharges.f90:193-224)
        lines.append("drnfile = sfincs.drn")
        xi, yi, xo, yo, ty, par = sc.drn
        with open(os.path.join(folder, "sfincs.drn"), "w") as f:
            for k in range(len(ty)):
                npar = 1 if int(ty[k]) <= 3 else 6
                f.write(" ".join(_fmt(v) for v in (xi[k], yi[k], xo[k], yo[k])) + f" {int(ty[k])} " +
                        " ".join(_fmt(v) for v in par[k][:npar]) + "\n")
    # thin dams and weirs: tekal polyline blocks (sfincs_structures.f90:113-160, 421-470)
    for key, fname, polys in (("thdfile", "sfincs.thd", getattr(sc, "thd", None)),
                              ("weirfile", "sfincs.weir", getattr(sc, "weirs", None))):
        if polys is None:
            continue
        lines.append(f"{key} = {fname}")
        with open(os.path.join(folder, fname), "w") as f:
            for k, pl in enumerate(polys):
                pl = np.asarray(pl, np.float32)
                f.write(f"line{k + 1:04d}\n{pl.shape[0]} {pl.shape[1]}\n")
                for row in pl:
                    f.write(" ".join(_fmt(v) for v in row) + "\n")
    for kind, g in (getattr(sc, "meteo", None) or {}).items():   # ASCII meteo files (sfincs_spiderweb.f90)
        fname = f"sfincs.{kind}"
        lines.append(f"{kind}file = {fname}")
        (_write_spw if kind == "spw" else _write_amuv)(os.path.join(folder, fname), kind, g)
    if getattr(sc, "qinf_map", None) is not None:   # float32 mm/h per active cell (sfincs_infiltration.f90:245-249)
        from . import _lib
        p = _lib.default_params(**sc.params)
        if sc.kind == "regular":
            mesh = _lib.Model.regular(p, sc.nmax, sc.mmax, sc.dx, sc.dy, sc.msk)
        else:
            mesh = _lib.Model.quadtree(p, sc.qt)
        lines.append("qinffile = sfincs.qinf")
        np.ascontiguousarray(sc.qinf_map(mesh), np.float32).tofile(os.path.join(folder, "sfincs.qinf"))
    if getattr(sc, "scs_map", None) is not None:    # float32 inches per active cell (sfincs_infiltration.f90:284-286)
        lines.append("scsfile = sfincs.scs")
        np.ascontiguousarray(sc.scs_map(_model_of(sc)), np.float32).tofile(os.path.join(folder, "sfincs.scs"))
    if sc.src is not None:
        x, y, t, q = sc.src
        lines += ["srcfile = sfincs.src", "disfile = sfincs.dis"]
        with open(os.path.join(folder, "sfincs.src"), "w") as f:
            for a, b in zip(x, y):
                f.write(f"{_fmt(a)} {_fmt(b)}\n")
        with open(os.path.join(folder, "sfincs.dis"), "w") as f:
            for it in range(len(t)):
                f.write(_fmt(t[it]) + " " + " ".join(_fmt(v) for v in q[it]) + "\n")
    if sc.rain is not None:
        t, v = sc.rain
        lines.append("prcfile = sfincs.prc")
        with open(os.path.join(folder, "sfincs.prc"), "w") as f:
            for a, b in zip(t, v):
                f.write(f"{_fmt(a)} {_fmt(b)}\n")
    if sc.wind is not None:
        t, mag, d = sc.wind
        lines.append("wndfile = sfincs.wnd")
        with open(os.path.join(folder, "sfincs.wnd"), "w") as f:
            for a, b, c in zip(t, mag, d):
                f.write(f"{_fmt(a)} {_fmt(b)} {_fmt(c)}\n")
    with open(os.path.join(folder, "sfincs.inp"), "w") as f:
        f.write("\n".join(lines) + "\n")
    return folder
