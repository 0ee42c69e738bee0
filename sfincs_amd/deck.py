"""Write a scenario as a SFINCS input deck in the reference's file formats.

sfincs.inp (key = value), inputformat = bin: indexfile (int32 np, then the column-major
indices (m-1)*nmax+n of the active cells), mskfile (int8 per index), depfile (float32 per
index) -- sfincs_domain.f90:244-260, 1670-1676; quadtree meshes as the legacy binary
qtrfile (sfincs_quadtree.F90:210-292) with a binary mskfile; the legacy binary sbgfile
(sfincs_subgrid.F90:458-903) for subgrid scenarios; ASCII bnd/bzs, src/dis,
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
    come from such a file (workloads.with_subgrid_binary) can be written: NetCDF subgrid
    files need a library this image does not have, and the two formats do not describe the
    same tables (huthresh, pwet and the fit differ)."""
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


def write(sc, folder):
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
    if sc.subgrid is not None:
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
