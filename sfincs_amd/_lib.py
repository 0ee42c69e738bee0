"""ctypes bindings of libsfx.so (declared in include/sfx.h).

The library is loaded from the package directory. There is no fallback: if the native
library is missing, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfx.so")

_lib = None


class SfxError(RuntimeError):
    pass


class sfx_params(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "t0", "t1", "tspinup", "alpha", "theta", "hmin_cfl", "manning", "manning_land",
        "manning_sea", "rgh_lev_land", "zsini", "dtmax", "huthresh", "huvmin", "rhoa", "rhow",
        "advlim", "slopelim", "btfilter", "btrelax", "wiggle_factor", "wiggle_threshold", "uvlim",
        "uvmax", "nuvisc", "nuviscfac", "latitude", "g", "pi")] + [(n, C.c_int32) for n in (
        "advection", "advection_scheme", "viscosity", "coriolis", "friction2d", "advection_mask",
        "wiggle_suppression", "store_zsmax", "bndtype", "h73table", "store_velmax", "store_fluxmax",
        "store_twet")] + [("twet_threshold", C.c_float), ("timestep_analysis", C.c_int32), ("use_qext", C.c_int32),
                                       ("structure_relax", C.c_float), ("qinf", C.c_float),
                                       ("qinf_zmin", C.c_float), ("dtwnd", C.c_float), ("baro", C.c_int32),
                                       ("pavbnd", C.c_float), ("gapres", C.c_float), ("spinup_meteo", C.c_int32),
                                       ("amprblock", C.c_int32), ("usespwprecip", C.c_int32),
                                       ("spw_merge_frac", C.c_float), ("factor_wind", C.c_float),
                                       ("factor_pres", C.c_float), ("factor_prcp", C.c_float),
                                       ("factor_spw_size", C.c_float), ("sfacinf", C.c_float)]


class sfx_quadtree_input(C.Structure):
    _fields_ = [("nr_points", C.c_int32), ("nr_levels", C.c_int32),
                ("x0", C.c_float), ("y0", C.c_float), ("dx", C.c_float), ("dy", C.c_float),
                ("rotation_deg", C.c_float)] + \
               [(n, C.c_void_p) for n in ("level", "md", "mu", "nd", "nu", "mask")] + \
               [(n, C.c_void_p) for n in ("n", "m", "md1", "md2", "mu1", "mu2", "nd1", "nd2", "nu1", "nu2")]


class sfx_subgrid_input(C.Structure):
    _fields_ = [("nlevels", C.c_int32)] + [(n, C.c_void_p) for n in (
        "z_zmin", "z_zmax", "z_volmax", "z_level", "uv_zmin", "uv_zmax", "uv_navg", "uv_ffit",
        "uv_havg", "uv_nrep", "uv_pwet")]


class sfx_sizes(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "np", "npuv", "ncuv", "nref", "nlevels", "nmax", "mmax", "ngbnd", "nkcuv2", "nsrc", "nbnd",
        "subgrid", "quadtree", "dense")] + [("dtmax", C.c_float), ("dtmin", C.c_float), ("ndrn", C.c_int32),
                                          ("nstruc", C.c_int32), ("nthd", C.c_int32)]


class sfx_step_stats(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_float), ("min_dt", C.c_float), ("nt", C.c_int64),
                ("error", C.c_int32), ("cell_updates", C.c_double), ("seconds", C.c_double)]


def _sig(lib):
    P = C.c_void_p
    I = C.c_int
    f = C.c_float
    d = C.c_double
    sz = C.c_size_t
    s = C.c_char_p
    i32 = C.c_int32
    table = {
        "sfx_params_default": (None, [C.POINTER(sfx_params)]),
        "sfx_model_create_regular": (I, [C.POINTER(sfx_params), i32, i32, f, f, f, f, f, P, C.POINTER(P)]),
        "sfx_model_create_quadtree": (I, [C.POINTER(sfx_params), C.POINTER(sfx_quadtree_input), C.POINTER(P)]),
        "sfx_model_read_subgrid_binary": (I, [P, C.c_char_p]),
        "sfx_model_read_subgrid_netcdf": (I, [P, C.c_char_p]),
        "sfx_model_create_quadtree_netcdf": (I, [C.POINTER(sfx_params), C.c_char_p, i32, C.POINTER(P)]),
        "sfx_model_set_drainage": (I, [P, i32, P, P, P, P, P, P]),
        "sfx_model_set_thin_dams": (I, [P, i32, P, P, P]),
        "sfx_model_set_infiltration": (I, [P, P]),
        "sfx_model_set_curve_number": (I, [P, P]),
        "sfx_model_set_meteo_grid": (I, [P, i32, i32, i32, i32, f, f, f, f, P, P]),
        "sfx_model_set_spiderweb": (I, [P, i32, i32, i32, f, P, P, P, P, P, P, P]),
        "sfx_model_read_meteo_file": (I, [P, s, s, s]),
        "sfx_model_find_cells": (I, [P, i32, P, P, P]),
        "sfx_model_set_weirs": (I, [P, i32, P, P, P, P]),
        "sfx_model_sizes": (I, [P, C.POINTER(sfx_sizes)]),
        "sfx_model_get_table": (I, [P, s, P, sz]),
        "sfx_model_table_bytes": (I, [P, s, C.POINTER(sz)]),
        "sfx_model_set_bathymetry": (I, [P, P]),
        "sfx_model_set_subgrid": (I, [P, C.POINTER(sfx_subgrid_input)]),
        "sfx_model_set_boundary": (I, [P, i32, P, P, i32, P, P]),
        "sfx_model_set_sources": (I, [P, i32, P, P, i32, P, P]),
        "sfx_model_set_rain": (I, [P, i32, P, P]),
        "sfx_model_set_river_boundary": (I, [P, i32, P, P, P, P, P, P]),
        "sfx_model_set_boundary_ig": (I, [P, P]),
        "sfx_model_set_storage_volume": (I, [P, P]),
        "sfx_model_timestep_analysis_per_cell": (I, [P, P, P, P, C.c_int64, P, P]),
        "sfx_model_set_wind": (I, [P, i32, P, P, P]),
        "sfx_model_initialize": (I, [P]),
        "sfx_model_set_restart": (I, [P, s]),
        "sfx_write_restart": (I, [P, s]),
        "sfx_model_get_state": (I, [P, s, P, sz]),
        "sfx_model_destroy": (None, [P]),
        "sfx_model_footprint": (I, [P, C.POINTER(sz)]),
        "sfx_create": (I, [P, i32, C.POINTER(P)]),
        "sfx_destroy": (I, [P]),
        "sfx_get_state": (I, [P, s, P, sz]),
        "sfx_set_state": (I, [P, s, P, sz]),
        "sfx_compute_fluxes": (I, [P, f, C.POINTER(f)]),
        "sfx_compute_water_levels": (I, [P, d, f]),
        "sfx_update_forcing": (I, [P, d, f]),
        "sfx_update": (I, [P, d, C.POINTER(sfx_step_stats)]),
        "sfx_update_to": (I, [P, d, C.POINTER(sfx_step_stats)]),
        "sfx_get_time": (I, [P, C.POINTER(d), C.POINTER(f), C.POINTER(C.c_int64)]),
        "sfx_set_counting": (I, [P, i32]),
        "sfx_get_counters": (I, [P, C.POINTER(C.c_int64)]),
        "sfx_get_store_counters": (I, [P, C.POINTER(C.c_int64)]),
        "sfx_total_volume": (I, [P, C.POINTER(d)]),
        "sfx_get_balance": (I, [P, C.POINTER(d)]),
        "sfx_reset_balance": (I, [P]),
        "sfx_update_zbuvmx": (I, [P]),
        "sfx_total_volume_noalloc": (I, [P, P]),
        "sfx_batch_create": (I, [i32, C.POINTER(P)]),
        "sfx_batch_destroy": (I, [P]),
        "sfx_batch_add": (I, [P, P, d]),
        "sfx_batch_remove": (I, [P, P]),
        "sfx_batch_advance": (I, [P, i32, C.POINTER(i32)]),
        "sfx_batch_done": (I, [P, P, C.POINTER(i32)]),
        "sfx_autotune": (I, [P, i32]),
        "sfx_get_tuning": (I, [P, C.POINTER(i32)]),
        "sfx_set_tuning": (I, [P, C.POINTER(i32)]),
        "sfx_set_work_order": (I, [P, i32, i32]),
        "sfx_last_error": (I, [C.c_char_p, sz]),
        "sfx_profile_steps": (I, [P, i32, C.POINTER(d)]),
        "sfx_create_ex": (I, [P, i32, i32, C.POINTER(P)]),
    }
    for name, (res, args) in table.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load libsfx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SfxError(f"{LIB_PATH} not built; run `python -m sfincs_amd.build`")
        _lib = C.CDLL(LIB_PATH)
        _sig(_lib)
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    lib().sfx_last_error(buf, 1024)
    return buf.value.decode()


def check(rc, what=""):
    if rc < 0:
        raise SfxError(f"{what}: {last_error()} (rc={rc})")
    return rc


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def default_params(**over):
    p = sfx_params()
    lib().sfx_params_default(C.byref(p))
    for k, v in over.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


TABLE_DTYPES = {
    "i8": ("kcs", "z_flags_iref", "uv_flags_iref", "uv_flags_type", "uv_flags_dir", "kcuv", "mask_adv",
           "kfuv", "ibuvdir"),
    "f64": ("zs", "z_volume"),
    "i32": ("z_index_z_n", "z_index_z_m", "z_index_uv_md", "z_index_uv_mu", "z_index_uv_nd", "z_index_uv_nu",
            "z_index_uv_md1", "z_index_uv_md2", "z_index_uv_mu1", "z_index_uv_mu2", "z_index_uv_nd1",
            "z_index_uv_nd2", "z_index_uv_nu1", "z_index_uv_nu2", "uv_index_z_nm", "uv_index_z_nmu",
            "uv_index_u_nmd", "uv_index_u_nmu", "uv_index_u_ndm", "uv_index_u_num", "uv_index_v_ndm",
            "uv_index_v_nm", "uv_index_v_ndmu", "uv_index_v_nmu", "cuv_index_uv", "cuv_index_uv1",
            "cuv_index_uv2", "nmindbnd", "index_kcuv2", "nmbkcuv2", "nmikcuv2", "ibkcuv2", "ind1_bnd_gbp",
            "ind2_bnd_gbp", "nmi_gbp", "nmindsrc", "index_bdr_gbp", "index_zsi_bdr", "drainage_type",
            "thindam_uv_index", "structure_uv_index"),
}


def dtype_of(name):
    for dt, names in TABLE_DTYPES.items():
        if name in names:
            return {"i8": np.int8, "f64": np.float64, "i32": np.int32}[dt]
    return np.float32


class Model:
    """Host-side model (tables + initial state); wraps sfx_model."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def regular(cls, params, nmax, mmax, dx, dy, msk, x0=0.0, y0=0.0, rotation=0.0):
        msk = np.ascontiguousarray(msk, dtype=np.int8).reshape(-1, order="C")
        if msk.size != nmax * mmax:
            raise ValueError("msk must have nmax*mmax entries (n fastest)")
        out = C.c_void_p()
        check(lib().sfx_model_create_regular(C.byref(params), nmax, mmax, x0, y0, dx, dy, rotation,
                                             ptr(msk), C.byref(out)), "create_regular")
        m = cls(out)
        m._keep = [msk]
        return m

    @classmethod
    def quadtree(cls, params, qt):
        """qt: dict with quadtree-file arrays (level, n, m, md, md1, ..., mask) and x0,y0,dx,dy."""
        arrs = {}
        s = sfx_quadtree_input()
        s.nr_points = len(qt["n"])
        s.nr_levels = int(qt["nr_levels"])
        s.x0, s.y0, s.dx, s.dy = qt["x0"], qt["y0"], qt["dx"], qt["dy"]
        s.rotation_deg = qt.get("rotation", 0.0)
        for k in ("level", "md", "mu", "nd", "nu", "mask"):
            arrs[k] = np.ascontiguousarray(qt[k], dtype=np.int8)
            setattr(s, k, arrs[k].ctypes.data)
        for k in ("n", "m", "md1", "md2", "mu1", "mu2", "nd1", "nd2", "nu1", "nu2"):
            arrs[k] = np.ascontiguousarray(qt[k], dtype=np.int32)
            setattr(s, k, arrs[k].ctypes.data)
        out = C.c_void_p()
        check(lib().sfx_model_create_quadtree(C.byref(params), C.byref(s), C.byref(out)), "create_quadtree")
        return cls(out)

    @classmethod
    def quadtree_netcdf(cls, params, path, with_bed=True):
        """NetCDF quadtree file (quadtree_read_file_netcdf), classic formats; with_bed sets
        the active cells' bed level from its z."""
        out = C.c_void_p()
        check(lib().sfx_model_create_quadtree_netcdf(C.byref(params), str(path).encode(), int(bool(with_bed)),
                                                     C.byref(out)), "create_quadtree_netcdf")
        return cls(out)

    def sizes(self):
        s = sfx_sizes()
        check(lib().sfx_model_sizes(self.h, C.byref(s)), "sizes")
        return {k: getattr(s, k) for k, _ in sfx_sizes._fields_}

    def table(self, name):
        n = C.c_size_t()
        check(lib().sfx_model_table_bytes(self.h, name.encode(), C.byref(n)), name)
        dt = dtype_of(name)
        a = np.empty(n.value // np.dtype(dt).itemsize, dtype=dt)
        check(lib().sfx_model_get_table(self.h, name.encode(), ptr(a), n.value), name)
        return a

    def params(self):
        """The model's sfx_params after set-up (huthresh is 0 once NetCDF-style subgrid
        tables are set, as the reference does; the legacy binary file keeps it)."""
        p = sfx_params()
        check(lib().sfx_model_get_table(self.h, b"params", C.byref(p), C.sizeof(p)), "params")
        return p

    def read_subgrid_netcdf(self, path):
        """NetCDF subgrid file (read_subgrid_file_netcdf), classic formats."""
        check(lib().sfx_model_read_subgrid_netcdf(self.h, str(path).encode()), "read_subgrid_netcdf")

    def read_subgrid_binary(self, path):
        """Legacy binary subgrid file (read_subgrid_file_original)."""
        check(lib().sfx_model_read_subgrid_binary(self.h, str(path).encode()), "read_subgrid_binary")

    def set_bathymetry(self, zb):
        zb = np.ascontiguousarray(zb, dtype=np.float32)
        check(lib().sfx_model_set_bathymetry(self.h, ptr(zb)), "set_bathymetry")

    def set_subgrid(self, sg):
        s = sfx_subgrid_input()
        s.nlevels = int(sg["nlevels"])
        keep = []
        for k in ("z_zmin", "z_zmax", "z_volmax", "z_level", "uv_zmin", "uv_zmax", "uv_navg", "uv_ffit",
                  "uv_havg", "uv_nrep", "uv_pwet"):
            a = np.ascontiguousarray(sg[k], dtype=np.float32)
            keep.append(a)
            setattr(s, k, a.ctypes.data)
        check(lib().sfx_model_set_subgrid(self.h, C.byref(s)), "set_subgrid")

    def set_boundary(self, x, y, t=None, zs=None):
        x = np.ascontiguousarray(x, np.float32); y = np.ascontiguousarray(y, np.float32)
        if t is None:
            check(lib().sfx_model_set_boundary(self.h, len(x), ptr(x), ptr(y), 0, None, None), "set_boundary")
            return
        t = np.ascontiguousarray(t, np.float32); zs = np.ascontiguousarray(zs, np.float32)
        check(lib().sfx_model_set_boundary(self.h, len(x), ptr(x), ptr(y), len(t), ptr(t), ptr(zs)), "set_boundary")

    def set_sources(self, x, y, t, q):
        x = np.ascontiguousarray(x, np.float32); y = np.ascontiguousarray(y, np.float32)
        t = np.ascontiguousarray(t, np.float32); q = np.ascontiguousarray(q, np.float32)
        check(lib().sfx_model_set_sources(self.h, len(x), ptr(x), ptr(y), len(t), ptr(t), ptr(q)), "set_sources")

    def set_drainage(self, x_in, y_in, x_out, y_out, types, params):
        """Drainage structures (drnfile): intake and outfall points, type 1-5, 6 params each."""
        a = [np.ascontiguousarray(v, np.float32) for v in (x_in, y_in, x_out, y_out)]
        ty = np.ascontiguousarray(types, np.int32)
        par = np.zeros((len(ty), 6), np.float32)
        pp = np.asarray(params, np.float32).reshape(len(ty), -1)
        par[:, :pp.shape[1]] = pp
        check(lib().sfx_model_set_drainage(self.h, len(ty), *[ptr(v) for v in a], ptr(ty), ptr(par)), "set_drainage")

    @staticmethod
    def _polylines(polylines, npar):
        nv = np.array([len(p) for p in polylines], np.int32)
        rows = np.concatenate([np.asarray(p, np.float32).reshape(len(p), -1) for p in polylines]) if len(nv) else \
            np.zeros((0, 2 + npar), np.float32)
        if rows.shape[1] != 2 + npar:
            raise ValueError(f"polyline vertices need {2 + npar} values (x, y{', parameters' if npar else ''})")
        x = np.ascontiguousarray(rows[:, 0]); y = np.ascontiguousarray(rows[:, 1])
        par = np.ascontiguousarray(rows[:, 2:]).reshape(-1)
        return nv, x, y, par

    def set_thin_dams(self, polylines):
        """Thin dams (thdfile): a list of polylines, each an (n, 2) array of x, y."""
        nv, x, y, _ = self._polylines(polylines, 0)
        check(lib().sfx_model_set_thin_dams(self.h, len(nv), ptr(nv), ptr(x), ptr(y)), "set_thin_dams")

    def set_weirs(self, polylines):
        """Weirs (weirfile): a list of polylines, each an (n, 4) array of x, y, crest level, Cd."""
        nv, x, y, par = self._polylines(polylines, 2)
        check(lib().sfx_model_set_weirs(self.h, len(nv), ptr(nv), ptr(x), ptr(y), ptr(par)), "set_weirs")

    METEO_QUANTITY = {"amu": 0, "amv": 1, "amp": 2, "ampr": 3}

    def set_meteo_grid(self, kind, times, values, x_llcorner, y_llcorner, dx, dy):
        """Gridded meteo (the contents of an amu, amv, amp or ampr file): values[nt, nrows,
        ncols] with row 0 the northernmost, times in s from tref; m/s, Pa or mm/h."""
        t = np.ascontiguousarray(times, np.float32)
        v = np.ascontiguousarray(values, np.float32)
        nt, nr, nc = v.shape
        self.meteo_inputs = getattr(self, "meteo_inputs", {})
        self.meteo_inputs[kind] = dict(times=t, values=v, x_llcorner=float(x_llcorner),
                                       y_llcorner=float(y_llcorner), dx=float(dx), dy=float(dy))
        check(lib().sfx_model_set_meteo_grid(self.h, self.METEO_QUANTITY[kind], nt, nr, nc, x_llcorner, y_llcorner,
                                             dx, dy, ptr(t), ptr(v)), "set_meteo_grid")

    def set_spiderweb(self, times, xe, ye, radius, vmag, vdir, pdrop, rain=None):
        """A spiderweb (spw file): eye track xe/ye per time, [nt, nrows, ncols] polar fields of
        wind speed (m/s), nautical direction (deg), pressure drop (Pa) and rain (mm/h or None)."""
        a = [np.ascontiguousarray(x, np.float32) for x in (times, xe, ye, vmag, vdir, pdrop)]
        r = None if rain is None else np.ascontiguousarray(rain, np.float32)
        nt, nr, nc = a[3].shape
        self.meteo_inputs = getattr(self, "meteo_inputs", {})
        self.meteo_inputs["spw"] = dict(times=a[0], xe=a[1], ye=a[2], radius=float(radius), vmag=a[3], vdir=a[4],
                                        pdrop=a[5], rain=r)
        check(lib().sfx_model_set_spiderweb(self.h, nt, nr, nc, radius, ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]),
                                            ptr(a[4]), ptr(a[5]), ptr(r) if r is not None else None),
              "set_spiderweb")

    def find_cells(self, x, y):
        """The active cell (1-based, 0 outside) holding each point (observation points)."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        nm = np.zeros(x.size, np.int32)
        check(lib().sfx_model_find_cells(self.h, x.size, ptr(x), ptr(y), ptr(nm)), "find_cells")
        return nm

    def read_meteo_file(self, kind, path, tref="20000101 000000"):
        """An ASCII amu / amv / amp / ampr / spw file, read as the reference reads it."""
        check(lib().sfx_model_read_meteo_file(self.h, kind.encode(), str(path).encode(), tref.encode()),
              "read_meteo_file")

    def set_infiltration(self, qinf_mmh):
        """Spatially varying constant infiltration (qinffile): mm/h per active cell. Acts only
        with rain; params.qinf > 0 (uniform) takes precedence."""
        a = np.ascontiguousarray(qinf_mmh, np.float32)
        self._keep_qinf = a
        check(lib().sfx_model_set_infiltration(self.h, ptr(a)), "set_infiltration")

    def set_curve_number(self, scs_inch):
        """Curve-number infiltration, method A (scsfile): maximum retention S per active
        cell in inches. Acts only with rain; qinf > 0 or an infiltration map takes
        precedence."""
        a = np.ascontiguousarray(scs_inch, np.float32)
        self._keep_scs = a
        check(lib().sfx_model_set_curve_number(self.h, ptr(a)), "set_curve_number")

    def set_rain(self, t, mmh):
        t = np.ascontiguousarray(t, np.float32); v = np.ascontiguousarray(mmh, np.float32)
        check(lib().sfx_model_set_rain(self.h, len(t), ptr(t), ptr(v)), "set_rain")

    def timestep_analysis_per_cell(self, tsa_sum, tsa_wet, tsa_lim, nt):
        """timestep_analysis_finalize: per cell (average required step, % of steps limiting)."""
        a = np.ascontiguousarray(tsa_sum, np.float32)
        b = np.ascontiguousarray(tsa_wet, np.int32)
        c = np.ascontiguousarray(tsa_lim, np.int32)
        n = self.sizes()["np"]
        dtc, pct = np.empty(n, np.float32), np.empty(n, np.float32)
        check(lib().sfx_model_timestep_analysis_per_cell(self.h, ptr(a), ptr(b), ptr(c), int(nt), ptr(dtc), ptr(pct)),
              "timestep_analysis_per_cell")
        return dtc, pct

    def set_storage_volume(self, vol):
        """Storage volume per cell in m3 (volfile); subgrid meshes."""
        a = np.ascontiguousarray(vol, np.float32)
        self._keep_vol = a
        check(lib().sfx_model_set_storage_volume(self.h, ptr(a)), "set_storage_volume")

    def set_boundary_ig(self, zsi):
        """Incoming infragravity levels (bzi file): [nt, nbnd] at the water-level series times."""
        a = np.ascontiguousarray(zsi, np.float32)
        self._keep_bzi = a
        check(lib().sfx_model_set_boundary_ig(self.h, ptr(a)), "set_boundary_ig")

    def set_river_boundary(self, x, y, x_in, y_in, slope, distance):
        """Downstream river boundary points (bdr file rows) for msk=5 cells."""
        a = [np.ascontiguousarray(v, np.float32) for v in (x, y, x_in, y_in, slope, distance)]
        self._keep_bdr = a
        check(lib().sfx_model_set_river_boundary(self.h, len(a[0]), *[ptr(v) for v in a]), "set_river_boundary")

    def set_wind(self, t, mag, direction):
        t = np.ascontiguousarray(t, np.float32); a = np.ascontiguousarray(mag, np.float32)
        b = np.ascontiguousarray(direction, np.float32)
        check(lib().sfx_model_set_wind(self.h, len(t), ptr(t), ptr(a), ptr(b)), "set_wind")

    def set_restart(self, path):
        """Binary restart file as initial condition (rstfile); before initialize()."""
        check(lib().sfx_model_set_restart(self.h, str(path).encode()), "set_restart")

    def initialize(self):
        check(lib().sfx_model_initialize(self.h), "initialize")
        return self

    def footprint(self):
        n = C.c_size_t()
        check(lib().sfx_model_footprint(self.h, C.byref(n)), "footprint")
        return n.value

    def close(self):
        if self.h:
            lib().sfx_model_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


STATE_FIELDS = {"zs": np.float64, "z_volume": np.float64, "q": np.float32, "uv": np.float32,
                "kfuv": np.int8, "zs0": np.float32, "zsderv": np.float32, "zsmax": np.float32,
                "uvmean": np.float32, "vmax": np.float32, "qmax": np.float32, "twet": np.float32,
                "tsa_sum": np.float32, "tsa_wet": np.int32, "tsa_lim": np.int32, "qext": np.float32,
                "storage_volume": np.float32, "zb": np.float32, "qtsrc": np.float32, "cumprcp": np.float32,
                "cuminf": np.float32}


class Instance:
    """A model resident in HBM (sfx_handle). One instance per scenario."""

    def __init__(self, model, device=0, fp64=False):
        out = C.c_void_p()
        check(lib().sfx_create_ex(model.h, device, 1 if fp64 else 0, C.byref(out)), "sfx_create")
        self.h = out
        self.sizes = model.sizes()
        self._nbatch = 0         # batches holding this instance (they keep a reference)

    def _n(self, name):
        s = self.sizes
        if name in ("zs", "z_volume", "zs0", "zsderv", "zsmax", "vmax", "qmax", "twet", "qext", "storage_volume",
                    "zb", "cumprcp", "cuminf"):
            return s["np"]
        if name in ("q", "uv"):
            return s["npuv"] + s["ncuv"] + 1
        if name in ("kfuv", "tsa_sum", "tsa_wet", "tsa_lim"):
            return s["npuv"]
        if name == "uvmean":
            return s["nkcuv2"]
        if name == "qtsrc":
            return s["nsrc"] + 2 * s["ndrn"]
        raise KeyError(name)

    def get(self, name):
        a = np.empty(self._n(name), dtype=STATE_FIELDS[name])
        check(lib().sfx_get_state(self.h, name.encode(), ptr(a), a.nbytes), f"get {name}")
        return a

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=STATE_FIELDS[name])
        check(lib().sfx_set_state(self.h, name.encode(), ptr(a), a.nbytes), f"set {name}")

    def write_restart(self, path):
        """Type-1 binary restart file of the current state (write_rst_file)."""
        check(lib().sfx_write_restart(self.h, str(path).encode()), "write_restart")

    def compute_fluxes(self, dt):
        out = C.c_float()
        check(lib().sfx_compute_fluxes(self.h, dt, C.byref(out)), "compute_fluxes")
        return out.value

    def compute_water_levels(self, t, dt):
        check(lib().sfx_compute_water_levels(self.h, t, dt), "compute_water_levels")

    def update_forcing(self, t, dt):
        check(lib().sfx_update_forcing(self.h, t, dt), "update_forcing")

    def update(self, dtrange=-1000.0):
        st = sfx_step_stats()
        rc = check(lib().sfx_update(self.h, dtrange, C.byref(st)), "update")
        return rc, {k: getattr(st, k) for k, _ in sfx_step_stats._fields_}

    def update_to(self, t_end):
        """Steps while t < t_end (absolute model time)."""
        st = sfx_step_stats()
        rc = check(lib().sfx_update_to(self.h, t_end, C.byref(st)), "update_to")
        return rc, {k: getattr(st, k) for k, _ in sfx_step_stats._fields_}

    def time(self):
        t = C.c_double(); dt = C.c_float(); nt = C.c_int64()
        check(lib().sfx_get_time(self.h, C.byref(t), C.byref(dt), C.byref(nt)), "get_time")
        return t.value, dt.value, nt.value

    def set_counting(self, on):
        check(lib().sfx_set_counting(self.h, int(on)), "set_counting")

    def counters(self):
        a = (C.c_int64 * 6)()
        check(lib().sfx_get_counters(self.h, a), "get_counters")
        return list(a)

    def store_counters(self):
        """Continuity since counting was enabled: cells that stored [zs, z_volume, zs0,
        zsderv, zsmax], and steady cells skipped."""
        a = (C.c_int64 * 6)()
        check(lib().sfx_get_store_counters(self.h, a), "get_store_counters")
        return list(a)

    def total_volume(self):
        v = C.c_double()
        check(lib().sfx_total_volume(self.h, C.byref(v)), "total_volume")
        return v.value

    def balance(self):
        """FP64 volume balance of the interior cells (sfx_get_balance): volume and storage
        left now, and the boundary, source, rain and qext inflows since creation or reset."""
        a = (C.c_double * 6)()
        check(lib().sfx_get_balance(self.h, a), "get_balance")
        return dict(zip(("volume", "storage", "boundary", "sources", "rain", "qext"), list(a)))

    def update_zbuvmx(self):
        """update_zbuv: zbuvmx from the device's bed levels (after a write of "zb")."""
        check(lib().sfx_update_zbuvmx(self.h), "update_zbuvmx")

    def reset_balance(self):
        check(lib().sfx_reset_balance(self.h), "reset_balance")

    def autotune(self, trials=3):
        check(lib().sfx_autotune(self.h, trials), "autotune")
        return self.tuning()

    def tuning(self):
        """The shapes and work orders in use, as autotune() reports them."""
        a = (C.c_int32 * 8)()
        check(lib().sfx_get_tuning(self.h, a), "get_tuning")
        return dict(dense=a[0], momentum_kernel=a[1], momentum_tile=[a[2], a[3]], continuity_tile=[a[4], a[5]],
                    xcd_run=[a[6], a[7]])

    def set_tuning(self, momentum_kernel=0, momentum_tile=(64, 4), continuity_tile=(64, 4)):
        """momentum_kernel: -1, -2, -4 one thread per 1, 2 or 4 uv-points (momentum_tile[1]
        rows apart), 0 one thread per cell, k > 0 strip march of k rows per wave."""
        a = (C.c_int32 * 5)(momentum_kernel, momentum_tile[0], momentum_tile[1], continuity_tile[0],
                            continuity_tile[1])
        check(lib().sfx_set_tuning(self.h, a), "set_tuning")

    def set_work_order(self, mom_run, cont_run):
        """Runs of tiles dealt to the XCDs in turn (0: one band per XCD), as autotune picks."""
        check(lib().sfx_set_work_order(self.h, int(mom_run), int(cont_run)), "set_work_order")

    def pin_tuning(self, tuning):
        """Pin a tuning dict returned by autotune() on this handle (same mesh shape)."""
        if tuning["dense"]:
            self.set_tuning(tuning["momentum_kernel"], tuning["momentum_tile"], tuning["continuity_tile"])
        self.set_work_order(*tuning["xcd_run"])

    def profile(self, nsteps):
        ms = (C.c_double * 3)()
        check(lib().sfx_profile_steps(self.h, nsteps, ms), "profile_steps")
        return list(ms)

    def close(self):
        if self.h:
            if self._nbatch:
                raise SfxError("close: the instance is still a member of a batch")
            lib().sfx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Members of uneven size advanced together (sfx_batch)."""

    def __init__(self, device=0):
        out = C.c_void_p()
        check(lib().sfx_batch_create(device, C.byref(out)), "batch_create")
        self.h = out
        self.members = []

    def add(self, inst, t_end):
        check(lib().sfx_batch_add(self.h, inst.h, t_end), "batch_add")
        self.members.append(inst)
        inst._nbatch += 1

    def remove(self, inst):
        check(lib().sfx_batch_remove(self.h, inst.h), "batch_remove")
        self.members.remove(inst)
        inst._nbatch -= 1

    def advance(self, nsteps):
        n = C.c_int32()
        check(lib().sfx_batch_advance(self.h, nsteps, C.byref(n)), "batch_advance")
        return n.value

    def done(self, inst):
        d = C.c_int32()
        check(lib().sfx_batch_done(self.h, inst.h, C.byref(d)), "batch_done")
        return bool(d.value)

    def close(self):
        if self.h:
            lib().sfx_batch_destroy(self.h)
            self.h = C.c_void_p()
            for m in self.members:
                m._nbatch -= 1
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
