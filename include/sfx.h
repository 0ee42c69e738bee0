/*
 * sfx.h -- C ABI of the MI355X-native SFINCS explicit time step.
 *
 * The library replaces the hot path of SFINCS (momentum flux update on uv-points,
 * continuity water-level update on cells, CFL minimum-dt reduction) together with
 * the per-step forcing it needs, behind plain C entry points: no C++ types, no torch
 * types, plain pointers and sizes. Every entry point returns an int status:
 *   0  = ok,
 *   1  = instability stop (the reference's error=1, sfincs_lib.f90:634-648),
 *  <0  = error; the message is available from sfx_last_error().
 *
 * Index tables exported by sfx_model_get_table() carry the reference's 1-based
 * values (dummy uv index = npuv+ncuv+1), stored in 0-based C arrays.
 * State fields are exchanged in the reference's element order (active cells
 * level-major then column-major; uv-points per cell right-u then above-v,
 * followed by the ncuv combined points).
 *
 * Threading: one host thread per handle; handles are independent.
 */
#ifndef SFX_H
#define SFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFX_OK            0
#define SFX_UNSTABLE      1
#define SFX_ERR_ARG      -1
#define SFX_ERR_STATE    -2
#define SFX_ERR_HIP      -3
#define SFX_ERR_MEMORY   -4
#define SFX_ERR_IO       -5

/* Physics switches and constants, named after the sfincs.inp keywords.
 * Defaults (sfx_params_default) are the reference's: sfincs_input.f90:50-343,
 * derived flags :347-745, constants g=9.81, pi=3.14159 (:372-373). */
typedef struct sfx_params {
  float t0, t1;              /* simulation start/stop in seconds (tstart/tstop - tref) */
  float tspinup;             /* seconds after t0 (the sfincs.inp key, default 0); boundary levels
                              * ramp from zsini until t0 + tspinup (sfincs_input.f90:60, 369) */
  float alpha;               /* CFL factor, default 0.5 */
  float theta;               /* 1.0 = no theta smoothing */
  float hmin_cfl;            /* 0.1 */
  float manning;             /* 0.04 */
  float manning_land;        /* -999 = use manning */
  float manning_sea;         /* -999 = use manning */
  float rgh_lev_land;        /* 0.0 */
  float zsini;               /* 0.0 */
  float dtmax;               /* 60.0 */
  float huthresh;            /* 0.05 (forced to 0 for subgrid tables, sfincs_subgrid.F90:40) */
  float huvmin;              /* 0.0 */
  float rhoa, rhow;          /* 1.25, 1024 */
  float advlim;              /* 1.0 */
  float slopelim;            /* 9999.9 */
  float btfilter, btrelax;   /* 60, 3600 */
  float wiggle_factor;       /* 0.1 */
  float wiggle_threshold;    /* 0.1 */
  float uvlim, uvmax;        /* 10, 1000 */
  float nuvisc;              /* nuviscdim 0.01 */
  float nuviscfac;           /* 100 */
  float latitude;            /* 0 -> no Coriolis */
  float g, pi;               /* 9.81, 3.14159 */
  int32_t advection;         /* 1 */
  int32_t advection_scheme;  /* 1 = upw1, 0 = original */
  int32_t viscosity;         /* 0 (code default, sfincs_input.f90:111) */
  int32_t coriolis;          /* 1, switched off at |latitude| < 0.01 */
  int32_t friction2d;        /* 1 */
  int32_t advection_mask;    /* 1 */
  int32_t wiggle_suppression;/* 1 (used by subgrid only) */
  int32_t store_zsmax;       /* 1 (dtmaxout > 0 by default) */
  int32_t bndtype;           /* 1 = weakly reflective */
  int32_t h73table;          /* 0 */
  /* compute_store_variables (sfincs_continuity.f90:626-720), all off by default; the
   * reference's switches derived at sfincs_input.f90:454-463, 501-504 */
  int32_t store_velmax;      /* vmax: storevelmax == 1 and dtmaxout > 0 */
  int32_t store_fluxmax;     /* qmax: storefluxmax == 1 and dtmaxout > 0 */
  int32_t store_twet;        /* twet: storetwet == 1 */
  float twet_threshold;      /* 0.01 (sfincs_input.f90:285) */
  int32_t timestep_analysis; /* 0 (sfincs_input.f90:288): per uv-point time step statistics,
                              * read as tsa_sum / tsa_wet / tsa_lim (sfx_get_state) */
  int32_t use_qext;          /* 0: external sources per cell (m/s), written with sfx_set_state
                              * "qext" (sfincs_continuity.f90:122-128, 473-497); the BMI
                              * host turns it on as sfincs_bmi.f90:60 does */
  float structure_relax;     /* 10 s: relaxation of drainage structure discharges
                              * (sfincs_input.f90:181, sfincs_discharges.f90:621-625) */
  float qinf;                /* 0 mm/h: uniform constant infiltration (sfincs_input.f90:78,376),
                              * on cells whose bed (zb, or the subgrid zmin) lies above qinf_zmin */
  float qinf_zmin;           /* 0 m (sfincs_input.f90:100) */
  /* gridded and spiderweb meteo (sfincs_input.f90:69,92-113,200-203; sfincs_meteo.f90) */
  float dtwnd;               /* 1800 s: interval of the meteo fields' updates */
  int32_t baro;              /* 1: atmospheric pressure from amp / spiderweb acts */
  float pavbnd;              /* 0 Pa: > 1 corrects boundary levels by (pavbnd - patm) / (rhow g) */
  float gapres;              /* 101200 Pa: background pressure */
  int32_t spinup_meteo;      /* 0: 1 ramps gridded meteo from zero over the spin-up time */
  int32_t amprblock;         /* 1: gridded rain held per interval (0: linear in time) */
  int32_t usespwprecip;      /* 1: rain from a spiderweb's fourth quantity */
  float spw_merge_frac;      /* 0.5 (spwmergefrac): the spiderweb fades into the background from
                              * this fraction of its radius outwards */
  float factor_wind, factor_pres, factor_prcp, factor_spw_size;   /* 1.0 each */
  float sfacinf;         /* initial abstraction of the curve-number method, fraction of S (0.2) */
} sfx_params;

void sfx_params_default(sfx_params* p);

/* Quadtree mesh in the reference's quadtree-file semantics (sfincs_quadtree.F90:210-292,
 * 295-442): points sorted level-major then by nm; md/mu/nd/nu flags (-1 coarser, 0 same,
 * 1 finer); md1..nu2 1-based point indices (0 = none); mask per point. */
typedef struct sfx_quadtree_input {
  int32_t nr_points, nr_levels;
  float x0, y0, dx, dy, rotation_deg;
  const int8_t  *level, *md, *mu, *nd, *nu, *mask;
  const int32_t *n, *m, *md1, *md2, *mu1, *mu2, *nd1, *nd2, *nu1, *nu2;
} sfx_quadtree_input;

/* Subgrid tables in the reference's NetCDF semantics (sfincs_subgrid.F90:57-455), given
 * per active cell / active uv-point in sfincs order; level-fastest 2-D arrays. */
typedef struct sfx_subgrid_input {
  int32_t nlevels;
  const float *z_zmin, *z_zmax, *z_volmax, *z_level;               /* np, np, np, np*nl */
  const float *uv_zmin, *uv_zmax, *uv_navg, *uv_ffit;              /* npuv each */
  const float *uv_havg, *uv_nrep, *uv_pwet;                        /* npuv*nl each */
} sfx_subgrid_input;

typedef struct sfx_sizes {
  int32_t np, npuv, ncuv, nref, nlevels, nmax, mmax;
  int32_t ngbnd, nkcuv2, nsrc, nbnd;
  int32_t subgrid, quadtree, dense;   /* dense = full rectangle, implicit neighbours on device */
  float dtmax, dtmin;
  int32_t ndrn;                       /* drainage structures (each adds an intake and an outfall point) */
  int32_t nstruc;                     /* weir uv-points (kcuv 3) */
  int32_t nthd;                       /* uv-points closed by thin dams (kcuv 0) */
} sfx_sizes;

/* ---- host-side model (CPU only; no GPU needed) -------------------------------------
 * Replaces the table/mask builders of sfincs_initialize: initialize_mesh
 * (sfincs_domain.f90:143-1586), make_quadtree_from_indices (sfincs_quadtree.F90:445-669),
 * initialize_bathymetry (:1588-1744), initialize_boundaries (:1765-1958),
 * initialize_roughness (:1960-2061), set_advection_mask (:2114-2180),
 * initialize_hydro (:2183-2410), read_subgrid_file sanitize (sfincs_subgrid.F90:57-455),
 * compute_initial_subgrid_volumes (:907-962), find_boundary_indices
 * (sfincs_boundaries.f90:455-679), read_discharges cell lookup (sfincs_discharges.f90:159-171),
 * the Cd table (sfincs_meteo.f90:316-328). */
typedef struct sfx_model sfx_model;

/* Regular grid: msk is nmax*mmax, n fastest (the reference's column-major kcsg(n,m)). */
int  sfx_model_create_regular(const sfx_params* p, int32_t nmax, int32_t mmax,
                              float x0, float y0, float dx, float dy, float rotation_deg,
                              const int8_t* msk, sfx_model** out);
int  sfx_model_create_quadtree(const sfx_params* p, const sfx_quadtree_input* qt, sfx_model** out);
/* The NetCDF quadtree file (a qtrfile whose name holds ".nc"; quadtree_read_file_netcdf,
 * sfincs_quadtree.F90:295-442): dimension mesh2d_nFaces; variables n, m, level, md, md1, md2,
 * mu, mu1, mu2, nd, nd1, nd2, nu, nu1, nu2, z, mask; global attributes x0, y0, dx, dy,
 * rotation (degrees), nmax, mmax, nr_levels. with_bed != 0 also sets the bed level of the
 * active cells from z (models without subgrid tables; sfincs_domain.f90:1635-1637). Classic
 * NetCDF formats only, as sfx_model_read_subgrid_netcdf. */
int  sfx_model_create_quadtree_netcdf(const sfx_params* p, const char* path, int32_t with_bed, sfx_model** out);
int  sfx_model_sizes(const sfx_model* m, sfx_sizes* s);
/* Copy a table by its reference name (e.g. "uv_index_z_nm", "kcuv", "mask_adv", "zbuvmx",
 * "z_index_uv_md", "cuv_index_uv1", "dxrinv"); nbytes must equal the table size.
 * "params" copies the model's sfx_params. */
int  sfx_model_get_table(const sfx_model* m, const char* name, void* dst, size_t nbytes);
int  sfx_model_table_bytes(const sfx_model* m, const char* name, size_t* nbytes);
int  sfx_model_set_bathymetry(sfx_model* m, const float* zb_active);           /* np values */
int  sfx_model_set_subgrid(sfx_model* m, const sfx_subgrid_input* sg);
/* The legacy binary subgrid file (sbgfile that is not NetCDF; read_subgrid_file_original,
 * sfincs_subgrid.F90:458-903, dispatched at :28-53): hrep/navg per level, huthresh kept,
 * pwet = 1, no fit. Regular and quadtree meshes (the quadtree file's layout over all its
 * points). Call after the mesh is built and before sfx_model_initialize. */
int  sfx_model_read_subgrid_binary(sfx_model* m, const char* path);
/* The NetCDF subgrid file (sbgfile that opens as NetCDF; read_subgrid_file_netcdf,
 * sfincs_subgrid.F90:57-455): dimensions levels, np, npuv; variables z_zmin, z_zmax,
 * z_volmax (np), z_level (np, levels), uv_zmin, uv_zmax, uv_navg, uv_ffit (npuv), uv_havg,
 * uv_nrep, uv_pwet (npuv, levels), C order. Regular meshes: records in sfincs order;
 * quadtree meshes: records over every quadtree point / quadtree uv-point, re-mapped to the
 * active cells and uv-points. Then as sfx_model_set_subgrid. The classic NetCDF formats
 * (CDF-1, CDF-2, CDF-5) are read; a NetCDF-4 (HDF5) file is refused (SFX_ERR_IO, the message
 * names the conversion). Call after the mesh is built and before sfx_model_initialize. */
int  sfx_model_read_subgrid_netcdf(sfx_model* m, const char* path);
/* Curve-number infiltration, method A (scsfile; initialize_infiltration 'cna',
 * sfincs_infiltration.f90:106-112,260-293; update_infiltration_map :678-722): the maximum
 * retention S per active cell in inches (converted to m by 0.0254). Each step every cell's
 * cumulative rain P grows by its rain times dt; while P <= sfacinf S all rain infiltrates,
 * after that the cumulative infiltration follows P - (P - sfacinf S)^2 / (P + (1 - sfacinf) S),
 * wet or dry, no recovery. Needs rain; a uniform qinf > 0 or an infiltration map takes
 * precedence (:92-112). Call before sfx_model_initialize. */
int  sfx_model_set_curve_number(sfx_model* m, const float* scs_inch);
/* bzs: water level boundaries (bnd/bzs files): zs is nt*nbnd, point fastest. */
int  sfx_model_set_boundary(sfx_model* m, int32_t nbnd, const float* x, const float* y,
                            int32_t nt, const float* t, const float* zs);
/* src/dis: point discharges; q is nt*nsrc, source fastest. */
int  sfx_model_set_sources(sfx_model* m, int32_t nsrc, const float* x, const float* y,
                           int32_t nt, const float* t, const float* q);
/* drnfile: drainage structures (read_discharges, sfincs_discharges.f90:46-57, 176-290;
 * update, :371-655): per structure the intake point (x_in, y_in), the outfall point
 * (x_out, y_out), the type (1 pump: params[0] = Q m3/s; 2 culvert and 3 check valve:
 * params[0] = coefficient of sqrt(dz); 4 gate opened by the intake level: width, sill
 * level, Manning's n, zmin, zmax, closing time; 5 gate closed and opened at set times:
 * width, sill level, Manning's n, t_close, t_open, closing time) and 6 params per
 * structure (params[6*k + j]). Both points must lie in active cells. The intake and
 * outfall discharges join the point sources in continuity. */
int  sfx_model_set_drainage(sfx_model* m, int32_t ndrn, const float* x_in, const float* y_in, const float* x_out,
                            const float* y_out, const int32_t* type, const float* params);
/* Thin dams (thdfile; read_thin_dams, sfincs_structures.f90:385-495): npoly polylines
 * with nv[i] vertices each, coordinates concatenated in x/y. At initialize every uv-point
 * whose cell-centre connection a polyline segment crosses
 * (find_uv_points_intersected_by_polyline, sfincs_quadtree.F90:853-1010) gets kcuv 0:
 * no flow through it. Table thindam_uv_index (1-based). */
int  sfx_model_set_thin_dams(sfx_model* m, int32_t npoly, const int32_t* nv, const float* x, const float* y);
/* The active cell (1-based; 0 outside the mesh) that holds each point (x[i], y[i]) -- how
 * the reference places observation points (sfincs_obspoints.f90:95-130). */
int  sfx_model_find_cells(sfx_model* m, int32_t n, const float* x, const float* y, int32_t* nm);
/* Gridded meteo on an equidistant grid: the contents of an amu, amv, amp or ampr file
 * (read_meteo_data, sfincs_meteo.f90:146-245). quantity: SFX_METEO_WIND_U / _WIND_V (m/s),
 * _PRESSURE (Pa), _RAIN (mm/h). values[nt][nrows][ncols], row 0 the northernmost (first in
 * the file); times in s from tref; x/y_llcorner the lower-left corner of the grid. Wind
 * needs both components on the same grid. The fields at each update time (every dtwnd from
 * t0) are interpolated to the cells at initialisation (update_amuv_data, update_amp_data,
 * update_ampr_data, :802-1244), and each step interpolates them in time per cell
 * (update_meteo_forcing, :1246-1419). Pressure acts only with baro = 1. */
enum { SFX_METEO_WIND_U = 0, SFX_METEO_WIND_V = 1, SFX_METEO_PRESSURE = 2, SFX_METEO_RAIN = 3 };
int  sfx_model_set_meteo_grid(sfx_model* m, int32_t quantity, int32_t nt, int32_t nrows, int32_t ncols,
                              float x_llcorner, float y_llcorner, float dx, float dy, const float* times,
                              const float* values);
/* A spiderweb (spwfile; read_spw_file, sfincs_spiderweb.f90:8-133): nt times (s from
 * tref), eye positions xe/ye[nt], and [nt][nrows][ncols] polar fields -- rows outwards from
 * the eye over `radius`, columns clockwise from north -- of wind speed (m/s), wind
 * direction (nautical degrees, where it comes from), pressure drop (Pa) and rain (mm/h, or
 * NULL). Merged over any gridded background (update_spiderweb_data, sfincs_meteo.f90:450-800). */
int  sfx_model_set_spiderweb(sfx_model* m, int32_t nt, int32_t nrows, int32_t ncols, float radius,
                             const float* times, const float* xe, const float* ye, const float* vmag,
                             const float* vdir, const float* pdrop, const float* rain);
/* Reads an ASCII meteo file as the reference does -- kind "amu", "amv", "amp", "ampr"
 * (read_amuv_dimensions / read_amuv_file, sfincs_spiderweb.f90:135-229, 320-408) or "spw"
 * (read_spw_dimensions / read_spw_file, :8-133, 231-318) -- with its TIME lines taken
 * relative to tref ("yyyymmdd HHMMSS"; compute_time_in_seconds, :411-484), and passes it
 * to sfx_model_set_meteo_grid / sfx_model_set_spiderweb. */
int  sfx_model_read_meteo_file(sfx_model* m, const char* kind, const char* path, const char* tref);
/* Spatially varying constant infiltration (qinffile; initialize_infiltration 'c2d',
 * sfincs_infiltration.f90:224-258): qinf_mmh[c] in mm/h per active cell, in sfincs order.
 * With sfx_params.qinf > 0 the uniform rate ('con', :195-222) takes precedence, as in the
 * reference (:92-104). Each step a wet cell (z_volume > 0 on subgrid meshes, zs > zb on
 * regular ones) loses its rate from the net precipitation
 * (update_infiltration_map, :632-676); it acts only where rain falls (continuity adds
 * netprcp only with precipitation). Table qinffield (m/s). */
int  sfx_model_set_infiltration(sfx_model* m, const float* qinf_mmh);
/* Weirs (weirfile; read_structure_file with type 1, sfincs_structures.f90:38-381): as thin
 * dams, with par[2*k] = crest level and par[2*k + 1] = discharge coefficient Cd at each
 * vertex k. A crossed uv-point whose interpolated crest lies above its bed (uv_zmin on
 * subgrid meshes, zbuvmx on regular ones) gets kcuv 3, and each step its discharge is the
 * broad-crested weir formula (compute_fluxes_over_structures, :592-697) times the crest
 * length across it. Tables structure_uv_index, structure_height, structure_cd,
 * structure_length (relative to the grid spacing). */
int  sfx_model_set_weirs(sfx_model* m, int32_t npoly, const int32_t* nv, const float* x, const float* y,
                         const float* par);
/* prcfile: uniform rain series in mm/h. */
/* Storage volume per cell, m3 (volfile; initialize_storage_volume, sfincs_domain.f90:2063-2112):
 * on subgrid meshes water entering a cell fills its storage before the cell's volume grows
 * (sfincs_continuity.f90:499-523). State name storage_volume (f32 per cell). */
int  sfx_model_set_storage_volume(sfx_model* m, const float* vol);
/* Incoming infragravity water levels at the boundary points (bzifile;
 * sfincs_boundaries.f90:126-140, 748-753, 829-866): zsi[it*nbnd + ib] at the times of the
 * series given to sfx_model_set_boundary (call after it). They are added to the total
 * boundary level but not to the still level, so they drive the boundary flux. */
int  sfx_model_set_boundary_ig(sfx_model* m, const float* zsi);
/* Downstream river boundary points for msk=5 cells (bdrfile; sfincs_boundaries.f90:217-285,
 * 587-613, 878-901): per point x, y, an interior point x_in, y_in, the slope and the
 * distance to use (< 0: the distance between the two points). Each river cell takes the
 * level of its nearest point's interior cell plus dzs = -slope*distance, raised to the bed.
 * The interior point must lie in an active interior cell (kcs 1). */
int  sfx_model_set_river_boundary(sfx_model* m, int32_t nbdr, const float* x, const float* y, const float* x_in,
                                  const float* y_in, const float* slope, const float* distance);
int  sfx_model_set_rain(sfx_model* m, int32_t nt, const float* t, const float* mmh);
/* wndfile: uniform wind series, speed m/s and nautical direction in degrees. */
int  sfx_model_set_wind(sfx_model* m, int32_t nt, const float* t, const float* mag, const float* dir);
/* Finish initialisation (kcuv, roughness, initial conditions, volumes, advection mask). */
/* Binary restart file as initial condition (rstfile; read_binary_restart_file,
 * sfincs_initial_conditions.F90:183-260, applied as :98-174 and sfincs_domain.f90:2272-2311
 * do): call after the mesh is built and before sfx_model_initialize. Types 1 and 2 give zs,
 * q and uvmean, type 3 zs only; levels are raised to the bed, combined uv-points take the
 * mean of their two points and uv starts at 0. An unknown type is skipped (SFX_OK, as the
 * reference does); types 4-6 (infiltration) and records sized for another mesh are errors. */
int  sfx_model_set_restart(sfx_model* m, const char* path);
int  sfx_model_initialize(sfx_model* m);
/* Initial state by name ("zs" f64, "q" f32, "uv" f32, "kfuv" i8, "z_volume" f64, ...). */
int  sfx_model_get_state(const sfx_model* m, const char* name, void* dst, size_t nbytes);
/* timestep_analysis_finalize (sfincs_timestep_analysis.f90:77-200), on the host: from the
 * per-uv accumulators (tsa_sum, tsa_wet, tsa_lim of sfx_get_state) after nt steps, per
 * cell the average required time step (alpha-scaled, capped at dtmax; -1 where no
 * uv-point of the cell was counted) and the percentage of steps it limited. np each. */
int  sfx_model_timestep_analysis_per_cell(const sfx_model* m, const float* tsa_sum, const int32_t* tsa_wet,
                                          const int32_t* tsa_lim, int64_t nt, float* dt_cell, float* pct_cell);
/* The active cell (1-based; 0 outside the mesh) holding point (x, y): the lookup behind
 * the BMI get_sfincs_cell_index / get_sfincs_cell_indices (sfincs_bmi.f90:299-349). */
int  sfx_model_find_cell(const sfx_model* m, float x, float y, int32_t* idx);
void sfx_model_destroy(sfx_model* m);

/* ---- device instance ---------------------------------------------------------------
 * sfx_create replaces initialize_openacc (sfincs_openacc.f90:9-45): all tables and state
 * move to HBM once; sfx_destroy replaces finalize_openacc (:47-79). */
typedef struct sfx_handle sfx_handle;

typedef struct sfx_step_stats {
  double t;          /* model time after the call */
  float  dt;         /* last dt */
  float  min_dt;     /* CFL minimum (without alpha) reduced in the last momentum sweep */
  int64_t nt;        /* steps taken since creation */
  int32_t error;     /* 1 = instability stop */
  double cell_updates; /* sum over steps of np */
  double seconds;      /* wall time of the call */
} sfx_step_stats;

int  sfx_create(const sfx_model* m, int32_t device, sfx_handle** out);
int  sfx_destroy(sfx_handle* h);
/* State exchange in reference order; the explicit form of `acc update host/device`.
 * Names: zs, z_volume (f64 per cell); q, uv (per uv-point); kfuv (i8); zs0, zsderv, zsmax,
 * uvmean; and, when their switch is on, vmax, qmax, twet (f32 per cell; initial values
 * -999, -999, 0 as sfincs_domain.f90:2300-2306 and sfincs_output.f90:210 reset them).
 * With timestep_analysis on (read only, per uv-point, npuv entries): tsa_sum (f32, the sum
 * of the required time steps of the sweeps in which the point was wet), tsa_wet and
 * tsa_lim (i32: those sweeps, and those in which it limited the step), the accumulators of
 * sfincs_timestep_analysis.f90:26-74; their average is tsa_sum / max(tsa_wet, 1). Read only
 * as well: qtsrc (f32, nsrc + 2 ndrn), the step's point-source discharges followed by each
 * drainage structure's intake (-Q) and outfall (+Q) discharge. */
int  sfx_get_state(sfx_handle* h, const char* name, void* dst, size_t nbytes);
/* Also "zb" (f32 per cell, regular meshes without subgrid tables): the bed level, as a BMI
 * coupler writes it through get_value_ptr; zbuvmx follows only with sfx_update_zbuvmx. */
int  sfx_set_state(sfx_handle* h, const char* name, const void* src, size_t nbytes);
/* update_zbuv (sfincs_bmi.f90:277-284 -> compute_zbuvmx, sfincs_domain.f90:1746-1762):
 * zbuvmx = max(zb(nm), zb(nmu)) + huthresh for every uv-point from the device's zb.
 * Regular meshes without subgrid tables. */
int  sfx_update_zbuvmx(sfx_handle* h);
/* write_rst_file (sfincs_output.f90:657-729): a type-1 restart file (zs as f32, q, uvmean)
 * of the current state at `path` (the reference names it sfincs.<time>.rst). */
int  sfx_write_restart(sfx_handle* h, const char* path);
/* compute_fluxes(dt) (sfincs_momentum.f90:9-784): one momentum sweep with the given dt,
 * including the combined-uv averaging; *min_dt receives the CFL reduction (without alpha). */
int  sfx_compute_fluxes(sfx_handle* h, float dt, float* min_dt);
/* compute_water_levels(t, dt) (sfincs_continuity.f90:5-43), point sources included. */
int  sfx_compute_water_levels(sfx_handle* h, double t, float dt);
/* The per-step companions of sfincs_update (sfincs_lib.f90:517-550): meteo series,
 * boundary points/conditions/fluxes, discharges, at time t with step dt. */
int  sfx_update_forcing(sfx_handle* h, double t, float dt);
/* sfincs_update(dtrange) (sfincs_lib.f90:337-683): dtrange < -999 runs to t1, dtrange <= 0
 * one step, dtrange > 0 runs to t + dtrange. Device-resident loop, graph-captured. */
int  sfx_update(sfx_handle* h, double dtrange, sfx_step_stats* st);
/* Steps while t < t_end (an absolute model time): the form a host uses to stop at the
 * first step whose time reaches an output time, as the reference's output triggers fire
 * (sfincs_lib.f90:419-489). Nothing is done when t >= t_end already. */
int  sfx_update_to(sfx_handle* h, double t_end, sfx_step_stats* st);
int  sfx_get_time(sfx_handle* h, double* t, float* dt, int64_t* nt);
/* Per-branch uv-point counters of the sweeps since counting was enabled by
 * sfx_set_counting(h,1): {copied (kcuv not 1/6, and the absent slots of a dense grid),
 * dry with zeros stored, wet regular or subgrid full, subgrid partial, dry with the zeros
 * already in both q/uv buffers (nothing stored), points of quiescent tiles skipped as a
 * whole}; used for the byte model. */
int  sfx_set_counting(sfx_handle* h, int32_t on);
int  sfx_get_counters(sfx_handle* h, int64_t counts[6]);
/* Continuity since counting was enabled: cells that stored {zs, z_volume, zs0, zsderv,
 * zsmax}, and steady cells skipped. A cell stores a field only when its value changes,
 * and a cell without inflow, rain or source whose last such update changed nothing is
 * skipped after reading its faces; used for the byte model. No counterpart in the
 * reference. */
int  sfx_get_store_counters(sfx_handle* h, int64_t counts[6]);
/* FP64 volume accumulator: total water volume (sum of zs*A or V) in double. */
int  sfx_total_volume(sfx_handle* h, double* vol);
/* sfx_total_volume without allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc). Writes the total volume to `vol_device`, a double in device
 * memory that the caller owns, asynchronously on the instance's stream (synchronize
 * before reading it). Its partial sums live in a static device array of 4096 doubles per
 * GPU, so meshes of at most 67,108,864 cells (larger: SFX_ERR_ARG); calls that overlap on
 * one GPU share that array and race. */
int  sfx_total_volume_noalloc(sfx_handle* h, double* vol_device);
/* FP64 volume balance of the interior cells (kcs 1). Every continuity sweep adds, in
 * double, the volume it moved into the interior cells through faces whose other side is
 * not interior (water-level, river and Neumann boundaries, inactive cells), from point
 * sources as applied (sfincs_continuity.f90:87-106 with its bed clamp, :311-325), rain
 * (:116-121, 473-497) less infiltration (sfincs_infiltration.f90:632-663), and qext.
 * out = {interior volume now (sum of V, or zs*A), storage volume left, boundary inflow,
 * source inflow, rain less infiltration, qext} in m3; the last four
 * are sums since creation or sfx_reset_balance. Between two calls the balance is
 * dV + d(storage used) = d(inflows). Host writes of cell state are not counted. No
 * counterpart in the reference, which keeps no volume balance. */
int  sfx_get_balance(sfx_handle* h, double out[6]);
int  sfx_reset_balance(sfx_handle* h);

/* ---- batched ensembles ---------------------------------------------------------------
 * No counterpart in the reference (it runs scenarios one at a time, docs/output.rst:19-20).
 * Members of uneven size advance together: each launch covers an (instance, tile) work
 * list, each member keeps its own adaptive dt and end time. A member must outlive the
 * batches that hold it (remove it, or destroy the batch, before sfx_destroy). */
typedef struct sfx_batch sfx_batch;
int  sfx_batch_create(int32_t device, sfx_batch** out);
int  sfx_batch_destroy(sfx_batch* b);
int  sfx_batch_add(sfx_batch* b, sfx_handle* h, double t_end);
int  sfx_batch_remove(sfx_batch* b, sfx_handle* h);
/* Advance every active member by up to nsteps steps (graph-captured). *n_active gets the
 * number of members still short of their end time. */
int  sfx_batch_advance(sfx_batch* b, int32_t nsteps, int32_t* n_active);
int  sfx_batch_done(sfx_batch* b, sfx_handle* h, int32_t* done);
/* HBM footprint of a model once on the device, for admission control. */
int  sfx_model_footprint(const sfx_model* m, size_t* bytes);

/* Profiling: nsteps ordinary (non-graph) steps with HIP events around each phase on the
 * launch stream; ms[0..2] = mean forcing, momentum, continuity launch durations. */
int  sfx_profile_steps(sfx_handle* h, int32_t nsteps, double ms[3]);
/* FP64 twin: the same kernels instantiated with double arithmetic and state (q, uv, zs0,
 * zsderv, zsmax), used to state the FP32 path's rounding error. */
int  sfx_create_ex(const sfx_model* m, int32_t device, int32_t fp64, sfx_handle** out);

/* Autotuner: on this handle's mesh, times the dense momentum kernels (one thread per cell,
 * or 1, 2 or 4 uv-points per thread, on tiles of several shapes; strip marches of several
 * lengths), the tile
 * shapes of the dense continuity kernel, and for both sweeps (dense or indexed) the runs
 * in which tiles are dealt to the XCDs, and keeps the fastest by the median of `trials`
 * timed launches per candidate. Call it before the handle joins a batch (SFX_ERR_STATE
 * otherwise). */
int  sfx_autotune(sfx_handle* h, int32_t trials);
/* out = {dense, momentum kernel (-1, -2, -4: per uv-point with 1, 2 or 4 points per
 *        thread, the points of a thread momentum-tile-m rows apart; 0: per cell; k > 0:
 *        strip march of k rows per wave), momentum tile n, momentum tile m, continuity tile n, continuity
 *        tile m, momentum and continuity work-list runs per XCD (0: one contiguous band
 *        per XCD)} */
int  sfx_get_tuning(sfx_handle* h, int32_t out[8]);
/* Pins the dense shapes without timing: in = {momentum kernel as above, momentum tile n,
 * momentum tile m, continuity tile n, continuity tile m}; tiles hold 256 cells. No effect
 * on indexed meshes. Call it before the handle joins a batch. */
int  sfx_set_tuning(sfx_handle* h, const int32_t in[5]);
/* Pins the work-list order without timing: the runs of tiles dealt to the XCDs in turn,
 * for the momentum and the continuity sweep (out[6], out[7] of sfx_get_tuning; 0: one
 * contiguous band per XCD). Call it before the handle joins a batch. */
int  sfx_set_work_order(sfx_handle* h, int32_t mom_run, int32_t cont_run);

int  sfx_last_error(char* buf, size_t n);

#ifdef __cplusplus
}
#endif
#endif
