/*
 * ORACLE -- test infrastructure only. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 *
 * A scalar C restatement of the SFINCS explicit time step, operation for operation and
 * type for type (real*4 arithmetic, real*8 zs / z_volume / t), so that it reproduces
 * the reference's own Fortran bit for bit on x86-64 (pinned against oracle/_ref, the
 * reference's sfincs_momentum/sfincs_continuity compiled from source):
 *   oracle_compute_fluxes        sfincs_momentum.f90:9-784, with timestep_analysis_update
 *                                (sfincs_timestep_analysis.f90:26-74), which sfincs_lib.f90:588
 *                                runs right after it
 *   oracle_compute_water_levels  sfincs_continuity.f90:5-720 (with compute_store_variables)
 *   oracle_update_meteo          sfincs_meteo.f90:1246-1558 (time-series branch)
 *   oracle_update_infiltration   sfincs_infiltration.f90:632-676 ('con' and 'c2d'), 678-722
 *                                ('cna', curve number A)
 *   oracle_update_boundaries     sfincs_boundaries.f90:682-1175
 *   oracle_update_discharges     sfincs_discharges.f90:325-655 (point sources and drainage
 *                                structures)
 *   oracle_compute_fluxes_over_structures
 *                                sfincs_structures.f90:592-697 (weirs)
 *   oracle_update                sfincs_lib.f90:337-683 (step order, lagged dt)
 * Arrays use 0-based storage; index tables hold the reference's 1-based values.
 *
 * What pins each function (tests/):
 *   oracle_compute_fluxes        bit for bit against the reference's compute_fluxes after
 *                                every step (test_oracle_vs_ref.py: regular variants, subgrid,
 *                                wind, quadtree, viscosity, sources; test_subgrid_binary.py,
 *                                test_drainage.py, test_structures.py); the CFL step on a
 *                                uniform depth and a dry domain by closed form
 *                                (test_forcing_pins.py)
 *   oracle_compute_water_levels  bit for bit against the reference's compute_water_levels as
 *                                above, with store variables, qext, storage volume and the
 *                                timestep analysis (test_oracle_vs_ref.py); rain volume R A T
 *                                on a closed basin (test_forcing_pins.py)
 *   oracle_update_meteo          wind stress against the Cd table, rain and wind series
 *                                interpolation at known times (test_forcing_pins.py)
 *   oracle_update_boundaries     a hand-computed weakly reflective flux and its dry guards,
 *                                the uvmean filter as a geometric series, series
 *                                interpolation, the bed floor, spin-up, Neumann copies
 *                                (test_forcing_pins.py); river cells (test_river.py);
 *                                infragravity levels (test_ig_waves.py)
 *   oracle_update_discharges     discharge series interpolation (test_forcing_pins.py);
 *                                pump, culvert, check valve and gate closed forms
 *                                (test_drainage.py)
 *   oracle_compute_fluxes_over_structures
 *                                the broad-crested weir formula in double, free, submerged
 *                                and below the crest, both directions (test_structures.py)
 *   oracle_update_infiltration   a flat basin's level falls by the rate less the rain to
 *                                the float step (regular and subgrid); dry cells keep their
 *                                rain; the removed volume against the rate times the wet
 *                                area (test_infiltration.py); 'cna': the cumulative
 *                                infiltration against the SCS runoff relation
 *                                P - (P - Ia)^2 / (P - Ia + S) at every step, all rain
 *                                infiltrating below Ia = sfacinf S, wet and dry cells alike
 *                                (test_curve_number.py)
 *   oracle_update                step counts and model time against the GPU and the
 *                                reference loop (test_oracle_vs_ref.py); the instability stop
 *                                by closed form (test_forcing_pins.py)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* A gridded meteo input: [nt][nrows][ncols], row 0 the northernmost (file order). */
typedef struct MG {
  int32_t nt, nrows, ncols;
  float xll, yll, dx, dy;
  float *times, *v;
} MG;

typedef struct OM {
  /* sizes and switches */
  int32_t np, npuv, ncuv, nref, nl;
  int32_t subgrid, use_quadtree, advection, advection_scheme, viscosity, coriolis, friction2d;
  int32_t thetasmoothing, wiggle_suppression, store_zsmax, wind, precip, h73table;
  int32_t store_velmax, store_fluxmax, store_twet;
  float twet_threshold;
  int32_t timestep_analysis, use_qext, use_storage_volume, infiltration;
  int32_t ngbnd, nkcuv2, nbnd, ntbnd, nsrc, ntsrc, ntprcp, ntwnd;
  float g, pi, huthresh, huvmin, advlim, slopelim, uvlim, nuviscfac, fcorio, theta;
  float wiggle_factor, wiggle_threshold, dtmax, dtmin, alpha, rhoa, rhow, btfilter, btrelax;
  float t0, t1, tspinup, zini, dx, dy;
  /* cell tables */
  const int8_t *kcs, *z_flags_iref;
  const int32_t *z_index_uv_md, *z_index_uv_mu, *z_index_uv_nd, *z_index_uv_nu;
  const float *zb, *cell_area, *dxrm, *dyrm, *dxrinv, *dyrinv, *dxr2inv, *dyr2inv, *dxrinvc, *dyrinvc, *nuvisc;
  /* uv tables */
  const int32_t *uv_index_z_nm, *uv_index_z_nmu, *uv_index_u_nmd, *uv_index_u_nmu, *uv_index_u_ndm, *uv_index_u_num;
  const int32_t *uv_index_v_ndm, *uv_index_v_nm, *uv_index_v_ndmu, *uv_index_v_nmu;
  const int32_t *cuv_index_uv, *cuv_index_uv1, *cuv_index_uv2;
  const int8_t *kcuv, *mask_adv, *uv_flags_iref, *uv_flags_type, *uv_flags_dir;
  const float *zbuv, *zbuvmx, *gn2uv;
  /* subgrid */
  const float *z_zmin, *z_zmax, *z_volmax, *z_dep;
  const float *uv_zmin, *uv_zmax, *uv_navg_w, *uv_fnfit, *uv_havg, *uv_nrep, *uv_pwet, *uv_havg_zmax, *uv_nrep_zmax;
  /* state */
  double *zs, *z_volume;
  float *q, *uv, *q0, *uv0, *zs0, *zsderv, *zsmax;
  float *vmax, *qmax, *twet;   /* compute_store_variables (np each, when switched on) */
  float *tsa_req, *tsa_sum;    /* timestep analysis (npuv each, when switched on) */
  float *qext;                 /* external sources, m/s per cell (use_qext) */
  float *storage_volume;       /* storage left per subgrid cell, m3 (use_storage_volume) */
  const float *qinffield;      /* infiltration rate per cell, m/s ('con' or 'c2d') */
  int32_t *tsa_wet, *tsa_lim;
  int8_t *kfuv;
  float *tauwu, *tauwv, *netprcp;   /* np each */
  int32_t cna;                       /* curve-number infiltration, method A */
  float sfacinf;
  const float *scs_m;                /* S per cell, m ('cna') */
  float *cumprcp, *cuminf;           /* cumulative rain and infiltration per cell, m ('cna') */
  /* boundaries */
  const int32_t *nmindbnd, *index_kcuv2, *nmbkcuv2, *nmikcuv2, *ibkcuv2, *ind1_bnd_gbp, *ind2_bnd_gbp, *nmi_gbp;
  const int8_t *ibuvdir;
  const float *fac_bnd_gbp, *t_bnd, *zs_bnd;
  const int32_t *index_bdr_gbp, *index_zsi_bdr;   /* river cells: bdr point per cell, interior cell per point */
  const float *dzs_bdr;
  const float *zsi_bnd;   /* bzi infragravity levels (NULL: none), as zs_bnd */
  float *zsit_bnd;        /* their value at the current time, nbnd */
  int32_t has_bzi;
  float *zsb, *zsb0, *uvmean, *zst_bnd;
  /* sources, meteo */
  const int32_t *nmindsrc;
  const float *tsrc, *qsrc, *tprcpt, *tprcpv, *twnd, *wndmag, *wnddir, *cdval;
  const float *x73;   /* h**(7/3) table, 1000 x 10, column-major (sfincs_data.f90:858-881) */
  float *qtsrc;   /* nsrc point sources, then each drainage structure's intake and outfall */
  /* drainage structures (sfincs_discharges.f90:176-290, 371-655) */
  int32_t ndrn;
  float structure_relax;
  const int32_t *drainage_type;
  const float *drainage_params, *drainage_distance;   /* params: 6 per structure */
  int32_t *drainage_status;
  float *drainage_fraction_open;
  /* weirs (sfincs_structures.f90:38-381, 592-697): uv-point (1-based), crest, Cd, length */
  int32_t nstruc;
  const int32_t *structure_uv_index;
  const float *structure_height, *structure_cd, *structure_length;
  /* gridded and spiderweb meteo (sfincs_meteo.f90): switches, parameters, the window,
   * the two fields of the window per cell, the inputs after read_meteo_data's conversions */
  int32_t meteo3d, patmos, meteo_wind, meteo_rain, spinup_meteo, amprblock, usespwprecip;
  float dtwnd, gapres, pavbnd, spw_merge_frac, cosrot, sinrot;
  float factor_wind, factor_pres, factor_prcp, factor_spw_size;
  float twindupd, meteo_t0, meteo_t1;
  const float *z_xz, *z_yz;
  float *patm;
  float *tauwu0, *tauwu1, *tauwv0, *tauwv1, *patm0, *patm1, *prcp0, *prcp1;
  MG amu, amv, amp, ampr;
  int32_t spw_nt, spw_nrows, spw_ncols, spw_rain;
  float spw_radius, dradspw, dphispw;
  float *spw_times, *spw_xe, *spw_ye, *spw_wu, *spw_wv, *spw_pabs, *spw_prcp;
  /* time state */
  double t;
  float dt, min_dt;
  int32_t nt, error, itbndlast, itsrclast, itwndlast, itprcplast;
} OM;

OM* oracle_new(void) {
  OM* m = (OM*)calloc(1, sizeof(OM));
  m->min_dt = 1.0e-6f;   /* sfincs_lib.f90:281-282 */
  m->dt = 1.0e-6f;
  m->itbndlast = 1; m->itsrclast = 1; m->itwndlast = 1; m->itprcplast = 1;
  return m;
}
static void mg_free(MG* g) { free(g->times); free(g->v); g->times = g->v = NULL; g->nt = 0; }
void oracle_free(OM* m) {
  mg_free(&m->amu); mg_free(&m->amv); mg_free(&m->amp); mg_free(&m->ampr);
  free(m->spw_times); free(m->spw_xe); free(m->spw_ye); free(m->spw_wu); free(m->spw_wv); free(m->spw_pabs);
  free(m->spw_prcp);
  free(m);
}

#define BIND(field) if (!strcmp(name, #field)) { *(void**)&m->field = ptr; return 0; }
int oracle_bind(OM* m, const char* name, void* ptr) {
  BIND(kcs) BIND(z_flags_iref) BIND(z_index_uv_md) BIND(z_index_uv_mu) BIND(z_index_uv_nd) BIND(z_index_uv_nu)
  BIND(zb) BIND(cell_area) BIND(dxrm) BIND(dyrm) BIND(dxrinv) BIND(dyrinv) BIND(dxr2inv) BIND(dyr2inv)
  BIND(dxrinvc) BIND(dyrinvc) BIND(nuvisc)
  BIND(uv_index_z_nm) BIND(uv_index_z_nmu) BIND(uv_index_u_nmd) BIND(uv_index_u_nmu) BIND(uv_index_u_ndm)
  BIND(uv_index_u_num) BIND(uv_index_v_ndm) BIND(uv_index_v_nm) BIND(uv_index_v_ndmu) BIND(uv_index_v_nmu)
  BIND(cuv_index_uv) BIND(cuv_index_uv1) BIND(cuv_index_uv2)
  BIND(kcuv) BIND(mask_adv) BIND(uv_flags_iref) BIND(uv_flags_type) BIND(uv_flags_dir) BIND(zbuv) BIND(zbuvmx) BIND(gn2uv)
  BIND(z_zmin) BIND(z_zmax) BIND(z_volmax) BIND(z_dep) BIND(uv_zmin) BIND(uv_zmax) BIND(uv_navg_w) BIND(uv_fnfit)
  BIND(uv_havg) BIND(uv_nrep) BIND(uv_pwet) BIND(uv_havg_zmax) BIND(uv_nrep_zmax)
  BIND(zs) BIND(z_volume) BIND(q) BIND(uv) BIND(q0) BIND(uv0) BIND(zs0) BIND(zsderv) BIND(zsmax) BIND(kfuv)
  BIND(vmax) BIND(qmax) BIND(twet) BIND(tsa_req) BIND(tsa_sum) BIND(tsa_wet) BIND(tsa_lim) BIND(qext) BIND(storage_volume) BIND(qinffield) BIND(scs_m) BIND(cumprcp) BIND(cuminf)
  BIND(tauwu) BIND(tauwv) BIND(netprcp) BIND(z_xz) BIND(z_yz) BIND(patm)
  BIND(tauwu0) BIND(tauwu1) BIND(tauwv0) BIND(tauwv1) BIND(patm0) BIND(patm1) BIND(prcp0) BIND(prcp1)
  BIND(nmindbnd) BIND(index_kcuv2) BIND(nmbkcuv2) BIND(nmikcuv2) BIND(ibkcuv2) BIND(ind1_bnd_gbp) BIND(ind2_bnd_gbp)
  BIND(nmi_gbp) BIND(index_bdr_gbp) BIND(index_zsi_bdr) BIND(dzs_bdr) BIND(ibuvdir) BIND(fac_bnd_gbp) BIND(t_bnd) BIND(zs_bnd) BIND(zsi_bnd) BIND(zsit_bnd) BIND(zsb) BIND(zsb0) BIND(uvmean) BIND(zst_bnd)
  BIND(nmindsrc) BIND(tsrc) BIND(qsrc) BIND(tprcpt) BIND(tprcpv) BIND(twnd) BIND(wndmag) BIND(wnddir) BIND(cdval) BIND(qtsrc) BIND(x73)
  BIND(drainage_type) BIND(drainage_params) BIND(drainage_distance) BIND(drainage_status) BIND(drainage_fraction_open)
  BIND(structure_uv_index) BIND(structure_height) BIND(structure_cd) BIND(structure_length)
  return -1;
}
#undef BIND

#define SETI(field) if (!strcmp(name, #field)) { m->field = (int32_t)v; return 0; }
#define SETF(field) if (!strcmp(name, #field)) { m->field = (float)v; return 0; }
int oracle_set(OM* m, const char* name, double v) {
  SETI(np) SETI(npuv) SETI(ncuv) SETI(nref) SETI(nl) SETI(subgrid) SETI(use_quadtree) SETI(advection)
  SETI(advection_scheme) SETI(viscosity) SETI(coriolis) SETI(friction2d) SETI(thetasmoothing) SETI(wiggle_suppression)
  SETI(store_zsmax) SETI(wind) SETI(precip) SETI(ngbnd) SETI(nkcuv2) SETI(nbnd) SETI(ntbnd) SETI(nsrc) SETI(ntsrc)
  SETI(ntprcp) SETI(ntwnd) SETI(h73table) SETI(nt) SETI(error) SETI(ndrn) SETF(structure_relax) SETI(nstruc)
  SETI(store_velmax) SETI(store_fluxmax) SETI(store_twet) SETF(twet_threshold) SETI(has_bzi) SETI(timestep_analysis) SETI(use_qext) SETI(use_storage_volume) SETI(infiltration) SETI(cna) SETF(sfacinf)
  SETI(meteo3d) SETI(patmos) SETI(meteo_wind) SETI(meteo_rain) SETI(spinup_meteo) SETI(amprblock) SETI(usespwprecip)
  SETF(dtwnd) SETF(gapres) SETF(pavbnd) SETF(spw_merge_frac) SETF(cosrot) SETF(sinrot) SETF(factor_wind)
  SETF(factor_pres) SETF(factor_prcp) SETF(factor_spw_size) SETF(twindupd) SETF(meteo_t0) SETF(meteo_t1)
  SETF(g) SETF(pi) SETF(huthresh) SETF(huvmin) SETF(advlim) SETF(slopelim) SETF(uvlim) SETF(nuviscfac) SETF(fcorio)
  SETF(theta) SETF(wiggle_factor) SETF(wiggle_threshold) SETF(dtmax) SETF(dtmin) SETF(alpha) SETF(rhoa) SETF(rhow)
  SETF(btfilter) SETF(btrelax) SETF(t0) SETF(t1) SETF(tspinup) SETF(zini) SETF(dx) SETF(dy) SETF(dt) SETF(min_dt)
  if (!strcmp(name, "t")) { m->t = v; return 0; }
  return -1;
}
#undef SETI
#undef SETF

double oracle_get(const OM* m, const char* name) {
  if (!strcmp(name, "t")) return m->t;
  if (!strcmp(name, "dt")) return m->dt;
  if (!strcmp(name, "min_dt")) return m->min_dt;
  if (!strcmp(name, "nt")) return m->nt;
  if (!strcmp(name, "error")) return m->error;
  if (!strcmp(name, "twindupd")) return m->twindupd;
  if (!strcmp(name, "meteo_t0")) return m->meteo_t0;
  if (!strcmp(name, "meteo_t1")) return m->meteo_t1;
  return NAN;
}

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline float fmx(float a, float b) { return a > b ? a : b; }
static inline float fmn(float a, float b) { return a < b ? a : b; }

/* power7over3 (sfincs_momentum.f90:787-820) */
static float power7over3(const float* x73, float hu) {
  if (hu < 0.00001f) return x73[0 * 1000 + (int)(1e8f * hu) - 1];
  if (hu < 0.0001f) return x73[1 * 1000 + (int)(1e7f * hu) - 1];
  if (hu < 0.001f) return x73[2 * 1000 + (int)(1e6f * hu) - 1];
  if (hu < 0.01f) return x73[3 * 1000 + (int)(1e5f * hu) - 1];
  if (hu < 0.1f) return x73[4 * 1000 + (int)(1e4f * hu) - 1];
  if (hu < 1.0f) return x73[5 * 1000 + (int)(1e3f * hu) - 1];
  if (hu < 10.0f) return x73[6 * 1000 + (int)(100 * hu) - 1];
  if (hu < 100.0f) return x73[7 * 1000 + (int)(10 * hu) - 1];
  if (hu < 1000.0f) return x73[8 * 1000 + (int)(1 * hu) - 1];
  if (hu < 10000.0f) return x73[9 * 1000 + (int)(0.1f * hu) - 1];
  return powf(10000.0f, 7.0f / 3.0f);
}

/* ---- momentum: sfincs_momentum.f90:9-784 ------------------------------------------- */
float oracle_compute_fluxes(OM* m, float dt) {
  const float expo = 1.0f / 3.0f;   /* :92 */
  float min_dt = m->dtmax;           /* :99 */
  const int32_t nuv = m->npuv + m->ncuv;
  memcpy(m->q0, m->q, sizeof(float) * nuv);   /* :125-138 */
  memcpy(m->uv0, m->uv, sizeof(float) * nuv);
  if (m->timestep_analysis)   /* :101-119 */
    for (int32_t i = 0; i < m->npuv; ++i) m->tsa_req[i] = m->dtmax;
  const int nbrs = m->advection || m->coriolis || m->viscosity || m->friction2d || m->thetasmoothing;
  /* every uv-point reads the snapshot q0/uv0 and writes only its own entries, so the
   * points are independent (the reference's own OpenMP loop, :142-752) */
#pragma omp parallel for schedule(static) reduction(min : min_dt)
  for (int32_t i = 0; i < m->npuv; ++i) {
    const int kc = m->kcuv[i];
    if (!(kc == 1 || kc == 6)) continue;
    const int32_t nm = m->uv_index_z_nm[i] - 1, nmu = m->uv_index_z_nmu[i] - 1;
    int iok = 0;
    const float zsu = (float)dmax(m->zs[nm], m->zs[nmu]);
    float zmin = 0.f, zmax = 0.f;
    if (m->subgrid) {
      zmin = m->uv_zmin[i]; zmax = m->uv_zmax[i];
      if (zsu > zmin + m->huthresh) iok = 1;
    } else {
      if (zsu > m->zbuvmx[i]) iok = 1;
    }
    if (!iok) { m->q[i] = 0.0f; m->uv[i] = 0.0f; m->kfuv[i] = 0; continue; }
    int iref = 1, itype = 0;
    if (m->use_quadtree) { iref = m->uv_flags_iref[i]; itype = m->uv_flags_type[i]; }
    const int idir = m->uv_flags_dir[i];
    float dxuvinv, dyuvinv, dxuv2inv, dyuv2inv;
    if (itype == 0) {
      if (idir == 0) { dxuvinv = m->dxrinv[iref - 1]; dyuvinv = m->dyrinv[iref - 1]; dxuv2inv = m->dxr2inv[iref - 1]; dyuv2inv = m->dyr2inv[iref - 1]; }
      else { dxuvinv = m->dyrinv[iref - 1]; dyuvinv = m->dxrinv[iref - 1]; dxuv2inv = m->dyr2inv[iref - 1]; dyuv2inv = m->dxr2inv[iref - 1]; }
    } else {
      if (idir == 0) { dxuvinv = m->dxrinvc[iref - 1]; dyuvinv = m->dyrinv[iref - 1]; dxuv2inv = 0.0f; dyuv2inv = m->dyr2inv[iref - 1]; }
      else { dxuvinv = m->dyrinvc[iref - 1]; dyuvinv = m->dxrinv[iref - 1]; dxuv2inv = 0.0f; dyuv2inv = m->dxr2inv[iref - 1]; }
    }
    (void)dyuvinv;
    const float fcoriouv = m->fcorio;
    const float qx_nm = m->q0[i];
    float qx_nmd = 0, qx_nmu = 0, qy_nm = 0, qy_nmu = 0, qy_ndm = 0, qy_ndmu = 0;
    float uu_nm = 0, uu_nmd = 0, uu_nmu = 0, uu_ndm = 0, uu_num = 0, vu = 0;
    if (nbrs) {
      qx_nmd = m->q0[m->uv_index_u_nmd[i] - 1];
      qx_nmu = m->q0[m->uv_index_u_nmu[i] - 1];
      qy_nm = m->q0[m->uv_index_v_nm[i] - 1];
      qy_nmu = m->q0[m->uv_index_v_nmu[i] - 1];
      qy_ndm = m->q0[m->uv_index_v_ndm[i] - 1];
      qy_ndmu = m->q0[m->uv_index_v_ndmu[i] - 1];
      uu_nm = m->uv0[i];
      uu_nmd = m->uv0[m->uv_index_u_nmd[i] - 1];
      uu_nmu = m->uv0[m->uv_index_u_nmu[i] - 1];
      uu_ndm = m->uv0[m->uv_index_u_ndm[i] - 1];
      uu_num = m->uv0[m->uv_index_u_num[i] - 1];
      vu = (m->uv0[m->uv_index_v_ndm[i] - 1] + m->uv0[m->uv_index_v_ndmu[i] - 1] + m->uv0[m->uv_index_v_nm[i] - 1] +
            m->uv0[m->uv_index_v_nmu[i] - 1]) / 4;
    }
    float phi = 1.0f, hu, gnavg2;
    if (m->subgrid) {
      const int nl = m->nl;
      if (zsu > zmax) {
        hu = m->uv_havg_zmax[i] + zsu;
        gnavg2 = m->uv_navg_w[i] - (m->uv_navg_w[i] - m->uv_nrep_zmax[i]) / (m->uv_fnfit[i] * (zsu - zmax) + 1.0f);
      } else {
        const float dzuv = (zmax - zmin) / (float)(nl - 1);
        int iuv = (int)((zsu - zmin) / dzuv) + 1;
        if (iuv > nl - 1) iuv = nl - 1;
        const float facint = (zsu - (zmin + (float)(iuv - 1) * dzuv)) / dzuv;
        const float* hv = m->uv_havg + (size_t)i * nl;
        const float* nr = m->uv_nrep + (size_t)i * nl;
        const float* pw = m->uv_pwet + (size_t)i * nl;
        hu = hv[iuv - 1] + (hv[iuv] - hv[iuv - 1]) * facint;
        gnavg2 = nr[iuv - 1] + (nr[iuv] - nr[iuv - 1]) * facint;
        phi = pw[iuv - 1] + (pw[iuv] - pw[iuv - 1]) * facint;
      }
    } else {
      hu = fmx(zsu - m->zbuvmx[i], m->huthresh);
      gnavg2 = m->gn2uv[i];
    }
    const float hwet = hu / phi;
    float dzdx;
    if (m->slopelim < 9999.0f) {
      double d = (m->zs[nmu] - m->zs[nm]) * (double)dxuvinv;
      if (d < (double)(-m->slopelim)) d = (double)(-m->slopelim);
      if (d > (double)m->slopelim) d = (double)m->slopelim;
      dzdx = (float)d;
    } else {
      dzdx = (float)((m->zs[nmu] - m->zs[nm]) * (double)dxuvinv);
    }
    float frc = -(m->g * hu * dzdx);
    if (m->advection && m->mask_adv[i] == 1) {
      float dqxudx = 0.0f, dqyudy = 0.0f;
      if (m->advection_scheme == 0) {
        if (qx_nm > 1.0e-6f) dqxudx = (qx_nm * uu_nm - qx_nmd * uu_nmd) * dxuvinv;
        else if (qx_nm < -1.0e-6f) dqxudx = (qx_nmu * uu_nmu - qx_nm * uu_nm) * dxuvinv;
        const float qy = qy_ndm + qy_ndmu + qy_nm + qy_nmu;
        if (qy > 1.0e-6f) dqyudy = (qy_ndm + qy_ndmu) * (uu_nm - uu_ndm) * dyuvinv / 2;
        else if (qy < -1.0e-6f) dqyudy = (qy_nm + qy_nmu) * (uu_num - uu_nm) * dyuvinv / 2;
      } else if (m->advection_scheme == 1) {
        float qd = (qx_nmd + qx_nm) / 2;
        float qu = (qx_nm + qx_nmu) / 2;
        if (qd > 1.0e-6f) {
          const float ud = (uu_nmd + uu_nm) / 2;
          dqxudx = (qd * (uu_nm - uu_nmd) + ud * (qx_nm - qx_nmd)) * dxuvinv;
        }
        if (qu < -1.0e-6f) {
          const float uu = (uu_nm + uu_nmu) / 2;
          dqxudx = dqxudx + (qu * (uu_nmu - uu_nm) + uu * (qx_nmu - qx_nm)) * dxuvinv;
        }
        qu = (qy_nm + qy_nmu) / 2;
        qd = (qy_ndm + qy_ndmu) / 2;
        if (qd > 1.0e-6f) dqyudy = qd * (uu_nm - uu_ndm) * dyuvinv;
        if (qu < -1.0e-6f) dqyudy = dqyudy + qu * (uu_num - uu_nm) * dyuvinv;
        const float ud = (uu_nmd + uu_nm) / 2;
        const float uu = (uu_nm + uu_nmu) / 2;
        if (ud > 1.0e-6f) dqyudy = dqyudy + ud * (qy_nm - qy_ndm) * dyuvinv;
        if (uu < -1.0e-6f) dqyudy = dqyudy + uu * (qy_nmu - qy_ndmu) * dyuvinv;
      }
      float adv = -(phi * (dqxudx + dqyudy));
      adv = fmn(fmx(adv, -(m->advlim * hu)), m->advlim * hu);
      frc = frc + adv;
    }
    if (m->viscosity) {
      const float lap = (uu_nmu - 2 * uu_nm + uu_nmd) * dxuv2inv + (uu_num - 2 * uu_nm + uu_ndm) * dyuv2inv;
      if (itype == 0) frc = frc + m->nuvisc[iref - 1] * hu * lap;
      else frc = frc + m->nuviscfac * m->nuvisc[iref - 1] * hu * lap;
    }
    if (m->coriolis) {
      if (idir == 0) frc = frc + fcoriouv * hu * vu;
      else frc = frc - fcoriouv * hu * vu;
    }
    if (m->wind) {
      if (hwet > 0.25f) frc = frc + phi * (idir == 0 ? m->tauwu[nm] : m->tauwv[nm]);
      else frc = frc + (idir == 0 ? m->tauwu[nm] : m->tauwv[nm]) * hu * 4;
    }
    if (m->patmos) {   /* :585-589 */
      frc = frc + hu * (m->patm[nm] - m->patm[nmu]) * dxuvinv / m->rhow;
    }
    float qfr;
    if (m->kfuv[i] == 0) {
      qfr = sqrtf(fabsf(dzdx) / (fmx(gnavg2, 1.0e-5f) / 10)) * powf(hu, 5.0f / 3.0f);
    } else if (m->friction2d) {
      qfr = sqrtf(qx_nm * qx_nm + (hu * vu) * (hu * vu));
    } else {
      qfr = fabsf(qx_nm);
    }
    float qsm = qx_nm;
    if (m->thetasmoothing) {
      if (m->kcuv[m->uv_index_u_nmd[i] - 1] == 1 && m->kcuv[m->uv_index_u_nmu[i] - 1] == 1)
        if (fabsf(qx_nmd) > 1.0e-6f && fabsf(qx_nmu) > 1.0e-6f)
          qsm = m->theta * qx_nm + 0.5f * (1.0f - m->theta) * (qx_nmu + qx_nmd);
    }
    const float hu73 = m->h73table ? power7over3(m->x73, fmx(hu, 1.0e-6f)) : hu * hu * powf(hu, expo);
    float qn = (qsm + frc * dt) / (1.0f + gnavg2 * dt * qfr / hu73);
    if (m->subgrid && m->wiggle_suppression) {
      const float mdrv = fabsf(m->zsderv[nm] - m->zsderv[nmu]) - m->wiggle_threshold;
      if (mdrv > 0.0f) qn = qn * m->wiggle_threshold / (m->wiggle_factor * mdrv + m->wiggle_threshold);
    }
    if (m->subgrid) {
      if (m->z_volume[nm] < 0.0) qn = fmn(qn, 0.0f);
      if (m->z_volume[nmu] < 0.0) qn = fmx(qn, 0.0f);
    } else {
      if (m->zs[nm] < (double)m->zb[nm]) qn = fmn(qn, 0.0f);
      if (m->zs[nmu] < (double)m->zb[nmu]) qn = fmx(qn, 0.0f);
    }
    qn = fmn(fmx(qn, -(hu * m->uvlim)), hu * m->uvlim);
    m->q[i] = qn;
    const float uvn = qn / fmx(hu, m->huvmin);
    m->uv[i] = uvn;
    m->kfuv[i] = 1;
    const float min_dt_ip = 1.0f / (fmx(sqrtf(m->g * hu), fabsf(uvn)) * dxuvinv);
    min_dt = fmn(min_dt, min_dt_ip);
    if (m->timestep_analysis) m->tsa_req[i] = min_dt_ip;   /* :737-741 */
  }
  for (int32_t j = 0; j < m->ncuv; ++j) {   /* :757-779 */
    const int32_t c = m->cuv_index_uv[j] - 1, a = m->cuv_index_uv1[j] - 1, b = m->cuv_index_uv2[j] - 1;
    m->q[c] = (m->q[a] + m->q[b]) / 2;
    m->uv[c] = (m->uv[a] + m->uv[b]) / 2;
  }
  m->min_dt = min_dt;
  if (m->timestep_analysis) {   /* timestep_analysis_update (sfincs_timestep_analysis.f90:26-74) */
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < m->npuv; ++i) {
      if (!(m->kcuv[i] == 1 || m->kcuv[i] == 6) || m->kfuv[i] != 1) continue;
      m->tsa_sum[i] = m->tsa_sum[i] + m->tsa_req[i];
      m->tsa_wet[i] = m->tsa_wet[i] + 1;
      if (m->tsa_req[i] <= min_dt + 1.0e-6f) m->tsa_lim[i] = m->tsa_lim[i] + 1;
    }
  }
  return min_dt;
}

/* ---- store variables: sfincs_continuity.f90:626-720 ------------------------------------ */
static void compute_store_variables(OM* m, float dt) {
#pragma omp parallel for schedule(static)
  for (int32_t nm = 0; nm < m->np; ++nm) {
    if (!(m->kcs[nm] == 1 || m->kcs[nm] == 4)) continue;
    const int32_t nmd = m->z_index_uv_md[nm] - 1, nmu = m->z_index_uv_mu[nm] - 1;
    const int32_t ndm = m->z_index_uv_nd[nm] - 1, num = m->z_index_uv_nu[nm] - 1;
    float uvz = 0.0f, qz = 0.0f;
    if (m->store_velmax) {
      const float quz = (m->uv[nmd] + m->uv[nmu]) / 2, qvz = (m->uv[ndm] + m->uv[num]) / 2;
      uvz = sqrtf(quz * quz + qvz * qvz);
    }
    if (m->store_fluxmax) {
      const float quz = (m->q[nmd] + m->q[nmu]) / 2, qvz = (m->q[ndm] + m->q[num]) / 2;
      qz = sqrtf(quz * quz + qvz * qvz);
    }
    if (m->store_velmax) m->vmax[nm] = fmx(m->vmax[nm], uvz);
    if (m->store_fluxmax) m->qmax[nm] = fmx(m->qmax[nm], qz);
    if (m->store_twet) {
      const double base = m->subgrid ? (double)m->z_zmin[nm] : (double)m->zb[nm];
      if ((m->zs[nm] - base) > (double)m->twet_threshold) m->twet[nm] = m->twet[nm] + dt;
    }
  }
}

static void water_levels(OM* m, float dt);

/* ---- continuity: sfincs_continuity.f90:5-43 ------------------------------------------ */
void oracle_compute_water_levels(OM* m, double t, float dt) {
  (void)t;
  water_levels(m, dt);
  if (m->store_velmax || m->store_fluxmax || m->store_twet) compute_store_variables(m, dt);
}

/* ---- continuity: sfincs_continuity.f90:46-262 and :265-624 -------------------------- */
static void water_levels(OM* m, float dt) {
  if (!m->subgrid) {
    for (int32_t s = 0; s < m->nsrc + 2 * m->ndrn; ++s) {   /* :87-106, nsrcdrn */
      const int32_t nm = m->nmindsrc[s] - 1;
      const float inc = m->qtsrc[s] * dt / m->cell_area[m->z_flags_iref[nm] - 1];
      m->zs[nm] = dmax(m->zs[nm] + (double)inc, (double)m->zb[nm]);
    }
#pragma omp parallel for schedule(static)
    for (int32_t nm = 0; nm < m->np; ++nm) {
      if (m->kcs[nm] == 1) {
        if (m->precip) m->zs[nm] = m->zs[nm] + (double)(m->netprcp[nm] * dt);
        if (m->use_qext) m->zs[nm] = m->zs[nm] + (double)(m->qext[nm] * dt);   /* :122-128 */
        const int32_t nmd = m->z_index_uv_md[nm] - 1, nmu = m->z_index_uv_mu[nm] - 1;
        const int32_t ndm = m->z_index_uv_nd[nm] - 1, num = m->z_index_uv_nu[nm] - 1;
        const int ir = m->z_flags_iref[nm] - 1;
        const float d = ((m->q[nmd] - m->q[nmu]) * m->dxrinv[ir] + (m->q[ndm] - m->q[num]) * m->dyrinv[ir]) * dt;
        m->zs[nm] = m->zs[nm] + (double)d;
      }
      if (m->store_zsmax) m->zsmax[nm] = (float)dmax((double)m->zsmax[nm], m->zs[nm]);
    }
    return;
  }
  for (int32_t s = 0; s < m->nsrc + 2 * m->ndrn; ++s) {   /* :311-325, nsrcdrn */
    const int32_t nm = m->nmindsrc[s] - 1;
    if ((m->z_volume[nm] >= 0) || ((m->qtsrc[s] < 0.0f) && (m->z_volume[nm] >= 0)))
      m->z_volume[nm] = m->z_volume[nm] + (double)(m->qtsrc[s] * dt);
  }
  const int nl = m->nl;
  /* cells read the new q and write only their own state: independent */
#pragma omp parallel for schedule(static)
  for (int32_t nm = 0; nm < m->np; ++nm) {
    float dvol = 0.0f;
    const int k = m->kcs[nm];
    const int ir = m->z_flags_iref[nm] - 1;
    if (k == 1) {
      const int32_t nmd = m->z_index_uv_md[nm] - 1, nmu = m->z_index_uv_mu[nm] - 1;
      const int32_t ndm = m->z_index_uv_nd[nm] - 1, num = m->z_index_uv_nu[nm] - 1;
      if (m->use_quadtree)
        dvol = dvol + ((m->q[nmd] - m->q[nmu]) * m->dyrm[ir] + (m->q[ndm] - m->q[num]) * m->dxrm[ir]) * dt;
      else
        dvol = dvol + ((m->q[nmd] - m->q[nmu]) * m->dy + (m->q[ndm] - m->q[num]) * m->dx) * dt;
    }
    if (k == 1 || k == 4) {
      const float a = m->cell_area[ir];
      if (m->precip || m->use_qext) {   /* :473-497 */
        float dzsdt = 0.0f;
        if (m->precip) dzsdt = dzsdt + m->netprcp[nm];
        if (m->use_qext) dzsdt = dzsdt + m->qext[nm];
        dvol = dvol + dzsdt * a * dt;
      }
      if (m->use_storage_volume && m->storage_volume[nm] > 1.0e-6f && dvol > 0.0f) {   /* :499-523 */
        const float dv = m->storage_volume[nm] - dvol;
        m->storage_volume[nm] = fmx(dv, 0.0f);
        dvol = dv < 0.0f ? -dv : 0.0f;
      }
      m->z_volume[nm] = m->z_volume[nm] + (double)dvol;
      float zs00 = 0.f, zs11 = 0.f;
      if (m->wiggle_suppression) {
        zs00 = m->zs0[nm];
        zs11 = (float)m->zs[nm];
        m->zs0[nm] = zs11;
      }
      const float vmax = m->z_volmax[nm];
      const double V = m->z_volume[nm];
      if (V >= (double)(vmax * 0.999f)) {
        m->zs[nm] = (double)fmx(m->z_zmax[nm], -20.0f) + (V - (double)vmax) / (double)a;
      } else if (V <= (double)1.0e-6f) {
        m->zs[nm] = (double)fmx(m->z_zmin[nm], -20.0f);
      } else {
        const float dzvol = vmax / (float)(nl - 1);
        const int iuv = (int)(V / (double)dzvol) + 1;
        const float facint = (float)((V - (double)((float)(iuv - 1) * dzvol)) / (double)dzvol);
        const float* dep = m->z_dep + (size_t)nm * nl;
        m->zs[nm] = (double)(dep[iuv - 1] + (dep[iuv] - dep[iuv - 1]) * facint);
      }
      if (m->wiggle_suppression)
        m->zsderv[nm] = (float)((m->zs[nm] - (double)(2 * zs11)) + (double)zs00);
    }
    if (m->store_zsmax) m->zsmax[nm] = (float)dmax((double)m->zsmax[nm], m->zs[nm]);
  }
}

/* ---- meteo time series: sfincs_meteo.f90:1423-1437, 1445-1558 ----------------------- */
static float f_modulo(float a, float p) {
  float r = fmodf(a, p);
  if (r != 0.0f && ((r < 0.0f) != (p < 0.0f))) r += p;
  return r;
}

static void meteo3d_forcing(OM* m, double t);
void oracle_update_meteo(OM* m, double t, float dt) {
  (void)dt;
  if (m->meteo3d) meteo3d_forcing(m, t);
  if (m->wind && m->ntwnd > 0) {
    for (int32_t itw = m->itwndlast; itw <= m->ntwnd; ++itw) {
      if ((double)m->twnd[itw - 1] > t) {
        const float twfac = (float)((t - (double)m->twnd[itw - 2]) / (double)(m->twnd[itw - 1] - m->twnd[itw - 2]));
        const float vmag = m->wndmag[itw - 2] * (1.0f - twfac) + m->wndmag[itw - 1] * twfac;
        float dr0 = f_modulo(m->wnddir[itw - 2], 2 * m->pi);
        float dr1 = f_modulo(m->wnddir[itw - 1], 2 * m->pi);
        if (dr1 > dr0 + m->pi) dr0 = dr0 + 2 * m->pi;
        else if (dr0 > dr1 + m->pi) dr1 = dr1 + 2 * m->pi;
        const float vdir = dr0 * (1.0f - twfac) + dr1 * twfac;
        const float cd = m->cdval[(int)(vmag * 10) + 1 - 1];
        const double twu = (double)(vmag * vmag * cosf(vdir) * m->rhoa * cd / m->rhow);
        const double twv = (double)(vmag * vmag * sinf(vdir) * m->rhoa * cd / m->rhow);
        for (int32_t nm = 0; nm < m->np; ++nm) { m->tauwu[nm] = (float)twu; m->tauwv[nm] = (float)twv; }
        m->itwndlast = itw - 1;
        break;
      }
    }
  }
  if (m->precip && m->ntprcp > 0) {
    int32_t itp;
    for (itp = m->itprcplast; itp <= m->ntprcp; ++itp)
      if ((double)m->tprcpt[itp - 1] > t) break;
    if (itp > m->ntprcp) itp = m->ntprcp;
    if (itp < 1) itp = 1;
    const float twfac = (float)((t - (double)m->tprcpt[itp - 2]) / (double)(m->tprcpt[itp - 1] - m->tprcpt[itp - 2]));
    const float ptmp = (m->tprcpv[itp - 2] * (1.0f - twfac) + m->tprcpv[itp - 1] * twfac) / (float)(1000 * 3600);
    for (int32_t nm = 0; nm < m->np; ++nm) m->netprcp[nm] = ptmp;
    m->itprcplast = itp - 1;
  }
}

/* ---- gridded and spiderweb meteo: sfincs_meteo.f90 ----------------------------------- */
static float* fdup(const float* a, size_t n) {
  float* b = (float*)malloc((n ? n : 1) * sizeof(float));
  if (n) memcpy(b, a, n * sizeof(float));
  return b;
}

/* An amu (q 0), amv (1), amp (2) or ampr (3) grid as read (read_amuv_file). */
int oracle_set_grid(OM* m, int32_t q, int32_t nt, int32_t nrows, int32_t ncols, float xll, float yll, float dx,
                    float dy, const float* times, const float* v) {
  MG* g = q == 0 ? &m->amu : q == 1 ? &m->amv : q == 2 ? &m->amp : q == 3 ? &m->ampr : NULL;
  if (!g) return -1;
  mg_free(g);
  g->nt = nt; g->nrows = nrows; g->ncols = ncols; g->xll = xll; g->yll = yll; g->dx = dx; g->dy = dy;
  g->times = fdup(times, (size_t)nt);
  g->v = fdup(v, (size_t)nt * nrows * ncols);
  return 0;
}

/* A spiderweb as read (read_spw_file), converted as read_meteo_data does (:57-95, 330-414):
 * u, v from speed and nautical direction, absolute pressure, the factors. Call after the
 * scalars are set. */
int oracle_set_spw(OM* m, int32_t nt, int32_t nrows, int32_t ncols, float radius, const float* times,
                   const float* xe, const float* ye, const float* vmag, const float* vdir, const float* pdrop,
                   const float* prcp) {
  const size_t n = (size_t)nt * nrows * ncols;
  m->spw_nt = nt; m->spw_nrows = nrows; m->spw_ncols = ncols;
  m->spw_times = fdup(times, (size_t)nt); m->spw_xe = fdup(xe, (size_t)nt); m->spw_ye = fdup(ye, (size_t)nt);
  m->spw_wu = (float*)malloc(n * sizeof(float));
  m->spw_wv = (float*)malloc(n * sizeof(float));
  m->spw_pabs = (float*)malloc(n * sizeof(float));
  m->spw_rain = prcp != NULL && m->usespwprecip;
  m->spw_prcp = m->spw_rain ? fdup(prcp, n) : NULL;
  for (size_t k = 0; k < n; ++k) {
    const float ang = m->pi * (270.0f - vdir[k]) / 180;
    m->spw_wu[k] = vmag[k] * cosf(ang);
    m->spw_wv[k] = vmag[k] * sinf(ang);
    m->spw_pabs[k] = m->gapres - pdrop[k];
  }
  if (m->factor_wind != 1.0f)
    for (size_t k = 0; k < n; ++k) { m->spw_wu[k] = m->factor_wind * m->spw_wu[k]; m->spw_wv[k] = m->factor_wind * m->spw_wv[k]; }
  if (m->factor_pres != 1.0f)
    for (size_t k = 0; k < n; ++k) m->spw_pabs[k] = m->gapres - m->factor_pres * (m->gapres - m->spw_pabs[k]);
  if (m->factor_prcp != 1.0f && m->spw_rain)
    for (size_t k = 0; k < n; ++k) m->spw_prcp[k] = m->factor_prcp * m->spw_prcp[k];
  m->spw_radius = radius;
  if (m->factor_spw_size != 1.0f) m->spw_radius = m->factor_spw_size * m->spw_radius;
  m->dradspw = m->spw_radius / nrows;
  m->dphispw = 2 * m->pi / ncols;
  return 0;
}

/* The factors on the grids (:347-381), after oracle_set_grid and the scalars. */
void oracle_meteo_factors(OM* m) {
  if (m->factor_wind != 1.0f) {
    for (size_t k = 0; k < (size_t)m->amu.nt * m->amu.nrows * m->amu.ncols; ++k) m->amu.v[k] = m->factor_wind * m->amu.v[k];
    for (size_t k = 0; k < (size_t)m->amv.nt * m->amv.nrows * m->amv.ncols; ++k) m->amv.v[k] = m->factor_wind * m->amv.v[k];
  }
  if (m->factor_pres != 1.0f)
    for (size_t k = 0; k < (size_t)m->amp.nt * m->amp.nrows * m->amp.ncols; ++k)
      m->amp.v[k] = m->gapres - m->factor_pres * (m->gapres - m->amp.v[k]);
  if (m->factor_prcp != 1.0f)
    for (size_t k = 0; k < (size_t)m->ampr.nt * m->ampr.nrows * m->ampr.ncols; ++k) m->ampr.v[k] = m->factor_prcp * m->ampr.v[k];
}

/* Row iy (1-based, from the north), column ix of a grid: value of frame blend `b`. */
#define GV(b, g, iy, ix) ((b)[((iy) - 1) * (g)->ncols + ((ix) - 1)])

/* One grid quantity at meteo time mt on every cell (update_amuv_data :821-975,
 * update_amp_data :996-1106, update_ampr_data :1126-1242): the two file frames around mt
 * blended (block: the earlier one), then bilinear weights on the cell centre; a cell
 * outside (the first row and column index 0 included) gets `outside`. kind 0 wind (both
 * components, to stresses), 1 pressure, 2 rain (to m/s). */
static void grid_to_cells(OM* m, const MG* g, const MG* g2, int kind, int block, float mt_in, float* o1, float* o2,
                          float outside) {
  float mt = mt_in;
  mt = mt > g->times[0] ? mt : g->times[0];
  mt = mt < g->times[g->nt - 1] ? mt : g->times[g->nt - 1];
  int itw0 = 1;
  for (int k = 1; k <= g->nt; ++k)
    if (g->times[k - 1] <= mt) itw0 = k;
  int itw1 = itw0 + 1 < g->nt ? itw0 + 1 : g->nt;
  mt = mt < g->times[itw1 - 1] ? mt : g->times[itw1 - 1];
  float den = g->times[itw1 - 1] - g->times[itw0 - 1];
  if (den < 1.0e-6f) den = 1.0e-6f;
  float twfac = (mt - g->times[itw0 - 1]) / den;
  if (block) twfac = 0.0f;
  const size_t nf = (size_t)g->nrows * g->ncols;
  float* b1 = (float*)malloc(nf * sizeof(float));
  float* b2 = g2 ? (float*)malloc(nf * sizeof(float)) : NULL;
  for (size_t k = 0; k < nf; ++k) {
    b1[k] = g->v[(size_t)(itw0 - 1) * nf + k] * (1.0f - twfac) + g->v[(size_t)(itw1 - 1) * nf + k] * twfac;
    if (g2) b2[k] = g2->v[(size_t)(itw0 - 1) * nf + k] * (1.0f - twfac) + g2->v[(size_t)(itw1 - 1) * nf + k] * twfac;
  }
  const float yul = g->yll + (g->nrows - 1) * g->dy + 0.5f * g->dy;
  const float xll = g->xll + 0.5f * g->dx;
  for (int32_t nm = 0; nm < m->np; ++nm) {
    const float x = m->z_xz[nm], y = m->z_yz[nm];
    const int iy = (int)((yul - y) / g->dy) + 1;
    const float dj1 = (yul - (iy - 1) * g->dy - y) / g->dy;
    const int ix = (int)((x - xll) / g->dx) + 1;
    const float di1 = (x - (xll + (ix - 1) * g->dx)) / g->dx;
    if (iy <= 1 || iy >= g->nrows || ix <= 0 || ix >= g->ncols) {
      o1[nm] = outside;
      if (o2) o2[nm] = outside;
      continue;
    }
    const int r[4] = {iy, iy, iy + 1, iy + 1}, c[4] = {ix, ix + 1, ix + 1, ix};
    const float f[4] = {(1.0f - di1) * (1.0f - dj1), di1 * (1.0f - dj1), di1 * dj1, (1.0f - di1) * dj1};
    float a = 0.0f, bsum = 0.0f;
    for (int ip = 0; ip < 4; ++ip) {
      a = a + f[ip] * GV(b1, g, r[ip], c[ip]);
      if (b2) bsum = bsum + f[ip] * GV(b2, g, r[ip], c[ip]);
    }
    if (kind == 0) {
      const float vmag = sqrtf(a * a + bsum * bsum);
      const float cd = m->cdval[(int)(vmag * 10) + 1 - 1];
      o1[nm] = vmag * (m->cosrot * a + m->sinrot * bsum) * m->rhoa * cd / m->rhow;
      o2[nm] = vmag * (-m->sinrot * a + m->cosrot * bsum) * m->rhoa * cd / m->rhow;
    } else if (kind == 1) {
      o1[nm] = a;
    } else {
      o1[nm] = a / (1000 * 3600);
    }
  }
  free(b1);
  free(b2);
}

/* update_spiderweb_data (:450-800) at meteo time mt onto the fields tu, tv, pa, pr. */
static void spw_to_cells(OM* m, float mt, float* tu, float* tv, float* pa, float* pr) {
  int itw0, itw1;
  float twfac, tspinup_fac;
  const int nt = m->spw_nt;
  if (mt < m->spw_times[0]) {
    itw0 = itw1 = 1; twfac = 1.0f;
    tspinup_fac = 1.0f - (m->spw_times[0] - mt) / 21600.0f;
    if (tspinup_fac < 0.0f) tspinup_fac = 0.0f;
  } else if (mt > m->spw_times[nt - 1]) {
    itw0 = itw1 = nt; twfac = 1.0f;
    tspinup_fac = 1.0f - (mt - m->spw_times[nt - 1]) / 21600.0f;
    if (tspinup_fac < 0.0f) tspinup_fac = 0.0f;
  } else {
    itw0 = 1;
    for (int k = 1; k <= nt; ++k)
      if (m->spw_times[k - 1] <= mt) itw0 = k;
    itw1 = itw0 + 1 < nt ? itw0 + 1 : nt;
    float den = m->spw_times[itw1 - 1] - m->spw_times[itw0 - 1];
    if (den < 1.0e-6f) den = 1.0e-6f;
    twfac = (mt - m->spw_times[itw0 - 1]) / den;
    tspinup_fac = 1.0f;
  }
  const float xe01 = m->spw_xe[itw0 - 1] * (1.0f - twfac) + m->spw_xe[itw1 - 1] * twfac;
  const float ye01 = m->spw_ye[itw0 - 1] * (1.0f - twfac) + m->spw_ye[itw1 - 1] * twfac;
  const int nrows = m->spw_nrows, ncols = m->spw_ncols;
  const size_t nf = (size_t)nrows * ncols;
  float* wu = (float*)malloc(nf * sizeof(float));
  float* wv = (float*)malloc(nf * sizeof(float));
  float* pb = (float*)malloc(nf * sizeof(float));
  float* rn = (float*)malloc(nf * sizeof(float));
  for (size_t k = 0; k < nf; ++k) {
    const size_t a = (size_t)(itw0 - 1) * nf + k, b = (size_t)(itw1 - 1) * nf + k;
    wu[k] = m->spw_wu[a] * (1.0f - twfac) + m->spw_wu[b] * twfac;
    wv[k] = m->spw_wv[a] * (1.0f - twfac) + m->spw_wv[b] * twfac;
    pb[k] = m->spw_pabs[a] * (1.0f - twfac) + m->spw_pabs[b] * twfac;
    rn[k] = m->spw_rain ? m->spw_prcp[a] * (1.0f - twfac) + m->spw_prcp[b] * twfac : 0.0f;
  }
  for (int32_t nm = 0; nm < m->np; ++nm) {
    const float dxe = m->z_xz[nm] - xe01, dye = m->z_yz[nm] - ye01;
    const float dst = sqrtf(dxe * dxe + dye * dye);
    if (dst > m->spw_radius) continue;
    int id = (int)(dst / m->dradspw);
    if (id < 1) id = 1;
    if (id + 1 > nrows) continue;
    float dj1 = (dst - m->dradspw * id) / m->dradspw;
    if (dj1 < 0.0f) dj1 = 0.0f;
    float phi = 0.5f * m->pi - atan2f(dye, dxe);
    phi = f_modulo(phi, 2 * m->pi);
    int ip = (int)(phi / m->dphispw) + 1;
    if (ip > ncols) ip = ncols;
    const int r[4] = {id, id, id + 1, id + 1};
    int c[4] = {ip, ip + 1, ip + 1, ip};
    if (c[1] > ncols) c[1] = 1;
    if (c[2] > ncols) c[2] = 1;
    const float di1 = (phi - m->dphispw * (ip - 1)) / m->dphispw;
    const float f[4] = {(1.0f - di1) * (1.0f - dj1), di1 * (1.0f - dj1), di1 * dj1, (1.0f - di1) * dj1};
    float wup = 0.0f, wvp = 0.0f, pcp = 0.0f, prp = 0.0f;
    for (int k = 0; k < 4; ++k) {
      const size_t at = (size_t)(r[k] - 1) * ncols + (c[k] - 1);
      wup = wup + f[k] * wu[at];
      wvp = wvp + f[k] * wv[at];
      if (m->patmos) pcp = pcp + f[k] * pb[at];
      if (m->precip && m->spw_rain) prp = prp + f[k] * rn[at];
    }
    const float vmag = sqrtf(wup * wup + wvp * wvp);
    const float cd = m->cdval[(int)(vmag * 10) + 1 - 1];
    float merge;
    if (dst > m->spw_merge_frac * m->spw_radius) {
      const float mf = m->spw_merge_frac < 0.999f ? m->spw_merge_frac : 0.999f;
      merge = (1.0f / (1.0f - mf)) * (m->spw_radius - dst) / m->spw_radius;
    } else {
      merge = 1.0f;
    }
    merge = merge * tspinup_fac;
    tu[nm] = (1.0f - merge) * tu[nm] + merge * vmag * (m->cosrot * wup + m->sinrot * wvp) * m->rhoa * cd / m->rhow;
    tv[nm] = (1.0f - merge) * tv[nm] + merge * vmag * (-m->sinrot * wup + m->cosrot * wvp) * m->rhoa * cd / m->rhow;
    if (m->patmos) pa[nm] = (1.0f - merge) * pa[nm] + merge * pcp;
    if (m->precip && m->spw_rain) pr[nm] = (1.0f - merge) * pr[nm] + merge * prp / (1000 * 3600);
  }
  free(wu); free(wv); free(pb); free(rn);
}

/* update_meteo_fields (:1561-1624): the fields at meteo_t0 and meteo_t1. */
void oracle_meteo_fields(OM* m) {
  for (int itw = 0; itw < 2; ++itw) {
    const float mt = itw == 0 ? m->meteo_t0 : m->meteo_t1;
    float* tu = itw == 0 ? m->tauwu0 : m->tauwu1;
    float* tv = itw == 0 ? m->tauwv0 : m->tauwv1;
    float* pa = itw == 0 ? m->patm0 : m->patm1;
    float* pr = itw == 0 ? m->prcp0 : m->prcp1;
    if (m->amu.nt > 0) grid_to_cells(m, &m->amu, &m->amv, 0, 0, mt, tu, tv, 0.0f);
    else for (int32_t nm = 0; nm < m->np; ++nm) { tu[nm] = 0.0f; tv[nm] = 0.0f; }
    if (m->patmos) {
      if (m->amp.nt > 0) grid_to_cells(m, &m->amp, NULL, 1, 0, mt, pa, NULL, m->gapres);
      else for (int32_t nm = 0; nm < m->np; ++nm) pa[nm] = m->gapres;
    }
    if (m->meteo_rain) {
      if (m->ampr.nt > 0) grid_to_cells(m, &m->ampr, NULL, 2, m->amprblock, mt, pr, NULL, 0.0f);
      else for (int32_t nm = 0; nm < m->np; ++nm) pr[nm] = 0.0f;
    }
    if (m->spw_nt > 0) spw_to_cells(m, mt, tu, tv, pa, pr);
  }
}

/* sfincs_lib.f90:493-503 with update_meteo_fields: a new window at the first step with
 * t >= twindupd. */
void oracle_meteo_window(OM* m) {
  if (m->meteo3d && m->t >= (double)m->twindupd) {
    m->meteo_t0 = m->twindupd;
    m->meteo_t1 = m->twindupd + m->dtwnd;
    m->twindupd = m->twindupd + m->dtwnd;
    oracle_meteo_fields(m);
  }
}

/* update_meteo_forcing's gridded branch (:1270-1419): every cell between the window's two
 * fields, negative rain dropped on dry cells, the spin-up. */
static void meteo3d_forcing(OM* m, double t) {
  const float twfact = (float)((t - (double)m->meteo_t0) / (double)(m->meteo_t1 - m->meteo_t0));
  const float onemintwfact = 1.0f - twfact;
  for (int32_t nm = 0; nm < m->np; ++nm) {
    const int dry = m->subgrid ? (m->z_volume[nm] <= 0.0) : (m->zs[nm] <= (double)m->zb[nm]);
    if (m->meteo_wind) {
      m->tauwu[nm] = m->tauwu0[nm] * onemintwfact + m->tauwu1[nm] * twfact;
      m->tauwv[nm] = m->tauwv0[nm] * onemintwfact + m->tauwv1[nm] * twfact;
    }
    if (m->patmos) m->patm[nm] = m->patm0[nm] * onemintwfact + m->patm1[nm] * twfact;
    if (m->meteo_rain) {
      float prcp = m->prcp0[nm] * onemintwfact + m->prcp1[nm] * twfact;
      if (prcp < 0.0f && dry) prcp = 0.0f;
      m->netprcp[nm] = prcp;
    }
  }
  if (t < (double)(m->tspinup - 1.0e-3f) && m->spinup_meteo) {
    const float smfac = (float)((t - (double)m->t0) / (double)(m->tspinup - m->t0));
    const float oneminsmfac = 1.0f - smfac;
    for (int32_t nm = 0; nm < m->np; ++nm) {
      if (m->meteo_wind) { m->tauwu[nm] = m->tauwu[nm] * smfac; m->tauwv[nm] = m->tauwv[nm] * smfac; }
      if (m->patmos) m->patm[nm] = m->patm[nm] * smfac + m->gapres * oneminsmfac;
      if (m->meteo_rain) {
        m->netprcp[nm] = m->netprcp[nm] * smfac;
        if (m->netprcp[nm] < 0.0f) {
          const int dry = m->subgrid ? (m->z_volume[nm] <= 0.0) : (m->zs[nm] <= (double)m->zb[nm]);
          if (dry) m->netprcp[nm] = 0.0f;
        }
      }
    }
  }
}

/* ---- infiltration: update_infiltration_map, sfincs_infiltration.f90:632-676 ----------- */
/* Constant rates ('con', 'c2d'): a cell with water keeps its rate, a dry one (z_volume <= 0
 * on subgrid meshes, zs <= zb on regular ones) none; the rate leaves the net
 * precipitation. The cumulative infiltration (:665-670) is output only and not kept. */
void oracle_update_infiltration(OM* m, float dt) {
  if (m->cna) {
    /* 'cna' (:678-722): every cell, wet or dry; the cumulative rain grows first, as
     * update_meteo_forcing adds it (sfincs_meteo.f90:1334-1336,1547-1549); netprcp holds
     * the cell's rain here */
    for (int32_t nm = 0; nm < m->np; ++nm) {
      const float prcp = m->netprcp[nm], S = m->scs_m[nm];
      m->cumprcp[nm] = m->cumprcp[nm] + prcp * dt;
      float qinfmap;
      if (m->cumprcp[nm] > m->sfacinf * S) {
        const float e = m->cumprcp[nm] - m->sfacinf * S;
        const float Qq = e * e / (m->cumprcp[nm] + (1.0f - m->sfacinf) * S);
        const float I = m->cumprcp[nm] - Qq;
        qinfmap = (I - m->cuminf[nm]) / dt;
      } else {
        qinfmap = prcp;
      }
      m->netprcp[nm] = m->netprcp[nm] - qinfmap;
      m->cuminf[nm] = m->cuminf[nm] + qinfmap * dt;
    }
    return;
  }
  for (int32_t nm = 0; nm < m->np; ++nm) {
    float qinfmap = m->qinffield[nm];
    if (m->subgrid) {
      if (m->z_volume[nm] <= 0.0) qinfmap = 0.0f;
    } else {
      if (m->zs[nm] <= (double)m->zb[nm]) qinfmap = 0.0f;
    }
    m->netprcp[nm] = m->netprcp[nm] - qinfmap;
  }
}

/* ---- boundaries: sfincs_boundaries.f90:682-1175 ------------------------------------- */
static void update_boundary_points(OM* m, double t) {
  int32_t itb0 = 1, itb1 = 1;
  float tb = 0.f;
  const int32_t nt = m->ntbnd;
  if ((double)m->t_bnd[0] > t - (double)1.0e-3f) { itb0 = 1; itb1 = 1; tb = m->t_bnd[0]; }
  else if ((double)m->t_bnd[nt - 1] < t + (double)1.0e-3f) { itb0 = nt; itb1 = nt; tb = m->t_bnd[nt - 1]; }
  else {
    for (int32_t itb = m->itbndlast; itb <= nt; ++itb) {
      if ((double)m->t_bnd[itb - 1] > t + (double)1.0e-6f) {
        itb0 = itb - 1; itb1 = itb; tb = (float)t; m->itbndlast = itb - 1;
        break;
      }
    }
  }
  const float tbfac = (tb - m->t_bnd[itb0 - 1]) / fmx(m->t_bnd[itb1 - 1] - m->t_bnd[itb0 - 1], 1.0e-6f);
  for (int32_t ib = 0; ib < m->nbnd; ++ib) {
    const float a = m->zs_bnd[(size_t)(itb0 - 1) * m->nbnd + ib], b = m->zs_bnd[(size_t)(itb1 - 1) * m->nbnd + ib];
    m->zst_bnd[ib] = a + (b - a) * tbfac;
    if (m->has_bzi) {   /* :748-753 */
      const float ia = m->zsi_bnd[(size_t)(itb0 - 1) * m->nbnd + ib], ibv = m->zsi_bnd[(size_t)(itb1 - 1) * m->nbnd + ib];
      m->zsit_bnd[ib] = ia + (ibv - ia) * tbfac;
    }
  }
}

static void update_boundary_conditions(OM* m, double t, float dt) {
  (void)dt;
  for (int32_t ib = 0; ib < m->ngbnd; ++ib) {
    const int32_t nmb = m->nmindbnd[ib] - 1;
    const int k = m->kcs[nmb];
    if (k == 2) {
      double zst;
      if (m->nbnd > 1) {
        const float f = m->fac_bnd_gbp[ib];
        zst = (double)(m->zst_bnd[m->ind1_bnd_gbp[ib] - 1] * f + m->zst_bnd[m->ind2_bnd_gbp[ib] - 1] * (1.0f - f));
      } else {
        zst = (double)m->zst_bnd[0];
      }
      if (m->patmos && m->pavbnd > 1.0f) zst = zst + (m->pavbnd - m->patm[nmb]) / (m->rhow * 9.81f);   /* :820-826 */
      const float zsetup = 0.0f;
      float zig = 0.0f;
      if (m->has_bzi) {   /* :829-848 */
        if (m->nbnd > 1) {
          const float f = m->fac_bnd_gbp[ib];
          zig = m->zsit_bnd[m->ind1_bnd_gbp[ib] - 1] * f + m->zsit_bnd[m->ind2_bnd_gbp[ib] - 1] * (1.0f - f);
        } else {
          zig = m->zsit_bnd[0];
        }
      }
      if (t < (double)(m->tspinup - 1.0e-3f)) {
        const float smfac = (float)(1.0 - (t - (double)m->t0) / (double)(m->tspinup - m->t0));
        float zs0act = (float)(zst + (double)zsetup);
        m->zsb0[ib] = m->zini * smfac + zs0act * (1.0f - smfac);
        zs0act = (float)(zst + (double)zsetup + (double)zig);
        m->zsb[ib] = m->zini * smfac + zs0act * (1.0f - smfac);
      } else {
        m->zsb0[ib] = (float)(zst + (double)zsetup);
        m->zsb[ib] = (float)(zst + (double)zsetup + (double)zig);
      }
      if (m->subgrid) {
        m->zsb0[ib] = fmx(m->zsb0[ib], m->z_zmin[nmb]);
        m->zsb[ib] = fmx(m->zsb[ib], m->z_zmin[nmb]);
      } else {
        m->zsb0[ib] = fmx(m->zsb0[ib], m->zb[nmb]);
        m->zsb[ib] = fmx(m->zsb[ib], m->zb[nmb]);
      }
    } else if (k == 5) {   /* downstream river point (sfincs_boundaries.f90:878-901) */
      const int32_t j = m->index_bdr_gbp[ib] - 1;
      double zst = m->zs[m->index_zsi_bdr[j] - 1] + (double)m->dzs_bdr[j];
      zst = dmax(zst, m->subgrid ? (double)m->z_zmin[nmb] : (double)m->zb[nmb]);
      m->zsb[ib] = (float)zst;
      m->zsb0[ib] = (float)zst;
    } else if (k == 6) {
      m->zs[nmb] = m->zs[m->nmi_gbp[ib] - 1];
    }
  }
}

static void update_boundary_fluxes(OM* m, float dt, double t) {
  (void)t;
  const float factime = fmn(dt / m->btfilter, 1.0f);
  const float one_minus_factime = 1.0f - factime;
  const float facrel = 1.0f - fmn(dt / m->btrelax, 1.0f);
  const int nl = m->nl;
  for (int32_t ib = 0; ib < m->nkcuv2; ++ib) {
    const int32_t indb = m->ibkcuv2[ib] - 1;
    const int32_t nmb = m->nmbkcuv2[ib] - 1;
    if (m->kcs[nmb] == 6) continue;
    const int32_t nmi = m->nmikcuv2[ib] - 1;
    const int32_t ip = m->index_kcuv2[ib] - 1;
    const float zsnmi = (float)m->zs[nmi];
    float zsnmb = m->zsb[indb];
    float zs0nmb = m->zsb0[indb];
    float hnmb;
    if (m->subgrid) {
      const float zsuv = fmx(zsnmb, zsnmi);
      float depthuv;
      if (zsuv >= m->uv_zmax[ip] - 1.0e-4f) {
        depthuv = m->uv_havg_zmax[ip] + zsuv;
      } else if (zsuv > m->uv_zmin[ip]) {
        const float dzuv = (m->uv_zmax[ip] - m->uv_zmin[ip]) / (float)(nl - 1);
        const int iuv = (int)((zsuv - m->uv_zmin[ip]) / dzuv) + 1;
        const float facint = (zsuv - (m->uv_zmin[ip] + (float)(iuv - 1) * dzuv)) / dzuv;
        const float* hv = m->uv_havg + (size_t)ip * nl;
        depthuv = hv[iuv - 1] + (hv[iuv] - hv[iuv - 1]) * facint;
      } else {
        depthuv = 0.0f;
      }
      hnmb = fmx(depthuv, m->huthresh);
      zsnmb = fmx(zsnmb, m->z_zmin[nmb]);
      zs0nmb = fmx(zs0nmb, m->z_zmin[nmb]);
    } else {
      hnmb = fmx(0.5f * (zsnmb + zsnmi) - m->zbuv[ip], m->huthresh);
      zsnmb = fmx(zsnmb, m->zb[nmb]);
      zs0nmb = fmx(zs0nmb, m->zb[nmb]);
    }
    if (hnmb < m->huthresh + 1.0e-6f || m->kcuv[ip] == 3) {
      m->q[ip] = 0.0f; m->uv[ip] = 0.0f; m->uvmean[ib] = 0.0f;
    } else {
      const float ui = sqrtf(m->g / hnmb) * (zsnmb - zs0nmb);
      const float ub = (float)m->ibuvdir[ib] * (2 * ui - sqrtf(m->g / hnmb) * (zsnmi - zs0nmb));
      float qb = ub * hnmb + m->uvmean[ib];
      const int dir = m->ibuvdir[ib];
      if (m->subgrid) {
        if (m->z_volume[nmi] <= 0.0) qb = (dir == 1) ? fmx(qb, 0.0f) : fmn(qb, 0.0f);
        if (zsnmb - m->z_zmin[nmb] < m->huthresh) qb = (dir == 1) ? fmn(qb, 0.0f) : fmx(qb, 0.0f);
      } else {
        if (zsnmi - m->zb[nmi] <= m->huthresh) qb = (dir == 1) ? fmx(qb, 0.0f) : fmn(qb, 0.0f);
        if (zsnmb - m->zb[nmb] < m->huthresh) qb = (dir == 1) ? fmn(qb, 0.0f) : fmx(qb, 0.0f);
      }
      m->q[ip] = qb;
      m->uv[ip] = fmx(fmn(qb / hnmb, 4.0f), -4.0f);
    }
    if (m->btfilter >= -1.0e-6f) m->uvmean[ib] = factime * m->q[ip] + facrel * one_minus_factime * m->uvmean[ib];
    else m->uvmean[ib] = 0.0f;
    m->zs[nmb] = (double)m->zsb[indb];
    if (m->store_zsmax) m->zsmax[nmb] = (float)dmax((double)m->zsmax[nmb], m->zs[nmb]);
  }
}

void oracle_update_boundaries(OM* m, double t, float dt) {
  if (m->ngbnd == 0) return;   /* boundaries_in_mask */
  if (m->nbnd > 0) update_boundary_points(m, t);
  update_boundary_conditions(m, t, dt);
  update_boundary_fluxes(m, dt, t);
}

/* ---- discharges: sfincs_discharges.f90:353-369 ---------------------------------------- */
static void update_drainage(OM* m, double t, float dt);

void oracle_update_discharges(OM* m, double t, float dt) {
  if (m->ndrn > 0) update_drainage(m, t, dt);
  if (m->nsrc <= 0) return;
  for (int32_t it = m->itsrclast; it <= m->ntsrc; ++it) {
    if ((double)m->tsrc[it - 1] > t) {
      for (int32_t s = 0; s < m->nsrc; ++s) {
        const float q0v = m->qsrc[(size_t)(it - 2) * m->nsrc + s], q1v = m->qsrc[(size_t)(it - 1) * m->nsrc + s];
        const double v = (double)q0v + (double)(q1v - q0v) * (t - (double)m->tsrc[it - 2]) /
                                          (double)(m->tsrc[it - 1] - m->tsrc[it - 2]);
        m->qtsrc[s] = (float)v;
      }
      m->itsrclast = it - 1;
      break;
    }
  }
}

/* ---- drainage structures: sfincs_discharges.f90:371-655 --------------------------------- */
/* Runs after the point sources' series (:353-369) in the reference; the oracle's
 * oracle_update_discharges calls it first because the series write only qtsrc[0..nsrc). */
static void update_drainage(OM* m, double t, float dt) {
  for (int32_t idrn = 0; idrn < m->ndrn; ++idrn) {
    const int32_t jin = m->nsrc + 2 * idrn, jout = jin + 1;   /* 0-based of nsrc + idrn*2 - 1, nsrc + idrn*2 */
    const int32_t nmin = m->nmindsrc[jin] - 1, nmout = m->nmindsrc[jout] - 1;
    if (nmin < 0 || nmout < 0) continue;
    const float* par = m->drainage_params + 6 * idrn;
    float qq = 0.0f;
    switch (m->drainage_type[idrn]) {
      case 1:   /* pump */
        qq = par[0];
        break;
      case 2:   /* culvert */
      case 3:   /* check valve */
        if (m->zs[nmin] > m->zs[nmout]) qq = (float)((double)par[0] * sqrt(m->zs[nmin] - m->zs[nmout]));
        else qq = (float)((double)(-par[0]) * sqrt(m->zs[nmout] - m->zs[nmin]));
        if (m->drainage_type[idrn] == 3) qq = fmx(qq, 0.0f);
        break;
      case 4:
      case 5: {
        const float wdt = par[0], zsill = par[1], mng = par[2], tcls = par[5];
        const float dzds = (float)((m->zs[nmout] - m->zs[nmin]) / (double)m->drainage_distance[idrn]);
        float frac = m->drainage_fraction_open[idrn];
        const float hgate = (float)dmax(dmax(m->zs[nmin], m->zs[nmout]) - (double)zsill, 0.0);
        const float dfrac = dt / tcls;
        const float qq0 = -m->qtsrc[jin] / (wdt * fmx(frac, 0.001f));
        int32_t* st = &m->drainage_status[idrn];
        if (m->drainage_type[idrn] == 4) {
          const float zmin = par[3], zmax = par[4];
          if (*st == 0) {
            if (m->zs[nmin] > (double)zmin && m->zs[nmin] < (double)zmax) *st = 3;
          } else if (*st == 1) {
            if (m->zs[nmin] <= (double)zmin || m->zs[nmin] >= (double)zmax) *st = 2;
          }
        } else {
          const float tclose = par[3], topen = par[4];
          if (*st == 0) {
            if (t >= (double)topen) *st = 3;
          } else if (*st == 1) {
            if (t >= (double)tclose && t < (double)topen) *st = 2;
          }
        }
        if (*st == 2) {
          frac = frac - dfrac;
          if (frac < 0.0f) { frac = 0.0f; *st = 0; }
        } else if (*st == 3) {
          frac = frac + dfrac;
          if (frac > 1.0f) { frac = 1.0f; *st = 1; }
        }
        m->drainage_fraction_open[idrn] = frac;
        qq = (qq0 - m->g * hgate * dzds * dt) / (1.0f + m->g * (mng * mng) * dt * fabsf(qq0) / powf(hgate, 7.0f / 3.0f));
        qq = qq * wdt * frac;
        break;
      }
    }
    qq = 1.0f / (m->structure_relax / dt) * qq + (1.0f - (1.0f / (m->structure_relax / dt))) * -m->qtsrc[jin];
    if (m->subgrid) {
      if (qq > 0.0f) qq = (float)fmin((double)qq, dmax(m->z_volume[nmin], 0.0) / (double)dt);
      else qq = (float)fmax((double)qq, -(dmax(m->z_volume[nmout], 0.0) / (double)dt));
    } else {
      if (qq > 0.0f)
        qq = (float)fmin((double)qq, dmax((m->zs[nmin] - (double)m->zb[nmin]) *
                                            (double)m->cell_area[m->z_flags_iref[nmin] - 1], 0.0) / (double)dt);
      else
        qq = (float)fmax((double)qq, -(dmax((m->zs[nmout] - (double)m->zb[nmout]) *
                                              (double)m->cell_area[m->z_flags_iref[nmout] - 1], 0.0) / (double)dt));
    }
    m->qtsrc[jin] = -qq;
    m->qtsrc[jout] = qq;
  }
}

/* ---- weirs: compute_fluxes_over_structures, sfincs_structures.f90:592-697 ------------- */
/* Broad-crested weir at each structure uv-point from the levels of its two cells, as
 * real*4: no flow with both below the crest; submerged (h2 > 2/3 h1) Cd h2 sqrt(2 g (h1 - h2)),
 * else free Cd 1.7049 h1**1.5 (the C library's powf); times the crest length; uv = 0. */
void oracle_compute_fluxes_over_structures(OM* m) {
  for (int32_t k = 0; k < m->nstruc; ++k) {
    const int32_t ip = m->structure_uv_index[k] - 1;
    m->q[ip] = 0.0f;
    m->uv[ip] = 0.0f;
    const int32_t nm = m->uv_index_z_nm[ip] - 1, nmu = m->uv_index_z_nmu[ip] - 1;
    const float zsnm = (float)m->zs[nm], zsnmu = (float)m->zs[nmu];
    const float zc = m->structure_height[k];
    if (zsnm < zc && zsnmu < zc) continue;
    const float cweir = 1.7049f, mlim = 0.0f, cd = m->structure_cd[k];
    int idir;
    float h1, h2, qstruc;
    if (zsnm > zsnmu) { idir = 1; h1 = zsnm - zc; h2 = zsnmu - zc; }
    else { idir = -1; h1 = zsnmu - zc; h2 = zsnm - zc; }
    if (h2 > 2.0f / 3.0f * h1) qstruc = cd * h2 * sqrtf(2.0f * 9.81f * (h1 - h2) / (1.0f - mlim));
    else qstruc = cd * cweir * powf(h1, 1.5f);
    qstruc = qstruc * m->structure_length[k];
    m->q[ip] = qstruc * (float)idir;
  }
}

/* ---- step driver: sfincs_lib.f90:337-683 ---------------------------------------------- */
/* Runs while t < tend; single step when max_steps == 1. Returns the number of steps. */
int32_t oracle_update(OM* m, double tend, int32_t max_steps) {
  int32_t steps = 0;
  while (m->t < tend) {
    if (max_steps > 0 && steps >= max_steps) break;
    m->nt = m->nt + 1;
    m->dt = m->alpha * m->min_dt;
    const float dtchk = m->alpha * m->min_dt;
    m->t = m->t + (double)m->dt;
    oracle_meteo_window(m);
    if (m->wind || m->patmos || m->precip) {
      oracle_update_meteo(m, m->t, m->dt);
      if (m->infiltration) oracle_update_infiltration(m, m->dt);   /* sfincs_lib.f90:534-540 */
    }
    oracle_update_boundaries(m, m->t, m->dt);
    oracle_update_discharges(m, m->t, m->dt);
    oracle_compute_fluxes(m, m->dt);
    if (m->nstruc > 0) oracle_compute_fluxes_over_structures(m);
    oracle_compute_water_levels(m, m->t, m->dt);
    ++steps;
    if (dtchk < m->dtmin && m->nt > 1) {
      m->error = 1;
      m->t = (double)m->t1 + 1.0;
    }
  }
  return steps;
}
