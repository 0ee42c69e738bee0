// Internal declarations shared by the host model builder and the device runtime.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sfx.h"

struct sfx_model;

namespace sfx {

void set_error(const std::string& msg);
// Writes a type-1 binary restart file (sfincs_output.f90:657-729) from state arrays in
// reference order; defined in sfx_model.cpp.
int write_restart_arrays(const char* path, const double* zs, int32_t np, const float* q, int32_t nq,
                         const float* uvmean, int32_t nuvmean);

// The gridded / spiderweb meteo fields of every update time (sfx_meteo.cpp); called by
// sfx_model_initialize after the Cd table is built.
int build_meteo(sfx_model* M);

// Where the subgrid files of a quadtree mesh keep each active cell and uv-point: the files
// hold every quadtree point and, per point, its right then its upper uv-points
// (sfincs_subgrid.F90:156-240, 530-600). zsrc[c] = the file record of active cell c,
// uvsrc[i] = the file record of active uv-point i, nuvq = the file's uv-point count.
// Returns false when the active uv-points do not all appear once. Defined in sfx_model.cpp.
bool quadtree_subgrid_order(const sfx_model* M, std::vector<int32_t>& zsrc, std::vector<int32_t>& uvsrc,
                            int32_t& nuvq);

// 1-based reference index -> 0-based storage slot.
inline int32_t z0(int32_t one_based) { return one_based - 1; }

}  // namespace sfx

// Host model: the reference's sfincs_data arrays after sfincs_initialize, in the
// reference's element order. Index tables hold 1-based values; storage is 0-based.
struct sfx_model {
  sfx_params p{};
  // derived switches (sfincs_input.f90:347-745)
  bool subgrid = false, use_quadtree = false, dense = false;
  bool coriolis = false, thetasmoothing = false, precip = false, wind = false;
  bool mesh_built = false, initialized = false;
  float fcorio = 0.f, rotation = 0.f, cosrot = 1.f, sinrot = 0.f;
  float x0 = 0.f, y0 = 0.f, dx = 0.f, dy = 0.f;
  int32_t nmax = 0, mmax = 0;        // regular: grid size; quadtree: level-1 extent
  int32_t np = 0, npuv = 0, ncuv = 0, nref = 1, nl = 0;
  int32_t qt_nmax = 0;                // quadtree_nmax used by point lookup

  // quadtree point list (for point lookup: level-major, nm sorted)
  std::vector<int8_t> qt_level;
  std::vector<int32_t> qt_nm, qt_first, qt_last, index_sfincs_in_quadtree;
  // quadtree-file right/above neighbour flags and points, for the legacy binary subgrid file
  std::vector<int8_t> qt_mu, qt_nu;
  std::vector<int32_t> qt_mu1, qt_mu2, qt_nu1, qt_nu2;

  // cells
  std::vector<int8_t> kcs, z_flags_iref;
  std::vector<int32_t> z_index_z_n, z_index_z_m;
  std::vector<int32_t> z_index_uv_md, z_index_uv_mu, z_index_uv_nd, z_index_uv_nu;
  std::vector<int32_t> z_index_uv_md1, z_index_uv_md2, z_index_uv_mu1, z_index_uv_mu2;
  std::vector<int32_t> z_index_uv_nd1, z_index_uv_nd2, z_index_uv_nu1, z_index_uv_nu2;
  std::vector<float> z_xz, z_yz, zb;

  // uv-points
  std::vector<int32_t> uv_index_z_nm, uv_index_z_nmu;
  std::vector<int32_t> uv_index_u_nmd, uv_index_u_nmu, uv_index_u_ndm, uv_index_u_num;
  std::vector<int32_t> uv_index_v_ndm, uv_index_v_nm, uv_index_v_ndmu, uv_index_v_nmu;
  std::vector<int8_t> uv_flags_iref, uv_flags_type, uv_flags_dir;
  std::vector<int8_t> kcuv;       // npuv + ncuv + 1
  std::vector<int8_t> mask_adv;   // npuv
  std::vector<int32_t> cuv_index_uv, cuv_index_uv1, cuv_index_uv2;
  std::vector<float> zbuv, zbuvmx, gn2uv;

  // refinement levels
  std::vector<float> dxr, dyr, dxyr, dxrm, dyrm, dxrinv, dyrinv, dxr2inv, dyr2inv;
  std::vector<float> dxrinvc, dyrinvc, cell_area, nuvisc;
  float dxymin = 0.f, dtmax = 0.f, dtmin = 0.f;

  // subgrid tables after sanitising (level fastest)
  std::vector<float> z_zmin, z_zmax, z_volmax, z_dep;
  std::vector<float> uv_zmin, uv_zmax, uv_navg_w, uv_fnfit, uv_havg, uv_nrep, uv_pwet;
  std::vector<float> uv_havg_zmax, uv_nrep_zmax;

  // initial state
  std::vector<double> zs, z_volume;
  std::vector<float> q, uv, zs0, zsderv, zsmax;
  std::vector<float> vmax, qmax, twet;   // compute_store_variables fields (used when switched on)
  std::vector<float> storage_volume;     // storage per subgrid cell (empty: none; volfile)
  std::vector<int8_t> kfuv;

  // boundaries
  int32_t ngbnd = 0, nkcuv2 = 0, nbnd = 0, ntbnd = 0;
  std::vector<int32_t> nmindbnd;
  std::vector<float> zsb, zsb0;
  std::vector<int32_t> index_kcuv2, nmbkcuv2, nmikcuv2, ibkcuv2;
  std::vector<int8_t> ibuvdir;
  std::vector<float> uvmean;
  // binary restart file read before initialisation (applied by sfx_model_initialize)
  int32_t rst_type = 0;                  // 0 = none, 1/2 = zs, q, uvmean, 3 = zs
  std::vector<float> rst_zs, rst_q, rst_uvmean;
  std::vector<float> x_bnd, y_bnd, t_bnd, zs_bnd;   // zs_bnd[it*nbnd + ib]
  std::vector<float> zsi_bnd;                       // bzi: infragravity levels, as zs_bnd (empty: none)
  std::vector<int32_t> ind1_bnd_gbp, ind2_bnd_gbp, nmi_gbp;
  std::vector<float> fac_bnd_gbp;
  // downstream river boundaries (bdr file, sfincs_boundaries.f90:217-285, 587-613)
  int32_t nbdr = 0;
  std::vector<float> x_bdr, y_bdr, x_bdr_in, y_bdr_in, dzs_bdr;
  std::vector<int32_t> index_zsi_bdr, index_bdr_gbp;   // 1-based cell / bdr point (0 = none)

  // point sources; nmindsrc holds nsrc + 2 ndrn cells (each drainage structure's intake
  // then outfall after the point sources, sfincs_discharges.f90:258-290)
  int32_t nsrc = 0, ntsrc = 0;
  std::vector<float> x_src, y_src;
  std::vector<int32_t> nmindsrc;
  // drainage structures (drnfile)
  int32_t ndrn = 0;
  std::vector<float> x_drn_in, y_drn_in, x_drn_out, y_drn_out, drainage_params, drainage_distance;
  std::vector<int32_t> drainage_type;
  std::vector<float> tsrc, qsrc;   // qsrc[it*nsrc + isrc]

  // thin dams and weirs (thdfile, weirfile; sfincs_structures.f90): polylines, vertex
  // count per polyline; weirs carry (crest height, discharge coefficient) per vertex
  std::vector<int32_t> thd_nv, weir_nv;
  std::vector<float> thd_x, thd_y, weir_x, weir_y, weir_par;
  // derived at initialize: the thin-dam uv-points (set to kcuv 0) and the weir uv-points
  // (kcuv 3) with crest height, coefficient and length relative to the grid spacing
  std::vector<int32_t> thindam_uv_index, structure_uv_index;   // 1-based
  std::vector<float> structure_height, structure_cd, structure_length;

  // gridded meteo (amu/amv/amp/ampr contents) and a spiderweb, as set; and at initialize the
  // fields of each update time T_k = t0 + k dtwnd (float sums, as twindupd) on the cells,
  // frame-major [nframes][np]: wind stress, pressure, rain (m/s)
  struct MeteoGrid {
    int32_t nt = 0, nrows = 0, ncols = 0;
    float xll = 0.f, yll = 0.f, dx = 0.f, dy = 0.f;
    std::vector<float> times, v;   // v[nt][nrows][ncols]
  };
  MeteoGrid amu, amv, amp, ampr;
  struct Spiderweb {
    int32_t nt = 0, nrows = 0, ncols = 0;
    float radius = 0.f;
    bool rain = false;
    std::vector<float> times, xe, ye, vmag, vdir, pdrop, prcp;
  };
  Spiderweb spw;
  bool meteo3d = false, patmos = false, meteo_wind = false, meteo_rain = false;
  int32_t meteo_nframes = 0;
  std::vector<float> meteo_times, meteo_tauwu, meteo_tauwv, meteo_patm, meteo_prcp;

  // infiltration ('con' from params.qinf, or 'c2d' from a per-cell map in mm/h): the rate
  // per cell in m/s after initialize (empty: no infiltration)
  std::vector<float> scs_inch;                  // curve-number retention S per cell, inches ('cna')
  std::vector<float> scs_m;                     // S in m when curve-number infiltration is on ('cna')
  std::vector<float> qinf_map_mmh, qinffield;

  // meteo series
  int32_t ntprcp = 0, ntwnd = 0;
  std::vector<float> tprcpt, tprcpv, twnd, wndmag, wnddir;
  std::vector<float> cdval;   // 1000 entries
  std::vector<float> x73;     // h**(7/3) table, 1000 x 10 column-major

  // helpers
  int32_t dummy_uv() const { return npuv + ncuv + 1; }
  int32_t find_cell_by_index(int32_t n, int32_t m, int32_t iref) const;   // sfincs cell (1-based) or 0
  int32_t find_cell_xy(float x, float y) const;                          // sfincs cell (1-based) or 0
};
