// Host-side model builder: the init-time part of sfincs_initialize that produces the
// tables the time step reads. Everything here runs once on the CPU; the resulting arrays
// are moved to HBM by sfx_create. Element orders and integer conventions are the
// reference's so that every index table and mask is bit-exact with it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "sfx_internal.hpp"

namespace sfx {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace sfx

extern "C" int sfx_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return SFX_ERR_ARG;
  std::strncpy(buf, sfx::g_last_error.c_str(), n - 1);
  buf[n - 1] = 0;
  return SFX_OK;
}

// Defaults of read_sfincs_input (sfincs_input.f90:50-343) and constants (:372-373).
extern "C" void sfx_params_default(sfx_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->t0 = 0.f; p->t1 = 0.f; p->tspinup = 0.f;
  p->alpha = 0.5f; p->theta = 1.0f; p->hmin_cfl = 0.1f; p->manning = 0.04f;
  p->manning_land = -999.0f; p->manning_sea = -999.0f; p->rgh_lev_land = 0.0f;
  p->zsini = 0.0f; p->dtmax = 60.0f; p->huthresh = 0.05f; p->huvmin = 0.0f;
  p->rhoa = 1.25f; p->rhow = 1024.0f; p->advlim = 1.0f; p->slopelim = 9999.9f;
  p->btfilter = 60.0f; p->btrelax = 3600.0f; p->wiggle_factor = 0.1f; p->wiggle_threshold = 0.1f;
  p->uvlim = 10.0f; p->uvmax = 1000.0f; p->nuvisc = 0.01f; p->nuviscfac = 100.0f;
  p->latitude = 0.0f; p->g = 9.81f; p->pi = 3.14159f;
  p->advection = 1; p->advection_scheme = 1; p->viscosity = 0; p->coriolis = 1;
  p->friction2d = 1; p->advection_mask = 1; p->wiggle_suppression = 1; p->store_zsmax = 1;
  p->store_velmax = 0; p->store_fluxmax = 0; p->store_twet = 0; p->twet_threshold = 0.01f;
  p->timestep_analysis = 0; p->use_qext = 0;
  p->structure_relax = 10.0f;
  p->qinf = 0.0f; p->qinf_zmin = 0.0f;
  p->dtwnd = 1800.0f; p->baro = 1; p->pavbnd = 0.0f; p->gapres = 101200.0f; p->spinup_meteo = 0;
  p->amprblock = 1; p->usespwprecip = 1; p->spw_merge_frac = 0.5f;
  p->factor_wind = 1.0f; p->factor_pres = 1.0f; p->factor_prcp = 1.0f; p->factor_spw_size = 1.0f;
  p->sfacinf = 0.2f;   // sfincs_input.f90:102
  p->bndtype = 1; p->h73table = 0;
}

namespace {

using sfx::set_error;

// Derived switches of read_sfincs_input (sfincs_input.f90:369-419, 611-619).
void derive_switches(sfx_model* M) {
  const sfx_params& p = M->p;
  M->rotation = M->rotation;  // set by the mesh constructor (radians)
  M->coriolis = p.coriolis != 0;
  M->fcorio = 2 * 7.2921e-05f * std::sin(p.latitude * p.pi / 180);
  if (p.latitude < 0.01f && p.latitude > -0.01f) M->coriolis = false;
  M->thetasmoothing = p.theta < 0.9999f;
}

bool need_uv_neighbours(const sfx_model* M) {
  const sfx_params& p = M->p;
  return p.advection || M->coriolis || p.viscosity || p.friction2d || M->thetasmoothing;
}

// Quadtree description in file semantics (what quadtree_read_file or
// make_quadtree_from_indices leave behind).
struct QT {
  int32_t npts = 0, nlev = 1, nmax = 0, mmax = 0;
  float x0 = 0, y0 = 0, dx = 0, dy = 0, rot = 0;
  std::vector<int8_t> level, md, mu, nd, nu, mask;
  std::vector<int32_t> n, m, md1, md2, mu1, mu2, nd1, nd2, nu1, nu2;
};

// make_quadtree_from_indices (sfincs_quadtree.F90:445-669), non-global branch.
void regular_to_qt(int32_t nmax, int32_t mmax, const int8_t* msk, QT& q) {
  std::vector<int32_t> indices;
  for (int32_t m = 1; m <= mmax; ++m)
    for (int32_t n = 1; n <= nmax; ++n)
      if (msk[(size_t)(m - 1) * nmax + (n - 1)] > 0) indices.push_back((m - 1) * nmax + n);
  const int32_t np = (int32_t)indices.size();
  std::vector<int32_t> g((size_t)nmax * mmax, 0);   // index_g_nm(n, m)
  q.npts = np; q.nlev = 1; q.nmax = nmax; q.mmax = mmax;
  q.level.assign(np, 1); q.md.assign(np, 0); q.mu.assign(np, 0); q.nd.assign(np, 0); q.nu.assign(np, 0);
  q.n.resize(np); q.m.resize(np);
  q.md1.assign(np, 0); q.md2.assign(np, 0); q.mu1.assign(np, 0); q.mu2.assign(np, 0);
  q.nd1.assign(np, 0); q.nd2.assign(np, 0); q.nu1.assign(np, 0); q.nu2.assign(np, 0);
  q.mask.resize(np);
  for (int32_t k = 0; k < np; ++k) {
    int32_t m = (indices[k] - 1) / nmax + 1;
    int32_t n = indices[k] - (m - 1) * nmax;
    q.n[k] = n; q.m[k] = m;
    g[(size_t)(m - 1) * nmax + (n - 1)] = k + 1;
    q.mask[k] = msk[(size_t)(m - 1) * nmax + (n - 1)];
  }
  auto G = [&](int32_t n, int32_t m) { return g[(size_t)(m - 1) * nmax + (n - 1)]; };
  for (int32_t k = 0; k < np; ++k) {
    const int32_t n = q.n[k], m = q.m[k];
    if (m < mmax) q.mu1[k] = G(n, m + 1);
    if (m > 1) q.md1[k] = G(n, m - 1);
    if (n < nmax) q.nu1[k] = G(n + 1, m);
    if (n > 1) q.nd1[k] = G(n - 1, m);
  }
}

// initialize_mesh (sfincs_domain.f90:143-1586) for projected coordinates.
int build_mesh(sfx_model* M, const QT& Q) {
  const int32_t nq = Q.npts;
  M->nref = Q.nlev;
  // point lookup arrays (sfincs_quadtree.F90:118-186)
  M->qt_level = Q.level;
  M->qt_nmax = 0;
  for (int32_t k = 0; k < nq; ++k) {
    int32_t iref = Q.level[k];
    int32_t p2 = 1 << (iref - 1);
    M->qt_nmax = std::max(M->qt_nmax, (int32_t)((1.0f * (Q.n[k] - 1) + 0.01f) / p2) + 1);
  }
  if (!M->use_quadtree) M->qt_nmax = Q.nmax;
  M->qt_nm.resize(nq);
  M->qt_first.assign(M->nref + 1, 0);
  M->qt_last.assign(M->nref + 1, 0);
  int32_t ireflast = 0;
  for (int32_t k = 0; k < nq; ++k) {
    int32_t iref = Q.level[k];
    int32_t nmx = M->qt_nmax * (1 << (iref - 1));
    M->qt_nm[k] = (Q.m[k] - 1) * nmx + Q.n[k];
    if (iref > ireflast) { M->qt_first[iref] = k + 1; ireflast = iref; }
    M->qt_last[iref] = k + 1;
  }

  // compaction of msk > 0 points (:327-482)
  M->index_sfincs_in_quadtree.assign(nq, 0);
  std::vector<int32_t> q_of_s;
  for (int32_t k = 0; k < nq; ++k)
    if (Q.mask[k] > 0) { q_of_s.push_back(k + 1); M->index_sfincs_in_quadtree[k] = (int32_t)q_of_s.size(); }
  const int32_t np = (int32_t)q_of_s.size();
  if (np == 0) { set_error("mesh has no active cells"); return SFX_ERR_ARG; }
  M->np = np;
  auto S = [&](int32_t qidx) { return qidx > 0 ? M->index_sfincs_in_quadtree[qidx - 1] : 0; };
  std::vector<int8_t> fmd(np), fmu(np), fnd(np), fnu(np);
  std::vector<int32_t> md1(np), md2(np), mu1(np), mu2(np), nd1(np), nd2(np), nu1(np), nu2(np);
  M->kcs.resize(np); M->z_flags_iref.resize(np); M->z_index_z_n.resize(np); M->z_index_z_m.resize(np);
  for (int32_t ip = 0; ip < np; ++ip) {
    const int32_t iq = q_of_s[ip] - 1;
    M->kcs[ip] = Q.mask[iq];
    M->z_flags_iref[ip] = Q.level[iq];
    M->z_index_z_n[ip] = Q.n[iq];
    M->z_index_z_m[ip] = Q.m[iq];
    fnu[ip] = Q.nu[iq]; fnd[ip] = Q.nd[iq]; fmu[ip] = Q.mu[iq]; fmd[ip] = Q.md[iq];
    nu1[ip] = S(Q.nu1[iq]); nu2[ip] = S(Q.nu2[iq]); nd1[ip] = S(Q.nd1[iq]); nd2[ip] = S(Q.nd2[iq]);
    mu1[ip] = S(Q.mu1[iq]); mu2[ip] = S(Q.mu2[iq]); md1[ip] = S(Q.md1[iq]); md2[ip] = S(Q.md2[iq]);
  }

  // count u and v points (:486-533)
  int32_t npu = 0, npv = 0;
  for (int32_t nm = 0; nm < np; ++nm) {
    if (fmu[nm] < 1) { if (mu1[nm] > 0) ++npu; }
    else { if (mu1[nm] > 0) ++npu; if (mu2[nm] > 0) ++npu; }
    if (fnu[nm] < 1) { if (nu1[nm] > 0) ++npv; }
    else { if (nu1[nm] > 0) ++npv; if (nu2[nm] > 0) ++npv; }
  }
  const int32_t npuv = npu + npv;
  // combined uv points at refinement transitions (:535-596)
  int32_t ncuv = 0;
  if (Q.nlev > 1) {
    for (int32_t nm = 0; nm < np; ++nm) {
      if (fmd[nm] == 1 && (md1[nm] > 0 || md2[nm] > 0)) ++ncuv;
      if (fmu[nm] == 1 && (mu1[nm] > 0 || mu2[nm] > 0)) ++ncuv;
      if (fnd[nm] == 1 && (nd1[nm] > 0 || nd2[nm] > 0)) ++ncuv;
      if (fnu[nm] == 1 && (nu1[nm] > 0 || nu2[nm] > 0)) ++ncuv;
    }
  }
  M->npuv = npuv; M->ncuv = ncuv;
  const int32_t dummy = npuv + ncuv + 1;
  M->z_index_uv_md.assign(np, dummy); M->z_index_uv_mu.assign(np, dummy);
  M->z_index_uv_nd.assign(np, dummy); M->z_index_uv_nu.assign(np, dummy);
  M->z_index_uv_md1.assign(np, 0); M->z_index_uv_md2.assign(np, 0);
  M->z_index_uv_mu1.assign(np, 0); M->z_index_uv_mu2.assign(np, 0);
  M->z_index_uv_nd1.assign(np, 0); M->z_index_uv_nd2.assign(np, 0);
  M->z_index_uv_nu1.assign(np, 0); M->z_index_uv_nu2.assign(np, 0);
  M->cuv_index_uv.assign(ncuv, 0);
  M->cuv_index_uv1.assign(ncuv, dummy); M->cuv_index_uv2.assign(ncuv, dummy);
  std::vector<int32_t> cmd(np, 0), cmu(np, 0), cnd(np, 0), cnu(np, 0);
  {
    int32_t irefuv = 0;
    for (int32_t nm = 0; nm < np; ++nm) {
      if (fmd[nm] == 1 && (md1[nm] > 0 || md2[nm] > 0)) { ++irefuv; M->cuv_index_uv[irefuv - 1] = npuv + irefuv; cmd[nm] = irefuv; }
      if (fmu[nm] == 1 && (mu1[nm] > 0 || mu2[nm] > 0)) { ++irefuv; M->cuv_index_uv[irefuv - 1] = npuv + irefuv; cmu[nm] = irefuv; }
      if (fnd[nm] == 1 && (nd1[nm] > 0 || nd2[nm] > 0)) { ++irefuv; M->cuv_index_uv[irefuv - 1] = npuv + irefuv; cnd[nm] = irefuv; }
      if (fnu[nm] == 1 && (nu1[nm] > 0 || nu2[nm] > 0)) { ++irefuv; M->cuv_index_uv[irefuv - 1] = npuv + irefuv; cnu[nm] = irefuv; }
    }
  }
  // uv points (:680-1003); all values below are 1-based reference indices
  M->kcuv.assign(dummy, 0);
  M->uv_index_z_nm.assign(npuv, 0); M->uv_index_z_nmu.assign(npuv, 0);
  M->uv_flags_iref.assign(npuv, 0); M->uv_flags_dir.assign(npuv, 0); M->uv_flags_type.assign(npuv, 0);
  auto& zmd = M->z_index_uv_md; auto& zmu = M->z_index_uv_mu; auto& znd = M->z_index_uv_nd; auto& znu = M->z_index_uv_nu;
  auto& zmd1 = M->z_index_uv_md1; auto& zmd2 = M->z_index_uv_md2; auto& zmu1 = M->z_index_uv_mu1; auto& zmu2 = M->z_index_uv_mu2;
  auto& znd1 = M->z_index_uv_nd1; auto& znd2 = M->z_index_uv_nd2; auto& znu1 = M->z_index_uv_nu1; auto& znu2 = M->z_index_uv_nu2;
  auto& cuv1 = M->cuv_index_uv1; auto& cuv2 = M->cuv_index_uv2; auto& cuv = M->cuv_index_uv;
  int32_t ip = 0;
  auto set_uv = [&](int32_t nmc, int32_t nbr, int dir, int type, int iref) {
    ++ip;
    M->uv_index_z_nm[ip - 1] = nmc; M->uv_index_z_nmu[ip - 1] = nbr;
    M->uv_flags_iref[ip - 1] = (int8_t)iref; M->uv_flags_dir[ip - 1] = (int8_t)dir; M->uv_flags_type[ip - 1] = (int8_t)type;
  };
  for (int32_t c = 0; c < np; ++c) {
    const int32_t nm = c + 1;
    const int iref = M->z_flags_iref[c];
    // right
    if (fmu[c] == 0) {
      if (mu1[c] > 0) {
        const int32_t nmu = mu1[c];
        set_uv(nm, nmu, 0, 0, iref);
        zmu[c] = ip; zmd[nmu - 1] = ip; zmu1[c] = ip; zmd1[nmu - 1] = ip;
      }
    } else if (fmu[c] == -1) {
      if (mu1[c] > 0) {
        const int32_t nmu = mu1[c];
        set_uv(nm, nmu, 0, -1, iref);
        zmu[c] = ip; zmu1[c] = ip;
        const int32_t icuv = cmd[nmu - 1];
        if (md1[mu1[c] - 1] == nm) { cuv1[icuv - 1] = ip; zmd1[nmu - 1] = ip; }
        else { cuv2[icuv - 1] = ip; zmd2[nmu - 1] = ip; }
        zmd[nmu - 1] = cuv[icuv - 1];
      }
    } else {
      if (mu1[c] > 0) {
        const int32_t nmu = mu1[c];
        set_uv(nm, nmu, 0, 1, iref + 1);
        zmu1[c] = ip; zmd[nmu - 1] = ip; zmd1[nmu - 1] = ip;
        const int32_t icuv = cmu[c];
        zmu[c] = cuv[icuv - 1];
        cuv1[icuv - 1] = ip;
      }
      if (mu2[c] > 0) {
        const int32_t nmu = mu2[c];
        set_uv(nm, nmu, 0, 1, iref + 1);
        zmu2[c] = ip; zmd[nmu - 1] = ip; zmd1[nmu - 1] = ip;
        const int32_t icuv = cmu[c];
        zmu[c] = cuv[icuv - 1];
        cuv2[icuv - 1] = ip;
      }
    }
    // above
    if (fnu[c] == 0) {
      if (nu1[c] > 0) {
        const int32_t num = nu1[c];
        set_uv(nm, num, 1, 0, iref);
        znu[c] = ip; znd[num - 1] = ip; znu1[c] = ip; znd1[num - 1] = ip;
      }
    } else if (fnu[c] == -1) {
      if (nu1[c] > 0) {
        const int32_t num = nu1[c];
        set_uv(nm, num, 1, -1, iref);
        znu[c] = ip; znu1[c] = ip;
        const int32_t icuv = cnd[num - 1];
        if (nd1[nu1[c] - 1] == nm) { cuv1[icuv - 1] = ip; znd1[num - 1] = ip; }
        else { cuv2[icuv - 1] = ip; znd2[num - 1] = ip; }
        znd[num - 1] = cuv[icuv - 1];
      }
    } else {
      if (nu1[c] > 0) {
        const int32_t num = nu1[c];
        set_uv(nm, num, 1, 1, iref + 1);
        znd[num - 1] = ip; znu1[c] = ip; znd1[num - 1] = ip;
        const int32_t icuv = cnu[c];
        znu[c] = cuv[icuv - 1];
        cuv1[icuv - 1] = ip;
      }
      if (nu2[c] > 0) {
        const int32_t num = nu2[c];
        set_uv(nm, num, 1, 1, iref + 1);
        znd[num - 1] = ip; znu2[c] = ip; znd1[num - 1] = ip;
        const int32_t icuv = cnu[c];
        znu[c] = cuv[icuv - 1];
        cuv2[icuv - 1] = ip;
      }
    }
  }
  if (ip != npuv) { set_error("inconsistent quadtree neighbour description"); return SFX_ERR_ARG; }

  // eight uv neighbours of every uv point (:1005-1261)
  M->uv_index_u_nmd.assign(npuv, 0); M->uv_index_u_nmu.assign(npuv, 0);
  M->uv_index_u_ndm.assign(npuv, 0); M->uv_index_u_num.assign(npuv, 0);
  M->uv_index_v_ndm.assign(npuv, 0); M->uv_index_v_nm.assign(npuv, 0);
  M->uv_index_v_ndmu.assign(npuv, 0); M->uv_index_v_nmu.assign(npuv, 0);
  if (need_uv_neighbours(M)) {
    for (int32_t i = 0; i < npuv; ++i) {
      const int32_t ipu = i + 1;
      if (M->uv_flags_dir[i] == 0) {
        const int32_t nm = M->uv_index_z_nm[i], nmu = M->uv_index_z_nmu[i];
        M->uv_index_u_nmd[i] = zmd[nm - 1];
        M->uv_index_u_nmu[i] = zmu[nmu - 1];
        if (fnd[nm - 1] > -1) {
          const int32_t ndm = nd1[nm - 1];
          if (ndm > 0) M->uv_index_u_ndm[i] = zmu[ndm - 1];
        } else {
          const int32_t ndm = nd2[nm - 1];
          if (ndm > 0) {
            if (fmu[ndm - 1] == 0) M->uv_index_u_ndm[i] = zmu[ndm - 1];
            else {
              const int32_t ndmu = mu1[ndm - 1];
              if (ndmu > 0) M->uv_index_u_ndm[i] = zmd[ndmu - 1];
            }
          }
        }
        if (fnu[nm - 1] > -1) {
          const int32_t num = nu1[nm - 1];
          if (num > 0) M->uv_index_u_num[i] = zmu[num - 1];
        } else {
          const int32_t num = nu2[nm - 1];
          if (num > 0) {
            if (fmu[num - 1] == 0) M->uv_index_u_num[i] = zmu[num - 1];
            else {
              const int32_t numu = mu1[num - 1];
              if (numu > 0) M->uv_index_u_num[i] = zmd[numu - 1];
            }
          }
        }
        M->uv_index_v_nm[i] = znu[nm - 1];
        M->uv_index_v_ndm[i] = znd[nm - 1];
        M->uv_index_v_nmu[i] = znu[nmu - 1];
        M->uv_index_v_ndmu[i] = znd[nmu - 1];
      } else {
        const int32_t nm = M->uv_index_z_nm[i], num = M->uv_index_z_nmu[i];
        M->uv_index_u_nmd[i] = znd[nm - 1];
        M->uv_index_u_nmu[i] = znu[num - 1];
        if (fmd[nm - 1] > -1) {
          const int32_t nmd = md1[nm - 1];
          if (nmd > 0) M->uv_index_u_ndm[i] = znu[nmd - 1];
        } else {
          const int32_t nmd = md2[nm - 1];
          if (nmd > 0) {
            if (fnu[nmd - 1] == 0) M->uv_index_u_ndm[i] = znu[nmd - 1];
            else {
              const int32_t nmdu = nu1[nmd - 1];
              if (nmdu > 0) M->uv_index_u_ndm[i] = znd[nmdu - 1];
            }
          }
        }
        if (fmu[nm - 1] > -1) {
          const int32_t nmu = mu1[nm - 1];
          if (nmu > 0) M->uv_index_u_num[i] = znu[nmu - 1];
        } else {
          const int32_t nmu = mu2[nm - 1];
          if (nmu > 0) {
            if (fnu[nmu - 1] == 0) M->uv_index_u_num[i] = znu[nmu - 1];
            else {
              const int32_t numu = nu1[nmu - 1];
              if (numu > 0) M->uv_index_u_num[i] = znd[numu - 1];
            }
          }
        }
        M->uv_index_v_nm[i] = zmu[nm - 1];
        M->uv_index_v_ndm[i] = zmd[nm - 1];
        M->uv_index_v_nmu[i] = zmu[num - 1];
        M->uv_index_v_ndmu[i] = zmd[num - 1];
      }
      if (M->uv_index_u_nmd[i] == 0) M->uv_index_u_nmd[i] = ipu;
      if (M->uv_index_u_nmu[i] == 0) M->uv_index_u_nmu[i] = ipu;
      if (M->uv_index_u_ndm[i] == 0) M->uv_index_u_ndm[i] = ipu;
      if (M->uv_index_u_num[i] == 0) M->uv_index_u_num[i] = ipu;
      if (M->uv_index_v_ndm[i] == 0) M->uv_index_v_ndm[i] = dummy;
      if (M->uv_index_v_ndmu[i] == 0) M->uv_index_v_ndmu[i] = dummy;
      if (M->uv_index_v_nm[i] == 0) M->uv_index_v_nm[i] = dummy;
      if (M->uv_index_v_nmu[i] == 0) M->uv_index_v_nmu[i] = dummy;
    }
  }

  // geometry per refinement level, projected coordinates (:1369-1541)
  const int nref = M->nref;
  M->dxr.resize(nref); M->dyr.resize(nref); M->dxyr.resize(nref);
  for (int iref = 1; iref <= nref; ++iref) {
    M->dxr[iref - 1] = M->dx / (float)(1 << (iref - 1));
    M->dyr[iref - 1] = M->dy / (float)(1 << (iref - 1));
    M->dxyr[iref - 1] = std::min(M->dxr[iref - 1], M->dyr[iref - 1]);
  }
  M->z_xz.resize(np); M->z_yz.resize(np);
  for (int32_t c = 0; c < np; ++c) {
    const int32_t n = M->z_index_z_n[c], m = M->z_index_z_m[c];
    const int iref = M->z_flags_iref[c];
    M->z_xz[c] = M->x0 + M->cosrot * (1.0f * ((float)m - 0.5f)) * M->dxr[iref - 1]
               - M->sinrot * (1.0f * ((float)n - 0.5f)) * M->dyr[iref - 1];
    M->z_yz[c] = M->y0 + M->sinrot * (1.0f * ((float)m - 0.5f)) * M->dxr[iref - 1]
               + M->cosrot * (1.0f * ((float)n - 0.5f)) * M->dyr[iref - 1];
  }
  M->dxymin = 1.0e6f;
  M->dxrm.resize(nref); M->dxrinv.resize(nref); M->dxr2inv.resize(nref); M->dxrinvc.assign(nref, 0.f);
  M->dyrm.resize(nref); M->dyrinv.resize(nref); M->dyr2inv.resize(nref); M->dyrinvc.assign(nref, 0.f);
  M->cell_area.resize(nref);
  for (int i = 0; i < nref; ++i) {
    M->dxrm[i] = M->dxr[i];
    M->dxrinv[i] = 1.0f / M->dxrm[i];
    M->dxr2inv[i] = M->dxrinv[i] * M->dxrinv[i];
    M->dyrm[i] = M->dyr[i];
    M->dyrinv[i] = 1.0f / M->dyrm[i];
    M->dyr2inv[i] = M->dyrinv[i] * M->dyrinv[i];
    M->cell_area[i] = M->dxrm[i] * M->dyrm[i];
    if (i + 1 > 1) {
      M->dxrinvc[i] = 1.0f / ((3 * M->dxrm[i]) / 2);
      M->dyrinvc[i] = 1.0f / ((3 * M->dyrm[i]) / 2);
    }
    M->dxymin = std::min(M->dxymin, M->dxr[i]);
    M->dxymin = std::min(M->dxymin, M->dyr[i]);
  }
  M->dtmax = std::min(M->p.dtmax, M->p.alpha * M->dxymin / std::sqrt(9.81f * M->p.hmin_cfl));
  M->dtmin = M->p.alpha * M->dxymin / (1.25f * std::fabs(M->p.uvmax));
  M->nuvisc.assign(nref, 0.f);
  if (M->p.viscosity)
    for (int i = 0; i < nref; ++i) M->nuvisc[i] = std::max(M->p.nuvisc * M->dxyr[i], 0.0f);
  M->mesh_built = true;
  return SFX_OK;
}

// initialize_mesh tail (:42-50): a single-level quadtree reverts to the regular path.
void finish_mesh_flags(sfx_model* M) {
  if (M->nref == 1) M->use_quadtree = false;
  // Full rectangle with no holes: the device may use implicit neighbours.
  M->dense = (!M->use_quadtree && M->nmax > 0 && (int64_t)M->np == (int64_t)M->nmax * M->mmax);
}

}  // namespace

int32_t sfx_model::find_cell_by_index(int32_t n, int32_t m, int32_t iref) const {
  // find_quadtree_cell_by_index (sfincs_quadtree.F90:822-845) + remap to sfincs cells
  if (iref < 1 || iref > nref || qt_first[iref] == 0) return 0;
  const int32_t nmx = qt_nmax * (1 << (iref - 1));
  const int32_t nm = (m - 1) * nmx + n;
  if (nm <= 0) return 0;
  const int32_t i1 = qt_first[iref], i2 = qt_last[iref];
  auto b = qt_nm.begin() + (i1 - 1), e = qt_nm.begin() + i2;
  auto it = std::lower_bound(b, e, nm);
  if (it == e || *it != nm) return 0;
  const int32_t k = (int32_t)(it - qt_nm.begin());
  if (qt_level[k] != iref) return 0;
  return index_sfincs_in_quadtree[k];
}

int32_t sfx_model::find_cell_xy(float x, float y) const {
  // find_quadtree_cell (sfincs_quadtree.F90:764-820)
  const float xt = cosrot * (x - x0) + sinrot * (y - y0);
  const float yt = -sinrot * (x - x0) + cosrot * (y - y0);
  for (int iref = nref; iref >= 1; --iref) {
    const int32_t n = (int32_t)(yt / dyr[iref - 1]) + 1;
    const int32_t m = (int32_t)(xt / dxr[iref - 1]) + 1;
    if (n < 1 || m < 1) continue;
    const int32_t c = find_cell_by_index(n, m, iref);
    if (c > 0) return c;
  }
  return 0;
}

extern "C" int sfx_model_create_regular(const sfx_params* p, int32_t nmax, int32_t mmax, float x0,
                                        float y0, float dx, float dy, float rotation_deg,
                                        const int8_t* msk, sfx_model** out) {
  if (!p || !msk || !out || nmax < 1 || mmax < 1 || !(dx > 0) || !(dy > 0)) {
    set_error("sfx_model_create_regular: invalid arguments");
    return SFX_ERR_ARG;
  }
  sfx_model* M = new sfx_model();
  M->p = *p;
  M->nmax = nmax; M->mmax = mmax; M->x0 = x0; M->y0 = y0; M->dx = dx; M->dy = dy;
  M->rotation = rotation_deg * p->pi / 180;   // sfincs_input.f90:378
  M->cosrot = std::cos(M->rotation); M->sinrot = std::sin(M->rotation);
  derive_switches(M);
  M->use_quadtree = false;
  QT Q;
  regular_to_qt(nmax, mmax, msk, Q);
  int rc = build_mesh(M, Q);
  if (rc != SFX_OK) { delete M; return rc; }
  finish_mesh_flags(M);
  *out = M;
  return SFX_OK;
}

extern "C" int sfx_model_create_quadtree(const sfx_params* p, const sfx_quadtree_input* qt, sfx_model** out) {
  if (!p || !qt || !out || qt->nr_points < 1 || qt->nr_levels < 1 || !(qt->dx > 0) || !(qt->dy > 0)) {
    set_error("sfx_model_create_quadtree: invalid arguments");
    return SFX_ERR_ARG;
  }
  const int32_t n = qt->nr_points;
  QT Q;
  Q.npts = n; Q.nlev = qt->nr_levels; Q.x0 = qt->x0; Q.y0 = qt->y0; Q.dx = qt->dx; Q.dy = qt->dy;
  auto cp8 = [&](const int8_t* s, std::vector<int8_t>& d) { d.assign(s, s + n); };
  auto cp32 = [&](const int32_t* s, std::vector<int32_t>& d) { d.assign(s, s + n); };
  cp8(qt->level, Q.level); cp8(qt->md, Q.md); cp8(qt->mu, Q.mu); cp8(qt->nd, Q.nd); cp8(qt->nu, Q.nu);
  cp8(qt->mask, Q.mask);
  cp32(qt->n, Q.n); cp32(qt->m, Q.m); cp32(qt->md1, Q.md1); cp32(qt->md2, Q.md2); cp32(qt->mu1, Q.mu1);
  cp32(qt->mu2, Q.mu2); cp32(qt->nd1, Q.nd1); cp32(qt->nd2, Q.nd2); cp32(qt->nu1, Q.nu1); cp32(qt->nu2, Q.nu2);
  for (int32_t k = 0; k < n; ++k) {
    if (Q.level[k] < 1 || Q.level[k] > Q.nlev || (k > 0 && Q.level[k] < Q.level[k - 1])) {
      set_error("quadtree points must be sorted by level with levels in 1..nr_levels");
      return SFX_ERR_ARG;
    }
    const int32_t nb[8] = {Q.md1[k], Q.md2[k], Q.mu1[k], Q.mu2[k], Q.nd1[k], Q.nd2[k], Q.nu1[k], Q.nu2[k]};
    for (int j = 0; j < 8; ++j)
      if (nb[j] < 0 || nb[j] > n) { set_error("quadtree neighbour index out of range"); return SFX_ERR_ARG; }
  }
  sfx_model* M = new sfx_model();
  M->p = *p;
  M->x0 = qt->x0; M->y0 = qt->y0; M->dx = qt->dx; M->dy = qt->dy;
  // quadtree_read_file converts the rotation with its own pi (sfincs_quadtree.F90:81,99)
  M->rotation = qt->rotation_deg * 3.141592653589793f / 180;
  M->cosrot = std::cos(M->rotation); M->sinrot = std::sin(M->rotation);
  derive_switches(M);
  M->use_quadtree = true;
  int rc = build_mesh(M, Q);
  if (rc != SFX_OK) { delete M; return rc; }
  M->qt_mu = Q.mu; M->qt_nu = Q.nu; M->qt_mu1 = Q.mu1; M->qt_mu2 = Q.mu2; M->qt_nu1 = Q.nu1; M->qt_nu2 = Q.nu2;
  M->nmax = 0; M->mmax = 0;
  for (int32_t k = 0; k < n; ++k) {
    if (Q.level[k] != 1) continue;
    M->nmax = std::max(M->nmax, Q.n[k]); M->mmax = std::max(M->mmax, Q.m[k]);
  }
  finish_mesh_flags(M);
  *out = M;
  return SFX_OK;
}

extern "C" void sfx_model_destroy(sfx_model* m) { delete m; }

extern "C" int sfx_model_sizes(const sfx_model* M, sfx_sizes* s) {
  if (!M || !s) { set_error("sfx_model_sizes: null"); return SFX_ERR_ARG; }
  std::memset(s, 0, sizeof(*s));
  s->np = M->np; s->npuv = M->npuv; s->ncuv = M->ncuv; s->nref = M->nref; s->nlevels = M->nl;
  s->nmax = M->nmax; s->mmax = M->mmax; s->ngbnd = M->ngbnd; s->nkcuv2 = M->nkcuv2;
  s->nsrc = M->nsrc; s->nbnd = M->nbnd;
  s->subgrid = M->subgrid; s->quadtree = M->use_quadtree; s->dense = M->dense;
  s->dtmax = M->dtmax; s->dtmin = M->dtmin;
  s->ndrn = M->ndrn;
  s->nstruc = (int32_t)M->structure_uv_index.size();
  s->nthd = (int32_t)M->thindam_uv_index.size();
  return SFX_OK;
}

// initialize_bathymetry, regular branch (sfincs_domain.f90:1608-1710)
extern "C" int sfx_model_set_bathymetry(sfx_model* M, const float* zb) {
  if (!M || !zb || !M->mesh_built || M->initialized) { set_error("sfx_model_set_bathymetry: bad state"); return SFX_ERR_STATE; }
  M->subgrid = false;
  M->zb.assign(zb, zb + M->np);
  M->zbuv.resize(M->npuv); M->zbuvmx.resize(M->npuv);
  for (int32_t i = 0; i < M->npuv; ++i) {
    const float a = M->zb[M->uv_index_z_nm[i] - 1], b = M->zb[M->uv_index_z_nmu[i] - 1];
    if (M->uv_flags_type[i] == 0) M->zbuv[i] = 0.5f * (a + b);
    else if (M->uv_flags_type[i] == -1) M->zbuv[i] = 0.3333f * a + 0.6667f * b;
    else M->zbuv[i] = 0.6667f * a + 0.3333f * b;
    M->zbuvmx[i] = std::max(a, b) + M->p.huthresh;
  }
  return SFX_OK;
}

// read_subgrid_file_netcdf sanitising (sfincs_subgrid.F90:316-446) with huthresh = 0 (:40).
extern "C" int sfx_model_set_subgrid(sfx_model* M, const sfx_subgrid_input* s) {
  if (!M || !s || !M->mesh_built || M->initialized) { set_error("sfx_model_set_subgrid: bad state"); return SFX_ERR_STATE; }
  if (s->nlevels < 2) { set_error("subgrid tables need at least 2 levels"); return SFX_ERR_ARG; }
  const int32_t np = M->np, npuv = M->npuv, nl = s->nlevels;
  const float g = M->p.g;
  M->subgrid = true;
  M->p.huthresh = 0.0f;
  M->nl = nl;
  M->z_zmin.assign(s->z_zmin, s->z_zmin + np);
  M->z_zmax.assign(s->z_zmax, s->z_zmax + np);
  M->z_volmax.assign(s->z_volmax, s->z_volmax + np);
  M->z_dep.assign(s->z_level, s->z_level + (size_t)np * nl);
  M->uv_zmin.assign(s->uv_zmin, s->uv_zmin + npuv);
  M->uv_zmax.resize(npuv);
  for (int32_t i = 0; i < npuv; ++i) M->uv_zmax[i] = std::max(s->uv_zmax[i], M->uv_zmin[i] + 0.01f);
  M->uv_fnfit.assign(s->uv_ffit, s->uv_ffit + npuv);
  M->uv_navg_w.resize(npuv);
  for (int32_t i = 0; i < npuv; ++i) { float v = std::max(s->uv_navg[i], 0.0001f); M->uv_navg_w[i] = g * (v * v); }
  M->uv_havg.assign(s->uv_havg, s->uv_havg + (size_t)npuv * nl);
  M->uv_nrep.resize((size_t)npuv * nl);
  for (size_t k = 0; k < (size_t)npuv * nl; ++k) { float v = std::max(s->uv_nrep[k], 0.0001f); M->uv_nrep[k] = g * (v * v); }
  M->uv_pwet.assign(s->uv_pwet, s->uv_pwet + (size_t)npuv * nl);
  for (int32_t i = 0; i < npuv; ++i) {
    const int32_t nm = M->uv_index_z_nm[i] - 1, nmu = M->uv_index_z_nmu[i] - 1;
    if (M->z_zmin[nm] > M->uv_zmin[i]) {
      M->uv_zmin[i] = std::max(M->uv_zmin[i], M->z_zmin[nm]);
      M->uv_zmax[i] = std::max(M->uv_zmax[i], M->uv_zmin[i] + 0.01f);
    }
    if (M->z_zmin[nmu] > M->uv_zmin[i]) {
      M->uv_zmin[i] = std::max(M->uv_zmin[i], M->z_zmin[nmu]);
      M->uv_zmax[i] = std::max(M->uv_zmax[i], M->uv_zmin[i] + 0.01f);
    }
  }
  for (int32_t c = 0; c < np; ++c)
    if (M->z_zmax[c] - M->z_zmin[c] <= 0.0f) M->z_zmax[c] = M->z_zmax[c] + 0.01f;
  for (int32_t i = 0; i < npuv; ++i)
    if (M->uv_zmax[i] - M->uv_zmin[i] < 1.0e-7f) M->uv_zmax[i] = M->uv_zmax[i] + 0.01f;
  M->uv_havg_zmax.resize(npuv); M->uv_nrep_zmax.resize(npuv);
  for (int32_t i = 0; i < npuv; ++i) {
    M->uv_havg_zmax[i] = M->uv_havg[(size_t)i * nl + (nl - 1)] - M->uv_zmax[i];
    M->uv_nrep_zmax[i] = M->uv_nrep[(size_t)i * nl + (nl - 1)];
  }
  return SFX_OK;
}

bool sfx::quadtree_subgrid_order(const sfx_model* M, std::vector<int32_t>& zsrc, std::vector<int32_t>& uvsrc,
                                 int32_t& nuvq) {
  const int32_t np = M->np, npuv = M->npuv, nzq = (int32_t)M->index_sfincs_in_quadtree.size();
  zsrc.assign(np, -1);
  uvsrc.assign(npuv, -1);
  for (int32_t k = 0; k < nzq; ++k)
    if (M->index_sfincs_in_quadtree[k] > 0) zsrc[M->index_sfincs_in_quadtree[k] - 1] = k;
  // one record per neighbour, right then above; the active ones are the sfincs uv-points
  // in the order met
  int32_t nq = 0, ns = 0;
  auto take = [&](int32_t nm, const std::vector<int32_t>& idx) {
    ++nq;
    if (nm > 0 && idx[nm - 1] > 0 && ns < npuv) uvsrc[ns++] = nq - 1;
  };
  for (int32_t ip = 0; ip < nzq; ++ip) {
    const int32_t nm = M->index_sfincs_in_quadtree[ip];
    if (M->qt_mu1[ip] > 0) take(nm, M->z_index_uv_mu1);
    if (M->qt_mu[ip] >= 1 && M->qt_mu2[ip] > 0) take(nm, M->z_index_uv_mu2);
    if (M->qt_nu1[ip] > 0) take(nm, M->z_index_uv_nu1);
    if (M->qt_nu[ip] >= 1 && M->qt_nu2[ip] > 0) take(nm, M->z_index_uv_nu2);
  }
  nuvq = nq;
  if (ns != npuv) return false;
  for (int32_t c = 0; c < np; ++c)
    if (zsrc[c] < 0) return false;
  return true;
}

// read_subgrid_file_original (sfincs_subgrid.F90:458-903): the legacy binary subgrid file
// (stream of int32 headers and float32 records, no NetCDF). It holds hrep and navg per
// level: levels 2..nl come from the file, level 1 of havg is huthresh (which keeps its
// input value on this path, unlike the NetCDF path), nrep = g max(navg, nmin)^2 with level 1
// a copy of level 2, pwet = 1 and no fit (fnfit = 0). Regular meshes: records of np values
// per cell, the uv records per cell through its right (u) and upper (v) point. Quadtree
// meshes: records over every quadtree point / every quadtree uv-point, re-mapped to the
// active cells and uv-points.
extern "C" int sfx_model_read_subgrid_binary(sfx_model* M, const char* path) {
  if (!M || !path || !M->mesh_built || M->initialized) { set_error("sfx_model_read_subgrid_binary: bad state"); return SFX_ERR_STATE; }
  FILE* f = std::fopen(path, "rb");
  if (!f) { set_error(std::string("cannot open subgrid file ") + path); return SFX_ERR_IO; }
  struct Closer { FILE* f; ~Closer() { std::fclose(f); } } closer{f};
  bool ok = true;
  auto rd_i = [&](int32_t& v) { ok = ok && std::fread(&v, 4, 1, f) == 1; };
  auto rd_f = [&](std::vector<float>& v, size_t n) { v.resize(n); ok = ok && std::fread(v.data(), 4, n, f) == n; };
  const int32_t np = M->np, npuv = M->npuv;
  const float g = M->p.g, huth = M->p.huthresh;
  // regular: cells (not checked, as the reference does not), option, bins;
  // quadtree: version, quadtree points, quadtree uv-points, bins
  int32_t h0 = 0, h1 = 0, h2 = 0, nbins = 0, nzq = np, nuvq = 0;
  rd_i(h0); rd_i(h1); rd_i(h2);
  if (M->use_quadtree) { nzq = h1; nuvq = h2; rd_i(nbins); } else { nbins = h2; }
  if (!ok || nbins < 1 || nbins > 1000) { set_error(std::string("subgrid file ") + path + ": bad header"); return SFX_ERR_IO; }
  {   // the records must fill the file exactly
    const long hdr = M->use_quadtree ? 16 : 12;
    const long expect = M->use_quadtree
                            ? hdr + 4L * ((long)nzq * (4 + nbins) + (long)nuvq * (2 + 2 * nbins))
                            : hdr + 4L * (long)np * ((3 + nbins) + 2 * (3 + 2 * nbins));
    const long here = std::ftell(f);
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, here, SEEK_SET);
    if (size != expect) {
      set_error(std::string("subgrid file ") + path + ": " + std::to_string(size) + " bytes, the mesh needs " +
                std::to_string(expect));
      return SFX_ERR_IO;
    }
  }
  (void)h0;
  // source index of each active cell and uv-point in the file's records
  std::vector<int32_t> zsrc(np), uvsrc(npuv, -1);
  if (M->use_quadtree) {
    if (nzq != (int32_t)M->index_sfincs_in_quadtree.size()) {
      set_error(std::string("subgrid file ") + path + ": header does not match the quadtree");
      return SFX_ERR_IO;
    }
    int32_t nq = 0;
    if (!sfx::quadtree_subgrid_order(M, zsrc, uvsrc, nq) || nq != nuvq) {
      set_error(std::string("subgrid file ") + path + ": uv-points do not match the quadtree");
      return SFX_ERR_IO;
    }
  } else {
    for (int32_t c = 0; c < np; ++c) zsrc[c] = c;
  }
  const int32_t nlev = nbins + 1;
  M->subgrid = true;
  M->nl = nlev;
  M->z_zmin.assign(np, 0.f); M->z_zmax.assign(np, 0.f); M->z_volmax.assign(np, 0.f);
  M->z_dep.assign((size_t)np * nlev, 0.f);
  M->uv_zmin.assign(npuv, 0.f); M->uv_zmax.assign(npuv, 0.f);
  M->uv_havg.assign((size_t)npuv * nlev, huth);
  M->uv_nrep.assign((size_t)npuv * nlev, 0.f);
  M->uv_pwet.assign((size_t)npuv * nlev, 1.0f);
  M->uv_fnfit.assign(npuv, 0.0f);
  std::vector<float> r;
  auto cells = [&](std::vector<float>& dst, size_t stride, size_t off) {
    rd_f(r, nzq);
    if (!ok) return;
    for (int32_t c = 0; c < np; ++c) dst[(size_t)c * stride + off] = r[zsrc[c]];
  };
  cells(M->z_zmin, 1, 0);
  cells(M->z_zmax, 1, 0);
  if (M->use_quadtree) rd_f(r, nzq);   // zmean (not used)
  cells(M->z_volmax, 1, 0);
  for (int32_t c = 0; c < np; ++c) M->z_dep[(size_t)c * nlev] = std::max(M->z_zmin[c], -20.0f);
  for (int32_t l = 1; l < nlev; ++l) cells(M->z_dep, nlev, l);
  if (M->use_quadtree) {
    auto uvs = [&](std::vector<float>& dst, size_t stride, size_t off, int kind) {
      rd_f(r, nuvq);
      if (!ok) return;
      for (int32_t i = 0; i < npuv; ++i) {
        const float v = r[uvsrc[i]];
        float& d = dst[(size_t)i * stride + off];
        if (kind == 0) d = v;
        else if (kind == 1) d = std::max(v, huth);
        else { const float w = std::max(v, 0.005f); d = g * (w * w); }
      }
    };
    uvs(M->uv_zmin, 1, 0, 0);
    uvs(M->uv_zmax, 1, 0, 0);
    for (int32_t l = 1; l < nlev; ++l) uvs(M->uv_havg, nlev, l, 1);
    for (int32_t l = 1; l < nlev; ++l) uvs(M->uv_nrep, nlev, l, 2);
  } else {
    // u records then v records, each per cell through its right / upper point
    for (int dir = 0; dir < 2 && ok; ++dir) {
      const std::vector<int32_t>& idx = dir == 0 ? M->z_index_uv_mu1 : M->z_index_uv_nu1;
      auto uvs = [&](std::vector<float>& dst, size_t stride, size_t off, int kind) {
        rd_f(r, np);
        if (!ok) return;
        for (int32_t c = 0; c < np; ++c) {
          const int32_t ip = idx[c];
          if (ip <= 0) continue;
          float& d = dst[(size_t)(ip - 1) * stride + off];
          if (kind == 0) d = r[c];
          else { const float w = std::max(r[c], 0.0001f); d = g * (w * w); }
        }
      };
      uvs(M->uv_zmin, 1, 0, 0);
      uvs(M->uv_zmax, 1, 0, 0);
      rd_f(r, np);   // dhdz (not used)
      for (int32_t l = 1; l < nlev; ++l) uvs(M->uv_havg, nlev, l, 0);
      for (int32_t l = 1; l < nlev; ++l) uvs(M->uv_nrep, nlev, l, 2);
    }
  }
  if (!ok) { set_error(std::string("subgrid file ") + path + ": file ends early"); return SFX_ERR_IO; }
  for (int32_t i = 0; i < npuv; ++i) M->uv_nrep[(size_t)i * nlev] = M->uv_nrep[(size_t)i * nlev + 1];
  // sanitising of this path (:873-900)
  for (int32_t i = 0; i < npuv; ++i) {
    const int32_t nm = M->uv_index_z_nm[i] - 1, nmu = M->uv_index_z_nmu[i] - 1;
    M->uv_zmin[i] = std::max(M->uv_zmin[i], M->z_zmin[nm]);
    M->uv_zmin[i] = std::max(M->uv_zmin[i], M->z_zmin[nmu]);
  }
  for (int32_t c = 0; c < np; ++c)
    if (M->z_zmax[c] - M->z_zmin[c] < 0.001f) M->z_zmax[c] = M->z_zmax[c] + 0.001f;
  for (int32_t i = 0; i < npuv; ++i)
    if (M->uv_zmax[i] - M->uv_zmin[i] < 0.001f) M->uv_zmax[i] = M->uv_zmax[i] + 0.001f;
  M->uv_havg_zmax.resize(npuv); M->uv_nrep_zmax.resize(npuv); M->uv_navg_w.resize(npuv);
  for (int32_t i = 0; i < npuv; ++i) {
    M->uv_havg_zmax[i] = M->uv_havg[(size_t)i * nlev + (nlev - 1)] - M->uv_zmax[i];
    M->uv_nrep_zmax[i] = M->uv_nrep[(size_t)i * nlev + (nlev - 1)];
    M->uv_navg_w[i] = M->uv_nrep_zmax[i];
  }
  return SFX_OK;
}

extern "C" int sfx_model_set_boundary(sfx_model* M, int32_t nbnd, const float* x, const float* y,
                                      int32_t nt, const float* t, const float* zs) {
  if (!M || M->initialized || nbnd < 0 || (nbnd > 0 && (!x || !y))) { set_error("sfx_model_set_boundary: bad args"); return SFX_ERR_ARG; }
  M->nbnd = nbnd;
  M->x_bnd.assign(x, x + nbnd); M->y_bnd.assign(y, y + nbnd);
  if (nt > 0 && t && zs) {
    M->ntbnd = nt;
    M->t_bnd.assign(t, t + nt);
    M->zs_bnd.assign(zs, zs + (size_t)nt * nbnd);
  } else {
    // bnd without bzs: zero levels over [t0, t1] (sfincs_boundaries.f90:94-112)
    M->ntbnd = 2;
    M->t_bnd = {M->p.t0, M->p.t1};
    M->zs_bnd.assign((size_t)2 * nbnd, 0.0f);
  }
  return SFX_OK;
}

// Storage volume per cell (volfile; initialize_storage_volume, sfincs_domain.f90:2063-2112):
// water entering a subgrid cell fills it before the cell's volume grows.
extern "C" int sfx_model_set_storage_volume(sfx_model* M, const float* vol) {
  if (!M || M->initialized || !M->mesh_built || !vol) { set_error("sfx_model_set_storage_volume: bad args"); return SFX_ERR_ARG; }
  M->storage_volume.assign(vol, vol + M->np);
  return SFX_OK;
}

// Incoming infragravity levels at the boundary points (bzifile, sfincs_boundaries.f90:126-140):
// one value per boundary point at each time of the water-level series, same layout.
extern "C" int sfx_model_set_boundary_ig(sfx_model* M, const float* zsi) {
  if (!M || M->initialized || !zsi || M->nbnd <= 0) {
    set_error("sfx_model_set_boundary_ig: set the boundary points and series first");
    return SFX_ERR_ARG;
  }
  M->zsi_bnd.assign(zsi, zsi + (size_t)M->ntbnd * M->nbnd);
  return SFX_OK;
}

// Downstream river boundary points (bdrfile, sfincs_boundaries.f90:217-285): per point the
// boundary location, an interior point, the slope and the distance between them (< 0: the
// distance of the two points). dzs = -slope * distance.
extern "C" int sfx_model_set_river_boundary(sfx_model* M, int32_t nbdr, const float* x, const float* y,
                                            const float* x_in, const float* y_in, const float* slope,
                                            const float* distance) {
  if (!M || M->initialized || nbdr < 0 || (nbdr > 0 && (!x || !y || !x_in || !y_in || !slope || !distance))) {
    set_error("sfx_model_set_river_boundary: bad args");
    return SFX_ERR_ARG;
  }
  M->nbdr = nbdr;
  M->x_bdr.assign(x, x + nbdr); M->y_bdr.assign(y, y + nbdr);
  M->x_bdr_in.assign(x_in, x_in + nbdr); M->y_bdr_in.assign(y_in, y_in + nbdr);
  M->dzs_bdr.resize(nbdr);
  for (int32_t j = 0; j < nbdr; ++j) {
    if (distance[j] < 0.0f) {
      const float ddx = x_in[j] - x[j], ddy = y_in[j] - y[j];
      M->dzs_bdr[j] = -slope[j] * std::sqrt(ddx * ddx + ddy * ddy);
    } else {
      M->dzs_bdr[j] = -slope[j] * distance[j];
    }
  }
  return SFX_OK;
}

extern "C" int sfx_model_set_drainage(sfx_model* M, int32_t ndrn, const float* x_in, const float* y_in,
                                      const float* x_out, const float* y_out, const int32_t* type,
                                      const float* params) {
  if (!M || M->initialized || ndrn < 0 || (ndrn > 0 && (!x_in || !y_in || !x_out || !y_out || !type || !params))) {
    set_error("sfx_model_set_drainage: bad arguments");
    return SFX_ERR_ARG;
  }
  for (int32_t k = 0; k < ndrn; ++k) {
    if (type[k] < 1 || type[k] > 5) {
      set_error("drainage type " + std::to_string(type[k]) + " not recognized");
      return SFX_ERR_ARG;
    }
    if ((type[k] == 4 || type[k] == 5) && !(params[6 * k + 5] > 0.0f && params[6 * k] > 0.0f)) {
      set_error("a drainage gate needs a positive width and closing time");
      return SFX_ERR_ARG;
    }
  }
  M->ndrn = ndrn;
  M->x_drn_in.assign(x_in, x_in + ndrn); M->y_drn_in.assign(y_in, y_in + ndrn);
  M->x_drn_out.assign(x_out, x_out + ndrn); M->y_drn_out.assign(y_out, y_out + ndrn);
  M->drainage_type.assign(type, type + ndrn);
  M->drainage_params.assign(params, params + (size_t)6 * ndrn);
  for (int32_t k = 0; k < ndrn; ++k) {   // unused parameters of pumps, culverts and valves: 0
    if (type[k] <= 3)
      for (int j = 1; j < 6; ++j) M->drainage_params[(size_t)6 * k + j] = 0.0f;
  }
  return SFX_OK;
}

namespace {
bool polylines_ok(int32_t npoly, const int32_t* nv, const float* x, const float* y, size_t& nvert) {
  nvert = 0;
  if (npoly < 0 || (npoly > 0 && (!nv || !x || !y))) return false;
  for (int32_t i = 0; i < npoly; ++i) {
    if (nv[i] < 2) return false;
    nvert += (size_t)nv[i];
  }
  return true;
}
}  // namespace

// The active cell (1-based, 0 outside the mesh) holding each point, as the reference finds
// observation points (find_quadtree_cell with index_sfincs_in_quadtree,
// sfincs_obspoints.f90:95-130).
extern "C" int sfx_model_find_cells(sfx_model* M, int32_t n, const float* x, const float* y, int32_t* nm) {
  if (!M || !M->mesh_built || n < 0 || (n > 0 && (!x || !y || !nm))) { set_error("sfx_model_find_cells: bad args"); return SFX_ERR_ARG; }
  for (int32_t i = 0; i < n; ++i) nm[i] = M->find_cell_xy(x[i], y[i]);
  return SFX_OK;
}

// Spatially varying constant infiltration (qinffile, sfincs_infiltration.f90:224-258), in
// mm/h per active cell; converted to m/s at initialize.
extern "C" int sfx_model_set_infiltration(sfx_model* M, const float* qinf_mmh) {
  if (!M || M->initialized || !M->mesh_built || !qinf_mmh) { set_error("sfx_model_set_infiltration: bad args"); return SFX_ERR_ARG; }
  M->qinf_map_mmh.assign(qinf_mmh, qinf_mmh + M->np);
  return SFX_OK;
}

extern "C" int sfx_model_set_curve_number(sfx_model* M, const float* scs_inch) {
  if (!M || M->initialized || !M->mesh_built || !scs_inch) { set_error("sfx_model_set_curve_number: bad args"); return SFX_ERR_ARG; }
  M->scs_inch.assign(scs_inch, scs_inch + M->np);
  return SFX_OK;
}

extern "C" int sfx_model_set_thin_dams(sfx_model* M, int32_t npoly, const int32_t* nv, const float* x, const float* y) {
  size_t nvert = 0;
  if (!M || M->initialized || !polylines_ok(npoly, nv, x, y, nvert)) {
    set_error("sfx_model_set_thin_dams: bad arguments (each polyline needs >= 2 vertices)");
    return SFX_ERR_ARG;
  }
  M->thd_nv.assign(nv, nv + npoly);
  M->thd_x.assign(x, x + nvert); M->thd_y.assign(y, y + nvert);
  return SFX_OK;
}

extern "C" int sfx_model_set_weirs(sfx_model* M, int32_t npoly, const int32_t* nv, const float* x, const float* y,
                                   const float* par) {
  size_t nvert = 0;
  if (!M || M->initialized || !polylines_ok(npoly, nv, x, y, nvert) || (npoly > 0 && !par)) {
    set_error("sfx_model_set_weirs: bad arguments (each polyline needs >= 2 vertices)");
    return SFX_ERR_ARG;
  }
  M->weir_nv.assign(nv, nv + npoly);
  M->weir_x.assign(x, x + nvert); M->weir_y.assign(y, y + nvert);
  M->weir_par.assign(par, par + 2 * nvert);
  return SFX_OK;
}

extern "C" int sfx_model_set_sources(sfx_model* M, int32_t nsrc, const float* x, const float* y,
                                     int32_t nt, const float* t, const float* q) {
  if (!M || M->initialized || nsrc < 0 || (nsrc > 0 && (!x || !y || !t || !q || nt < 2))) {
    set_error("sfx_model_set_sources: bad args (need >= 2 times)");
    return SFX_ERR_ARG;
  }
  M->nsrc = nsrc; M->ntsrc = nt;
  M->x_src.assign(x, x + nsrc); M->y_src.assign(y, y + nsrc);
  M->tsrc.assign(t, t + nt);
  M->qsrc.assign(q, q + (size_t)nt * nsrc);
  return SFX_OK;
}

extern "C" int sfx_model_set_rain(sfx_model* M, int32_t nt, const float* t, const float* mmh) {
  if (!M || M->initialized || nt < 2 || !t || !mmh) { set_error("sfx_model_set_rain: need >= 2 times"); return SFX_ERR_ARG; }
  M->ntprcp = nt; M->tprcpt.assign(t, t + nt); M->tprcpv.assign(mmh, mmh + nt);
  M->precip = true;
  return SFX_OK;
}

extern "C" int sfx_model_set_wind(sfx_model* M, int32_t nt, const float* t, const float* mag, const float* dir) {
  if (!M || M->initialized || nt < 2 || !t || !mag || !dir) { set_error("sfx_model_set_wind: need >= 2 times"); return SFX_ERR_ARG; }
  M->ntwnd = nt; M->twnd.assign(t, t + nt); M->wndmag.assign(mag, mag + nt);
  M->wnddir.resize(nt);
  // Cartesian direction the wind blows to, with grid rotation (sfincs_meteo.f90:275)
  for (int32_t i = 0; i < nt; ++i) M->wnddir[i] = M->p.pi * (270.0f - dir[i]) / 180 - M->rotation;
  M->wind = true;
  return SFX_OK;
}

// ---- thin dams and weirs -----------------------------------------------------------------
namespace {

// direction (geometry.f90:124-139): the side of c from the line a -> b, 0 within 1e-6.
int side_of(float ax, float ay, float bx, float by, float cx, float cy) {
  const float val = (by - ay) * (cx - bx) - (bx - ax) * (cy - by);
  if (val < -1.0e-6f) return -1;
  if (val > 1.0e-6f) return 1;
  return 0;
}

// cross (geometry.f90:157-194): segments 1 and 2 cross when each separates the other's ends.
bool segments_cross(float x11, float y11, float x12, float y12, float x21, float y21, float x22, float y22) {
  return side_of(x11, y11, x12, y12, x21, y21) != side_of(x11, y11, x12, y12, x22, y22) &&
         side_of(x21, y21, x22, y22, x11, y11) != side_of(x21, y21, x22, y22, x12, y12);
}

// distance_between_points_projected_on_line_segment (geometry.f90:215-296): the length of
// the part of segment 1-2 between the projections of points a and b onto it.
float projected_length(float x0a, float y0a, float x0b, float y0b, float x1, float y1, float x2, float y2,
                       float mxdst) {
  auto position = [&](float px, float py, float& b) {
    b = std::fabs((x2 - x1) * (y1 - py) - (x1 - px) * (y2 - y1)) / std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    const float c1 = std::sqrt((px - x1) * (px - x1) + (py - y1) * (py - y1));
    const float c2 = std::sqrt((px - x2) * (px - x2) + (py - y2) * (py - y2));
    const float aa = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    const float a1 = std::sqrt(std::max(c1 * c1 - b * b, 0.0f));
    const float a2 = std::sqrt(std::max(c2 * c2 - b * b, 0.0f));
    if (a1 > aa || a2 > aa) return a1 < a2 ? 0.0f : 1.0f;
    return a1 / aa;
  };
  float b1, b2;
  const float p1 = position(x0a, y0a, b1);
  const float p2 = position(x0b, y0b, b2);
  const float aa = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
  return (b1 < mxdst && b2 < mxdst) ? aa * std::fabs(p2 - p1) : 0.0f;
}

// find_uv_points_intersected_by_polyline (sfincs_quadtree.F90:853-1010): for each segment,
// the cells of each level in the segment's index box widened by one cell, and of each the
// uv-points to its right and top (same level or coarser, and finer); a uv-point is crossed
// when the segment crosses the line between its two cell centres. Returns, per uv-point,
// the last crossing segment (1-based, 0 = not crossed).
std::vector<int32_t> crossed_uv_points(const sfx_model* M, const float* xp, const float* yp, int32_t npol) {
  std::vector<int32_t> ind(M->npuv, 0);
  std::vector<int64_t> mmax_level(M->nref + 1, 0);
  for (size_t k = 0; k < M->qt_nm.size(); ++k) {
    const int iref = M->qt_level[k];
    const int64_t nmx = (int64_t)M->qt_nmax << (iref - 1);
    mmax_level[iref] = std::max<int64_t>(mmax_level[iref], (M->qt_nm[k] - 1) / nmx + 1);
  }
  const std::vector<int32_t>* right_top[4] = {&M->z_index_uv_mu1, &M->z_index_uv_mu2, &M->z_index_uv_nu1,
                                              &M->z_index_uv_nu2};
  for (int32_t ipol = 1; ipol < npol; ++ipol) {
    const float xpa = xp[ipol - 1], ypa = yp[ipol - 1], xpb = xp[ipol], ypb = yp[ipol];
    const float xta = M->cosrot * (xpa - M->x0) + M->sinrot * (ypa - M->y0);
    const float yta = -M->sinrot * (xpa - M->x0) + M->cosrot * (ypa - M->y0);
    const float xtb = M->cosrot * (xpb - M->x0) + M->sinrot * (ypb - M->y0);
    const float ytb = -M->sinrot * (xpb - M->x0) + M->cosrot * (ypb - M->y0);
    for (int iref = M->nref; iref >= 1; --iref) {
      // n runs to nmx at this level, m to the largest m of the level's points
      const int64_t nmx = (int64_t)M->qt_nmax << (iref - 1);
      const int64_t mmx = mmax_level[iref];
      const double dyr = M->dyr[iref - 1], dxr = M->dxr[iref - 1];
      // int() of the reference truncates toward zero; cells outside [1, nmx] x [1, mmx]
      // hold no uv-point, so the box is clipped to the mesh
      const int64_t na = (int64_t)std::trunc(std::max(std::min((double)(yta / (float)dyr), 1e15), -1e15)) + 1;
      const int64_t ma = (int64_t)std::trunc(std::max(std::min((double)(xta / (float)dxr), 1e15), -1e15)) + 1;
      const int64_t nb = (int64_t)std::trunc(std::max(std::min((double)(ytb / (float)dyr), 1e15), -1e15)) + 1;
      const int64_t mb = (int64_t)std::trunc(std::max(std::min((double)(xtb / (float)dxr), 1e15), -1e15)) + 1;
      const int64_t n0 = std::max<int64_t>(std::min(na, nb) - 1, 1), n1 = std::min<int64_t>(std::max(na, nb) + 1, nmx);
      const int64_t m0 = std::max<int64_t>(std::min(ma, mb) - 1, 1), m1 = std::min<int64_t>(std::max(ma, mb) + 1, mmx);
      for (int64_t m = m0; m <= m1; ++m)
        for (int64_t n = n0; n <= n1; ++n) {
          const int32_t nm = M->find_cell_by_index((int32_t)n, (int32_t)m, iref);
          if (nm == 0) continue;
          for (const auto* tab : right_top) {
            const int32_t ip = (*tab)[nm - 1];
            if (ip <= 0) continue;
            const int32_t nmu = M->uv_index_z_nmu[ip - 1];
            if (segments_cross(xpa, ypa, xpb, ypb, M->z_xz[nm - 1], M->z_yz[nm - 1], M->z_xz[nmu - 1], M->z_yz[nmu - 1]))
              ind[ip - 1] = ipol;
          }
        }
    }
  }
  return ind;
}

// read_thin_dams (sfincs_structures.f90:385-495) and read_structure_file for weirs
// (:38-381), run after the domain tables, the boundaries and the advection mask are set
// (sfincs_lib.f90:154-156), so neither the boundary list nor mask_adv sees them.
void apply_structures(sfx_model* M) {
  M->thindam_uv_index.clear();
  M->structure_uv_index.clear(); M->structure_height.clear(); M->structure_cd.clear(); M->structure_length.clear();
  const int32_t npuv = M->npuv;
  if (!M->thd_nv.empty()) {
    std::vector<char> closed(npuv, 0);
    size_t v0 = 0;
    for (int32_t nv : M->thd_nv) {
      const std::vector<int32_t> ind = crossed_uv_points(M, &M->thd_x[v0], &M->thd_y[v0], nv);
      for (int32_t ip = 0; ip < npuv; ++ip)
        if (ind[ip] > 0) { M->kcuv[ip] = 0; closed[ip] = 1; }
      v0 += (size_t)nv;
    }
    for (int32_t ip = 0; ip < npuv; ++ip)
      if (closed[ip]) M->thindam_uv_index.push_back(ip + 1);
  }
  if (M->weir_nv.empty()) return;
  std::vector<char> istruc(npuv, 0);
  std::vector<float> height(npuv, -9999.0f), cd(npuv, 0.0f), lngth(npuv, 0.0f);
  const float pi = M->p.pi, rot = M->rotation;
  size_t v0 = 0;
  for (int32_t nv : M->weir_nv) {
    const float* xt = &M->weir_x[v0];
    const float* yt = &M->weir_y[v0];
    const float* pr = &M->weir_par[2 * v0];
    const std::vector<int32_t> ind = crossed_uv_points(M, xt, yt, nv);
    for (int32_t ip = 0; ip < npuv; ++ip) {
      const int32_t irow = ind[ip];
      if (irow == 0) continue;
      istruc[ip] = 1;
      // the uv-point's face: its centre and its two corners (:169-242)
      const int32_t nm = M->uv_index_z_nm[ip], nmu = M->uv_index_z_nmu[ip];
      const int iref = M->uv_flags_iref[ip], idir = M->uv_flags_dir[ip];
      const int irefnm = M->z_flags_iref[nm - 1], irefnmu = M->z_flags_iref[nmu - 1];
      const float dxs = idir == 0 ? M->dyrm[iref - 1] : M->dxrm[iref - 1];
      float xuv, yuv;
      if (irefnm >= irefnmu) {
        const float a = idir == 0 ? rot : rot + 0.5f * pi;
        xuv = M->z_xz[nm - 1] + 0.5f * dxs * std::cos(a);
        yuv = M->z_yz[nm - 1] + 0.5f * dxs * std::sin(a);
      } else {
        const float a = idir == 0 ? rot - pi : rot - 0.5f * pi;
        xuv = M->z_xz[nmu - 1] + 0.5f * dxs * std::cos(a);
        yuv = M->z_yz[nmu - 1] + 0.5f * dxs * std::sin(a);
      }
      const float a1 = idir == 0 ? rot - 0.5f * pi : rot - pi;
      const float a2 = idir == 0 ? rot + 0.5f * pi : rot;
      const float xuv1 = xuv + 0.5f * dxs * std::cos(a1), yuv1 = yuv + 0.5f * dxs * std::sin(a1);
      const float xuv2 = xuv + 0.5f * dxs * std::cos(a2), yuv2 = yuv + 0.5f * dxs * std::sin(a2);
      const int32_t k = irow - 1;   // the crossing segment runs from vertex k to k + 1
      lngth[ip] = lngth[ip] + projected_length(xuv1, yuv1, xuv2, yuv2, xt[k], yt[k], xt[k + 1], yt[k + 1], 999999.0f);
      const float dst1 = std::sqrt((xuv - xt[k]) * (xuv - xt[k]) + (yuv - yt[k]) * (yuv - yt[k]));
      const float dst2 = std::sqrt((xuv - xt[k + 1]) * (xuv - xt[k + 1]) + (yuv - yt[k + 1]) * (yuv - yt[k + 1]));
      const float wfac = dst1 / (dst1 + dst2);
      const float h = pr[2 * k] * (1.0f - wfac) + pr[2 * (k + 1)] * wfac;
      height[ip] = std::max(height[ip], h);
      cd[ip] = pr[2 * k + 1] * (1.0f - wfac) + pr[2 * (k + 1) + 1] * wfac;
    }
    v0 += (size_t)nv;
  }
  // a crest at or below the uv-point's bed is no structure (:293-325)
  for (int32_t ip = 0; ip < npuv; ++ip) {
    if (!istruc[ip]) continue;
    const float bed = M->subgrid ? M->uv_zmin[ip] : M->zbuvmx[ip];
    if (!(height[ip] > bed)) continue;
    const int iref = M->uv_flags_iref[ip], idir = M->uv_flags_dir[ip];
    const float dxs = idir == 0 ? M->dyrm[iref - 1] : M->dxrm[iref - 1];
    M->structure_uv_index.push_back(ip + 1);
    M->structure_height.push_back(height[ip]);
    M->structure_cd.push_back(cd[ip]);
    M->structure_length.push_back(lngth[ip] / dxs);
    M->kcuv[ip] = 3;
  }
}

}  // namespace

extern "C" int sfx_model_initialize(sfx_model* M) {
  if (!M || !M->mesh_built || M->initialized) { set_error("sfx_model_initialize: bad state"); return SFX_ERR_STATE; }
  if (!M->subgrid && M->zb.empty()) { set_error("bathymetry or subgrid tables must be set first"); return SFX_ERR_STATE; }
  const int32_t np = M->np, npuv = M->npuv, ncuv = M->ncuv;
  const sfx_params& p = M->p;

  // initialize_boundaries (sfincs_domain.f90:1786-1956)
  M->ngbnd = 0; M->nmindbnd.clear();
  bool has2 = false, has5 = false;
  for (int32_t c = 0; c < np; ++c) {
    const int k = M->kcs[c];
    if (k == 2 || k == 3 || k == 5 || k == 6) {
      M->nmindbnd.push_back(c + 1);
      if (k == 2) has2 = true;
      if (k == 5) has5 = true;
    }
    if (k == 4) { set_error("wave maker cells (msk=4) are not supported"); return SFX_ERR_ARG; }
  }
  M->ngbnd = (int32_t)M->nmindbnd.size();
  M->zsb.assign(M->ngbnd, 0.f); M->zsb0.assign(M->ngbnd, 0.f);
  M->nkcuv2 = 0;
  for (int32_t i = 0; i < npuv; ++i) {
    const int a = M->kcs[M->uv_index_z_nm[i] - 1], b = M->kcs[M->uv_index_z_nmu[i] - 1];
    int8_t kv;
    if (a > 1 && b > 1) kv = 0;
    else if (a * b == 0) kv = 0;
    else if ((a == 1 && b == 2) || (a == 2 && b == 1)) { kv = 2; ++M->nkcuv2; }
    else if ((a == 1 && b == 3) || (a == 3 && b == 1)) { kv = 2; ++M->nkcuv2; }
    else if ((a == 1 && b == 5) || (a == 5 && b == 1)) { kv = 2; ++M->nkcuv2; }
    else if ((a == 1 && b == 6) || (a == 6 && b == 1)) { kv = 6; ++M->nkcuv2; }
    else kv = 1;
    M->kcuv[i] = kv;
  }
  M->index_kcuv2.clear(); M->nmbkcuv2.clear(); M->nmikcuv2.clear(); M->ibuvdir.clear(); M->ibkcuv2.clear();
  for (int32_t i = 0; i < npuv; ++i) {
    if (M->kcuv[i] != 2 && M->kcuv[i] != 6) continue;
    const int32_t nm = M->uv_index_z_nm[i], nmu = M->uv_index_z_nmu[i];
    M->index_kcuv2.push_back(i + 1);
    int32_t nmb, nmi; int8_t dir;
    if (M->kcs[nm - 1] == 1) { nmb = nmu; nmi = nm; dir = -1; } else { nmb = nm; nmi = nmu; dir = 1; }
    M->nmbkcuv2.push_back(nmb); M->nmikcuv2.push_back(nmi); M->ibuvdir.push_back(dir);
    int32_t ibk = 0;
    for (int32_t ib = 0; ib < M->ngbnd; ++ib) {
      if (M->nmindbnd[ib] == nmb) {
        ibk = ib + 1;
        if (M->kcs[nmb - 1] == 3) {
          M->zsb0[ib] = M->subgrid ? M->z_zmin[nmb - 1] : M->zb[nmb - 1];
          M->zsb[ib] = M->zsb0[ib];
        }
        break;
      }
    }
    M->ibkcuv2.push_back(ibk);
  }
  M->uvmean.assign(M->nkcuv2, 0.f);

  // initialize_roughness (:1978-2047): uniform manning (land/sea split by rgh_lev_land)
  if (!M->subgrid) {
    M->gn2uv.assign(npuv, 9.81f * 0.02f * 0.02f);
    const float gn2land = (p.manning_land < 0.0f) ? 9.81f * (p.manning * p.manning) : 9.81f * (p.manning_land * p.manning_land);
    const float gn2sea = (p.manning_sea < 0.0f) ? 9.81f * (p.manning * p.manning) : 9.81f * (p.manning_sea * p.manning_sea);
    for (int32_t i = 0; i < npuv; ++i)
      if (M->kcuv[i] > 0) M->gn2uv[i] = (M->zbuv[i] > p.rgh_lev_land) ? gn2land : gn2sea;
  }

  // initialize_hydro + set_initial_conditions (:2195-2398; sfincs_initial_conditions.F90:22-174).
  // Initial levels are zsini, or a restart file's (f32) levels, raised to the bed; a
  // restart also gives q on the uv-points (not the combined ones) and uvmean.
  if (M->rst_type && M->rst_zs.size() != (size_t)np) { set_error("restart file: zs has the wrong size"); return SFX_ERR_ARG; }
  const bool rst_q = M->rst_type == 1 || M->rst_type == 2;
  if (rst_q && M->rst_q.size() != (size_t)(npuv + ncuv + 1)) { set_error("restart file: q has the wrong size"); return SFX_ERR_ARG; }
  if (rst_q && M->rst_uvmean.size() != (size_t)M->nkcuv2) { set_error("restart file: uvmean has the wrong size"); return SFX_ERR_ARG; }
  M->zs.resize(np);
  const double zini = (double)p.zsini;
  for (int32_t c = 0; c < np; ++c) {
    const double z = M->rst_type ? (double)M->rst_zs[c] : zini;
    M->zs[c] = M->subgrid ? std::max((double)M->z_zmin[c], z) : std::max((double)M->zb[c], z);
  }
  M->q.assign(npuv + ncuv + 1, 0.f);
  M->uv.assign(npuv + ncuv + 1, 0.f);
  if (rst_q) {
    for (int32_t i = 0; i < npuv; ++i) M->q[i] = M->rst_q[i];
    M->uvmean = M->rst_uvmean;
  }
  M->kfuv.assign(npuv, 0);
  if (p.wiggle_suppression) {
    M->zs0.resize(np); M->zsderv.assign(np, 0.f);
    for (int32_t c = 0; c < np; ++c) M->zs0[c] = (float)M->zs[c];
  } else {
    M->zs0.assign(np, 0.f); M->zsderv.assign(np, 0.f);
  }
  // combined uv-points take the mean of their two points (sfincs_domain.f90:2272-2280).
  // The uv that set_initial_conditions derives from a restart's q is then discarded:
  // uv = 0 at :2311 runs after it, so every run starts with uv = 0.
  for (int32_t j = 0; j < ncuv; ++j) {
    const int32_t c = M->cuv_index_uv[j] - 1, a = M->cuv_index_uv1[j] - 1, b = M->cuv_index_uv2[j] - 1;
    M->q[c] = 0.5f * (M->q[a] + M->q[b]);
  }
  M->zsmax.assign(np, -999.0f);
  // sfincs_domain.f90:2300-2306 (qmax is allocated without a value there, :2251; -999 is
  // what sfincs_output.f90:210 resets it to)
  M->vmax.assign(np, -999.0f); M->qmax.assign(np, -999.0f); M->twet.assign(np, 0.0f);
  if (M->subgrid) {
    // compute_initial_subgrid_volumes (sfincs_subgrid.F90:907-962)
    const int nl = M->nl;
    M->z_volume.resize(np);
    for (int32_t c = 0; c < np; ++c) {
      const double zsc = M->zs[c];
      const float area = M->cell_area[M->z_flags_iref[c] - 1];
      if (zsc >= (double)M->z_zmax[c]) {
        M->z_volume[c] = (double)M->z_volmax[c] + (double)area * (zsc - (double)std::max(M->z_zmax[c], -20.0f));
      } else {
        int ivol = 1;
        for (int il = 2; il <= nl; ++il)
          if ((double)M->z_dep[(size_t)c * nl + il - 1] > zsc) { ivol = il - 1; break; }
        const float dzvol = M->z_volmax[c] / (float)(nl - 1);
        const float d0 = M->z_dep[(size_t)c * nl + ivol - 1], d1 = M->z_dep[(size_t)c * nl + ivol];
        const float facint = (float)((zsc - (double)d0) / (double)std::max(d1 - d0, 0.001f));
        M->z_volume[c] = (double)((float)(ivol - 1) * dzvol + facint * dzvol);
      }
      // a cell that starts wet has no storage left for the water it already holds
      // (sfincs_subgrid.F90:952-956; f64 difference, stored as f32)
      if (!M->storage_volume.empty())
        M->storage_volume[c] = (float)std::max((double)M->storage_volume[c] - M->z_volume[c], 0.0);
    }
  } else {
    M->z_volume.clear();
  }

  // find_boundary_indices (sfincs_boundaries.f90:473-679)
  M->ind1_bnd_gbp.assign(M->ngbnd, 0); M->ind2_bnd_gbp.assign(M->ngbnd, 0);
  M->fac_bnd_gbp.assign(M->ngbnd, 0.f); M->nmi_gbp.assign(M->ngbnd, 0);
  if (has2 && M->nbnd == 0) { set_error("water level boundary cells (msk=2) need boundary points"); return SFX_ERR_ARG; }
  // (sfincs_boundaries.f90:272-275: river cells without a bdr file stop the model)
  if (has5 && M->nbdr == 0) { set_error("downstream river boundary cells (msk=5) need a bdr file"); return SFX_ERR_ARG; }
  M->index_bdr_gbp.assign(M->ngbnd, 0);
  for (int32_t ib = 0; ib < M->ngbnd; ++ib) {
    const int32_t nm = M->nmindbnd[ib];
    const float x = M->z_xz[nm - 1], y = M->z_yz[nm - 1];
    const int k = M->kcs[nm - 1];
    if (k == 2) {
      if (M->nbnd > 1) {
        float dst1 = 1.0e10f, dst2 = 1.0e10f; int32_t ib1 = 0, ib2 = 0;
        for (int32_t j = 0; j < M->nbnd; ++j) {
          const float ddx = M->x_bnd[j] - x, ddy = M->y_bnd[j] - y;
          const float dst = std::sqrt(ddx * ddx + ddy * ddy);
          if (dst < dst1) { dst2 = dst1; ib2 = ib1; dst1 = dst; ib1 = j + 1; }
          else if (dst < dst2) { dst2 = dst; ib2 = j + 1; }
        }
        M->ind1_bnd_gbp[ib] = ib1; M->ind2_bnd_gbp[ib] = ib2;
        M->fac_bnd_gbp[ib] = dst2 / std::max(dst1 + dst2, 1.0e-9f);
      } else {
        M->ind1_bnd_gbp[ib] = 1; M->ind2_bnd_gbp[ib] = 1; M->fac_bnd_gbp[ib] = 1.0f;
      }
    } else if (k == 5) {   // nearest bdr point (:587-613)
      float dst1 = 1.0e10f; int32_t ib1 = 0;
      for (int32_t j = 0; j < M->nbdr; ++j) {
        const float ddx = M->x_bdr[j] - x, ddy = M->y_bdr[j] - y;
        const float dst = std::sqrt(ddx * ddx + ddy * ddy);
        if (dst < dst1) { dst1 = dst; ib1 = j + 1; }
      }
      if (ib1 == 0) { set_error("a downstream river boundary cell has no bdr point within 1e10 m"); return SFX_ERR_ARG; }
      M->index_bdr_gbp[ib] = ib1;
    } else if (k == 6) {
      const int32_t n = M->z_index_z_n[nm - 1], m = M->z_index_z_m[nm - 1];
      const int iref = M->z_flags_iref[nm - 1];
      const int32_t cand[4][2] = {{n, m + 1}, {n + 1, m}, {n, m - 1}, {n - 1, m}};
      for (int j = 0; j < 4; ++j) {
        const int32_t nmi = M->find_cell_by_index(cand[j][0], cand[j][1], iref);
        if (nmi > 0 && M->kcs[nmi - 1] == 1) M->nmi_gbp[ib] = nmi;
      }
      if (M->nmi_gbp[ib] == 0) M->kcs[nm - 1] = 0;
    }
  }

  // the cell holding each bdr interior point (find_quadtree_cell, :248). Its level is read
  // while the boundary levels are set, so it must be an ordinary cell, which no boundary
  // step writes (the reference does not check this; outside the mesh it reads zs(0)).
  M->index_zsi_bdr.assign(M->nbdr, 0);
  for (int32_t j = 0; j < M->nbdr; ++j) {
    const int32_t c = M->find_cell_xy(M->x_bdr_in[j], M->y_bdr_in[j]);
    if (c == 0 || M->kcs[c - 1] != 1) {
      set_error("a river boundary's interior point must lie in an active interior cell (kcs 1)");
      return SFX_ERR_ARG;
    }
    M->index_zsi_bdr[j] = c;
  }

  // read_discharges cell lookup (sfincs_discharges.f90:159-171), then each drainage
  // structure's intake and outfall cells and the distance between their centres (:258-280)
  M->nmindsrc.assign(M->nsrc + 2 * M->ndrn, 0);
  for (int32_t i = 0; i < M->nsrc; ++i) {
    M->nmindsrc[i] = M->find_cell_xy(M->x_src[i], M->y_src[i]);
    if (M->nmindsrc[i] == 0) { set_error("point source outside the active mesh"); return SFX_ERR_ARG; }
  }
  M->drainage_distance.assign(M->ndrn, 0.f);
  for (int32_t k = 0; k < M->ndrn; ++k) {
    const int32_t a = M->find_cell_xy(M->x_drn_in[k], M->y_drn_in[k]);
    const int32_t b = M->find_cell_xy(M->x_drn_out[k], M->y_drn_out[k]);
    if (a == 0 || b == 0) { set_error("a drainage structure's intake or outfall lies outside the active mesh"); return SFX_ERR_ARG; }
    M->nmindsrc[M->nsrc + 2 * k] = a;
    M->nmindsrc[M->nsrc + 2 * k + 1] = b;
    const float dxs = M->z_xz[b - 1] - M->z_xz[a - 1], dys = M->z_yz[b - 1] - M->z_yz[a - 1];
    M->drainage_distance[k] = std::sqrt(dxs * dxs + dys * dys);
  }
  // discharge times must cover [t0, t1] (sfincs_discharges.f90:134-155)
  if (M->nsrc > 0) {
    if (M->tsrc[0] > M->p.t0 + 1.0f || M->tsrc[M->ntsrc - 1] < M->p.t1 - 1.0f) {
      if (M->tsrc[0] > M->p.t0 + 1.0f) M->tsrc[0] = M->p.t0 - 1.0f;
      else M->tsrc[M->ntsrc - 1] = M->p.t1 + 1.0f;
    }
  }
  if (M->precip && M->tprcpt[0] > M->p.t0) { set_error("rain series must start at or before t0"); return SFX_ERR_ARG; }
  if (M->wind && M->twnd[0] > M->p.t0) { set_error("wind series must start at or before t0"); return SFX_ERR_ARG; }

  // Cd table (sfincs_meteo.f90:316-328), default cd_nr = 3 (sfincs_input.f90:314-328)
  {
    const float cd_wnd[3] = {0.0f, 28.0f, 50.0f}, cd_val[3] = {0.0010f, 0.0025f, 0.0015f};
    const int cd_nr = 3;
    M->cdval.assign(1000, 0.f);
    for (int iw = 1; iw <= 1000; ++iw) {
      const float wnd = (float)iw * 0.1f;
      for (int icd = 1; icd <= cd_nr - 1; ++icd) {
        M->cdval[iw - 1] = cd_val[icd - 1];
        if (wnd >= cd_wnd[icd - 1] && wnd < cd_wnd[icd]) {
          M->cdval[iw - 1] = cd_val[icd - 1] + (wnd - cd_wnd[icd - 1]) * (cd_val[icd] - cd_val[icd - 1]) / (cd_wnd[icd] - cd_wnd[icd - 1]);
          break;
        } else if (wnd >= cd_wnd[cd_nr - 1]) {
          M->cdval[iw - 1] = cd_val[cd_nr - 1];
          break;
        }
      }
    }
  }

  // fill_h73_tables (sfincs_data.f90:858-881): hh = 0.001*j*10**(k-6), x73 = hh**2*hh**(1/3)
  M->x73.assign(10000, 0.f);
  for (int k = 1; k <= 10; ++k) {
    float p10 = 1.0f;
    if (k >= 6) { for (int i = 0; i < k - 6; ++i) p10 *= 10.0f; }
    else { float d = 1.0f; for (int i = 0; i < 6 - k; ++i) d *= 10.0f; p10 = 1.0f / d; }
    for (int j = 1; j <= 1000; ++j) {
      const float hh = 0.001f * (float)j * p10;
      M->x73[(k - 1) * 1000 + (j - 1)] = hh * hh * std::pow(hh, 1.0f / 3.0f);
    }
  }

  // set_advection_mask (sfincs_domain.f90:2114-2180)
  for (int32_t j = 0; j < ncuv; ++j) {
    const int a = M->kcuv[M->cuv_index_uv1[j] - 1], b = M->kcuv[M->cuv_index_uv2[j] - 1];
    int8_t kv;
    if (a == 0 || b == 0) kv = 0;
    else if (a == 2 || b == 2) kv = 2;
    else if (a == 4 || b == 4) kv = 4;
    else kv = 1;
    M->kcuv[M->cuv_index_uv[j] - 1] = kv;
  }
  M->mask_adv.assign(npuv, 1);
  if (p.advection && p.advection_mask && (M->use_quadtree || M->nmax > 1) && need_uv_neighbours(M)) {
    for (int32_t i = 0; i < npuv; ++i) {
      const int32_t nb[8] = {M->uv_index_u_nmu[i], M->uv_index_u_nmd[i], M->uv_index_v_nm[i], M->uv_index_v_nmu[i],
                             M->uv_index_v_ndm[i], M->uv_index_v_ndmu[i], M->uv_index_u_ndm[i], M->uv_index_u_num[i]};
      for (int j = 0; j < 8; ++j)
        if (M->kcuv[nb[j] - 1] != 1) { M->mask_adv[i] = 0; break; }
    }
  }
  apply_structures(M);
  {
    const int rc = sfx::build_meteo(M);
    if (rc) return rc;
  }

  // initialize_infiltration (sfincs_infiltration.f90:8-258): only with rain (:63); the
  // uniform rate ('con', qinf > 0, converted by sfincs_input.f90:376) before the map
  // ('c2d', :224-258), as the reference tries them (:92-104)
  M->qinffield.clear();
  M->scs_m.clear();
  if (M->precip && p.qinf <= 0.0f && M->qinf_map_mmh.empty() && !M->scs_inch.empty()) {
    // 'cna' (:260-293): S in m (the reference keeps it in qinffield)
    M->scs_m.assign(np, 0.0f);
    for (int32_t nm = 0; nm < np; ++nm) M->scs_m[nm] = M->scs_inch[nm] * 0.0254f;
  } else if (M->precip && (p.qinf > 0.0f || !M->qinf_map_mmh.empty())) {
    M->qinffield.assign(np, 0.0f);
    if (p.qinf > 0.0f) {
      const float qinf = p.qinf / (float)(3600 * 1000);
      for (int32_t nm = 0; nm < np; ++nm) {
        const float floorz = M->subgrid ? M->z_zmin[nm] : M->zb[nm];
        M->qinffield[nm] = floorz > p.qinf_zmin ? qinf : 0.0f;
      }
    } else {
      for (int32_t nm = 0; nm < np; ++nm) M->qinffield[nm] = M->qinf_map_mmh[nm] / 3600.0f / 1000.0f;
    }
  }
  M->initialized = true;
  return SFX_OK;
}

// ---- table / state export ----------------------------------------------------------
namespace {
struct View { const void* ptr; size_t bytes; };

template <typename T> View V(const std::vector<T>& v) { return View{v.data(), v.size() * sizeof(T)}; }

bool lookup(const sfx_model* M, const std::string& k, View& out) {
  if (k == "params") { out = View{&M->p, sizeof(sfx_params)}; return true; }
  if (k == "cosrot") { out = View{&M->cosrot, sizeof(float)}; return true; }
  if (k == "sinrot") { out = View{&M->sinrot, sizeof(float)}; return true; }
  // the meteo inputs as set or read: <kind>_times, <kind>_values and <kind>_geo (x_llcorner,
  // y_llcorner, dx, dy) for amu, amv, amp, ampr; spw_times, spw_xe, spw_ye, spw_vmag,
  // spw_vdir, spw_pdrop, spw_prcp, spw_radius
  {
    const sfx_model::MeteoGrid* g[4] = {&M->amu, &M->amv, &M->amp, &M->ampr};
    const char* nm[4] = {"amu", "amv", "amp", "ampr"};
    for (int i = 0; i < 4; ++i) {
      const std::string b(nm[i]);
      if (k == b + "_times") { out = V(g[i]->times); return true; }
      if (k == b + "_values") { out = V(g[i]->v); return true; }
      if (k == b + "_geo") { out = View{&g[i]->xll, 4 * sizeof(float)}; return true; }
    }
    if (k == "spw_times") { out = V(M->spw.times); return true; }
    if (k == "spw_xe") { out = V(M->spw.xe); return true; }
    if (k == "spw_ye") { out = V(M->spw.ye); return true; }
    if (k == "spw_vmag") { out = V(M->spw.vmag); return true; }
    if (k == "spw_vdir") { out = V(M->spw.vdir); return true; }
    if (k == "spw_pdrop") { out = V(M->spw.pdrop); return true; }
    if (k == "spw_prcp") { out = V(M->spw.prcp); return true; }
    if (k == "spw_radius") { out = View{&M->spw.radius, sizeof(float)}; return true; }
  }
#define T(name) if (k == #name) { out = V(M->name); return true; }
  T(kcs) T(z_flags_iref) T(z_index_z_n) T(z_index_z_m)
  T(z_index_uv_md) T(z_index_uv_mu) T(z_index_uv_nd) T(z_index_uv_nu)
  T(z_index_uv_md1) T(z_index_uv_md2) T(z_index_uv_mu1) T(z_index_uv_mu2)
  T(z_index_uv_nd1) T(z_index_uv_nd2) T(z_index_uv_nu1) T(z_index_uv_nu2)
  T(z_xz) T(z_yz) T(zb)
  T(uv_index_z_nm) T(uv_index_z_nmu) T(uv_index_u_nmd) T(uv_index_u_nmu) T(uv_index_u_ndm) T(uv_index_u_num)
  T(uv_index_v_ndm) T(uv_index_v_nm) T(uv_index_v_ndmu) T(uv_index_v_nmu)
  T(uv_flags_iref) T(uv_flags_type) T(uv_flags_dir) T(kcuv) T(mask_adv)
  T(thindam_uv_index) T(structure_uv_index) T(structure_height) T(structure_cd) T(structure_length)
  T(qinffield) T(scs_m) T(meteo_times) T(meteo_tauwu) T(meteo_tauwv) T(meteo_patm) T(meteo_prcp)
  T(cuv_index_uv) T(cuv_index_uv1) T(cuv_index_uv2) T(zbuv) T(zbuvmx) T(gn2uv)
  T(dxr) T(dyr) T(dxrm) T(dyrm) T(dxrinv) T(dyrinv) T(dxr2inv) T(dyr2inv) T(dxrinvc) T(dyrinvc) T(cell_area) T(nuvisc)
  T(z_zmin) T(z_zmax) T(z_volmax) T(z_dep) T(uv_zmin) T(uv_zmax) T(uv_navg_w) T(uv_fnfit)
  T(uv_havg) T(uv_nrep) T(uv_pwet) T(uv_havg_zmax) T(uv_nrep_zmax)
  T(zs) T(z_volume) T(q) T(uv) T(zs0) T(zsderv) T(zsmax) T(kfuv) T(vmax) T(qmax) T(twet)
  T(nmindbnd) T(zsb) T(zsb0) T(index_kcuv2) T(nmbkcuv2) T(nmikcuv2) T(ibkcuv2) T(ibuvdir) T(uvmean)
  T(ind1_bnd_gbp) T(ind2_bnd_gbp) T(fac_bnd_gbp) T(nmi_gbp) T(nmindsrc) T(cdval) T(x73)
  T(drainage_type) T(drainage_params) T(drainage_distance)
  T(index_bdr_gbp) T(index_zsi_bdr) T(dzs_bdr) T(zsi_bnd) T(storage_volume)
  T(t_bnd) T(zs_bnd) T(tsrc) T(qsrc) T(tprcpt) T(tprcpv) T(twnd) T(wndmag) T(wnddir)
#undef T
  return false;
}
}  // namespace

extern "C" int sfx_model_table_bytes(const sfx_model* M, const char* name, size_t* nbytes) {
  View v{};
  if (!M || !name || !nbytes || !lookup(M, name, v)) { set_error(std::string("unknown table ") + (name ? name : "")); return SFX_ERR_ARG; }
  *nbytes = v.bytes;
  return SFX_OK;
}

extern "C" int sfx_model_get_table(const sfx_model* M, const char* name, void* dst, size_t nbytes) {
  View v{};
  if (!M || !name || !dst || !lookup(M, name, v)) { set_error(std::string("unknown table ") + (name ? name : "")); return SFX_ERR_ARG; }
  if (nbytes != v.bytes) { set_error(std::string("size mismatch for ") + name); return SFX_ERR_ARG; }
  if (nbytes) std::memcpy(dst, v.ptr, nbytes);
  return SFX_OK;
}

// ---- binary restart files (sfincs_output.f90:657-729, sfincs_initial_conditions.F90:183-260)
// Fortran sequential unformatted records: a 4-byte length, the payload, the length again.
// The records are the type (int32), zs (f32, np), q (f32, npuv+ncuv+1) and uvmean (f32,
// nkcuv2); type 3 holds zs only, types 4-6 add infiltration fields.
namespace {
bool read_record(std::FILE* f, std::vector<char>& out) {
  int32_t n1 = 0, n2 = 0;
  if (std::fread(&n1, 4, 1, f) != 1 || n1 < 0) return false;
  out.resize((size_t)n1);
  if (n1 && std::fread(out.data(), 1, (size_t)n1, f) != (size_t)n1) return false;
  return std::fread(&n2, 4, 1, f) == 1 && n2 == n1;
}
bool write_record(std::FILE* f, const void* p, size_t nbytes) {
  const int32_t n = (int32_t)nbytes;
  return std::fwrite(&n, 4, 1, f) == 1 && (nbytes == 0 || std::fwrite(p, 1, nbytes, f) == nbytes) &&
         std::fwrite(&n, 4, 1, f) == 1;
}
std::vector<float> as_floats(const std::vector<char>& r) {
  std::vector<float> v(r.size() / 4);
  if (!v.empty()) std::memcpy(v.data(), r.data(), v.size() * 4);
  return v;
}
}  // namespace

extern "C" int sfx_model_set_restart(sfx_model* M, const char* path) {
  if (!M || !path || !M->mesh_built || M->initialized) {
    set_error("sfx_model_set_restart: needs a built mesh, before sfx_model_initialize");
    return SFX_ERR_STATE;
  }
  std::FILE* f = std::fopen(path, "rb");
  if (!f) { set_error(std::string("restart file not found: ") + path); return SFX_ERR_IO; }
  std::vector<char> r;
  int32_t type = 0;
  bool ok = read_record(f, r) && r.size() == 4;
  if (ok) std::memcpy(&type, r.data(), 4);
  if (ok && (type < 1 || type > 6)) {   // as the reference: an unknown type is skipped
    std::fclose(f);
    M->rst_type = 0;
    return SFX_OK;
  }
  if (ok && type >= 4) {
    std::fclose(f);
    set_error("restart file types 4-6 carry infiltration fields, which this build does not model");
    return SFX_ERR_ARG;
  }
  std::vector<float> zs, q, uvm;
  ok = ok && read_record(f, r) && r.size() == (size_t)M->np * 4;
  if (ok) zs = as_floats(r);
  if (ok && type != 3) {
    ok = read_record(f, r) && r.size() == (size_t)(M->npuv + M->ncuv + 1) * 4;
    if (ok) q = as_floats(r);
    ok = ok && read_record(f, r) && r.size() % 4 == 0;
    if (ok) uvm = as_floats(r);
  }
  std::fclose(f);
  if (!ok) { set_error(std::string("restart file unreadable or sized for another mesh: ") + path); return SFX_ERR_IO; }
  M->rst_type = type;
  M->rst_zs = std::move(zs); M->rst_q = std::move(q); M->rst_uvmean = std::move(uvm);
  return SFX_OK;
}

// Writes a type-1 restart file from state arrays in reference order (zs f64, np; q f32,
// npuv+ncuv+1; uvmean f32, nkcuv2); zs is stored as f32, as write_rst_file does.
int sfx::write_restart_arrays(const char* path, const double* zs, int32_t np, const float* q, int32_t nq,
                               const float* uvmean, int32_t nuvmean) {
  std::vector<float> zs4((size_t)np);
  for (int32_t c = 0; c < np; ++c) zs4[c] = (float)zs[c];
  std::FILE* f = std::fopen(path, "wb");
  if (!f) { set_error(std::string("cannot write restart file: ") + path); return SFX_ERR_IO; }
  const int32_t type = 1;
  const bool ok = write_record(f, &type, 4) && write_record(f, zs4.data(), zs4.size() * 4) &&
                  write_record(f, q, (size_t)nq * 4) && write_record(f, uvmean, (size_t)nuvmean * 4);
  if (std::fclose(f) != 0 || !ok) { set_error(std::string("cannot write restart file: ") + path); return SFX_ERR_IO; }
  return SFX_OK;
}

// timestep_analysis_finalize (sfincs_timestep_analysis.f90:77-200): per active cell, over
// its (up to 8) uv-points with kcuv 1 or 6, the smallest average required step
// (sum / max(wet, 1)) times alpha, capped at dtmax (-1 when the cell has no such point),
// and the largest count of limiting sweeps in percent of nt; inactive cells get 0.
extern "C" int sfx_model_timestep_analysis_per_cell(const sfx_model* M, const float* tsa_sum, const int32_t* tsa_wet,
                                                    const int32_t* tsa_lim, int64_t nt, float* dt_cell,
                                                    float* pct_cell) {
  if (!M || !M->initialized || !tsa_sum || !tsa_wet || !tsa_lim || !dt_cell || !pct_cell || nt <= 0) {
    set_error("sfx_model_timestep_analysis_per_cell: bad args");
    return SFX_ERR_ARG;
  }
  const std::vector<int32_t>* nbr[8] = {&M->z_index_uv_md1, &M->z_index_uv_md2, &M->z_index_uv_mu1, &M->z_index_uv_mu2,
                                        &M->z_index_uv_nd1, &M->z_index_uv_nd2, &M->z_index_uv_nu1, &M->z_index_uv_nu2};
  for (int32_t nm = 0; nm < M->np; ++nm) {
    dt_cell[nm] = 0.0f; pct_cell[nm] = 0.0f;
    if (M->kcs[nm] <= 0) continue;
    float dtm = 1.0e6f;
    int32_t tmsl = 0;
    for (int j = 0; j < 8; ++j) {
      const int32_t ip = (*nbr[j])[nm];
      if (ip <= 0 || !(M->kcuv[ip - 1] == 1 || M->kcuv[ip - 1] == 6)) continue;
      dtm = std::min(dtm, tsa_sum[ip - 1] / (float)std::max(tsa_wet[ip - 1], 1));
      tmsl = std::max(tmsl, tsa_lim[ip - 1]);
    }
    dt_cell[nm] = dtm < 1.0e5f ? std::min(dtm * M->p.alpha, M->dtmax) : -1.0f;
    pct_cell[nm] = 100.0f * (float)tmsl / (float)nt;
  }
  return SFX_OK;
}

// The active cell holding point (x, y), 1-based, 0 outside the mesh: the lookup of the
// BMI get_sfincs_cell_index (sfincs_bmi.f90:299-316, find_quadtree_cell).
extern "C" int sfx_model_find_cell(const sfx_model* M, float x, float y, int32_t* idx) {
  if (!M || !M->mesh_built || !idx) { set_error("sfx_model_find_cell: bad args"); return SFX_ERR_ARG; }
  *idx = M->find_cell_xy(x, y);
  return SFX_OK;
}

extern "C" int sfx_model_get_state(const sfx_model* M, const char* name, void* dst, size_t nbytes) {
  return sfx_model_get_table(M, name, dst, nbytes);
}
