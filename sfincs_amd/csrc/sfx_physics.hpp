// Per-point physics of the explicit step, shared by every kernel variant (dense/indexed,
// regular/subgrid, FP32/FP64). Operation order and types follow the reference exactly:
// real*4 arithmetic in R, zs and z_volume in double, mixed expressions promoted where
// the reference promotes them.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "sfx_libm.hpp"

namespace sfx {

enum : uint32_t {
  F_ADVECTION = 1u << 0, F_UPW1 = 1u << 1, F_VISCOSITY = 1u << 2, F_CORIOLIS = 1u << 3,
  F_FRICTION2D = 1u << 4, F_THETA = 1u << 5, F_WIGGLE = 1u << 6, F_ZSMAX = 1u << 7,
  F_WIND = 1u << 8, F_PRECIP = 1u << 9, F_H73 = 1u << 10, F_QUADTREE = 1u << 11,
  F_SUBGRID = 1u << 12, F_NBRS = 1u << 13,
  // compute_store_variables switches (sfincs_continuity.f90:626-720)
  F_VELMAX = 1u << 14, F_FLUXMAX = 1u << 15, F_TWET = 1u << 16,
  F_STORE = F_VELMAX | F_FLUXMAX | F_TWET,
  F_TSA = 1u << 17,  // timestep_analysis (sfincs_timestep_analysis.f90)
  F_QEXT = 1u << 18, // external sources per cell (use_qext, BMI coupling)
  F_STORAGE = 1u << 19, // storage volume per subgrid cell (use_storage_volume)
  F_INFIL = 1u << 20,   // constant infiltration per cell, uniform or mapped (sfincs_infiltration.f90)
  // gridded / spiderweb meteo (sfincs_meteo.f90): update windows, wind stress per cell,
  // atmospheric pressure (baro), rain per cell, spin-up of the gridded fields
  F_METEO3D = 1u << 21, F_MWIND = 1u << 22, F_PATM = 1u << 23, F_MPRCP = 1u << 24, F_MSPIN = 1u << 25,
  F_CNA = 1u << 26,     // curve-number infiltration, method A (qinffield holds S; cumulative rain and infiltration)
};

constexpr int MAXREF = 8;

// Device arrays in the global address space, indexed by 32-bit unsigned element numbers:
// an access compiles to a global load or store with the array base in scalar registers
// and one 32-bit byte offset per lane (every array is smaller than 4 GiB; sfx_create
// checks). Used for all device memory reached through pointers held in structures,
// which the compiler would otherwise access with flat instructions and 64-bit per-lane
// addresses.
template <typename T> using gptr = __attribute__((address_space(1))) T*;
template <typename T>
struct GA {
  gptr<T> p;
  __device__ __forceinline__ __attribute__((address_space(1))) T& operator[](uint32_t i) const {
    return *(gptr<T>)((gptr<char>)(gptr<void>)p + i * (uint32_t)sizeof(T));
  }
  __device__ __forceinline__ GA operator+(uint32_t i) const {
    return GA{(gptr<T>)((gptr<char>)(gptr<void>)p + i * (uint32_t)sizeof(T))};
  }
  __device__ __forceinline__ gptr<T> operator->() const { return p; }
};
template <typename T> __device__ __forceinline__ GA<T> ga(T* p) { return GA<T>{(gptr<T>)p}; }

// A row of a level table (z_dep, uv_lev): element `row + i` of the array, so that the
// access keeps the scalar base and a 32-bit offset.
template <typename T>
struct GRow {
  GA<T> base;
  uint32_t row;
  __device__ __forceinline__ T operator[](uint32_t i) const { return base[row + i]; }
};

// 16-byte loads of float4 tables.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldf4(const float4* p, size_t i) { return p[i]; }
template <typename T>
__device__ __forceinline__ float4 ldf4(GA<T> p, uint32_t i) {
  const f4v v = *(gptr<const f4v>)(gptr<const void>)(p + i).p;
  return make_float4(v.x, v.y, v.z, v.w);
}
template <typename T>
__device__ __forceinline__ float4 ldf4(GRow<T> r, uint32_t i) { return ldf4(r.base, r.row + i); }

// Physics constants of one scenario (sfincs_input.f90 defaults + derived values).
struct Consts {
  float g, huthresh, huvmin, advlim, slopelim, uvlim, nuviscfac, fcorio, theta;
  float wiggle_factor, wiggle_threshold, dtmax, dtmin, alpha, rhoa, rhow, btfilter, btrelax;
  float t0, t1, tspinup, zini, dx, dy, pi, twet_threshold;
  float dtwnd, gapres, pavbnd;   // meteo update interval, background pressure, boundary pressure
  float sfacinf;                 // curve-number initial abstraction, fraction of S
  float dxrinv[MAXREF], dyrinv[MAXREF], dxr2inv[MAXREF], dyr2inv[MAXREF], dxrinvc[MAXREF], dyrinvc[MAXREF];
  float cell_area[MAXREF], dxrm[MAXREF], dyrm[MAXREF], nuvisc[MAXREF];
  uint32_t flags;
  int32_t nl;
};

template <typename T> __device__ __forceinline__ T vmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T vmin(T a, T b) { return a < b ? a : b; }

template <typename R> __device__ __forceinline__ R rpow(R x, R y);
template <> __device__ __forceinline__ float rpow<float>(float x, float y) { return powf(x, y); }
template <> __device__ __forceinline__ double rpow<double>(double x, double y) { return pow(x, y); }
// x**(float)(1/3), x**(float)(5/3), x**(float)(7/3) and x**1.5: the C library's exact
// powf results in FP32, ordinary double pow in the FP64 twin.
// Device view of the powf tables: 64-bit entries {key, value} in the global address space.
struct PowTablesG { GA<const uint64_t> t13, t53, t73, t15; };
__device__ __forceinline__ PowTablesG pow_tables_g(const PowTables& T) {
  return PowTablesG{ga((const uint64_t*)T.t13), ga((const uint64_t*)T.t53), ga((const uint64_t*)T.t73),
                    ga((const uint64_t*)T.t15)};
}
template <typename R> __device__ __forceinline__ R pow_third(R x, int which, const PowTablesG& T);
template <> __device__ __forceinline__ float pow_third<float>(float x, int which, const PowTablesG& T) {
  return libm_pow_dev(x, which, T.t13, T.t53, T.t73, T.t15);
}
template <> __device__ __forceinline__ double pow_third<double>(double x, int which, const PowTablesG& T) {
  (void)T;
  return pow(x, which == 0 ? (double)(1.0f / 3.0f)
                           : (which == 1 ? (double)(5.0f / 3.0f) : (which == 2 ? (double)(7.0f / 3.0f) : 1.5)));
}
template <typename R> __device__ __forceinline__ R rsqrt_(R x);
template <> __device__ __forceinline__ float rsqrt_<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double rsqrt_<double>(double x) { return sqrt(x); }
template <typename R> __device__ __forceinline__ R rabs(R x) { return x < R(0) ? -x : x; }

// Inputs of one uv-point update gathered by the caller (sfincs_momentum.f90:159-340).
template <typename R>
struct UvIn {
  double zs_nm, zs_nmu;        // water levels of the two cells (real*8)
  double vol_nm, vol_nmu;      // subgrid cell volumes (real*8)
  R zsderv_nm, zsderv_nmu;     // subgrid wiggle detector
  float zb_nm, zb_nmu;         // regular bed levels (dry guard)
  R qx_nm, qx_nmd, qx_nmu, qy_nm, qy_nmu, qy_ndm, qy_ndmu;
  R uu_nm, uu_nmd, uu_nmu, uu_ndm, uu_num, v_ndm, v_ndmu, v_nm, v_nmu;
  int kcuv_nmd, kcuv_nmu;      // theta smoothing test
  int kfuv, mask_adv, idir, iref, itype;
  R tauwu, tauwv;              // wind stress at cell nm
  R dpatm;                     // patm(nm) - patm(nmu), the pressure difference across the point (F_PATM)
};

struct UvTables {              // per-point tables (regular or subgrid)
  float zbuvmx, gn2uv;         // regular
  float zmin, zmax, havg_zmax, nrep_zmax, navg_w, fnfit;   // subgrid
};

template <typename R>
struct UvOut {
  R q, uv;
  R den;        // the point's CFL denominator: its required step is 1 / den (:731-733)
  int kfuv;
  int branch;   // 1 dry, 2 wet regular / subgrid full, 3 subgrid partial
};

// Wet test (:168-187). Returns true when the point is wet; zsu is real*4 max(zs(nm),zs(nmu)).
template <typename R, bool SUBGRID>
__device__ __forceinline__ bool uv_wet(const UvIn<R>& in, const UvTables& tb, const Consts& c, R& zsu) {
  zsu = (R)vmax(in.zs_nm, in.zs_nmu);
  if (SUBGRID) return zsu > (R)tb.zmin + (R)c.huthresh;
  return zsu > (R)tb.zbuvmx;
}

// h**(7/3) table lookup (power7over3, sfincs_momentum.f90:787-820)
template <typename R, typename P>
__device__ __forceinline__ R power7over3(P x73, R hu) {
  const float h = (float)hu;
  if (h < 0.00001f) return x73[0 * 1000 + (int)(1e8f * h) - 1];
  if (h < 0.0001f) return x73[1 * 1000 + (int)(1e7f * h) - 1];
  if (h < 0.001f) return x73[2 * 1000 + (int)(1e6f * h) - 1];
  if (h < 0.01f) return x73[3 * 1000 + (int)(1e5f * h) - 1];
  if (h < 0.1f) return x73[4 * 1000 + (int)(1e4f * h) - 1];
  if (h < 1.0f) return x73[5 * 1000 + (int)(1e3f * h) - 1];
  if (h < 10.0f) return x73[6 * 1000 + (int)(100 * h) - 1];
  if (h < 100.0f) return x73[7 * 1000 + (int)(10 * h) - 1];
  if (h < 1000.0f) return x73[8 * 1000 + (int)(1 * h) - 1];
  if (h < 10000.0f) return x73[9 * 1000 + (int)(0.1f * h) - 1];
  return rpow<R>(R(10000.0f), R(7.0f / 3.0f));
}

// One wet uv-point (sfincs_momentum.f90:189-742). `lev` points at the 2 levels
// {havg,nrep,pwet,-} bracketing zsu for partially wet subgrid points.
template <typename R, bool SUBGRID, typename LevP, typename X73P>
__device__ __forceinline__ UvOut<R> uv_update(const UvIn<R>& in, const UvTables& tb, const Consts& c, R zsu,
                                              LevP lev_row, X73P x73, const PowTablesG& pw, R dt) {
  const uint32_t F = c.flags;
  UvOut<R> o;
  const int iref = in.iref - 1;
  R dxuvinv, dxuv2inv, dyuv2inv;
  R dyuvinv;
  if (in.itype == 0) {
    if (in.idir == 0) { dxuvinv = c.dxrinv[iref]; dyuvinv = c.dyrinv[iref]; dxuv2inv = c.dxr2inv[iref]; dyuv2inv = c.dyr2inv[iref]; }
    else { dxuvinv = c.dyrinv[iref]; dyuvinv = c.dxrinv[iref]; dxuv2inv = c.dyr2inv[iref]; dyuv2inv = c.dxr2inv[iref]; }
  } else {
    if (in.idir == 0) { dxuvinv = c.dxrinvc[iref]; dyuvinv = c.dyrinv[iref]; dxuv2inv = R(0); dyuv2inv = c.dyr2inv[iref]; }
    else { dxuvinv = c.dyrinvc[iref]; dyuvinv = c.dxrinv[iref]; dxuv2inv = R(0); dyuv2inv = c.dxr2inv[iref]; }
  }
  const R qx_nm = in.qx_nm;
  R vu = R(0);
  if (F & F_NBRS) vu = (in.v_ndm + in.v_ndmu + in.v_nm + in.v_nmu) / R(4);
  R phi = R(1), hu, gnavg2;
  o.branch = 2;
  if (SUBGRID) {
    const R zmin = tb.zmin, zmax = tb.zmax;
    if (zsu > zmax) {
      hu = (R)tb.havg_zmax + zsu;
      const R navg = tb.navg_w;
      gnavg2 = navg - (navg - (R)tb.nrep_zmax) / ((R)tb.fnfit * (zsu - zmax) + R(1));
    } else {
      o.branch = 3;
      const int nl = c.nl;
      const R dzuv = (zmax - zmin) / (R)(nl - 1);
      int iuv = (int)((zsu - zmin) / dzuv) + 1;
      if (iuv > nl - 1) iuv = nl - 1;
      const R facint = (zsu - (zmin + (R)(iuv - 1) * dzuv)) / dzuv;
      const float4 a = ldf4(lev_row, iuv - 1);
      const float4 b = ldf4(lev_row, iuv);
      hu = (R)a.x + ((R)b.x - (R)a.x) * facint;
      gnavg2 = (R)a.y + ((R)b.y - (R)a.y) * facint;
      phi = (R)a.z + ((R)b.z - (R)a.z) * facint;
    }
  } else {
    hu = vmax<R>(zsu - (R)tb.zbuvmx, (R)c.huthresh);
    gnavg2 = tb.gn2uv;
  }
  R dzdx;
  if (c.slopelim < 9999.0f) {
    double d = (in.zs_nmu - in.zs_nm) * (double)dxuvinv;
    if (d < (double)(-c.slopelim)) d = (double)(-c.slopelim);
    if (d > (double)c.slopelim) d = (double)c.slopelim;
    dzdx = (R)d;
  } else {
    dzdx = (R)((in.zs_nmu - in.zs_nm) * (double)dxuvinv);
  }
  const R g = c.g;
  R frc = -(g * hu * dzdx);
  if ((F & F_ADVECTION) && in.mask_adv == 1) {
    R dqxudx = R(0), dqyudy = R(0);
    const R qx_nmd = in.qx_nmd, qx_nmu = in.qx_nmu, qy_nm = in.qy_nm, qy_nmu = in.qy_nmu;
    const R qy_ndm = in.qy_ndm, qy_ndmu = in.qy_ndmu;
    const R uu_nm = in.uu_nm, uu_nmd = in.uu_nmd, uu_nmu = in.uu_nmu, uu_ndm = in.uu_ndm, uu_num = in.uu_num;
    const R eps = R(1.0e-6f);
    if (F & F_UPW1) {
      R qd = (qx_nmd + qx_nm) / R(2);
      R qu = (qx_nm + qx_nmu) / R(2);
      if (qd > eps) {
        const R ud = (uu_nmd + uu_nm) / R(2);
        dqxudx = (qd * (uu_nm - uu_nmd) + ud * (qx_nm - qx_nmd)) * dxuvinv;
      }
      if (qu < -eps) {
        const R uu = (uu_nm + uu_nmu) / R(2);
        dqxudx = dqxudx + (qu * (uu_nmu - uu_nm) + uu * (qx_nmu - qx_nm)) * dxuvinv;
      }
      qu = (qy_nm + qy_nmu) / R(2);
      qd = (qy_ndm + qy_ndmu) / R(2);
      if (qd > eps) dqyudy = qd * (uu_nm - uu_ndm) * dyuvinv;
      if (qu < -eps) dqyudy = dqyudy + qu * (uu_num - uu_nm) * dyuvinv;
      const R ud = (uu_nmd + uu_nm) / R(2);
      const R uu = (uu_nm + uu_nmu) / R(2);
      if (ud > eps) dqyudy = dqyudy + ud * (qy_nm - qy_ndm) * dyuvinv;
      if (uu < -eps) dqyudy = dqyudy + uu * (qy_nmu - qy_ndmu) * dyuvinv;
    } else {
      if (qx_nm > eps) dqxudx = (qx_nm * uu_nm - qx_nmd * uu_nmd) * dxuvinv;
      else if (qx_nm < -eps) dqxudx = (qx_nmu * uu_nmu - qx_nm * uu_nm) * dxuvinv;
      const R qy = qy_ndm + qy_ndmu + qy_nm + qy_nmu;
      if (qy > eps) dqyudy = (qy_ndm + qy_ndmu) * (uu_nm - uu_ndm) * dyuvinv / R(2);
      else if (qy < -eps) dqyudy = (qy_nm + qy_nmu) * (uu_num - uu_nm) * dyuvinv / R(2);
    }
    R adv = -(phi * (dqxudx + dqyudy));
    adv = vmin<R>(vmax<R>(adv, -((R)c.advlim * hu)), (R)c.advlim * hu);
    frc = frc + adv;
  }
  if (F & F_VISCOSITY) {
    const R lap = (in.uu_nmu - R(2) * in.uu_nm + in.uu_nmd) * dxuv2inv + (in.uu_num - R(2) * in.uu_nm + in.uu_ndm) * dyuv2inv;
    if (in.itype == 0) frc = frc + (R)c.nuvisc[iref] * hu * lap;
    else frc = frc + (R)c.nuviscfac * (R)c.nuvisc[iref] * hu * lap;
  }
  if (F & F_CORIOLIS) {
    if (in.idir == 0) frc = frc + (R)c.fcorio * hu * vu;
    else frc = frc - (R)c.fcorio * hu * vu;
  }
  if (F & F_WIND) {
    const R tw = in.idir == 0 ? in.tauwu : in.tauwv;
    const R hwet = hu / phi;
    if (hwet > R(0.25f)) frc = frc + phi * tw;
    else frc = frc + tw * hu * R(4);
  }
  if (F & F_PATM) frc = frc + hu * in.dpatm * dxuvinv / (R)c.rhow;   // :585-589
  R qfr;
  if (in.kfuv == 0) {
    qfr = rsqrt_<R>(rabs<R>(dzdx) / (vmax<R>(gnavg2, R(1.0e-5f)) / R(10))) * pow_third<R>(hu, 1, pw);
  } else if (F & F_FRICTION2D) {
    qfr = rsqrt_<R>(qx_nm * qx_nm + (hu * vu) * (hu * vu));
  } else {
    qfr = rabs<R>(qx_nm);
  }
  R qsm = qx_nm;
  if (F & F_THETA) {
    if (in.kcuv_nmd == 1 && in.kcuv_nmu == 1)
      if (rabs<R>(in.qx_nmd) > R(1.0e-6f) && rabs<R>(in.qx_nmu) > R(1.0e-6f))
        qsm = (R)c.theta * qx_nm + R(0.5f) * (R(1) - (R)c.theta) * (in.qx_nmu + in.qx_nmd);
  }
  const R hu73 = (F & F_H73) ? power7over3<R>(x73, vmax<R>(hu, R(1.0e-6f))) : hu * hu * pow_third<R>(hu, 0, pw);
  R qn = (qsm + frc * dt) / (R(1) + gnavg2 * dt * qfr / hu73);
  if (SUBGRID) {
    if (F & F_WIGGLE) {
      const R mdrv = rabs<R>(in.zsderv_nm - in.zsderv_nmu) - (R)c.wiggle_threshold;
      if (mdrv > R(0)) qn = qn * (R)c.wiggle_threshold / ((R)c.wiggle_factor * mdrv + (R)c.wiggle_threshold);
    }
    if (in.vol_nm < 0.0) qn = vmin<R>(qn, R(0));
    if (in.vol_nmu < 0.0) qn = vmax<R>(qn, R(0));
  } else {
    if (in.zs_nm < (double)in.zb_nm) qn = vmin<R>(qn, R(0));
    if (in.zs_nmu < (double)in.zb_nmu) qn = vmax<R>(qn, R(0));
  }
  qn = vmin<R>(vmax<R>(qn, -(hu * (R)c.uvlim)), hu * (R)c.uvlim);
  o.q = qn;
  o.uv = qn / vmax<R>(hu, (R)c.huvmin);
  o.kfuv = 1;
  o.den = vmax<R>(rsqrt_<R>(g * hu), rabs<R>(o.uv)) * dxuvinv;
  return o;
}

// Cell update of the regular continuity (sfincs_continuity.f90:112-155, 239-255).
template <typename R>
__device__ __forceinline__ void cell_regular(double& zs, R& zsmax, int kcs, R q_md, R q_mu, R q_nd, R q_nu,
                                             R netprcp, R qext, int iref, const Consts& c, R dt) {
  if (kcs == 1) {
    if (c.flags & F_PRECIP) zs = zs + (double)(netprcp * dt);
    if (c.flags & F_QEXT) zs = zs + (double)(qext * dt);   // :122-128
    const R d = ((q_md - q_mu) * (R)c.dxrinv[iref - 1] + (q_nd - q_nu) * (R)c.dyrinv[iref - 1]) * dt;
    zs = zs + (double)d;
  }
  if (c.flags & F_ZSMAX) zsmax = (R)vmax<double>((double)zsmax, zs);
}

// Cell update of the subgrid continuity (sfincs_continuity.f90:337-618).
template <typename R, typename DepP>
__device__ __forceinline__ void cell_subgrid(double& zs, double& vol, R& zs0, R& zsderv, R& zsmax, int kcs,
                                             R q_md, R q_mu, R q_nd, R q_nu, R netprcp, R qext, R& stor, int iref,
                                             float zmin, float zmax, float volmax, DepP dep,
                                             const Consts& c, R dt) {
  R dvol = R(0);
  if (kcs == 1) {
    if (c.flags & F_QUADTREE)
      dvol = dvol + ((q_md - q_mu) * (R)c.dyrm[iref - 1] + (q_nd - q_nu) * (R)c.dxrm[iref - 1]) * dt;
    else
      dvol = dvol + ((q_md - q_mu) * (R)c.dy + (q_nd - q_nu) * (R)c.dx) * dt;
  }
  if (kcs == 1 || kcs == 4) {
    const R a = c.cell_area[iref - 1];
    if (c.flags & (F_PRECIP | F_QEXT)) {   // :473-497
      R dzsdt = R(0);
      if (c.flags & F_PRECIP) dzsdt = dzsdt + netprcp;
      if (c.flags & F_QEXT) dzsdt = dzsdt + qext;
      dvol = dvol + dzsdt * a * dt;
    }
    if ((c.flags & F_STORAGE) && stor > R(1.0e-6f) && dvol > R(0)) {   // :499-523: storage fills first
      const R dv = stor - dvol;
      stor = vmax<R>(dv, R(0));
      dvol = dv < R(0) ? -dv : R(0);
    }
    vol = vol + (double)dvol;
    R zs00 = R(0), zs11 = R(0);
    if (c.flags & F_WIGGLE) { zs00 = zs0; zs11 = (R)zs; zs0 = zs11; }
    const float vmx = volmax;
    if (vol >= (double)(vmx * 0.999f)) {
      zs = (double)vmax<float>(zmax, -20.0f) + (vol - (double)vmx) / (double)a;
    } else if (vol <= (double)1.0e-6f) {
      zs = (double)vmax<float>(zmin, -20.0f);
    } else {
      const int nl = c.nl;
      const R dzvol = (R)vmx / (R)(nl - 1);
      const int iuv = (int)(vol / (double)dzvol) + 1;
      const R facint = (R)((vol - (double)((R)(iuv - 1) * dzvol)) / (double)dzvol);
      const R d0 = dep[iuv - 1], d1 = dep[iuv];
      zs = (double)(d0 + (d1 - d0) * facint);
    }
    if (c.flags & F_WIGGLE) zsderv = (R)((zs - (double)(R(2) * zs11)) + (double)zs00);
  }
  if (c.flags & F_ZSMAX) zsmax = (R)vmax<double>((double)zsmax, zs);
}

}  // namespace sfx
