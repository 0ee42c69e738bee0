// x**(1/3), x**(5/3), x**(7/3) and x**1.5 with the exact float results of the C library's
// powf.
//
// The reference computes hu**expo (expo = 1.0/3.0 as real*4) and hu**(5.0/3.0)
// (sfincs_momentum.f90:92,616,673), hgate**(7.0/3.0) for drainage gates
// (sfincs_discharges.f90:518,610) and h1**1.5 for weirs (sfincs_structures.f90:672), with
// the C library's powf. The momentum and continuity
// updates branch on thresholds (wet/dry tests, the table inversion switch at
// V = 0.999*volmax), so a one-ulp difference in powf can move a cell to the other branch.
// These routines return powf's exact result: a double-precision estimate (pow_estimate)
// is rounded to float unless it lies within 2^-32 (relative) of a rounding midpoint;
// those inputs (about 0.5% of calls) are looked up in a table of powf's own results that
// sfx_libm.hip enumerates over every float in [2^-40, 2^14) and verifies. Outside that
// range the rounded estimate is returned.
#pragma once
#include <cmath>
#include <cstdint>

#ifndef SFX_HD
#define SFX_HD __host__ __device__
#endif

namespace sfx {

constexpr int POW_HASH_BITS = 20;
constexpr uint32_t POW_HASH_SIZE = 1u << POW_HASH_BITS;
constexpr float POW_XLO = 9.094947017729282e-13f;   // 2^-40
constexpr float POW_XHI = 16384.0f;                 // 2^14

constexpr int NPOW = 4;   // exponents (float)(1/3), (float)(5/3), (float)(7/3), 1.5

// cosf and sinf of the wind direction (sfincs_meteo.f90:1475-1478), exact like powf for
// every float in [0, TRIG_XHI): the direction is the interpolation of two angles reduced
// modulo 2 pi, one of them raised by 2 pi, so it lies in [0, 4 pi]. glibc's cosf/sinf
// round up to 2^-27.9 (relative) away from a midpoint the other way, so their band is
// 2^-27 and their tables hold about 0.6M and 1.1M entries: hashes of 2^22 slots.
constexpr float TRIG_XHI = 12.6f;
constexpr int TRIG_HASH_BITS = 22;
constexpr uint32_t TRIG_HASH_SIZE = 1u << TRIG_HASH_BITS;

struct PowTables {
  const uint2* t13;   // open-addressing hash: {x bits, powf(x, 1/3f) bits}; key 0 = empty
  const uint2* t53;
  const uint2* t73;
  const uint2* t15;
  const uint2* tcos;  // TRIG_HASH_SIZE slots each
  const uint2* tsin;
};

SFX_HD inline uint32_t hash_bits(uint32_t k) {
  k ^= k >> 16; k *= 0x7feb352dU; k ^= k >> 15; k *= 0x846ca68bU; k ^= k >> 16;
  return k;
}
SFX_HD inline uint32_t pow_hash(uint32_t k) { return hash_bits(k) & (POW_HASH_SIZE - 1); }

SFX_HD inline uint32_t f2u(float f) { union { float f; uint32_t u; } v; v.f = f; return v.u; }
SFX_HD inline float u2f(uint32_t u) { union { float f; uint32_t u; } v; v.u = u; return v.f; }

// Double estimate of x**e for e = (float)(1/3) (which = 0), (float)(5/3) (which = 1),
// (float)(7/3) (which = 2) or 1.5 (which = 3, x sqrt(x) in double),
// relative error below 2^-46 (2^-44 for 7/3) for normal positive x (checked for every float of the table
// range when the tables are built). Only IEEE-754 basic operations in double are used,
// evaluated in the order written, so host and device compute bit-identical estimates.
//   x = m * 2^e, m in [1, 2), e = 3q + r:  x^(1/3) = (m * 2^r)^(1/3) * 2^q, from a
//   polynomial guess of (m 2^r)^(-1/3) refined by two Newton steps z += z (1 - a z^3) / 3
//   (no division); the float exponents differ from 1/3 and 5/3 by d, applied as
//   x^d = 1 + t + t^2/2 with t = d ln x, ln x = e ln 2 + ln m from a polynomial.
SFX_HD inline double pow_estimate(float x, int which) {
  if (which == 3) return (double)x * sqrt((double)x);
  const uint32_t b = f2u(x);
  if (b < 0x00800000u) {   // subnormal: outside the table range, no exactness claimed
    const double c = (double)cbrtf(x);
    return which == 0 ? c : (which == 1 ? (double)x * c * c : (double)x * (double)x * c);
  }
  const int e = (int)(b >> 23) - 127;
  const double m = (double)u2f((b & 0x7fffffu) | 0x3f800000u);
  const int q = e >= 0 ? e / 3 : -((2 - e) / 3);
  const int r = e - 3 * q;
  const double a = m * (double)(1 << r);
  const double cr = r == 0 ? 1.0 : (r == 1 ? 0.79370052598409979 : 0.62996052494743658);   // 2^(-r/3)
  // m^(-1/3) on [1, 2], relative error 3.8e-5
  double z = (((0.0285876759712902 * m - 0.22121268212268402) * m + 0.6954153443691581) * m -
              1.1728037038550854) * m + 1.6699753569259383;
  z = z * cr;
  const double third = 1.0 / 3.0;
  z = z + z * (1.0 - a * z * z * z) * third;
  z = z + z * (1.0 - a * z * z * z) * third;
  // 2^q as a double (exact power of two; q is within [-43, 42] for normal floats)
  union { uint64_t u; double d; } p2;
  p2.u = (uint64_t)(q + 1023) << 52;
  const double y = a * z * z * p2.d;
  // ln x, with m folded into [sqrt(1/2), sqrt(2)); ln(1+u) on that range, abs error 1.9e-7
  double mm = m;
  int ee = e;
  if (mm > 1.4142135623730951) { mm = mm * 0.5; ee = ee + 1; }
  const double u = mm - 1.0;
  const double lnm = ((((((0.11860324847309349 * u - 0.1896650762497095) * u + 0.20616715842644762) * u -
                         0.24870626475368562) * u + 0.333012775052759) * u - 0.5000253637544767) * u +
                      1.0000038762934083) * u + 9.149265129190923e-08;
  const double lnx = (double)ee * 0.69314718055994530942 + lnm;
  // offsets of the float exponents from 1/3 and 5/3
  const double d13 = (double)(1.0f / 3.0f) - 1.0 / 3.0;
  const double d53 = (double)(5.0f / 3.0f) - 5.0 / 3.0;
  const double d73 = (double)(7.0f / 3.0f) - 7.0 / 3.0;
  const double t = (which == 0 ? d13 : (which == 1 ? d53 : d73)) * lnx;
  const double corr = 1.0 + t + 0.5 * t * t;
  if (which == 0) return y * corr;
  if (which == 1) return (double)x * y * y * corr;
  return (double)x * (double)x * y * corr;
}

// The device's estimate of the same powers: the cube root's polynomial and Newton steps as
// fused multiply-adds, and ln x for the float exponents' offsets from the hardware log2 of
// the mantissa (v_log_f32, absolute error near 2^-23 on [1, 2), which the offsets of at most
// 8e-8 turn into 2^-47 relative), about half the double operations of pow_estimate. Its
// rounding decisions need it within 2^-40 of the exact power: a device test checks it
// against pow_estimate (itself verified within 2^-44 of pow for every table input) for
// every float of [2^-40, 2^14) and each exponent.
__device__ inline double pow_estimate_dev(float x, int which) {
  if (which == 3) return (double)x * sqrt((double)x);
  const uint32_t b = f2u(x);
  if (b < 0x00800000u) return pow_estimate(x, which);
  const int e = (int)(b >> 23) - 127;
  const float mf = u2f((b & 0x7fffffu) | 0x3f800000u);
  const double m = (double)mf;
  const int q = e >= 0 ? e / 3 : -((2 - e) / 3);
  const int r = e - 3 * q;
  const double a = m * (double)(1 << r);
  const double cr = r == 0 ? 1.0 : (r == 1 ? 0.79370052598409979 : 0.62996052494743658);
  double z = fma(fma(fma(fma(0.0285876759712902, m, -0.22121268212268402), m, 0.6954153443691581), m,
                     -1.1728037038550854), m, 1.6699753569259383);
  z = z * cr;
  const double third = 1.0 / 3.0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double w = a * z * z;
    z = fma(z * third, fma(-w, z, 1.0), z);
  }
  union { uint64_t u; double d; } p2;
  p2.u = (uint64_t)(q + 1023) << 52;
  const double y = a * z * z * p2.d;
  const double lnx = ((double)e + (double)__builtin_amdgcn_logf(mf)) * 0.69314718055994530942;
  const double d13 = (double)(1.0f / 3.0f) - 1.0 / 3.0;
  const double d53 = (double)(5.0f / 3.0f) - 5.0 / 3.0;
  const double d73 = (double)(7.0f / 3.0f) - 7.0 / 3.0;
  const double t = (which == 0 ? d13 : (which == 1 ? d53 : d73)) * lnx;
  const double corr = fma(t, fma(0.5, t, 1.0), 1.0);
  if (which == 0) return y * corr;
  if (which == 1) return (double)x * y * y * corr;
  return (double)x * (double)x * y * corr;
}

// True when the double value r lies within 2^-32 (relative) of a midpoint between two
// adjacent floats, i.e. when its float rounding is not certain. r must be a positive
// double in the normal float range. On the bits of r = M 2^(E-52) (M the 53-bit
// significand): float midpoints sit at offset 2^28 of every block of 2^29 units of
// 2^(E-52) (the 29 significand bits below float precision), so |r - mid| =
// |low29 - 2^28| units, and 2^-32 r = M / 2^32 units = H + lo / 2^32 with H the top 21
// bits of M and lo the low word. An integer k is below H + lo / 2^32 when k < H, or
// k == H and lo != 0.
SFX_HD inline bool near_midpoint(double r) {
  union { double d; uint64_t u; } v;
  v.d = r;
  const uint32_t lo = (uint32_t)v.u, hi = (uint32_t)(v.u >> 32);
  const uint32_t low29 = lo & 0x1fffffffu;
  const uint32_t k = low29 >= 0x10000000u ? low29 - 0x10000000u : 0x10000000u - low29;
  const uint32_t H = 0x100000u | (hi & 0xfffffu);
  return k < H || (k == H && lo != 0u);
}

// The same test evaluated in double arithmetic (kept for the host-side check of the bit
// form above).
SFX_HD inline bool near_midpoint_fp(double r) {
  const float f = (float)r;
  const double fd = f;
  const float nb = r >= fd ? nextafterf(f, INFINITY) : nextafterf(f, 0.0f);
  const double mid = 0.5 * (fd + (double)nb);
  const double d = r > mid ? r - mid : mid - r;
  return d < 2.3283064365386963e-10 * r;             // 2^-32
}

// Table entries read as 64-bit words: key in the low half, powf bits in the high half
// (the little-endian layout of uint2{key, value}). `mask` = slots - 1.
template <typename P>
SFX_HD inline float pow_lookup(P t, uint32_t key, float fallback, uint32_t mask = POW_HASH_SIZE - 1) {
  uint32_t h = hash_bits(key) & mask;
  for (int probe = 0; probe < 64; ++probe) {
    const uint64_t e = t[h];
    const uint32_t k = (uint32_t)e;
    if (k == key) return u2f((uint32_t)(e >> 32));
    if (k == 0u) break;
    h = (h + 1) & mask;
  }
  return fallback;
}

// Double estimate of cos(x) (which = 0) or sin(x) (which = 1) for a float x >= 0 (below
// 2^23 pi/2; the tables cover [0, TRIG_XHI)), relative error below 2^-46 (checked for
// every float of the table range when the tables are built). x = k pi/2 + r with
// pi/2 = P1 + P2 + P3 (P1, P2 of 30 bits: k P1 and k P2 are exact and so is x - k P1,
// both being multiples of P1's last bit), |r| <= pi/4, then Taylor series of sin and
// cos in r^2 to the r^19 and r^20 terms (truncation below 2^-60 on |r| <= pi/4).
// IEEE basic operations in the order written, so host and device agree to the bit.
SFX_HD inline double trig_estimate(float x, int which) {
  const double xd = (double)x;
  const double k = floor(xd * 0.6366197723675814 + 0.5);
  const double P1 = 1.5707963276654482, P2 = -8.705515692000731e-10, P3 = -3.50343439808993e-19;
  const double r = ((xd - k * P1) - k * P2) - k * P3;
  const double r2 = r * r;
  const double sp = ((((((((-8.22063524662433e-18 * r2 + 2.8114572543455206e-15) * r2 - 7.647163731819816e-13) * r2 +
                          1.6059043836821613e-10) * r2 - 2.505210838544172e-08) * r2 + 2.7557319223985893e-06) * r2 -
                       0.0001984126984126984) * r2 + 0.008333333333333333) * r2 - 0.16666666666666666);
  const double sn = r + r * r2 * sp;
  const double cp = (((((((((4.110317623312165e-19 * r2 - 1.5619206968586225e-16) * r2 + 4.779477332387385e-14) * r2 -
                           1.1470745597729725e-11) * r2 + 2.08767569878681e-09) * r2 - 2.755731922398589e-07) * r2 +
                        2.48015873015873e-05) * r2 - 0.001388888888888889) * r2 + 0.041666666666666664) * r2 - 0.5);
  const double cs = 1.0 + r2 * cp;
  const int q = ((int)k + (which ? 3 : 0)) & 3;   // sin x = cos(x - pi/2)
  return q == 0 ? cs : (q == 1 ? -sn : (q == 2 ? -cs : sn));
}

// True when |r| lies within 2^-27 (relative) of a midpoint between two adjacent floats,
// on the double's bits as near_midpoint: the threshold is M / 2^27 units of 2^(E-52).
SFX_HD inline bool near_midpoint27(double r) {
  union { double d; uint64_t u; } v;
  v.d = r;
  const uint64_t M = (1ull << 52) | (v.u & ((1ull << 52) - 1));
  const uint32_t low29 = (uint32_t)v.u & 0x1fffffffu;
  const uint32_t k = low29 >= 0x10000000u ? low29 - 0x10000000u : 0x10000000u - low29;
  const uint64_t thr = M >> 27;
  return (uint64_t)k < thr || ((uint64_t)k == thr && (M & ((1ull << 27) - 1)) != 0);
}

// cosf(x) (which = 0) or sinf(x) (which = 1) as the C library computes them, for
// 0 <= x < TRIG_XHI (and -x by symmetry); other arguments get the rounded estimate.
template <typename P>
SFX_HD inline float libm_trig(float x, int which, P tcos, P tsin) {
  const bool neg = (f2u(x) >> 31) != 0u;   // -0.0 included: sinf(-0) = -0
  const float a = u2f(f2u(x) & 0x7fffffffu);
  const double r = trig_estimate(a, which);
  float f = (float)r;
  if (a < TRIG_XHI && near_midpoint27(r)) f = pow_lookup(which ? tsin : tcos, f2u(a), f, TRIG_HASH_SIZE - 1);
  return (neg && which) ? -f : f;
}

// powf(x, (float)(1/3)), powf(x, (float)(5/3)), powf(x, (float)(7/3)) or powf(x, 1.5f)
// as the C library computes it, from the estimate r.
template <typename P>
SFX_HD inline float libm_pow_from(float x, double r, int which, P t13, P t53, P t73, P t15) {
  const float f = (float)r;
  if (x >= POW_XLO && x < POW_XHI && near_midpoint(r))
    return pow_lookup(which == 0 ? t13 : (which == 1 ? t53 : (which == 2 ? t73 : t15)), f2u(x), f);
  return f;
}
template <typename P>
SFX_HD inline float libm_pow(float x, int which, P t13, P t53, P t73, P t15) {
  if (!(x > 0.0f)) return x == 0.0f ? 0.0f : NAN;
  return libm_pow_from(x, pow_estimate(x, which), which, t13, t53, t73, t15);
}
// The same on the device from pow_estimate_dev (what the kernels call).
template <typename P>
__device__ inline float libm_pow_dev(float x, int which, P t13, P t53, P t73, P t15) {
  if (!(x > 0.0f)) return x == 0.0f ? 0.0f : NAN;
  return libm_pow_from(x, pow_estimate_dev(x, which), which, t13, t53, t73, t15);
}
SFX_HD inline float libm_trig(float x, int which, const PowTables& T) {
  return libm_trig(x, which, (const uint64_t*)T.tcos, (const uint64_t*)T.tsin);
}
SFX_HD inline float libm_pow(float x, int which, const PowTables& T) {
  return libm_pow(x, which, (const uint64_t*)T.t13, (const uint64_t*)T.t53, (const uint64_t*)T.t73,
                  (const uint64_t*)T.t15);
}

}  // namespace sfx
