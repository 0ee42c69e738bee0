// Tables of the C library's powf results for the inputs whose float rounding the device
// estimate in sfx_libm.hpp cannot decide (see there). The tables are enumerated once over
// every float in [2^-40, 2^14) for each exponent, verified while they are built, cached
// in a file next to the library, and uploaded once per device.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sfx_internal.hpp"
#include "sfx_libm.hpp"

namespace sfx {

namespace {

constexpr uint32_t MAGIC = 0x5346504Du;   // "SFPM"
constexpr uint32_t VERSION = 9;
constexpr int NFUN = NPOW + 2;   // the four powf exponents, then cosf and sinf

struct Entry { uint32_t x, g; };

struct PowSet {
  std::vector<Entry> e[NFUN];
  std::vector<Entry> probe[NFUN];
  bool ok = false;
};

bool is_trig(int which) { return which >= NPOW; }
uint32_t slots_of(int which) { return is_trig(which) ? TRIG_HASH_SIZE : POW_HASH_SIZE; }
float libc_value(int which, float x);

std::mutex g_mu;
PowSet g_set;
std::map<int, PowTables> g_dev;

float exponent(int which) {
  return which == 0 ? 1.0f / 3.0f : (which == 1 ? 5.0f / 3.0f : (which == 2 ? 7.0f / 3.0f : 1.5f));
}

float libc_value(int which, float x) {
  if (which == NPOW) return cosf(x);
  if (which == NPOW + 1) return sinf(x);
  return powf(x, exponent(which));
}

// cosf / sinf over [lo, hi) bit patterns (non-negative floats), as enumerate below: the
// estimate within 2^-46 of the double result, every misrounding of the C library inside
// the 2^-27 band, and the table holding those and the x whose result lies within 2^-40
// of a midpoint.
bool enumerate_trig(int which, uint32_t lo, uint32_t hi, std::vector<Entry>& out, std::string& err) {
  const int fn = which - NPOW;
  for (uint32_t u = lo; u < hi; ++u) {
    const float x = u2f(u);
    const float g = fn ? sinf(x) : cosf(x);
    const double r = fn ? sin((double)x) : cos((double)x);
    const double est = trig_estimate(x, fn);
    const double ar = std::fabs(r);
    if (std::fabs(est - r) > 1.4210854715202004e-14 * ar) {
      err = "trig estimate error too large at x=" + std::to_string(x);
      return false;
    }
    const float f = (float)r;
    const float af = std::fabs(f);
    const float nb = ar >= (double)af ? nextafterf(af, INFINITY) : nextafterf(af, 0.0f);
    const double mid = 0.5 * ((double)af + (double)nb);
    const double dist = std::fabs(ar - mid);
    if (g != f) {
      if (!(dist < 5.268356063861754e-09 * ar)) {            // 2^-27.5: inside the device band of 2^-27
        err = "cosf/sinf differ outside the midpoint band at x=" + std::to_string(x);
        return false;
      }
      out.push_back(Entry{u, f2u(g)});
    } else if (dist < 9.094947017729282e-13 * ar) {         // 2^-40: estimate may straddle
      out.push_back(Entry{u, f2u(g)});
    }
  }
  return true;
}

// Enumerate one exponent over [lo, hi) bit patterns; returns false if a verification fails.
bool enumerate(int which, uint32_t lo, uint32_t hi, std::vector<Entry>& out, std::string& err) {
  const float y = exponent(which);
  for (uint32_t u = lo; u < hi; ++u) {
    const float x = u2f(u);
    const float g = powf(x, y);
    const double r = pow((double)x, (double)y);
    const float f = (float)r;
    const double est = pow_estimate(x, which);
    // 2^-46 (1/3, 5/3) or 2^-44 (7/3, whose exponent offset scales the error of ln x four
    // times more); rounding decisions only need the error below the 2^-40 straddle guard
    const double tol = which == 2 ? 5.6843418860808015e-14 : 1.4210854715202004e-14;
    if (std::fabs(est - r) > tol * r) {
      err = "estimate error too large at x=" + std::to_string(x);
      return false;
    }
    if (near_midpoint(est) != near_midpoint_fp(est)) {       // bit form of the band test
      err = "midpoint band test disagrees with its double form at x=" + std::to_string(x);
      return false;
    }
    const float fd = f;
    const float nb = r >= (double)fd ? nextafterf(f, INFINITY) : nextafterf(f, 0.0f);
    const double mid = 0.5 * ((double)fd + (double)nb);
    const double dist = std::fabs(r - mid);
    if (g != f) {
      if (!(dist < 1.6462644155254631e-10 * r)) {           // 2^-32.5: inside the device band of 2^-32
        err = "powf differs outside the midpoint band at x=" + std::to_string(x);
        return false;
      }
      out.push_back(Entry{u, f2u(g)});
    } else if (dist < 9.094947017729282e-13 * r) {         // 2^-40: estimate may straddle
      out.push_back(Entry{u, f2u(g)});
    }
  }
  return true;
}

bool generate(PowSet& S, std::string& err) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 4;
  nt = std::min(nt, 64u);
  for (int which = 0; which < NFUN; ++which) {
    const uint32_t lo = is_trig(which) ? 0u : f2u(POW_XLO), hi = is_trig(which) ? f2u(TRIG_XHI) : f2u(POW_XHI);
    std::vector<std::vector<Entry>> parts(nt);
    std::vector<std::string> errs(nt);
    std::vector<char> oks(nt, 1);
    std::vector<std::thread> th;
    const uint32_t span = (hi - lo + nt - 1) / nt;
    for (unsigned k = 0; k < nt; ++k) {
      const uint32_t a = lo + k * span, b = std::min(hi, a + span);
      th.emplace_back([&, k, a, b]() {
        oks[k] = is_trig(which) ? enumerate_trig(which, a, b, parts[k], errs[k]) : enumerate(which, a, b, parts[k], errs[k]);
      });
    }
    for (auto& t : th) t.join();
    S.e[which].clear();
    for (unsigned k = 0; k < nt; ++k) {
      if (!oks[k]) { err = errs[k]; return false; }
      S.e[which].insert(S.e[which].end(), parts[k].begin(), parts[k].end());
    }
    if (S.e[which].size() > slots_of(which) / 2) { err = "libm table larger than the hash capacity"; return false; }
    // probes: every 997th entry, re-checked against the live powf when the file is loaded
    S.probe[which].clear();
    for (size_t i = 0; i < S.e[which].size(); i += 997) S.probe[which].push_back(S.e[which][i]);
  }
  S.ok = true;
  return true;
}

bool probes_match(const PowSet& S) {
  for (int which = 0; which < NFUN; ++which)
    for (const Entry& p : S.probe[which])
      if (f2u(libc_value(which, u2f(p.x))) != p.g) return false;
  return true;
}

bool load_file(const std::string& path, PowSet& S) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  uint32_t hdr[4];
  bool ok = std::fread(hdr, 4, 4, f) == 4 && hdr[0] == MAGIC && hdr[1] == VERSION && hdr[2] == f2u(POW_XLO) &&
            hdr[3] == f2u(POW_XHI);
  for (int which = 0; ok && which < NFUN; ++which) {
    uint64_t n = 0, np = 0;
    ok = std::fread(&n, 8, 1, f) == 1 && std::fread(&np, 8, 1, f) == 1 && n <= slots_of(which) / 2 && np <= n;
    if (!ok) break;
    S.e[which].resize(n);
    S.probe[which].resize(np);
    ok = std::fread(S.e[which].data(), sizeof(Entry), n, f) == n &&
         std::fread(S.probe[which].data(), sizeof(Entry), np, f) == np;
  }
  std::fclose(f);
  S.ok = ok && probes_match(S);
  return S.ok;
}

void save_file(const std::string& path, const PowSet& S) {
  const std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  const uint32_t hdr[4] = {MAGIC, VERSION, f2u(POW_XLO), f2u(POW_XHI)};
  bool ok = std::fwrite(hdr, 4, 4, f) == 4;
  for (int which = 0; ok && which < NFUN; ++which) {
    const uint64_t n = S.e[which].size(), np = S.probe[which].size();
    ok = std::fwrite(&n, 8, 1, f) == 1 && std::fwrite(&np, 8, 1, f) == 1 &&
         std::fwrite(S.e[which].data(), sizeof(Entry), n, f) == n &&
         std::fwrite(S.probe[which].data(), sizeof(Entry), np, f) == np;
  }
  std::fclose(f);
  if (ok) std::rename(tmp.c_str(), path.c_str());
  else std::remove(tmp.c_str());
}

std::vector<uint2> hash_table(const std::vector<Entry>& e, uint32_t slots = POW_HASH_SIZE) {
  std::vector<uint2> t(slots, make_uint2(0u, 0u));
  for (const Entry& x : e) {
    uint32_t h = hash_bits(x.x) & (slots - 1);
    while (t[h].x != 0u) h = (h + 1) & (slots - 1);
    t[h] = make_uint2(x.x, x.g);
  }
  return t;
}

std::string table_path() {
  const char* env = std::getenv("SFX_POW_TABLE");
  if (env && *env) return env;
  Dl_info info;
  if (dladdr((void*)&table_path, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t k = p.find_last_of('/');
    return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/libm_pow.tbl";
  }
  return "libm_pow.tbl";
}

bool ensure_host(std::string& err) {
  if (g_set.ok) return true;
  const std::string path = table_path();
  if (load_file(path, g_set)) return true;
  if (!generate(g_set, err)) return false;
  save_file(path, g_set);
  return true;
}

}  // namespace

// Host copy of the tables (for the host-side checks of the emulation).
bool pow_tables_host(std::vector<uint2> (&t)[NFUN], std::string& err) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!ensure_host(err)) return false;
  for (int which = 0; which < NFUN; ++which) t[which] = hash_table(g_set.e[which], slots_of(which));
  return true;
}

// Device tables for the current device (uploaded once, kept until the process ends).
bool pow_tables_device(int device, PowTables& out, std::string& err) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_dev.find(device);
  if (it != g_dev.end()) { out = it->second; return true; }
  if (!ensure_host(err)) return false;
  PowTables T{};
  for (int which = 0; which < NFUN; ++which) {
    std::vector<uint2> h = hash_table(g_set.e[which], slots_of(which));
    void* p = nullptr;
    if (hipMalloc(&p, h.size() * sizeof(uint2)) != hipSuccess ||
        hipMemcpy(p, h.data(), h.size() * sizeof(uint2), hipMemcpyHostToDevice) != hipSuccess) {
      err = "upload of the powf tables failed";
      return false;
    }
    const uint2** slot[NFUN] = {&T.t13, &T.t53, &T.t73, &T.t15, &T.tcos, &T.tsin};
    *slot[which] = (const uint2*)p;
  }
  g_dev[device] = T;
  out = T;
  return true;
}

}  // namespace sfx

// Build (or load and verify) the powf tables; path NULL = next to the library.
extern "C" int sfx_libm_tables(const char* path, int64_t counts[6]) {
  std::lock_guard<std::mutex> lk(sfx::g_mu);
  std::string err;
  if (path && *path) {
    sfx::PowSet S;
    if (!sfx::load_file(path, S)) {
      if (!sfx::generate(S, err)) { sfx::set_error(err); return SFX_ERR_STATE; }
      sfx::save_file(path, S);
    }
    sfx::g_set = S;
  } else if (!sfx::ensure_host(err)) {
    sfx::set_error(err);
    return SFX_ERR_STATE;
  }
  if (counts)
    for (int k = 0; k < sfx::NFUN; ++k) counts[k] = (int64_t)sfx::g_set.e[k].size();
  return SFX_OK;
}

// Host evaluation of the emulation, for tests: out[i] = powf(x[i], exponent `which`) for
// which 0..3, cosf(x[i]) for 4, sinf(x[i]) for 5.
extern "C" int sfx_libm_pow_host(const float* x, int64_t n, int32_t which, float* out) {
  std::vector<uint2> t[sfx::NFUN];
  std::string err;
  if (!sfx::pow_tables_host(t, err)) { sfx::set_error(err); return SFX_ERR_STATE; }
  sfx::PowTables T{t[0].data(), t[1].data(), t[2].data(), t[3].data(), t[4].data(), t[5].data()};
  for (int64_t i = 0; i < n; ++i)
    out[i] = which >= sfx::NPOW ? sfx::libm_trig(x[i], which - sfx::NPOW, T) : sfx::libm_pow(x[i], which, T);
  return SFX_OK;
}

namespace {
__global__ void k_libm_pow(const float* __restrict__ x, int64_t n, int which, sfx::PowTables T, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = which >= sfx::NPOW ? sfx::libm_trig(x[i], which - sfx::NPOW, T)
                                : sfx::libm_pow_dev(x[i], which, (const uint64_t*)T.t13, (const uint64_t*)T.t53,
                                                    (const uint64_t*)T.t73, (const uint64_t*)T.t15);
}

// The device estimate against pow_estimate over float bit patterns [lo, lo + n): counts
// the inputs where they differ by more than 2^-44 (relative) and keeps the largest
// difference (as a double's bits, for an atomic max of non-negative values).
__global__ void k_estimate_check(uint32_t lo, uint32_t n, int which, unsigned long long* bad,
                                 unsigned long long* worst) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float x = sfx::u2f(lo + i);
    const double a = sfx::pow_estimate(x, which), b = sfx::pow_estimate_dev(x, which);
    const double d = fabs(a - b) / a;
    if (!(d <= 5.684341886080802e-14)) atomicAdd(bad, 1ull);   // 2^-44
    atomicMax(worst, (unsigned long long)__double_as_longlong(d));
  }
}
}  // namespace

// Device check of the device powf estimate for every float of [2^-40, 2^14) (exponent
// which 0..2): out[0] = inputs beyond 2^-44 of pow_estimate, out[1] = the largest relative
// difference's bits.
extern "C" int sfx_libm_estimate_check(int32_t which, uint64_t out[2]) {
  if (which < 0 || which > 2 || !out) { sfx::set_error("sfx_libm_estimate_check: which 0..2"); return SFX_ERR_ARG; }
  unsigned long long* d = nullptr;
  if (hipMalloc(&d, 2 * sizeof(unsigned long long)) != hipSuccess) { sfx::set_error("hipMalloc failed"); return SFX_ERR_MEMORY; }
  hipMemset(d, 0, 2 * sizeof(unsigned long long));
  const uint32_t lo = sfx::f2u(sfx::POW_XLO), n = sfx::f2u(sfx::POW_XHI) - lo;
  hipLaunchKernelGGL(k_estimate_check, dim3(8192), dim3(256), 0, 0, lo, n, which, d, d + 1);
  const hipError_t e = hipDeviceSynchronize();
  unsigned long long h[2] = {0, 0};
  if (e == hipSuccess) hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  hipFree(d);
  if (e != hipSuccess) { sfx::set_error(std::string("sfx_libm_estimate_check: ") + hipGetErrorString(e)); return SFX_ERR_HIP; }
  out[0] = h[0]; out[1] = h[1];
  return SFX_OK;
}

// Device evaluation of the emulation, for tests (host buffers in and out).
extern "C" int sfx_libm_pow_device(const float* x, int64_t n, int32_t which, float* out) {
  sfx::PowTables T{};
  std::string err;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { sfx::set_error("no HIP device"); return SFX_ERR_HIP; }
  if (!sfx::pow_tables_device(dev, T, err)) { sfx::set_error(err); return SFX_ERR_STATE; }
  float *dx = nullptr, *dy = nullptr;
  if (hipMalloc(&dx, n * 4) != hipSuccess || hipMalloc(&dy, n * 4) != hipSuccess) {
    sfx::set_error("hipMalloc failed");
    return SFX_ERR_MEMORY;
  }
  (void)hipMemcpy(dx, x, n * 4, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_libm_pow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx, n, which, T, dy);
  const hipError_t e = hipDeviceSynchronize();
  (void)hipMemcpy(out, dy, n * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(dy);
  if (e != hipSuccess) { sfx::set_error(hipGetErrorString(e)); return SFX_ERR_HIP; }
  return SFX_OK;
}
