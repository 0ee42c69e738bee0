// Gridded and spiderweb meteo on the host: the ASCII readers of the amu/amv/amp/ampr and
// spw files, the setters, and the meteo fields on the cells at each update time.
//
// The reference refreshes its fields at every dtwnd: at the first step with t >= twindupd
// it sets meteo_t0 = twindupd, meteo_t1 = twindupd + dtwnd and interpolates the meteo
// input to the cells at those two times (update_meteo_fields, sfincs_meteo.f90:1561-1624,
// scheduled by sfincs_lib.f90:493-503); each step then interpolates per cell between the
// two (update_meteo_forcing, :1246-1419). A field at a time depends on that time only, so
// the fields at every update time T_k (T_0 = t0, T_k+1 = T_k + dtwnd in float, the
// reference's twindupd sequence) are computed once here, frame-major, and the device
// picks frames k and k+1 for window k. The interpolation restates update_amuv_data,
// update_amp_data, update_ampr_data and update_spiderweb_data (:802-1244, 450-800)
// operation for operation in float, with the C library's sqrtf, atan2f, cosf and sinf.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "sfx_internal.hpp"

using sfx::set_error;

namespace {

// --------------------------------------------------------------------------- readers
// Days since 1970-01-01 of a proleptic Gregorian date (the civil-from-days inverse of
// H. Hinnant's "chrono-compatible low-level date algorithms").
int64_t days_from_civil(int64_t y, int64_t m, int64_t d) {
  y -= m <= 2;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const int64_t yoe = y - era * 400;
  const int64_t doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
  const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + doe - 719468;
}

// Integer field of `w` characters at `pos` of s, blanks ignored (Fortran I edit descriptor).
int64_t int_field(const std::string& s, size_t pos, size_t w) {
  int64_t v = 0;
  bool neg = false;
  for (size_t i = pos; i < pos + w && i < s.size(); ++i) {
    const char ch = s[i];
    if (ch == '-') neg = true;
    else if (ch >= '0' && ch <= '9') v = v * 10 + (ch - '0');
  }
  return neg ? -v : v;
}

// Seconds of "yyyymmdd HHMMSS" since 1970 (its fields read as time_difference reads them,
// sfincs_date.f90:183-204).
int64_t date_seconds(const std::string& d) {
  const int64_t days = days_from_civil(int_field(d, 0, 4), int_field(d, 4, 2), int_field(d, 6, 2));
  return days * 86400 + int_field(d, 9, 2) * 3600 + int_field(d, 11, 2) * 60 + int_field(d, 13, 2);
}

// First list-directed real of s (a Fortran 'd' exponent read as 'e').
bool first_real(std::string s, float& out) {
  for (char& ch : s)
    if (ch == 'd' || ch == 'D') ch = 'e';
  const char* p = s.c_str();
  while (*p == ' ' || *p == ',') ++p;
  char* end = nullptr;
  const float v = std::strtof(p, &end);
  if (end == p) return false;
  out = v;
  return true;
}

// compute_time_in_seconds (sfincs_spiderweb.f90:411-484): "TIME = <value> <unit> since
// yyyy-mm-dd HH:MM:SS": value * unit minus (tref - the file's date), in float, truncated to
// whole seconds.
bool time_line_seconds(const std::string& line, const std::string& tref, float& out) {
  const size_t j = line.find('=');
  const size_t js = line.find("since");
  size_t ju = std::string::npos;
  int ifac = 0;
  const char* units[4] = {"seconds", "minutes", "hours", "days"};
  const int facs[4] = {1, 60, 3600, 86400};
  for (int k = 0; k < 4; ++k) {
    const size_t at = line.find(units[k]);
    if (at != std::string::npos) { ju = at; ifac = facs[k]; break; }
  }
  if (j == std::string::npos || js == std::string::npos || ju == std::string::npos || ju <= j) return false;
  float tim;
  if (!first_real(line.substr(j + 1, ju - j - 1), tim)) return false;
  std::string ts = line.size() > js + 6 ? line.substr(js + 6, 19) : std::string();
  ts.resize(19, ' ');
  // datespw = yyyy mm dd ' ' HH MM SS from the positions of "yyyy-mm-dd HH:MM:SS"
  const std::string date = ts.substr(0, 4) + ts.substr(5, 2) + ts.substr(8, 2) + " " + ts.substr(11, 2) +
                           ts.substr(14, 2) + ts.substr(17, 2);
  const int64_t dtsec = date_seconds(tref) - date_seconds(date);
  const float r = tim * (float)ifac - (float)dtsec;
  out = (float)(int64_t)r;
  return true;
}

bool read_lines(const char* path, std::vector<std::string>& lines) {
  std::ifstream f(path);
  if (!f) return false;
  std::string s;
  while (std::getline(f, s)) {
    if (!s.empty() && s.back() == '\r') s.pop_back();
    lines.push_back(s);
  }
  return true;
}

// read_int_input / read_real_input on a meteo file (sfincs_read.f90:5-110, 192-247): the
// first line "key = value" whose key (text before '=', tabs as spaces, trailing blanks
// dropped) equals `key`; lines starting with #, ! or @ are comments.
bool header_value(const std::vector<std::string>& lines, const char* key, std::string& val) {
  for (std::string l : lines) {
    for (char& ch : l)
      if (ch == '\t') ch = ' ';
    if (l.empty() || l[0] == '#' || l[0] == '!' || l[0] == '@') continue;
    const size_t j = l.find('=');
    if (j == std::string::npos) continue;
    std::string k = l.substr(0, j);
    while (!k.empty() && k.back() == ' ') k.pop_back();
    if (k != key) continue;
    val = l.substr(j + 1);
    const size_t h = val.find('#');
    if (h != std::string::npos) val = val.substr(0, h);
    return true;
  }
  return false;
}

// Lines before the first that starts with TIME, searched in the first 30 lines; when none
// does, `fallback` (read_amuv_dimensions, read_spw_dimensions).
size_t header_lines(const std::vector<std::string>& lines, size_t fallback) {
  for (size_t i = 0; i < lines.size() && i < 30; ++i)
    if (lines[i].compare(0, 4, "TIME") == 0) return i;
  return fallback;
}

// A row of n list-directed reals.
bool read_row(const std::string& l, int n, float* out) {
  std::string s = l;
  for (char& ch : s)
    if (ch == ',' || ch == '\t') ch = ' ';
    else if (ch == 'd' || ch == 'D') ch = 'e';
  const char* p = s.c_str();
  for (int i = 0; i < n; ++i) {
    char* end = nullptr;
    const float v = std::strtof(p, &end);
    if (end == p) return false;
    out[i] = v;
    p = end;
  }
  return true;
}

int read_amuv(sfx_model* M, int quantity, const char* path, const std::string& tref) {
  std::vector<std::string> L;
  if (!read_lines(path, L)) { set_error(std::string("cannot open meteo file ") + path); return SFX_ERR_ARG; }
  auto hnum = [&](const char* k, float def) {
    std::string v;
    float x = def;
    if (header_value(L, k, v)) first_real(v, x);
    return x;
  };
  const int ncols = (int)hnum("n_cols", 0.f), nrows = (int)hnum("n_rows", 0.f), nquant = (int)hnum("n_quantity", 0.f);
  const float xll = hnum("x_llcorner", 0.f), yll = hnum("y_llcorner", 0.f), dx = hnum("dx", 0.f), dy = hnum("dy", 0.f);
  if (ncols < 2 || nrows < 2 || nquant < 1) { set_error(std::string("meteo file without a grid: ") + path); return SFX_ERR_ARG; }
  size_t i = header_lines(L, 13);
  std::vector<float> times, v;
  while (i < L.size()) {
    float tsec;
    if (!time_line_seconds(L[i], tref, tsec)) { set_error(std::string("bad TIME line in ") + path + ": " + L[i]); return SFX_ERR_ARG; }
    ++i;
    times.push_back(tsec);
    const size_t base = v.size();
    v.resize(base + (size_t)nrows * ncols);
    for (int r = 0; r < nrows; ++r, ++i)
      if (i >= L.size() || !read_row(L[i], ncols, &v[base + (size_t)r * ncols])) {
        set_error(std::string("short meteo file ") + path);
        return SFX_ERR_ARG;
      }
    i += (size_t)nrows * (nquant - 1);   // further quantities of the block are not read (:194-196)
    while (i < L.size() && L[i].find_first_not_of(" \t") == std::string::npos) ++i;
  }
  return sfx_model_set_meteo_grid(M, quantity, (int32_t)times.size(), nrows, ncols, xll, yll, dx, dy, times.data(),
                                  v.data());
}

int read_spw(sfx_model* M, const char* path, const std::string& tref) {
  std::vector<std::string> L;
  if (!read_lines(path, L)) { set_error(std::string("cannot open spiderweb file ") + path); return SFX_ERR_ARG; }
  auto hnum = [&](const char* k, float def) {
    std::string v;
    float x = def;
    if (header_value(L, k, v)) first_real(v, x);
    return x;
  };
  const int ncols = (int)hnum("n_cols", 0.f), nrows = (int)hnum("n_rows", 0.f), nquant = (int)hnum("n_quantity", 0.f);
  const float radius = hnum("spw_radius", 0.f);
  if (ncols < 1 || nrows < 2 || nquant < 3) { set_error(std::string("spiderweb file without a grid: ") + path); return SFX_ERR_ARG; }
  size_t i = header_lines(L, nquant == 4 ? 18 : 16);
  const size_t nf = (size_t)nrows * ncols;
  std::vector<float> times, xe, ye, vmag, vdir, pdrp, prcp;
  // the eye position is read from the 12 characters after '=' (read_spw_file, :85-96)
  auto eye = [&](const std::string& l, float& out) {
    const size_t j = l.find('=');
    return j != std::string::npos && first_real(l.substr(j + 1, 12), out);
  };
  while (i < L.size()) {
    float tsec, x, y;
    if (!time_line_seconds(L[i], tref, tsec)) { set_error(std::string("bad TIME line in ") + path + ": " + L[i]); return SFX_ERR_ARG; }
    if (i + 3 >= L.size() || !eye(L[i + 1], x) || !eye(L[i + 2], y)) { set_error(std::string("bad eye lines in ") + path); return SFX_ERR_ARG; }
    i += 4;   // TIME, x_spw_eye, y_spw_eye, p_drop_spw_eye
    times.push_back(tsec); xe.push_back(x); ye.push_back(y);
    std::vector<float>* q[4] = {&vmag, &vdir, &pdrp, &prcp};
    for (int k = 0; k < 4; ++k) {
      const size_t base = q[k]->size();
      q[k]->resize(base + nf, 0.0f);
      if (k == 3 && nquant < 4) continue;
      for (int r = 0; r < nrows; ++r, ++i)
        if (i >= L.size() || !read_row(L[i], ncols, &(*q[k])[base + (size_t)r * ncols])) {
          set_error(std::string("short spiderweb file ") + path);
          return SFX_ERR_ARG;
        }
    }
    while (i < L.size() && L[i].find_first_not_of(" \t") == std::string::npos) ++i;
  }
  return sfx_model_set_spiderweb(M, (int32_t)times.size(), nrows, ncols, radius, times.data(), xe.data(), ye.data(),
                                 vmag.data(), vdir.data(), pdrp.data(), nquant == 4 ? prcp.data() : nullptr);
}

// ----------------------------------------------------------------- fields at a time
float f_modulo(float a, float p) {   // Fortran MODULO for reals
  float r = std::fmod(a, p);
  if (r != 0.0f && ((r < 0.0f) != (p < 0.0f))) r += p;
  return r;
}

// The grid frame at meteo time T and the time weight, clamped to the grid's times
// (update_amuv_data :821-847; block rain holds the earlier frame, :1142-1149).
void grid_time(const sfx_model::MeteoGrid& g, float T, bool block, int& i0, int& i1, float& tw) {
  float mt = std::max(T, g.times[0]);
  mt = std::min(mt, g.times[g.nt - 1]);
  i0 = 0;
  for (int k = 0; k < g.nt; ++k)
    if (g.times[k] <= mt) i0 = k;
  i1 = std::min(i0 + 1, g.nt - 1);
  mt = std::min(mt, g.times[i1]);
  tw = (mt - g.times[i0]) / std::max(g.times[i1] - g.times[i0], 1.0e-6f);
  if (block) tw = 0.0f;
}

// The frame of a grid blended in time: a(i0) (1 - tw) + a(i1) tw.
std::vector<float> grid_blend(const sfx_model::MeteoGrid& g, int i0, int i1, float tw) {
  const size_t n = (size_t)g.nrows * g.ncols;
  std::vector<float> out(n);
  for (size_t k = 0; k < n; ++k) out[k] = g.v[(size_t)i0 * n + k] * (1.0f - tw) + g.v[(size_t)i1 * n + k] * tw;
  return out;
}

// Bilinear weights of cell centre (x, y) on a grid (:857-895); false outside -- which for
// the reference includes the first (northernmost) row and column index 0.
struct Bilin { uint32_t k[4]; float f[4]; };
bool grid_weights(const sfx_model::MeteoGrid& g, float x, float y, Bilin& b) {
  const float yul = g.yll + (float)(g.nrows - 1) * g.dy + 0.5f * g.dy;
  const float xll = g.xll + 0.5f * g.dx;
  const int iy = (int)((yul - y) / g.dy) + 1;
  const float dj1 = (yul - (float)(iy - 1) * g.dy - y) / g.dy;
  const int ix = (int)((x - xll) / g.dx) + 1;
  const float di1 = (x - (xll + (float)(ix - 1) * g.dx)) / g.dx;
  if (iy <= 1 || iy >= g.nrows || ix <= 0 || ix >= g.ncols) return false;
  const int r[4] = {iy, iy, iy + 1, iy + 1}, c[4] = {ix, ix + 1, ix + 1, ix};
  for (int i = 0; i < 4; ++i) b.k[i] = (uint32_t)((r[i] - 1) * g.ncols + (c[i] - 1));
  b.f[0] = (1.0f - di1) * (1.0f - dj1);
  b.f[1] = di1 * (1.0f - dj1);
  b.f[2] = di1 * dj1;
  b.f[3] = (1.0f - di1) * dj1;
  return true;
}

float cd_of(const sfx_model* M, float vmag) {
  const int i = std::min(std::max((int)(vmag * 10.0f), 0), 999);
  return M->cdval[i];
}

// The meteo input after read_meteo_data's conversions (sfincs_meteo.f90:57-95, 330-414):
// the grids with their factors applied, the spiderweb as u, v, absolute pressure and rain.
struct MeteoIn {
  sfx_model::MeteoGrid amu, amv, amp, ampr;
  int32_t snt = 0, snrows = 0, sncols = 0;
  bool srain = false;
  float radius = 0.f, drad = 0.f, dphi = 0.f;
  std::vector<float> stimes, xe, ye, wu, wv, pabs, prcp;
};

// The fields of one update time T on cells [c0, c1) (update_meteo_fields order: wind,
// pressure, rain, then the spiderweb merged over them).
void fields_at(const sfx_model* M, const MeteoIn& in, float T, size_t c0, size_t c1, float* tu, float* tv,
               float* pa, float* pr) {
  const sfx_params& p = M->p;
  const float rhoa = p.rhoa, rhow = p.rhow, cosrot = M->cosrot, sinrot = M->sinrot;
  if (in.amu.nt > 0) {   // update_amuv_data (:802-978)
    int i0, i1;
    float tw;
    grid_time(in.amu, T, false, i0, i1, tw);
    const auto wu = grid_blend(in.amu, i0, i1, tw), wv = grid_blend(in.amv, i0, i1, tw);
    for (size_t c = c0; c < c1; ++c) {
      Bilin b;
      if (!grid_weights(in.amu, M->z_xz[c], M->z_yz[c], b)) { tu[c] = 0.0f; tv[c] = 0.0f; continue; }
      float wup = 0.0f, wvp = 0.0f;
      for (int i = 0; i < 4; ++i) { wup = wup + b.f[i] * wu[b.k[i]]; wvp = wvp + b.f[i] * wv[b.k[i]]; }
      const float vmag = std::sqrt(wup * wup + wvp * wvp);
      const float cd = cd_of(M, vmag);
      tu[c] = vmag * (cosrot * wup + sinrot * wvp) * rhoa * cd / rhow;
      tv[c] = vmag * (-sinrot * wup + cosrot * wvp) * rhoa * cd / rhow;
    }
  } else if (tu) {
    for (size_t c = c0; c < c1; ++c) { tu[c] = 0.0f; tv[c] = 0.0f; }
  }
  if (pa) {
    if (in.amp.nt > 0) {   // update_amp_data (:980-1108)
      int i0, i1;
      float tw;
      grid_time(in.amp, T, false, i0, i1, tw);
      const auto pg = grid_blend(in.amp, i0, i1, tw);
      for (size_t c = c0; c < c1; ++c) {
        Bilin b;
        if (!grid_weights(in.amp, M->z_xz[c], M->z_yz[c], b)) { pa[c] = p.gapres; continue; }
        float s = 0.0f;
        for (int i = 0; i < 4; ++i) s = s + b.f[i] * pg[b.k[i]];
        pa[c] = s;
      }
    } else {
      for (size_t c = c0; c < c1; ++c) pa[c] = p.gapres;
    }
  }
  if (pr) {
    if (in.ampr.nt > 0) {   // update_ampr_data (:1110-1244)
      int i0, i1;
      float tw;
      grid_time(in.ampr, T, p.amprblock != 0, i0, i1, tw);
      const auto rg = grid_blend(in.ampr, i0, i1, tw);
      for (size_t c = c0; c < c1; ++c) {
        Bilin b;
        if (!grid_weights(in.ampr, M->z_xz[c], M->z_yz[c], b)) { pr[c] = 0.0f; continue; }
        float s = 0.0f;
        for (int i = 0; i < 4; ++i) s = s + b.f[i] * rg[b.k[i]];
        pr[c] = s / (float)(1000 * 3600);
      }
    } else {
      for (size_t c = c0; c < c1; ++c) pr[c] = 0.0f;
    }
  }
  if (in.snt == 0) return;
  // update_spiderweb_data (:450-800)
  const int snt = in.snt;
  int i0, i1;
  float tw, tsf;
  if (T < in.stimes[0]) {
    i0 = i1 = 0; tw = 1.0f;
    tsf = std::max(1.0f - (in.stimes[0] - T) / 21600.0f, 0.0f);
  } else if (T > in.stimes[snt - 1]) {
    i0 = i1 = snt - 1; tw = 1.0f;
    tsf = std::max(1.0f - (T - in.stimes[snt - 1]) / 21600.0f, 0.0f);
  } else {
    i0 = 0;
    for (int k = 0; k < snt; ++k)
      if (in.stimes[k] <= T) i0 = k;
    i1 = std::min(i0 + 1, snt - 1);
    tw = (T - in.stimes[i0]) / std::max(in.stimes[i1] - in.stimes[i0], 1.0e-6f);
    tsf = 1.0f;
  }
  const float xe = in.xe[i0] * (1.0f - tw) + in.xe[i1] * tw;
  const float ye = in.ye[i0] * (1.0f - tw) + in.ye[i1] * tw;
  const size_t n = (size_t)in.snrows * in.sncols;
  std::vector<float> wu(n), wv(n), pb(n), rn(n);
  for (size_t k = 0; k < n; ++k) {
    const size_t a = (size_t)i0 * n + k, b = (size_t)i1 * n + k;
    wu[k] = in.wu[a] * (1.0f - tw) + in.wu[b] * tw;
    wv[k] = in.wv[a] * (1.0f - tw) + in.wv[b] * tw;
    pb[k] = in.pabs[a] * (1.0f - tw) + in.pabs[b] * tw;
    if (in.srain) rn[k] = in.prcp[a] * (1.0f - tw) + in.prcp[b] * tw;
  }
  struct { int nrows, ncols; bool rain; } s{in.snrows, in.sncols, in.srain};
  const float radius = in.radius, drad = in.drad, dphi = in.dphi, pi = p.pi;
  const float mfrac = p.spw_merge_frac;
  for (size_t c = c0; c < c1; ++c) {
    const float dxe = M->z_xz[c] - xe, dye = M->z_yz[c] - ye;
    const float dst = std::sqrt(dxe * dxe + dye * dye);
    if (dst > radius) continue;
    int id = (int)(dst / drad);
    id = std::max(id, 1);
    if (id + 1 > s.nrows) continue;
    float dj1 = (dst - drad * (float)id) / drad;
    dj1 = std::max(dj1, 0.0f);
    float phi = 0.5f * pi - std::atan2(dye, dxe);
    phi = f_modulo(phi, 2 * pi);
    const int ip = std::min((int)(phi / dphi) + 1, s.ncols);
    int c2 = ip + 1;
    if (c2 > s.ncols) c2 = 1;
    const int r[4] = {id, id, id + 1, id + 1}, col[4] = {ip, c2, c2, ip};
    const float di1 = (phi - dphi * (float)(ip - 1)) / dphi;
    const float f[4] = {(1.0f - di1) * (1.0f - dj1), di1 * (1.0f - dj1), di1 * dj1, (1.0f - di1) * dj1};
    float wup = 0.0f, wvp = 0.0f, pcp = 0.0f, prp = 0.0f;
    for (int i = 0; i < 4; ++i) {
      const size_t k = (size_t)(r[i] - 1) * s.ncols + (col[i] - 1);
      wup = wup + f[i] * wu[k];
      wvp = wvp + f[i] * wv[k];
      if (pa) pcp = pcp + f[i] * pb[k];
      if (pr && s.rain) prp = prp + f[i] * rn[k];
    }
    const float vmag = std::sqrt(wup * wup + wvp * wvp);
    const float cd = cd_of(M, vmag);
    float merge;
    if (dst > mfrac * radius) merge = (1.0f / (1.0f - std::min(mfrac, 0.999f))) * (radius - dst) / radius;
    else merge = 1.0f;
    merge = merge * tsf;
    tu[c] = (1.0f - merge) * tu[c] + merge * vmag * (cosrot * wup + sinrot * wvp) * rhoa * cd / rhow;
    tv[c] = (1.0f - merge) * tv[c] + merge * vmag * (-sinrot * wup + cosrot * wvp) * rhoa * cd / rhow;
    if (pa) pa[c] = (1.0f - merge) * pa[c] + merge * pcp;
    if (pr && s.rain) pr[c] = (1.0f - merge) * pr[c] + merge * prp / (float)(1000 * 3600);
  }
}

}  // namespace

// ------------------------------------------------------------------------------ setters
extern "C" int sfx_model_set_meteo_grid(sfx_model* M, int32_t quantity, int32_t nt, int32_t nrows, int32_t ncols,
                                        float x_llcorner, float y_llcorner, float dx, float dy, const float* times,
                                        const float* values) {
  if (!M || M->initialized || quantity < 0 || quantity > 3 || nt < 1 || nrows < 2 || ncols < 2 || !(dx > 0.f) ||
      !(dy > 0.f) || !times || !values) {
    set_error("sfx_model_set_meteo_grid: bad args");
    return SFX_ERR_ARG;
  }
  for (int32_t k = 1; k < nt; ++k)
    if (!(times[k] > times[k - 1])) { set_error("sfx_model_set_meteo_grid: times must increase"); return SFX_ERR_ARG; }
  sfx_model::MeteoGrid* g[4] = {&M->amu, &M->amv, &M->amp, &M->ampr};
  sfx_model::MeteoGrid& m = *g[quantity];
  m.nt = nt; m.nrows = nrows; m.ncols = ncols;
  m.xll = x_llcorner; m.yll = y_llcorner; m.dx = dx; m.dy = dy;
  m.times.assign(times, times + nt);
  m.v.assign(values, values + (size_t)nt * nrows * ncols);
  return SFX_OK;
}

extern "C" int sfx_model_set_spiderweb(sfx_model* M, int32_t nt, int32_t nrows, int32_t ncols, float radius,
                                       const float* times, const float* xe, const float* ye, const float* vmag,
                                       const float* vdir, const float* pdrop, const float* rain) {
  if (!M || M->initialized || nt < 1 || nrows < 2 || ncols < 1 || !(radius > 0.f) || !times || !xe || !ye || !vmag ||
      !vdir || !pdrop) {
    set_error("sfx_model_set_spiderweb: bad args");
    return SFX_ERR_ARG;
  }
  const size_t n = (size_t)nt * nrows * ncols;
  auto& s = M->spw;
  s.nt = nt; s.nrows = nrows; s.ncols = ncols; s.radius = radius;
  s.times.assign(times, times + nt); s.xe.assign(xe, xe + nt); s.ye.assign(ye, ye + nt);
  s.vmag.assign(vmag, vmag + n); s.vdir.assign(vdir, vdir + n); s.pdrop.assign(pdrop, pdrop + n);
  s.rain = rain != nullptr;
  if (rain) s.prcp.assign(rain, rain + n);
  else s.prcp.clear();
  return SFX_OK;
}

extern "C" int sfx_model_read_meteo_file(sfx_model* M, const char* kind, const char* path, const char* tref) {
  if (!M || !kind || !path || !tref) { set_error("sfx_model_read_meteo_file: bad args"); return SFX_ERR_ARG; }
  const std::string k(kind), tr(tref);
  if (k == "amu") return read_amuv(M, SFX_METEO_WIND_U, path, tr);
  if (k == "amv") return read_amuv(M, SFX_METEO_WIND_V, path, tr);
  if (k == "amp") return read_amuv(M, SFX_METEO_PRESSURE, path, tr);
  if (k == "ampr") return read_amuv(M, SFX_METEO_RAIN, path, tr);
  if (k == "spw") return read_spw(M, path, tr);
  set_error("sfx_model_read_meteo_file: kind must be amu, amv, amp, ampr or spw");
  return SFX_ERR_ARG;
}

// --------------------------------------------------------------- fields at initialize
namespace sfx {

// read_meteo_data's conversions (sfincs_meteo.f90:57-95, 316-414) and the fields of every
// update time from t0 until past t1 (a window more than the run needs). Called by
// sfx_model_initialize after the Cd table is built.
int build_meteo(sfx_model* M) {
  const sfx_params& p = M->p;
  M->meteo_times.clear(); M->meteo_tauwu.clear(); M->meteo_tauwv.clear(); M->meteo_patm.clear(); M->meteo_prcp.clear();
  M->meteo_nframes = 0;
  const bool has_wind = M->amu.nt > 0 || M->amv.nt > 0;
  M->meteo3d = has_wind || M->amp.nt > 0 || M->ampr.nt > 0 || M->spw.nt > 0;
  M->meteo_wind = M->patmos = M->meteo_rain = false;
  if (!M->meteo3d) return SFX_OK;
  if (has_wind) {
    const auto& a = M->amu;
    const auto& b = M->amv;
    if (a.nt == 0 || b.nt == 0 || a.nt != b.nt || a.nrows != b.nrows || a.ncols != b.ncols || a.times != b.times) {
      set_error("gridded wind needs amu and amv on the same grid and times");
      return SFX_ERR_ARG;
    }
  }
  if (M->wind || M->precip) {   // a time series would overwrite the gridded fields every step
    if ((M->wind && (has_wind || M->spw.nt > 0)) || (M->precip && (M->ampr.nt > 0 || (M->spw.rain && p.usespwprecip)))) {
      set_error("wind or rain is given both as a time series and as gridded meteo");
      return SFX_ERR_ARG;
    }
  }
  if (!(p.dtwnd > 0.f)) { set_error("dtwnd must be positive"); return SFX_ERR_ARG; }
  // read_meteo_data's conversions, on copies
  MeteoIn in;
  in.amu = M->amu; in.amv = M->amv; in.amp = M->amp; in.ampr = M->ampr;
  const bool spw_rain = M->spw.nt > 0 && M->spw.rain && p.usespwprecip;
  if (M->spw.nt > 0) {
    const auto& s = M->spw;
    const size_t n = s.vmag.size();
    const float pi = p.pi;
    in.snt = s.nt; in.snrows = s.nrows; in.sncols = s.ncols; in.srain = spw_rain;
    in.stimes = s.times; in.xe = s.xe; in.ye = s.ye;
    in.wu.resize(n); in.wv.resize(n); in.pabs.resize(n);
    for (size_t k = 0; k < n; ++k) {
      in.wu[k] = s.vmag[k] * std::cos(pi * (270.0f - s.vdir[k]) / 180.0f);
      in.wv[k] = s.vmag[k] * std::sin(pi * (270.0f - s.vdir[k]) / 180.0f);
      in.pabs[k] = p.gapres - s.pdrop[k];
    }
    if (spw_rain) in.prcp = s.prcp;
    if (p.factor_wind != 1.0f)
      for (size_t k = 0; k < n; ++k) { in.wu[k] = p.factor_wind * in.wu[k]; in.wv[k] = p.factor_wind * in.wv[k]; }
    if (p.factor_pres != 1.0f)
      for (size_t k = 0; k < n; ++k) in.pabs[k] = p.gapres - p.factor_pres * (p.gapres - in.pabs[k]);
    if (p.factor_prcp != 1.0f && spw_rain)
      for (size_t k = 0; k < n; ++k) in.prcp[k] = p.factor_prcp * in.prcp[k];
    in.radius = s.radius;
    if (p.factor_spw_size != 1.0f) in.radius = p.factor_spw_size * in.radius;
    in.drad = in.radius / (float)s.nrows;
    in.dphi = 2 * p.pi / (float)s.ncols;
  }
  if (p.factor_wind != 1.0f && has_wind)
    for (auto* g : {&in.amu, &in.amv})
      for (float& v : g->v) v = p.factor_wind * v;
  if (p.factor_pres != 1.0f && in.amp.nt > 0)
    for (float& v : in.amp.v) v = p.gapres - p.factor_pres * (p.gapres - v);
  if (p.factor_prcp != 1.0f && in.ampr.nt > 0)
    for (float& v : in.ampr.v) v = p.factor_prcp * v;

  // switches (sfincs_domain.f90:99-123, sfincs_meteo.f90:75-90)
  M->meteo_wind = has_wind || M->spw.nt > 0;
  M->patmos = (M->amp.nt > 0 || M->spw.nt > 0) && p.baro == 1;
  M->meteo_rain = M->ampr.nt > 0 || spw_rain;
  if (M->meteo_wind) M->wind = true;
  if (M->meteo_rain) M->precip = true;

  // update times T_k (twindupd, sfincs_lib.f90:290, 496-501) until the second last lies
  // past the last step's time (t1 plus at most dtmax): windows 0 .. nframes - 2
  float T = p.t0;
  std::vector<float> times;
  while (true) {
    times.push_back(T);
    if (times.size() > 2 && (double)times[times.size() - 2] > (double)p.t1 + (double)p.dtmax) break;
    if (times.size() > 100000) { set_error("dtwnd too small for the simulated period"); return SFX_ERR_ARG; }
    T = T + p.dtwnd;
  }
  const size_t nf = times.size(), np = (size_t)M->np;
  M->meteo_nframes = (int32_t)nf;
  M->meteo_times = times;
  M->meteo_tauwu.assign(nf * np, 0.0f);
  M->meteo_tauwv.assign(nf * np, 0.0f);
  if (M->patmos) M->meteo_patm.assign(nf * np, 0.0f);
  if (M->meteo_rain) M->meteo_prcp.assign(nf * np, 0.0f);
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (size_t k = 0; k < nf; ++k) {
    float* tu = M->meteo_tauwu.data() + k * np;
    float* tv = M->meteo_tauwv.data() + k * np;
    float* pa = M->patmos ? M->meteo_patm.data() + k * np : nullptr;
    float* pr = M->meteo_rain ? M->meteo_prcp.data() + k * np : nullptr;
    std::vector<std::thread> th;
    const size_t chunk = (np + nth - 1) / nth;
    for (unsigned t = 0; t < nth; ++t) {
      const size_t c0 = t * chunk, c1 = std::min(np, c0 + chunk);
      if (c0 >= c1) break;
      th.emplace_back([=, &in] { fields_at(M, in, times[k], c0, c1, tu, tv, pa, pr); });
    }
    for (auto& x : th) x.join();
  }
  if (!M->meteo_wind) { M->meteo_tauwu.clear(); M->meteo_tauwv.clear(); }
  return SFX_OK;
}

}  // namespace sfx
