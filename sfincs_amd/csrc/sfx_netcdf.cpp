// NetCDF mesh and subgrid inputs:
//   sfx_model_create_quadtree_netcdf  quadtree_read_file_netcdf (sfincs_quadtree.F90:295-442)
//   sfx_model_read_subgrid_netcdf     read_subgrid_file_netcdf (sfincs_subgrid.F90:57-455)
// The reference reads these files through the NetCDF library, which this build does not
// link. The classic formats (CDF-1, CDF-2 with 64-bit offsets, CDF-5 with 64-bit data) are
// read here from their published layout: a big-endian header listing the dimensions, the
// global attributes and the variables (name, dimensions, type, offset), then each
// variable's values at its offset, C order, big-endian. Values are read whole and
// converted to the type the reference reads them into, with the library's range check.
// NetCDF-4 files (HDF5 containers) are refused with a message naming the conversion.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "sfx_internal.hpp"

using sfx::set_error;

namespace {

enum { NC_BYTE = 1, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE, NC_UBYTE, NC_USHORT, NC_UINT, NC_INT64, NC_UINT64 };

int type_size(int t) {
  switch (t) {
    case NC_BYTE: case NC_CHAR: case NC_UBYTE: return 1;
    case NC_SHORT: case NC_USHORT: return 2;
    case NC_INT: case NC_FLOAT: case NC_UINT: return 4;
    case NC_DOUBLE: case NC_INT64: case NC_UINT64: return 8;
    default: return 0;
  }
}

// One big-endian value of type t at p, as a double.
double be_value(const unsigned char* p, int t) {
  uint64_t u = 0;
  const int k = type_size(t);
  for (int i = 0; i < k; ++i) u = (u << 8) | p[i];
  switch (t) {
    case NC_BYTE: return (double)(int8_t)u;
    case NC_CHAR: case NC_UBYTE: return (double)(uint8_t)u;
    case NC_SHORT: return (double)(int16_t)u;
    case NC_USHORT: return (double)(uint16_t)u;
    case NC_INT: return (double)(int32_t)u;
    case NC_UINT: return (double)(uint32_t)u;
    case NC_INT64: return (double)(int64_t)u;
    case NC_UINT64: return (double)u;
    case NC_FLOAT: { uint32_t b = (uint32_t)u; float f; std::memcpy(&f, &b, 4); return f; }
    default: { double d; std::memcpy(&d, &u, 8); return d; }
  }
}

struct NcDim { std::string name; int64_t len; };
struct NcAtt { std::string name; int type; std::vector<double> num; };
struct NcVar {
  std::string name;
  std::vector<int64_t> dimids;
  int type = 0;
  int64_t begin = 0;
  bool record = false;
};

class NcFile {
 public:
  ~NcFile() { if (f_) std::fclose(f_); }
  bool open(const std::string& path, std::string& err);
  int64_t dim(const char* name) const {
    for (const NcDim& d : dims_) if (d.name == name) return d.len;
    return -1;
  }
  const NcVar* var(const char* name) const {
    for (const NcVar& v : vars_) if (v.name == name) return &v;
    return nullptr;
  }
  bool att(const char* name, double& v) const {
    for (const NcAtt& a : gatts_)
      if (a.name == name && !a.num.empty()) { v = a.num[0]; return true; }
    return false;
  }
  std::vector<int64_t> shape(const NcVar& v) const {
    std::vector<int64_t> s;
    for (size_t k = 0; k < v.dimids.size(); ++k) s.push_back(v.record && k == 0 ? numrecs_ : dims_[v.dimids[k]].len);
    return s;
  }
  // Every value of a numeric variable in file order (records concatenated).
  bool values(const NcVar& v, std::vector<double>& out, std::string& err) const;

 private:
  bool parse(const std::vector<unsigned char>& h, bool& short_header, std::string& err);
  // One record's bytes (all of a fixed variable); -1 when that exceeds the file.
  int64_t slab_bytes(const NcVar& v) const {
    int64_t n = type_size(v.type);
    for (size_t k = v.record ? 1 : 0; k < v.dimids.size(); ++k) {
      const int64_t d = dims_[v.dimids[k]].len;
      if (d < 0 || (d > 0 && n > fsize_ / d)) return -1;
      n *= d;
    }
    return n;
  }
  FILE* f_ = nullptr;
  std::string path_;
  int version_ = 0;
  int64_t numrecs_ = 0, recsize_ = 0, fsize_ = 0;
  std::vector<NcDim> dims_;
  std::vector<NcAtt> gatts_;
  std::vector<NcVar> vars_;
};

// Header cursor: big-endian fields, sizes by format version.
struct Cur {
  const std::vector<unsigned char>& b;
  size_t i = 0;
  bool out = false;
  explicit Cur(const std::vector<unsigned char>& buf) : b(buf) {}
  uint64_t be(int k) {
    if (out || i + (size_t)k > b.size()) { out = true; return 0; }
    uint64_t v = 0;
    for (int j = 0; j < k; ++j) v = (v << 8) | b[i++];
    return v;
  }
  void skip(uint64_t k) {
    if (out || k > b.size() - i) { out = true; return; }
    i += (size_t)k;
  }
};

bool NcFile::parse(const std::vector<unsigned char>& h, bool& short_header, std::string& err) {
  Cur c(h);
  const int nn = version_ == 5 ? 8 : 4;                // NON_NEG / nelems width
  const int off = version_ == 1 ? 4 : 8;               // OFFSET width
  c.skip(4);
  numrecs_ = (int64_t)c.be(nn);
  const bool streaming = version_ == 5 ? (uint64_t)numrecs_ == ~0ull : numrecs_ == 0xffffffffLL;
  auto name = [&](std::string& s) {
    const uint64_t n = c.be(nn);
    if (c.out || n > h.size()) { c.out = true; return; }
    const size_t at = c.i;
    c.skip((n + 3) & ~3ull);
    if (!c.out) s.assign((const char*)h.data() + at, (size_t)n);
  };
  auto list = [&](uint32_t tag, uint64_t& n) -> bool {
    const uint32_t t = (uint32_t)c.be(4);
    n = c.be(nn);
    if (c.out) return true;
    if (t == 0 && n == 0) return true;                   // ABSENT
    if (t != tag) { err = "bad header list tag"; return false; }
    return true;
  };
  auto attrs = [&](std::vector<NcAtt>* keep) -> bool {
    uint64_t n = 0;
    if (!list(12u, n)) return false;
    for (uint64_t a = 0; a < n && !c.out; ++a) {
      NcAtt at;
      name(at.name);
      at.type = (int)c.be(4);
      const uint64_t ne = c.be(nn);
      const int ts = type_size(at.type);
      if (c.out) break;
      if (ts == 0) { err = "attribute " + at.name + " of unknown type"; return false; }
      if (ne > h.size()) { c.out = true; break; }
      const size_t at0 = c.i;
      c.skip((ne * ts + 3) & ~3ull);
      if (c.out) break;
      if (at.type != NC_CHAR)
        for (uint64_t k = 0; k < ne; ++k) at.num.push_back(be_value(h.data() + at0 + k * ts, at.type));
      if (keep) keep->push_back(at);
    }
    return true;
  };
  uint64_t nd = 0;
  if (!list(10u, nd)) return false;
  for (uint64_t k = 0; k < nd && !c.out; ++k) {
    NcDim d;
    name(d.name);
    d.len = (int64_t)c.be(nn);
    dims_.push_back(d);
  }
  if (!c.out && !attrs(&gatts_)) return false;
  uint64_t nv = 0;
  if (!c.out && !list(11u, nv)) return false;
  for (uint64_t k = 0; k < nv && !c.out; ++k) {
    NcVar v;
    name(v.name);
    const uint64_t ndim = c.be(nn);
    if (c.out || ndim > 1024) { if (!c.out) err = "variable " + v.name + " has too many dimensions"; break; }
    for (uint64_t j = 0; j < ndim; ++j) v.dimids.push_back((int64_t)c.be(nn));
    if (!attrs(nullptr)) return false;
    v.type = (int)c.be(4);
    (void)c.be(nn);                                      // vsize (recomputed from the shape)
    v.begin = (int64_t)c.be(off);
    if (c.out) break;
    if (type_size(v.type) == 0) { err = "variable " + v.name + " of unknown type"; return false; }
    for (int64_t id : v.dimids)
      if (id < 0 || id >= (int64_t)dims_.size()) { err = "variable " + v.name + " names a missing dimension"; return false; }
    v.record = !v.dimids.empty() && dims_[v.dimids[0]].len == 0;
    vars_.push_back(v);
  }
  if (!err.empty()) return false;
  if (c.out) { short_header = true; return false; }
  // record layout: each record holds every record variable's slab, padded to 4 bytes
  // (one record variable alone is not padded)
  int nrec = 0;
  for (const NcVar& v : vars_) nrec += v.record;
  recsize_ = 0;
  for (const NcVar& v : vars_) {
    if (!v.record) continue;
    const int64_t sb = slab_bytes(v);
    if (sb < 0) { err = "variable " + v.name + " is larger than the file"; return false; }
    recsize_ += nrec == 1 ? sb : ((sb + 3) & ~3LL);
  }
  if (streaming) {
    numrecs_ = 0;
    for (const NcVar& v : vars_)
      if (v.record && recsize_ > 0) { numrecs_ = (fsize_ - v.begin) / recsize_; break; }
  }
  return true;
}

bool NcFile::open(const std::string& path, std::string& err) {
  path_ = path;
  f_ = std::fopen(path.c_str(), "rb");
  if (!f_) { err = "cannot open " + path; return false; }
  std::fseek(f_, 0, SEEK_END);
  fsize_ = (int64_t)std::ftell(f_);
  std::fseek(f_, 0, SEEK_SET);
  unsigned char magic[8] = {0};
  const size_t got = std::fread(magic, 1, 8, f_);
  static const unsigned char hdf[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  if (got == 8 && std::memcmp(magic, hdf, 8) == 0) {
    err = path + " is a NetCDF-4 (HDF5) file; this build reads the classic NetCDF formats (CDF-1, CDF-2, CDF-5): "
          "convert it with nccopy -k cdf5 (or write it with format='NETCDF3_64BIT')";
    return false;
  }
  if (got < 4 || magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F' ||
      (magic[3] != 1 && magic[3] != 2 && magic[3] != 5)) {
    err = path + " is not a NetCDF file";
    return false;
  }
  version_ = magic[3];
  // the header is read in growing chunks until it parses
  for (size_t want = 1 << 16;; want *= 4) {
    const size_t n = (size_t)std::min<int64_t>((int64_t)want, fsize_);
    std::vector<unsigned char> h(n);
    std::fseek(f_, 0, SEEK_SET);
    if (std::fread(h.data(), 1, n, f_) != n) { err = "cannot read " + path; return false; }
    dims_.clear(); gatts_.clear(); vars_.clear();
    bool short_header = false;
    if (parse(h, short_header, err)) return true;
    if (!short_header) { err = path + ": " + err; return false; }
    if ((int64_t)n >= fsize_) { err = path + ": truncated NetCDF header"; return false; }
  }
}

bool NcFile::values(const NcVar& v, std::vector<double>& out, std::string& err) const {
  if (v.type == NC_CHAR) { err = path_ + ": variable " + v.name + " is text"; return false; }
  const int ts = type_size(v.type);
  const int64_t slab = slab_bytes(v);
  const int64_t nslab = v.record ? numrecs_ : 1;
  if (slab < 0 || nslab < 0 || (slab > 0 && nslab > fsize_ / slab)) {
    err = path_ + ": variable " + v.name + " is larger than the file";
    return false;
  }
  out.resize((size_t)(slab / ts * nslab));
  std::vector<unsigned char> raw((size_t)slab);
  for (int64_t r = 0; r < nslab; ++r) {
    const int64_t at = v.begin + (v.record ? r * recsize_ : 0);
    if (at < 0 || at + slab > fsize_) { err = path_ + ": variable " + v.name + " lies beyond the end of the file"; return false; }
    if (std::fseek(f_, (long)at, SEEK_SET) != 0 || std::fread(raw.data(), 1, (size_t)slab, f_) != (size_t)slab) {
      err = path_ + ": cannot read variable " + v.name;
      return false;
    }
    double* o = out.data() + (size_t)(r * (slab / ts));
    for (int64_t k = 0; k < slab / ts; ++k) o[k] = be_value(raw.data() + k * ts, v.type);
  }
  return true;
}

// A variable read into type T with the NetCDF library's conversion: the value must lie in
// T's range (else NC_ERANGE); reals convert to integers by truncation. `shape` (C order)
// must match when given.
template <typename T>
bool get_var(const NcFile& f, const char* name, const std::vector<int64_t>& shape, std::vector<T>& out,
             std::string& err, const std::string& path) {
  const NcVar* v = f.var(name);
  if (!v) { err = path + ": variable " + name + " not found"; return false; }
  if (f.shape(*v) != shape) {
    std::string s;
    for (int64_t d : shape) s += (s.empty() ? "" : ", ") + std::to_string(d);
    err = path + ": variable " + name + " does not have the shape (" + s + ")";
    return false;
  }
  std::vector<double> d;
  if (!f.values(*v, d, err)) return false;
  out.resize(d.size());
  const double lo = (double)std::numeric_limits<T>::lowest(), hi = (double)std::numeric_limits<T>::max();
  // a float variable read as float passes unconverted (infinities and NaN included)
  const bool same = std::is_same<T, float>::value && v->type == NC_FLOAT;
  for (size_t k = 0; k < d.size(); ++k) {
    if (!same && !(d[k] >= lo && d[k] <= hi) && !(std::is_floating_point<T>::value && std::isnan(d[k]))) {
      err = path + ": variable " + name + " holds a value out of range";
      return false;
    }
    out[k] = (T)d[k];
  }
  return true;
}

}  // namespace

namespace {

int create_quadtree_netcdf(const sfx_params* p, const char* path, int32_t with_bed, sfx_model** out) {
  NcFile f;
  std::string err;
  if (!f.open(path, err)) { set_error(err); return SFX_ERR_IO; }
  const int64_t np = f.dim("mesh2d_nFaces");
  if (np < 1 || np > std::numeric_limits<int32_t>::max()) {
    set_error(std::string(path) + ": dimension mesh2d_nFaces missing or empty");
    return SFX_ERR_IO;
  }
  const std::vector<int64_t> sh{np};
  std::vector<int8_t> level, md, mu, nd, nu, mask;
  std::vector<int32_t> n, m, md1, md2, mu1, mu2, nd1, nd2, nu1, nu2;
  std::vector<float> z;
  bool ok = get_var(f, "n", sh, n, err, path) && get_var(f, "m", sh, m, err, path) &&
            get_var(f, "level", sh, level, err, path) && get_var(f, "md", sh, md, err, path) &&
            get_var(f, "md1", sh, md1, err, path) && get_var(f, "md2", sh, md2, err, path) &&
            get_var(f, "mu", sh, mu, err, path) && get_var(f, "mu1", sh, mu1, err, path) &&
            get_var(f, "mu2", sh, mu2, err, path) && get_var(f, "nd", sh, nd, err, path) &&
            get_var(f, "nd1", sh, nd1, err, path) && get_var(f, "nd2", sh, nd2, err, path) &&
            get_var(f, "nu", sh, nu, err, path) && get_var(f, "nu1", sh, nu1, err, path) &&
            get_var(f, "nu2", sh, nu2, err, path) && get_var(f, "z", sh, z, err, path) &&
            get_var(f, "mask", sh, mask, err, path);
  if (!ok) { set_error(err); return SFX_ERR_IO; }
  double a[8];
  const char* names[8] = {"x0", "y0", "dx", "dy", "rotation", "nmax", "mmax", "nr_levels"};
  for (int k = 0; k < 8; ++k)
    if (!f.att(names[k], a[k])) { set_error(std::string(path) + ": global attribute " + names[k] + " not found"); return SFX_ERR_IO; }
  if (!(a[7] >= 1.0 && a[7] <= 127.0)) { set_error(std::string(path) + ": nr_levels out of range"); return SFX_ERR_IO; }
  sfx_quadtree_input q{};
  q.nr_points = (int32_t)np;
  q.nr_levels = (int32_t)a[7];
  q.x0 = (float)a[0]; q.y0 = (float)a[1]; q.dx = (float)a[2]; q.dy = (float)a[3]; q.rotation_deg = (float)a[4];
  q.level = level.data(); q.md = md.data(); q.mu = mu.data(); q.nd = nd.data(); q.nu = nu.data(); q.mask = mask.data();
  q.n = n.data(); q.m = m.data(); q.md1 = md1.data(); q.md2 = md2.data(); q.mu1 = mu1.data(); q.mu2 = mu2.data();
  q.nd1 = nd1.data(); q.nd2 = nd2.data(); q.nu1 = nu1.data(); q.nu2 = nu2.data();
  int rc = sfx_model_create_quadtree(p, &q, out);
  if (rc != SFX_OK || !with_bed) return rc;
  // the bed level of the active cells from z (sfincs_domain.f90:1635-1637)
  std::vector<float> zb;
  for (int64_t k = 0; k < np; ++k)
    if (mask[k] > 0) zb.push_back(z[k]);
  rc = sfx_model_set_bathymetry(*out, zb.data());
  if (rc != SFX_OK) { sfx_model_destroy(*out); *out = nullptr; }
  return rc;
}

int read_subgrid_netcdf(sfx_model* M, const char* path) {
  NcFile f;
  std::string err;
  if (!f.open(path, err)) { set_error(err); return SFX_ERR_IO; }
  const int64_t nl = f.dim("levels"), npf = f.dim("np"), nuvf = f.dim("npuv");
  if (nl < 0 || npf < 0 || nuvf < 0) { set_error(std::string(path) + ": dimensions levels, np and npuv are needed"); return SFX_ERR_IO; }
  if (nl < 2 || nl > 1000) { set_error(std::string(path) + ": subgrid tables need 2 to 1000 levels"); return SFX_ERR_IO; }
  const int32_t np = M->np, npuv = M->npuv;
  // the file record of each active cell and uv-point (sfincs_subgrid.F90:156-240)
  std::vector<int32_t> zsrc, uvsrc;
  if (M->use_quadtree) {
    int32_t nq = 0;
    if (!sfx::quadtree_subgrid_order(M, zsrc, uvsrc, nq)) {
      set_error(std::string(path) + ": uv-points do not match the quadtree");
      return SFX_ERR_IO;
    }
  } else {
    zsrc.resize(np); uvsrc.resize(npuv);
    for (int32_t c = 0; c < np; ++c) zsrc[c] = c;
    for (int32_t i = 0; i < npuv; ++i) uvsrc[i] = i;
  }
  for (int32_t c = 0; c < np; ++c)
    if (zsrc[c] >= npf) { set_error(std::string(path) + ": fewer cells (np) than the mesh needs"); return SFX_ERR_IO; }
  for (int32_t i = 0; i < npuv; ++i)
    if (uvsrc[i] >= nuvf) { set_error(std::string(path) + ": fewer uv-points (npuv) than the mesh needs"); return SFX_ERR_IO; }
  const std::vector<int64_t> sz{npf}, suv{nuvf}, sz2{npf, nl}, suv2{nuvf, nl};
  std::vector<float> zmin, zmax, volmax, zlev, uzmin, uzmax, navg, ffit, havg, nrep, pwet;
  const bool ok = get_var(f, "z_zmin", sz, zmin, err, path) && get_var(f, "z_zmax", sz, zmax, err, path) &&
                  get_var(f, "z_volmax", sz, volmax, err, path) && get_var(f, "z_level", sz2, zlev, err, path) &&
                  get_var(f, "uv_zmin", suv, uzmin, err, path) && get_var(f, "uv_zmax", suv, uzmax, err, path) &&
                  get_var(f, "uv_navg", suv, navg, err, path) && get_var(f, "uv_ffit", suv, ffit, err, path) &&
                  get_var(f, "uv_havg", suv2, havg, err, path) && get_var(f, "uv_nrep", suv2, nrep, err, path) &&
                  get_var(f, "uv_pwet", suv2, pwet, err, path);
  if (!ok) { set_error(err); return SFX_ERR_IO; }
  // active-order copies, levels fastest (the reference's (level, point) arrays)
  auto pick = [](const std::vector<float>& a, const std::vector<int32_t>& src, int64_t w) {
    std::vector<float> o(src.size() * (size_t)w);
    for (size_t k = 0; k < src.size(); ++k)
      std::memcpy(o.data() + k * w, a.data() + (size_t)src[k] * w, (size_t)w * sizeof(float));
    return o;
  };
  const std::vector<float> a_zmin = pick(zmin, zsrc, 1), a_zmax = pick(zmax, zsrc, 1), a_vol = pick(volmax, zsrc, 1),
                           a_lev = pick(zlev, zsrc, nl), a_uzmin = pick(uzmin, uvsrc, 1),
                           a_uzmax = pick(uzmax, uvsrc, 1), a_navg = pick(navg, uvsrc, 1),
                           a_ffit = pick(ffit, uvsrc, 1), a_havg = pick(havg, uvsrc, nl),
                           a_nrep = pick(nrep, uvsrc, nl), a_pwet = pick(pwet, uvsrc, nl);
  sfx_subgrid_input s{};
  s.nlevels = (int32_t)nl;
  s.z_zmin = a_zmin.data(); s.z_zmax = a_zmax.data(); s.z_volmax = a_vol.data(); s.z_level = a_lev.data();
  s.uv_zmin = a_uzmin.data(); s.uv_zmax = a_uzmax.data(); s.uv_navg = a_navg.data(); s.uv_ffit = a_ffit.data();
  s.uv_havg = a_havg.data(); s.uv_nrep = a_nrep.data(); s.uv_pwet = a_pwet.data();
  return sfx_model_set_subgrid(M, &s);
}

}  // namespace

extern "C" int sfx_model_create_quadtree_netcdf(const sfx_params* p, const char* path, int32_t with_bed,
                                                sfx_model** out) {
  if (!p || !path || !out) { set_error("sfx_model_create_quadtree_netcdf: bad args"); return SFX_ERR_ARG; }
  *out = nullptr;
  try {
    return create_quadtree_netcdf(p, path, with_bed, out);
  } catch (const std::bad_alloc&) {
    set_error(std::string(path) + ": out of host memory");
    return SFX_ERR_MEMORY;
  }
}

extern "C" int sfx_model_read_subgrid_netcdf(sfx_model* M, const char* path) {
  if (!M || !path || !M->mesh_built || M->initialized) { set_error("sfx_model_read_subgrid_netcdf: bad state"); return SFX_ERR_STATE; }
  try {
    return read_subgrid_netcdf(M, path);
  } catch (const std::bad_alloc&) {
    set_error(std::string(path) + ": out of host memory");
    return SFX_ERR_MEMORY;
  }
}
