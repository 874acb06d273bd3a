// dxchange HDF5 containers of the reference pipeline (see include/mosaic_h5.h), on the HDF5 C library.
//
// File layout written and read here (iohdf5/dxchange_hdf5_chunks.py:182-330 defines it):
//   <dir>/<base><ext>                       master, libver latest:
//       /exchange/data   virtual (nproj, ny, nx) float32, fill -5, attributes
//                        nbanks_per_svchunk, stype ("proj" | "slice"), vchunks_0..2, chunks_0..2
//       /exchange/theta  float32 degrees (step2_radon.py:107,137)
//   <dir>/<base>/<base>_data_%06d<ext>      bank files, /exchange/data chunked float32, no fill;
//       the bank axis is 0 (proj) or 1 (slice); the axis is cut into vchunks of vchunks[axis] items and
//       every vchunk into nbanks banks of ceil(vchunks[axis] / nbanks) items (the last ones shorter or
//       empty; empty banks get no file).  Virtual sources name the banks relative to the master's
//       directory, so the set can be moved as a whole.
#include <hdf5.h>

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mosaic_h5.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Owns one HDF5 identifier; closes it with the matching H5?close.
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id() = default;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  void reset(hid_t i, herr_t (*c)(hid_t)) {
    if (id >= 0 && close) close(id);
    id = i;
    close = c;
  }
  bool ok() const { return id >= 0; }
  operator hid_t() const { return id; }
};

// HDF5's own error stack is silenced (failures come back as codes and messages); MM_H5_TRACE=1 prints it.
void quiet() {
  static const bool trace = std::getenv("MM_H5_TRACE") != nullptr;
  if (!trace) H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
}

#define H5_CHECK(expr, what)                                    \
  do {                                                          \
    if ((expr) < 0) return fail(MM_H5_ERR_IO, std::string(what)); \
  } while (0)

struct Plan {  // banking plan of iohdf5/dxchange_hdf5_chunks.py:182-221
  std::string dir, base, ext;
  int axis = 0;
  std::vector<long long> lo, hi;  // bank item ranges along axis (hi <= lo: empty)
  std::string file(size_t b) const {  // path of bank b
    char num[16];
    std::snprintf(num, sizeof num, "%06zu", b);
    return (dir.empty() ? std::string() : dir + "/") + base + "/" + base + "_data_" + num + ext;
  }
  std::string vsrc(size_t b) const {  // the same, relative to the master's directory
    char num[16];
    std::snprintf(num, sizeof num, "%06zu", b);
    return base + "/" + base + "_data_" + num + ext;
  }
};

Plan make_plan(const std::string& path, const long long* shape, const long long* vchunks, int axis, int nbanks) {
  Plan p;
  const size_t slash = path.find_last_of('/');
  p.dir = slash == std::string::npos ? "" : path.substr(0, slash);
  std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
  const size_t dot = name.find_last_of('.');
  p.base = dot == std::string::npos ? name : name.substr(0, dot);
  p.ext = dot == std::string::npos ? "" : name.substr(dot);
  p.axis = axis;
  const long long vc = vchunks[axis], len = shape[axis];
  const long long per = (vc + nbanks - 1) / nbanks;
  const long long nchunks = (len + vc - 1) / vc;
  for (long long c = 0; c < nchunks; ++c)
    for (int b = 0; b < nbanks; ++b) {
      const long long off = c * vc;
      const long long s = off + b * per;
      const long long e = std::min({off + (b + 1) * per, off + vc, len});
      p.lo.push_back(s);
      p.hi.push_back(e);
    }
  return p;
}

bool mkdir_p(const std::string& d) {
  if (d.empty()) return true;
  std::string cur;
  for (size_t i = 0; i <= d.size(); ++i) {
    if (i == d.size() || d[i] == '/') {
      if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
    if (i < d.size()) cur += d[i];
  }
  return true;
}

int write_int_attr(hid_t obj, const char* name, long long v) {
  Id sp(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(obj, name, H5T_STD_I64LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return fail(MM_H5_ERR_IO, std::string("H5Acreate2 ") + name);
  H5_CHECK(H5Awrite(a, H5T_NATIVE_LLONG, &v), std::string("H5Awrite ") + name);
  return MM_H5_OK;
}

int write_str_attr(hid_t obj, const char* name, const char* v) {  // variable-length UTF-8, as h5py writes str
  Id t(H5Tcopy(H5T_C_S1), H5Tclose);
  H5_CHECK(H5Tset_size(t, H5T_VARIABLE), "H5Tset_size");
  H5_CHECK(H5Tset_cset(t, H5T_CSET_UTF8), "H5Tset_cset");
  Id sp(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return fail(MM_H5_ERR_IO, std::string("H5Acreate2 ") + name);
  H5_CHECK(H5Awrite(a, t, &v), std::string("H5Awrite ") + name);
  return MM_H5_OK;
}

bool read_int_attr(hid_t obj, const char* name, long long* v) {
  if (H5Aexists(obj, name) <= 0) return false;
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  return a.ok() && H5Aread(a, H5T_NATIVE_LLONG, v) >= 0;
}

bool read_str_attr(hid_t obj, const char* name, std::string* v) {
  if (H5Aexists(obj, name) <= 0) return false;
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return false;
  Id ft(H5Aget_type(a), H5Tclose);
  if (H5Tget_class(ft) != H5T_STRING) return false;
  if (H5Tis_variable_str(ft) > 0) {
    Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, H5T_VARIABLE);
    H5Tset_cset(mt, H5Tget_cset(ft));  // no conversion between ASCII and UTF-8 strings
    char* s = nullptr;
    if (H5Aread(a, mt, &s) < 0 || !s) return false;
    *v = s;
    H5free_memory(s);
    return true;
  }
  const size_t n = H5Tget_size(ft);
  std::vector<char> buf(n + 1, 0);
  Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(mt, n + 1);
  H5Tset_cset(mt, H5Tget_cset(ft));
  if (H5Aread(a, mt, buf.data()) < 0) return false;
  *v = buf.data();
  return true;
}

bool box_ok(const hsize_t* dims, const long long* start, const long long* count) {
  for (int i = 0; i < 3; ++i)
    if (start[i] < 0 || count[i] < 1 || (hsize_t)(start[i] + count[i]) > dims[i]) return false;
  return true;
}

int open_data(const char* path, unsigned mode, Id& f, Id& d, hsize_t* dims) {
  quiet();
  f.reset(H5Fopen(path, mode, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(MM_H5_ERR_IO, std::string("cannot open ") + path);
  d.reset(H5Dopen2(f, "/exchange/data", H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(MM_H5_ERR_IO, std::string(path) + ": no /exchange/data");
  Id s(H5Dget_space(d), H5Sclose);
  if (H5Sget_simple_extent_ndims(s) != 3) return fail(MM_H5_ERR_ARG, std::string(path) + ": /exchange/data is not 3-D");
  H5Sget_simple_extent_dims(s, dims, nullptr);
  return MM_H5_OK;
}

}  // namespace

extern "C" {

const char* mm_h5_last_error(void) { return g_err.c_str(); }

int mm_h5_info_get(const char* path, mm_h5_info* info) {
  if (!path || !info) return fail(MM_H5_ERR_ARG, "mm_h5_info_get: null pointer");
  Id f, d;
  hsize_t dims[3];
  int rc = open_data(path, H5F_ACC_RDONLY, f, d, dims);
  if (rc) return rc;
  std::memset(info, 0, sizeof *info);
  for (int i = 0; i < 3; ++i) info->shape[i] = (long long)dims[i];
  Id dcpl(H5Dget_create_plist(d), H5Pclose);
  const H5D_layout_t lay = H5Pget_layout(dcpl);
  info->is_virtual = lay == H5D_VIRTUAL;
  long long nb = 1;
  read_int_attr(d, "nbanks_per_svchunk", &nb);
  info->nbanks_per_svchunk = (int)nb;
  for (int i = 0; i < 3; ++i) {
    char name[16];
    std::snprintf(name, sizeof name, "vchunks_%d", i);
    if (!read_int_attr(d, name, &info->vchunks[i])) info->vchunks[i] = info->shape[i];
  }
  std::string st = "proj";
  read_str_attr(d, "stype", &st);
  info->stype = (st.rfind("proj", 0) == 0) ? MM_H5_STYPE_PROJ : MM_H5_STYPE_SLICE;
  if (lay == H5D_CHUNKED) {
    hsize_t c[3];
    H5Pget_chunk(dcpl, 3, c);
    for (int i = 0; i < 3; ++i) info->chunks[i] = (long long)c[i];
  } else if (lay == H5D_VIRTUAL) {
    // a virtual dataset records its banks' chunk shape; without the attributes it follows stype
    bool have = true;
    for (int i = 0; i < 3; ++i) {
      char name[16];
      std::snprintf(name, sizeof name, "chunks_%d", i);
      have = have && read_int_attr(d, name, &info->chunks[i]);
    }
    if (!have) {
      if (info->stype == MM_H5_STYPE_PROJ) {
        info->chunks[0] = 1, info->chunks[1] = info->vchunks[1], info->chunks[2] = info->vchunks[2];
      } else {
        info->chunks[0] = info->vchunks[0], info->chunks[1] = 1, info->chunks[2] = info->vchunks[2];
      }
    }
  }
  return MM_H5_OK;
}

int mm_h5_read(const char* path, const long long* start, const long long* count, float* dst) {
  if (!path || !start || !count || !dst) return fail(MM_H5_ERR_ARG, "mm_h5_read: null pointer");
  Id f, d;
  hsize_t dims[3];
  int rc = open_data(path, H5F_ACC_RDONLY, f, d, dims);
  if (rc) return rc;
  if (!box_ok(dims, start, count)) return fail(MM_H5_ERR_ARG, "mm_h5_read: box outside /exchange/data");
  const hsize_t st[3] = {(hsize_t)start[0], (hsize_t)start[1], (hsize_t)start[2]};
  const hsize_t ct[3] = {(hsize_t)count[0], (hsize_t)count[1], (hsize_t)count[2]};
  Id fs(H5Dget_space(d), H5Sclose);
  H5_CHECK(H5Sselect_hyperslab(fs, H5S_SELECT_SET, st, nullptr, ct, nullptr), "H5Sselect_hyperslab");
  Id ms(H5Screate_simple(3, ct, nullptr), H5Sclose);
  H5_CHECK(H5Dread(d, H5T_NATIVE_FLOAT, ms, fs, H5P_DEFAULT, dst), std::string("H5Dread ") + path);
  return MM_H5_OK;
}

int mm_h5_create(const char* path, const long long* shape, const long long* vchunks_in, int stype, int nbanks,
                 const long long* chunks_in) {
  if (!path || !shape) return fail(MM_H5_ERR_ARG, "mm_h5_create: null pointer");
  if (stype != MM_H5_STYPE_PROJ && stype != MM_H5_STYPE_SLICE) return fail(MM_H5_ERR_ARG, "mm_h5_create: bad stype");
  if (nbanks < 1) return fail(MM_H5_ERR_ARG, "mm_h5_create: nbanks < 1");
  long long vchunks[3], chunks[3];
  for (int i = 0; i < 3; ++i) {
    if (shape[i] < 1) return fail(MM_H5_ERR_ARG, "mm_h5_create: empty shape");
    vchunks[i] = vchunks_in ? vchunks_in[i] : shape[i];
    if (vchunks[i] < 1) return fail(MM_H5_ERR_ARG, "mm_h5_create: vchunks must be positive");
  }
  quiet();
  const int axis = stype == MM_H5_STYPE_PROJ ? 0 : 1;
  const Plan plan = make_plan(path, shape, vchunks, axis, nbanks);
  const long long per = (vchunks[axis] + nbanks - 1) / nbanks;
  long long nominal[3] = {shape[0], shape[1], shape[2]};
  nominal[axis] = per;
  for (int i = 0; i < 3; ++i) {  // default: one item of the bank axis by a whole vchunk of the others
    chunks[i] = chunks_in ? chunks_in[i] : (i == axis ? 1 : vchunks[i]);
    if (chunks[i] < 1) return fail(MM_H5_ERR_ARG, "mm_h5_create: chunks must be positive");
  }
  if (!mkdir_p((plan.dir.empty() ? std::string() : plan.dir + "/") + plan.base))
    return fail(MM_H5_ERR_IO, "mm_h5_create: cannot create the bank directory");
  // bank files
  const hsize_t vdims[3] = {(hsize_t)shape[0], (hsize_t)shape[1], (hsize_t)shape[2]};
  Id dcpl_v(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
  const float fill = -5.0f;
  H5_CHECK(H5Pset_fill_value(dcpl_v, H5T_NATIVE_FLOAT, &fill), "H5Pset_fill_value");
  for (size_t b = 0; b < plan.lo.size(); ++b) {
    if (plan.hi[b] <= plan.lo[b]) continue;
    hsize_t bdims[3] = {vdims[0], vdims[1], vdims[2]};
    bdims[axis] = (hsize_t)(plan.hi[b] - plan.lo[b]);
    hsize_t bch[3];
    for (int i = 0; i < 3; ++i) bch[i] = (hsize_t)std::min<long long>(chunks[i], (long long)bdims[i]);
    {
      Id bf(H5Fcreate(plan.file(b).c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
      if (!bf.ok()) return fail(MM_H5_ERR_IO, "cannot create " + plan.file(b));
      Id g(H5Gcreate2(bf, "/exchange", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
      Id sp(H5Screate_simple(3, bdims, nullptr), H5Sclose);
      Id dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
      H5_CHECK(H5Pset_chunk(dcpl, 3, bch), "H5Pset_chunk");
      H5_CHECK(H5Pset_fill_time(dcpl, H5D_FILL_TIME_NEVER), "H5Pset_fill_time");
      Id ds(H5Dcreate2(bf, "/exchange/data", H5T_IEEE_F32LE, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT), H5Dclose);
      if (!ds.ok()) return fail(MM_H5_ERR_IO, "cannot create /exchange/data in " + plan.file(b));
    }
    Id vsp(H5Screate_simple(3, vdims, nullptr), H5Sclose);
    hsize_t vst[3] = {0, 0, 0};
    vst[axis] = (hsize_t)plan.lo[b];
    H5_CHECK(H5Sselect_hyperslab(vsp, H5S_SELECT_SET, vst, nullptr, bdims, nullptr), "H5Sselect_hyperslab");
    Id ssp(H5Screate_simple(3, bdims, nullptr), H5Sclose);
    H5_CHECK(H5Pset_virtual(dcpl_v, vsp, plan.vsrc(b).c_str(), "/exchange/data", ssp), "H5Pset_virtual");
  }
  // master
  Id fapl(H5Pcreate(H5P_FILE_ACCESS), H5Pclose);
  H5_CHECK(H5Pset_libver_bounds(fapl, H5F_LIBVER_LATEST, H5F_LIBVER_LATEST), "H5Pset_libver_bounds");
  Id mf(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, fapl), H5Fclose);
  if (!mf.ok()) return fail(MM_H5_ERR_IO, std::string("cannot create ") + path);
  Id g(H5Gcreate2(mf, "/exchange", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  Id sp(H5Screate_simple(3, vdims, nullptr), H5Sclose);
  Id ds(H5Dcreate2(mf, "/exchange/data", H5T_IEEE_F32LE, sp, H5P_DEFAULT, dcpl_v, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(MM_H5_ERR_IO, "cannot create the virtual /exchange/data");
  int rc;
  if ((rc = write_int_attr(ds, "nbanks_per_svchunk", nbanks)) ||
      (rc = write_str_attr(ds, "stype", stype == MM_H5_STYPE_PROJ ? "proj" : "slice")))
    return rc;
  for (int i = 0; i < 3; ++i) {
    char name[16];
    std::snprintf(name, sizeof name, "vchunks_%d", i);
    if ((rc = write_int_attr(ds, name, vchunks[i]))) return rc;
    std::snprintf(name, sizeof name, "chunks_%d", i);
    if ((rc = write_int_attr(ds, name, std::min(chunks[i], nominal[i])))) return rc;
  }
  return MM_H5_OK;
}

int mm_h5_write(const char* path, const long long* start, const long long* count, const float* src) {
  if (!path || !start || !count || !src) return fail(MM_H5_ERR_ARG, "mm_h5_write: null pointer");
  mm_h5_info info;
  int rc = mm_h5_info_get(path, &info);
  if (rc) return rc;
  hsize_t dims[3];
  for (int i = 0; i < 3; ++i) dims[i] = (hsize_t)info.shape[i];
  if (!box_ok(dims, start, count)) return fail(MM_H5_ERR_ARG, "mm_h5_write: box outside /exchange/data");
  const hsize_t ct[3] = {(hsize_t)count[0], (hsize_t)count[1], (hsize_t)count[2]};
  auto write_into = [&](const char* file, const hsize_t* fstart, const hsize_t* mstart, const hsize_t* cnt) {
    Id f, d;
    hsize_t fd[3];
    int r = open_data(file, H5F_ACC_RDWR, f, d, fd);
    if (r) return r;
    Id fs(H5Dget_space(d), H5Sclose);
    H5_CHECK(H5Sselect_hyperslab(fs, H5S_SELECT_SET, fstart, nullptr, cnt, nullptr), "H5Sselect_hyperslab");
    Id ms(H5Screate_simple(3, ct, nullptr), H5Sclose);
    H5_CHECK(H5Sselect_hyperslab(ms, H5S_SELECT_SET, mstart, nullptr, cnt, nullptr), "H5Sselect_hyperslab");
    H5_CHECK(H5Dwrite(d, H5T_NATIVE_FLOAT, ms, fs, H5P_DEFAULT, src), std::string("H5Dwrite ") + file);
    return MM_H5_OK;
  };
  if (!info.is_virtual) {
    const hsize_t st[3] = {(hsize_t)start[0], (hsize_t)start[1], (hsize_t)start[2]};
    const hsize_t zero[3] = {0, 0, 0};
    return write_into(path, st, zero, ct);
  }
  // the bank files that cover the box along the bank axis, as the writer's banking plan lays them out
  const int axis = info.stype == MM_H5_STYPE_PROJ ? 0 : 1;
  const Plan plan = make_plan(path, info.shape, info.vchunks, axis, info.nbanks_per_svchunk);
  const long long a0 = start[axis], a1 = start[axis] + count[axis];
  for (size_t b = 0; b < plan.lo.size(); ++b) {
    const long long lo = std::max(a0, plan.lo[b]), hi = std::min(a1, plan.hi[b]);
    if (hi <= lo) continue;
    hsize_t fst[3] = {(hsize_t)start[0], (hsize_t)start[1], (hsize_t)start[2]};
    hsize_t mst[3] = {0, 0, 0};
    hsize_t cnt[3] = {ct[0], ct[1], ct[2]};
    fst[axis] = (hsize_t)(lo - plan.lo[b]);
    mst[axis] = (hsize_t)(lo - a0);
    cnt[axis] = (hsize_t)(hi - lo);
    if ((rc = write_into(plan.file(b).c_str(), fst, mst, cnt))) return rc;
  }
  return MM_H5_OK;
}

int mm_h5_write_theta(const char* path, const float* theta_deg, long long n) {
  if (!path || !theta_deg || n < 1) return fail(MM_H5_ERR_ARG, "mm_h5_write_theta: bad argument");
  quiet();
  Id f(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(MM_H5_ERR_IO, std::string("cannot open ") + path);
  if (H5Lexists(f, "/exchange", H5P_DEFAULT) <= 0) {
    Id g(H5Gcreate2(f, "/exchange", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(MM_H5_ERR_IO, "cannot create /exchange");
  }
  if (H5Lexists(f, "/exchange/theta", H5P_DEFAULT) > 0) H5_CHECK(H5Ldelete(f, "/exchange/theta", H5P_DEFAULT), "H5Ldelete");
  const hsize_t dn = (hsize_t)n;
  Id sp(H5Screate_simple(1, &dn, nullptr), H5Sclose);
  Id ds(H5Dcreate2(f, "/exchange/theta", H5T_IEEE_F32LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(MM_H5_ERR_IO, "cannot create /exchange/theta");
  H5_CHECK(H5Dwrite(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, theta_deg), "H5Dwrite theta");
  return MM_H5_OK;
}

int mm_h5_read_theta(const char* path, float* theta_deg, long long* n) {
  if (!path || !n) return fail(MM_H5_ERR_ARG, "mm_h5_read_theta: null pointer");
  quiet();
  Id f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(MM_H5_ERR_IO, std::string("cannot open ") + path);
  Id d(H5Dopen2(f, "/exchange/theta", H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(MM_H5_ERR_IO, std::string(path) + ": no /exchange/theta");
  Id s(H5Dget_space(d), H5Sclose);
  const hssize_t len = H5Sget_simple_extent_npoints(s);
  if (!theta_deg) {
    *n = (long long)len;
    return MM_H5_OK;
  }
  if (*n < len) return fail(MM_H5_ERR_ARG, "mm_h5_read_theta: buffer of " + std::to_string(*n) + " for " +
                                               std::to_string(len) + " angles");
  H5_CHECK(H5Dread(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, theta_deg), "H5Dread theta");
  *n = (long long)len;
  return MM_H5_OK;
}

}  // extern "C"
