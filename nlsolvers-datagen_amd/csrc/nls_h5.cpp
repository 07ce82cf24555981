// HDF5 trajectory files (layout of scripts/real_launcher_2d.py:201-245,
// scripts/real_launcher_3d.py:206-252, scripts/complex_launcher_2d.py:193-236)
// written through libhdf5's C API, loaded with dlopen so the rest of libnls has no
// link-time dependency on it.  Types and constants below are those of the HDF5 1.10
// public API (H5public.h, H5Ipublic.h, H5Spublic.h, H5Tpublic.h, H5Fpublic.h).
//
// Python's h5py, which the reference launchers use, stores str attributes as
// variable-length UTF-8 strings, int as 64-bit integers, float as doubles and
// complex128 as a compound {r, i} of doubles; the files here use the same types, so
// readers of the reference's files read these unchanged.
#include "nls_h5.hpp"
#include "nls_common.hpp"

#include <dlfcn.h>

#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <deque>
#include <mutex>
#include <thread>

namespace nls {
namespace h5 {
namespace {

using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0, E_DEFAULT = 0;
constexpr unsigned ACC_RDONLY = 0u, ACC_TRUNC = 2u;
constexpr int S_SCALAR = 0, S_SELECT_SET = 0, T_STRING = 3, T_COMPOUND = 6, T_CSET_UTF8 = 1;
constexpr hsize_t S_UNLIMITED = ~0ull;
constexpr size_t T_VARIABLE = ~(size_t)0;

struct Api {
  void *lib = nullptr;
  std::string error, version;
  herr_t (*open)(void);
  herr_t (*get_libversion)(unsigned *, unsigned *, unsigned *);
  herr_t (*eset_auto)(hid_t, void *, void *);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  herr_t (*fflush)(hid_t, int);
  hid_t (*gcreate)(hid_t, const char *, hid_t, hid_t, hid_t);
  hid_t (*gopen)(hid_t, const char *, hid_t);
  herr_t (*gclose)(hid_t);
  hid_t (*screate)(int);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
  herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *,
                              const hsize_t *);
  herr_t (*sclose)(hid_t);
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen)(hid_t, const char *, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  herr_t (*dset_extent)(hid_t, const hsize_t *);
  hid_t (*dget_space)(hid_t);
  hid_t (*dget_type)(hid_t);
  herr_t (*dclose)(hid_t);
  hid_t (*pcreate)(hid_t);
  herr_t (*pset_chunk)(hid_t, int, const hsize_t *);
  herr_t (*pclose)(hid_t);
  hid_t (*acreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
  hid_t (*aopen_by_name)(hid_t, const char *, const char *, hid_t, hid_t);
  herr_t (*awrite)(hid_t, hid_t, const void *);
  herr_t (*aread)(hid_t, hid_t, void *);
  hid_t (*aget_type)(hid_t);
  herr_t (*aclose)(hid_t);
  hid_t (*tcopy)(hid_t);
  herr_t (*tset_size)(hid_t, size_t);
  herr_t (*tset_cset)(hid_t, int);
  hid_t (*tcreate)(int, size_t);
  herr_t (*tinsert)(hid_t, const char *, size_t, hid_t);
  int (*tget_class)(hid_t);
  size_t (*tget_size)(hid_t);
  htri_t (*tis_variable_str)(hid_t);
  herr_t (*tclose)(hid_t);
  herr_t (*free_memory)(void *);
  hid_t *native_double, *native_llong, *c_s1, *dataset_create;
  // NetCDF-4 class files
  hid_t *native_float, *native_int, *ieee_f32be;
  herr_t (*pset_fill_value)(hid_t, hid_t, const void *);
  void *hl = nullptr; // libhdf5_hl (dimension scales); optional
  herr_t (*ds_set_scale)(hid_t, const char *) = nullptr;
  herr_t (*ds_attach_scale)(hid_t, hid_t, unsigned) = nullptr;
};

std::mutex g_mu; // every HDF5 call runs under this lock

template <class F> bool bind(Api &a, F &f, const char *name) {
  f = reinterpret_cast<F>(dlsym(a.lib, name));
  if (!f) a.error = std::string("libhdf5 lacks ") + name;
  return f != nullptr;
}

Api *load() {
  static Api api;
  static bool tried = false;
  if (tried) return api.lib && api.error.empty() ? &api : nullptr;
  tried = true;
  std::vector<std::string> names;
  if (const char *e = std::getenv("NLS_HDF5_LIB")) names.push_back(e);
  names.insert(names.end(), {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"});
  for (auto &n : names) {
    api.lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (api.lib) break;
  }
  if (!api.lib) {
    api.error = "libhdf5 not found (set NLS_HDF5_LIB to its path)";
    return nullptr;
  }
  Api &a = api;
  bool ok = bind(a, a.open, "H5open") && bind(a, a.get_libversion, "H5get_libversion") &&
            bind(a, a.eset_auto, "H5Eset_auto2") && bind(a, a.fcreate, "H5Fcreate") &&
            bind(a, a.fopen, "H5Fopen") && bind(a, a.fclose, "H5Fclose") &&
            bind(a, a.fflush, "H5Fflush") && bind(a, a.gcreate, "H5Gcreate2") &&
            bind(a, a.gopen, "H5Gopen2") &&
            bind(a, a.gclose, "H5Gclose") && bind(a, a.screate, "H5Screate") &&
            bind(a, a.screate_simple, "H5Screate_simple") &&
            bind(a, a.sget_ndims, "H5Sget_simple_extent_ndims") &&
            bind(a, a.sget_dims, "H5Sget_simple_extent_dims") &&
            bind(a, a.sselect_hyperslab, "H5Sselect_hyperslab") && bind(a, a.sclose, "H5Sclose") &&
            bind(a, a.dcreate, "H5Dcreate2") && bind(a, a.dopen, "H5Dopen2") &&
            bind(a, a.dwrite, "H5Dwrite") && bind(a, a.dread, "H5Dread") &&
            bind(a, a.dset_extent, "H5Dset_extent") && bind(a, a.dget_space, "H5Dget_space") &&
            bind(a, a.dget_type, "H5Dget_type") && bind(a, a.dclose, "H5Dclose") &&
            bind(a, a.pcreate, "H5Pcreate") && bind(a, a.pset_chunk, "H5Pset_chunk") &&
            bind(a, a.pclose, "H5Pclose") && bind(a, a.acreate, "H5Acreate2") &&
            bind(a, a.aopen_by_name, "H5Aopen_by_name") && bind(a, a.awrite, "H5Awrite") &&
            bind(a, a.aread, "H5Aread") && bind(a, a.aget_type, "H5Aget_type") &&
            bind(a, a.aclose, "H5Aclose") && bind(a, a.tcopy, "H5Tcopy") &&
            bind(a, a.tset_size, "H5Tset_size") && bind(a, a.tset_cset, "H5Tset_cset") &&
            bind(a, a.tcreate, "H5Tcreate") && bind(a, a.tinsert, "H5Tinsert") &&
            bind(a, a.tget_class, "H5Tget_class") && bind(a, a.tget_size, "H5Tget_size") &&
            bind(a, a.tis_variable_str, "H5Tis_variable_str") && bind(a, a.tclose, "H5Tclose") &&
            bind(a, a.free_memory, "H5free_memory") &&
            bind(a, a.native_double, "H5T_NATIVE_DOUBLE_g") &&
            bind(a, a.native_llong, "H5T_NATIVE_LLONG_g") && bind(a, a.c_s1, "H5T_C_S1_g") &&
            bind(a, a.dataset_create, "H5P_CLS_DATASET_CREATE_ID_g") &&
            bind(a, a.native_float, "H5T_NATIVE_FLOAT_g") &&
            bind(a, a.native_int, "H5T_NATIVE_INT_g") &&
            bind(a, a.ieee_f32be, "H5T_IEEE_F32BE_g") &&
            bind(a, a.pset_fill_value, "H5Pset_fill_value");
  if (!ok) return nullptr;
  // the high-level library (dimension scales, for NetCDF-4 files), next to libhdf5
  std::vector<std::string> hl;
  if (const char *e = std::getenv("NLS_HDF5_HL_LIB")) hl.push_back(e);
  hl.insert(hl.end(), {"libhdf5_hl.so", "libhdf5_hl.so.100", "/opt/conda/lib/libhdf5_hl.so"});
  for (auto &n : hl) {
    a.hl = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (a.hl) break;
  }
  if (a.hl) {
    a.ds_set_scale = reinterpret_cast<herr_t (*)(hid_t, const char *)>(dlsym(a.hl, "H5DSset_scale"));
    a.ds_attach_scale =
        reinterpret_cast<herr_t (*)(hid_t, hid_t, unsigned)>(dlsym(a.hl, "H5DSattach_scale"));
  }
  if (a.open() < 0) {
    a.error = "H5open failed";
    return nullptr;
  }
  a.eset_auto(E_DEFAULT, nullptr, nullptr); // errors come back as status codes only
  unsigned ma = 0, mi = 0, re = 0;
  a.get_libversion(&ma, &mi, &re);
  a.version = std::to_string(ma) + "." + std::to_string(mi) + "." + std::to_string(re);
  return &api;
}

Api &need() {
  Api *a = load();
  if (!a) throw Error("HDF5 unavailable: libhdf5 could not be loaded (set NLS_HDF5_LIB)");
  a->eset_auto(E_DEFAULT, nullptr, nullptr); // per thread: errors come back as status codes
  return *a;
}

hid_t chk(hid_t id, const std::string &what) {
  if (id < 0) throw Error("HDF5: " + what + " failed");
  return id;
}

// Scoped id closed by the matching H5*close.
struct Id {
  hid_t id;
  herr_t (*close)(hid_t);
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Id() {
    if (id >= 0) close(id);
  }
  Id(const Id &) = delete;
  Id &operator=(const Id &) = delete;
};

hid_t complex_type(Api &a) {
  hid_t t = chk(a.tcreate(T_COMPOUND, 16), "H5Tcreate");
  a.tinsert(t, "r", 0, *a.native_double);
  a.tinsert(t, "i", 8, *a.native_double);
  return t;
}

void attr_str(Api &a, hid_t obj, const char *name, const std::string &v) {
  Id t(chk(a.tcopy(*a.c_s1), "H5Tcopy"), a.tclose);
  a.tset_size(t.id, T_VARIABLE);
  a.tset_cset(t.id, T_CSET_UTF8);
  Id sp(chk(a.screate(S_SCALAR), "H5Screate"), a.sclose);
  Id at(chk(a.acreate(obj, name, t.id, sp.id, P_DEFAULT, P_DEFAULT), std::string("attribute ") + name),
        a.aclose);
  const char *p = v.c_str();
  if (a.awrite(at.id, t.id, &p) < 0) throw Error(std::string("HDF5: writing attribute ") + name);
}

void attr_f64(Api &a, hid_t obj, const char *name, double v) {
  Id sp(chk(a.screate(S_SCALAR), "H5Screate"), a.sclose);
  Id at(chk(a.acreate(obj, name, *a.native_double, sp.id, P_DEFAULT, P_DEFAULT),
            std::string("attribute ") + name),
        a.aclose);
  if (a.awrite(at.id, *a.native_double, &v) < 0) throw Error(std::string("HDF5: writing ") + name);
}

void attr_i64(Api &a, hid_t obj, const char *name, long long v) {
  Id sp(chk(a.screate(S_SCALAR), "H5Screate"), a.sclose);
  Id at(chk(a.acreate(obj, name, *a.native_llong, sp.id, P_DEFAULT, P_DEFAULT),
            std::string("attribute ") + name),
        a.aclose);
  if (a.awrite(at.id, *a.native_llong, &v) < 0) throw Error(std::string("HDF5: writing ") + name);
}

void dataset(Api &a, hid_t loc, const char *name, const std::vector<hsize_t> &dims,
             const double *data, bool cplx) {
  Id sp(chk(a.screate_simple((int)dims.size(), dims.data(), nullptr), "H5Screate_simple"), a.sclose);
  Id ty(cplx ? complex_type(a) : a.tcopy(*a.native_double), a.tclose);
  Id ds(chk(a.dcreate(loc, name, ty.id, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT),
            std::string("dataset ") + name),
        a.dclose);
  if (a.dwrite(ds.id, ty.id, S_ALL, S_ALL, P_DEFAULT, data) < 0)
    throw Error(std::string("HDF5: writing dataset ") + name);
}

std::string now_string() {
  // str(datetime.datetime.now()): "YYYY-MM-DD HH:MM:SS.ffffff"
  const auto tp = std::chrono::system_clock::now();
  const std::time_t t = std::chrono::system_clock::to_time_t(tp);
  const long us = (long)(std::chrono::duration_cast<std::chrono::microseconds>(
                             tp.time_since_epoch()).count() % 1000000);
  std::tm lt{};
  localtime_r(&t, &lt);
  char buf[64];
  std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", &lt);
  char out[80];
  std::snprintf(out, sizeof out, "%s.%06ld", buf, us);
  return out;
}

std::vector<hsize_t> grid_dims(const Run &r) {
  if (r.dim == 3) return {(hsize_t)r.nx, (hsize_t)r.ny, (hsize_t)r.nz};
  return {(hsize_t)r.nx, (hsize_t)r.ny};
}

std::vector<hsize_t> traj_dims(const Run &r, hsize_t ns) {
  if (r.dim == 3) return {ns, (hsize_t)r.ox, (hsize_t)r.oy, (hsize_t)r.oz};
  return {ns, (hsize_t)r.ox, (hsize_t)r.oy};
}

// Everything except the trajectory datasets u, v (and elapsed_time when a stream
// writes it at close).
void write_header(Api &a, hid_t f, const Run &r, bool with_elapsed = true) {
  {
    Id g(chk(a.gcreate(f, "metadata", P_DEFAULT, P_DEFAULT, P_DEFAULT), "group metadata"), a.gclose);
    attr_str(a, g.id, "problem_type", r.problem_type);
    attr_str(a, g.id, "boundary_condition", "noflux");
    attr_str(a, g.id, "run_id", r.run_id);
    attr_i64(a, g.id, "run_index", r.run_index);
    attr_str(a, g.id, "timestamp", r.timestamp.empty() ? now_string() : r.timestamp);
    if (with_elapsed) attr_f64(a, g.id, "elapsed_time", r.elapsed_time);
    attr_str(a, g.id, "phenomenon", r.phenomenon);
    for (auto &kv : r.params) attr_str(a, g.id, ("phenomenon_" + kv.first).c_str(), kv.second);
    // the conductivity class (the reference keeps it only in the run directory's name,
    // c_<class>, parallel_generate_problem_configs.py:91-92)
    if (!r.c_type.empty()) attr_str(a, g.id, "c_type", r.c_type);
  }
  {
    Id g(chk(a.gcreate(f, "grid", P_DEFAULT, P_DEFAULT, P_DEFAULT), "group grid"), a.gclose);
    attr_i64(a, g.id, "nx", r.nx);
    attr_i64(a, g.id, "ny", r.ny);
    if (r.dim == 3) attr_i64(a, g.id, "nz", r.nz);
    attr_f64(a, g.id, "Lx", r.Lx);
    attr_f64(a, g.id, "Ly", r.Ly);
    if (r.dim == 3) attr_f64(a, g.id, "Lz", r.Lz);
  }
  {
    Id g(chk(a.gcreate(f, "time", P_DEFAULT, P_DEFAULT, P_DEFAULT), "group time"), a.gclose);
    attr_f64(a, g.id, "T", r.T);
    attr_i64(a, g.id, "nt", r.nt);
    attr_i64(a, g.id, "num_snapshots", r.num_snapshots);
  }
  const auto gd = grid_dims(r);
  {
    Id g(chk(a.gcreate(f, "initial_condition", P_DEFAULT, P_DEFAULT, P_DEFAULT), "group ic"),
         a.gclose);
    dataset(a, g.id, "u0", gd, r.u0.data(), r.complex_field);
    if (!r.complex_field) dataset(a, g.id, "v0", gd, r.v0.data(), false);
  }
  {
    Id g(chk(a.gcreate(f, "focusing", P_DEFAULT, P_DEFAULT, P_DEFAULT), "group focusing"), a.gclose);
    attr_str(a, g.id, "type", r.m_type);
    dataset(a, g.id, "m", gd, r.m.data(), false);
    if (!r.complex_field) dataset(a, g.id, "c", gd, r.c.data(), false);
  }
  if (r.complex_field) dataset(a, f, "c", gd, r.c.data(), false);
  // X, Y: meshgrid(linspace(-L, L, n) ..., indexing='ij') on the solver grid
  const size_t N = (size_t)r.nx * r.ny * (r.dim == 3 ? r.nz : 1);
  std::vector<double> X(N), Y(N);
  // numpy.linspace(-L, L, n): k * step + start with step = (stop - start) / (n - 1),
  // the last point set to stop
  auto lin = [](double L, int n, int k) {
    if (n <= 1) return -L;
    if (k == n - 1) return L;
    const double step = (L - (-L)) / (double)(n - 1);
    return (double)k * step + (-L);
  };
  const int nz = r.dim == 3 ? r.nz : 1;
  for (int i = 0; i < r.nx; ++i)
    for (int j = 0; j < r.ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const size_t idx = ((size_t)i * r.ny + j) * nz + k;
        X[idx] = lin(r.Lx, r.nx, i);
        Y[idx] = lin(r.Ly, r.ny, j);
      }
  dataset(a, f, "X", gd, X.data(), false);
  dataset(a, f, "Y", gd, Y.data(), false);
}

void check_run(const Run &r) {
  NLS_REQUIRE(!r.path.empty(), "HDF5 run: empty path");
  NLS_REQUIRE(r.dim == 2 || r.dim == 3, "HDF5 run: dim must be 2 or 3");
  NLS_REQUIRE(r.nx > 0 && r.ny > 0 && r.nz > 0, "HDF5 run: bad grid");
  const size_t N = (size_t)r.nx * r.ny * (r.dim == 3 ? r.nz : 1);
  const size_t cw = r.complex_field ? 2 : 1;
  NLS_REQUIRE(r.u0.size() == N * cw && r.m.size() == N && r.c.size() == N,
              "HDF5 run: initial condition / coefficient arrays missing");
  NLS_REQUIRE(r.complex_field || r.v0.size() == N, "HDF5 run: v0 missing");
}

} // namespace

bool available(std::string *version) {
  std::lock_guard<std::mutex> lk(g_mu);
  Api *a = load();
  if (a && version) *version = a->version;
  return a != nullptr;
}

Run from_abi(const nls_h5_run &r) {
  NLS_REQUIRE(r.path, "HDF5 run: null path");
  Run o;
  o.path = r.path;
  o.complex_field = r.complex_field != 0;
  o.dim = r.dim;
  o.nx = r.nx;
  o.ny = r.ny;
  o.nz = r.dim == 3 ? r.nz : 1;
  o.Lx = r.Lx;
  o.Ly = r.Ly;
  o.Lz = r.Lz;
  o.T = r.T;
  o.nt = r.nt;
  o.num_snapshots = r.num_snapshots;
  auto str = [](const char *s) { return s ? std::string(s) : std::string(); };
  o.problem_type = str(r.problem_type);
  o.run_id = str(r.run_id);
  o.phenomenon = str(r.phenomenon);
  o.m_type = str(r.m_type);
  o.c_type = str(r.c_type);
  o.timestamp = str(r.timestamp);
  o.run_index = r.run_index;
  o.elapsed_time = r.elapsed_time;
  NLS_REQUIRE(r.n_params == 0 || (r.param_keys && r.param_values), "HDF5 run: null params");
  for (int k = 0; k < r.n_params; ++k)
    o.params.emplace_back(str(r.param_keys[k]), str(r.param_values[k]));
  NLS_REQUIRE(o.dim == 2 || o.dim == 3, "HDF5 run: dim must be 2 or 3");
  NLS_REQUIRE(o.nx > 0 && o.ny > 0 && o.nz > 0, "HDF5 run: bad grid");
  const size_t N = (size_t)o.nx * o.ny * o.nz, cw = o.complex_field ? 2 : 1;
  NLS_REQUIRE(r.u0 && r.m && r.c && (o.complex_field || r.v0), "HDF5 run: null grid array");
  o.u0.assign(r.u0, r.u0 + N * cw);
  if (!o.complex_field) o.v0.assign(r.v0, r.v0 + N);
  o.m.assign(r.m, r.m + N);
  o.c.assign(r.c, r.c + N);
  o.ns_out = r.ns_out;
  o.ox = r.ox;
  o.oy = r.oy;
  o.oz = r.dim == 3 ? r.oz : 1;
  NLS_REQUIRE(o.ns_out >= 0 && o.ox > 0 && o.oy > 0 && o.oz > 0, "HDF5 run: bad trajectory shape");
  const size_t T = (size_t)o.ns_out * o.ox * o.oy * o.oz;
  NLS_REQUIRE(T == 0 || (r.u && (o.complex_field || r.v)), "HDF5 run: null trajectory");
  if (T) {
    o.u.assign(r.u, r.u + T * cw);
    if (!o.complex_field) o.v.assign(r.v, r.v + T);
  }
  return o;
}

void write_run(const Run &r) {
  check_run(r);
  const size_t cw = r.complex_field ? 2 : 1;
  const size_t T = (size_t)r.ns_out * r.ox * r.oy * r.oz;
  NLS_REQUIRE(r.u.size() == T * cw && (r.complex_field || r.v.size() == T),
              "HDF5 run: trajectory size does not match ns_out * ox * oy * oz");
  std::lock_guard<std::mutex> lk(g_mu);
  Api &a = need();
  Id f(chk(a.fcreate(r.path.c_str(), ACC_TRUNC, P_DEFAULT, P_DEFAULT), "create " + r.path), a.fclose);
  write_header(a, f.id, r);
  const auto td = traj_dims(r, (hsize_t)r.ns_out);
  dataset(a, f.id, "u", td, r.u.data(), r.complex_field);
  if (!r.complex_field) dataset(a, f.id, "v", td, r.v.data(), false);
}

Stream::Stream(const Run &h) {
  check_run(h);
  complex_ = h.complex_field;
  has_v_ = !h.complex_field;
  std::lock_guard<std::mutex> lk(g_mu);
  Api &a = need();
  file_ = chk(a.fcreate(h.path.c_str(), ACC_TRUNC, P_DEFAULT, P_DEFAULT), "create " + h.path);
  try {
    write_header(a, file_, h, false);
    auto td = traj_dims(h, 0);
    auto mx = td;
    mx[0] = S_UNLIMITED;
    auto ch = td;
    ch[0] = 1; // one snapshot per chunk
    dims_.assign(td.begin(), td.end());
    Id sp(chk(a.screate_simple((int)td.size(), td.data(), mx.data()), "H5Screate_simple"), a.sclose);
    Id pl(chk(a.pcreate(*a.dataset_create), "H5Pcreate"), a.pclose);
    a.pset_chunk(pl.id, (int)ch.size(), ch.data());
    Id ty(complex_ ? complex_type(a) : a.tcopy(*a.native_double), a.tclose);
    du_ = chk(a.dcreate(file_, "u", ty.id, sp.id, P_DEFAULT, pl.id, P_DEFAULT), "dataset u");
    if (has_v_)
      dv_ = chk(a.dcreate(file_, "v", ty.id, sp.id, P_DEFAULT, pl.id, P_DEFAULT), "dataset v");
    a.fflush(file_, 0);
  } catch (...) {
    if (du_ >= 0) a.dclose(du_);
    if (dv_ >= 0) a.dclose(dv_);
    a.fclose(file_);
    file_ = du_ = dv_ = -1;
    throw;
  }
}

void Stream::append(const double *u, const double *v) {
  std::lock_guard<std::mutex> lk(g_mu);
  NLS_REQUIRE(file_ >= 0, "HDF5 stream: closed");
  Api &a = need();
  auto put = [&](hid_t d, const double *x) {
    std::vector<hsize_t> ext(dims_.begin(), dims_.end());
    ext[0] = (hsize_t)count_ + 1;
    if (a.dset_extent(d, ext.data()) < 0) throw Error("HDF5: extending a trajectory dataset");
    Id fs(chk(a.dget_space(d), "H5Dget_space"), a.sclose);
    std::vector<hsize_t> start(ext.size(), 0), cnt(ext.begin(), ext.end());
    start[0] = (hsize_t)count_;
    cnt[0] = 1;
    a.sselect_hyperslab(fs.id, S_SELECT_SET, start.data(), nullptr, cnt.data(), nullptr);
    Id ms(chk(a.screate_simple((int)cnt.size(), cnt.data(), nullptr), "H5Screate_simple"), a.sclose);
    Id ty(complex_ ? complex_type(a) : a.tcopy(*a.native_double), a.tclose);
    if (a.dwrite(d, ty.id, ms.id, fs.id, P_DEFAULT, x) < 0)
      throw Error("HDF5: appending a snapshot");
  };
  put(du_, u);
  if (has_v_) put(dv_, v);
  ++count_;
  a.fflush(file_, 0);
}

void Stream::close(double elapsed_time) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (file_ < 0) return;
  Api &a = need();
  {
    Id g(a.gopen(file_, "metadata", P_DEFAULT), a.gclose); // elapsed_time is known only now
    if (g.id >= 0) attr_f64(a, g.id, "elapsed_time", elapsed_time);
  }
  a.dclose(du_);
  if (dv_ >= 0) a.dclose(dv_);
  a.fclose(file_);
  file_ = du_ = dv_ = -1;
}

Stream::~Stream() {
  try {
    close(0.0);
  } catch (...) {
  }
}

// ---------------------------------------------------------------- NetCDF-4 class files
namespace {
// NetCDF-4 conventions on top of HDF5: a dimension without a coordinate variable is a
// dataset of its length (big-endian float, never written) made a dimension scale whose
// NAME is the netCDF marker below, with an int attribute _Netcdf4Dimid; variables attach
// the scales of their dimensions; attributes are one-element arrays (numbers) or
// fixed-length strings.  netCDF's default fill value for float variables is
// 9.9692099683868690e+36 (NC_FILL_FLOAT).
constexpr float kFillFloat = 9.9692099683868690e+36f;

void nc_attr_i32(Api &a, hid_t obj, const char *name, int v) {
  const hsize_t one = 1;
  Id sp(chk(a.screate_simple(1, &one, nullptr), "H5Screate_simple"), a.sclose);
  Id at(chk(a.acreate(obj, name, *a.native_int, sp.id, P_DEFAULT, P_DEFAULT),
            std::string("attribute ") + name),
        a.aclose);
  if (a.awrite(at.id, *a.native_int, &v) < 0) throw Error(std::string("HDF5: writing ") + name);
}

void nc_attr_f32(Api &a, hid_t obj, const char *name, float v) {
  const hsize_t one = 1;
  Id sp(chk(a.screate_simple(1, &one, nullptr), "H5Screate_simple"), a.sclose);
  Id at(chk(a.acreate(obj, name, *a.native_float, sp.id, P_DEFAULT, P_DEFAULT),
            std::string("attribute ") + name),
        a.aclose);
  if (a.awrite(at.id, *a.native_float, &v) < 0) throw Error(std::string("HDF5: writing ") + name);
}

void nc_attr_text(Api &a, hid_t obj, const char *name, const std::string &v) {
  Id t(chk(a.tcopy(*a.c_s1), "H5Tcopy"), a.tclose);
  a.tset_size(t.id, v.empty() ? 1 : v.size());
  Id sp(chk(a.screate(S_SCALAR), "H5Screate"), a.sclose);
  Id at(chk(a.acreate(obj, name, t.id, sp.id, P_DEFAULT, P_DEFAULT), std::string("attribute ") + name),
        a.aclose);
  const std::string buf = v.empty() ? std::string(1, '\0') : v;
  if (a.awrite(at.id, t.id, buf.data()) < 0) throw Error(std::string("HDF5: writing ") + name);
}

hid_t nc_dimension(Api &a, hid_t f, const char *name, hsize_t len, int dimid) {
  Id sp(chk(a.screate_simple(1, &len, nullptr), "H5Screate_simple"), a.sclose);
  const hid_t d = chk(a.dcreate(f, name, *a.ieee_f32be, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT),
                      std::string("dimension ") + name);
  char marker[96];
  std::snprintf(marker, sizeof marker, "This is a netCDF dimension but not a netCDF variable.%10llu",
                len);
  if (a.ds_set_scale(d, marker) < 0) throw Error(std::string("HDF5: dimension scale ") + name);
  nc_attr_i32(a, d, "_Netcdf4Dimid", dimid);
  return d;
}

hid_t nc_variable(Api &a, hid_t f, const char *name, hid_t dim, hsize_t len, bool chunked) {
  Id sp(chk(a.screate_simple(1, &len, nullptr), "H5Screate_simple"), a.sclose);
  Id pl(chk(a.pcreate(*a.dataset_create), "H5Pcreate"), a.pclose);
  if (chunked) {
    // chunksizes=(65536,) (aggregate_single_core.py:180-188), at most the length
    const hsize_t ch = len < 65536 ? (len > 0 ? len : 1) : 65536;
    a.pset_chunk(pl.id, 1, &ch);
  }
  a.pset_fill_value(pl.id, *a.native_float, &kFillFloat);
  const hid_t d = chk(a.dcreate(f, name, *a.native_float, sp.id, P_DEFAULT, pl.id, P_DEFAULT),
                      std::string("variable ") + name);
  if (a.ds_attach_scale(d, dim, 0) < 0) throw Error(std::string("HDF5: attaching ") + name);
  return d;
}

void nc_put(Api &a, hid_t d, hsize_t off, hsize_t count, const float *x) {
  if (!x || count == 0) return;
  Id fs(chk(a.dget_space(d), "H5Dget_space"), a.sclose);
  a.sselect_hyperslab(fs.id, S_SELECT_SET, &off, nullptr, &count, nullptr);
  Id ms(chk(a.screate_simple(1, &count, nullptr), "H5Screate_simple"), a.sclose);
  if (a.dwrite(d, *a.native_float, ms.id, fs.id, P_DEFAULT, x) < 0)
    throw Error("HDF5: writing a trajectory slab");
}
} // namespace

NcClass::NcClass(const std::string &path, int dims, const std::vector<int> &shape, int ns,
                 int ntraj, bool is_complex, double Lx, double Ly, double Lz)
    : ns_(ns), ntraj_(ntraj), complex_(is_complex) {
  NLS_REQUIRE(dims == 2 || dims == 3, "NetCDF class: dims must be 2 or 3");
  NLS_REQUIRE((int)shape.size() == dims, "NetCDF class: shape must have dims entries");
  NLS_REQUIRE(ns > 0 && ntraj > 0, "NetCDF class: need snapshots and trajectories");
  spat_ = 1;
  for (int x : shape) {
    NLS_REQUIRE(x > 0, "NetCDF class: bad shape");
    spat_ *= (hsize_t)x;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  Api &a = need();
  NLS_REQUIRE(a.ds_set_scale && a.ds_attach_scale,
              "NetCDF class files need libhdf5_hl (dimension scales); set NLS_HDF5_HL_LIB");
  file_ = chk(a.fcreate(path.c_str(), ACC_TRUNC, P_DEFAULT, P_DEFAULT), "create " + path);
  try {
    const hsize_t temp = (hsize_t)ns * spat_;
    dt_ = nc_dimension(a, file_, "total_temporal", (hsize_t)ntraj * temp, 0);
    ds_ = nc_dimension(a, file_, "total_spatial", (hsize_t)ntraj * spat_, 1);
    const char *names_t[2] = {complex_ ? "re_u" : "u", complex_ ? "im_u" : "v"};
    const char *names_s[2] = {complex_ ? "re_u0" : "u0", complex_ ? "im_u0" : "v0"};
    for (int k = 0; k < 2; ++k) {
      vt_[k] = nc_variable(a, file_, names_t[k], dt_, (hsize_t)ntraj * temp, true);
      vs_[k] = nc_variable(a, file_, names_s[k], ds_, (hsize_t)ntraj * spat_, true);
    }
    vm_ = nc_variable(a, file_, "m", ds_, (hsize_t)ntraj * spat_, false);
    vc_ = nc_variable(a, file_, "c", ds_, (hsize_t)ntraj * spat_, false);
    // global attributes (aggregate_single_core.py:193-201)
    nc_attr_i32(a, file_, "num_snapshots", ns);
    nc_attr_i32(a, file_, "num_trajectories", ntraj);
    std::string sh = "(";
    for (int k = 0; k < dims; ++k) sh += std::to_string(shape[k]) + (k + 1 < dims ? ", " : "");
    sh += dims == 1 ? ",)" : ")";
    nc_attr_text(a, file_, "spatial_shape", sh);
    nc_attr_f32(a, file_, "Lx", (float)Lx);
    nc_attr_f32(a, file_, "Ly", (float)Ly);
    if (dims == 3) nc_attr_f32(a, file_, "Lz", (float)Lz);
    nc_attr_i32(a, file_, "is_complex", complex_ ? 1 : 0);
    nc_attr_i32(a, file_, "dims", dims);
    a.fflush(file_, 0);
  } catch (...) {
    close_locked();
    throw;
  }
}

void NcClass::put(int traj, const float *u, const float *v, const float *u0, const float *v0,
                  const float *m, const float *c) {
  std::lock_guard<std::mutex> lk(g_mu);
  NLS_REQUIRE(file_ >= 0, "NetCDF class: closed");
  NLS_REQUIRE(traj >= 0 && traj < ntraj_, "NetCDF class: trajectory index out of range");
  Api &a = need();
  const hsize_t temp = (hsize_t)ns_ * spat_;
  nc_put(a, vt_[0], (hsize_t)traj * temp, temp, u);
  nc_put(a, vt_[1], (hsize_t)traj * temp, temp, v);
  nc_put(a, vs_[0], (hsize_t)traj * spat_, spat_, u0);
  nc_put(a, vs_[1], (hsize_t)traj * spat_, spat_, v0);
  nc_put(a, vm_, (hsize_t)traj * spat_, spat_, m);
  nc_put(a, vc_, (hsize_t)traj * spat_, spat_, c);
}

void NcClass::close_locked() {
  if (file_ < 0) return;
  Api &a = need();
  for (int64_t *d : {&vt_[0], &vt_[1], &vs_[0], &vs_[1], &vm_, &vc_, &dt_, &ds_})
    if (*d >= 0) {
      a.dclose(*d);
      *d = -1;
    }
  a.fclose(file_);
  file_ = -1;
}

void NcClass::close() {
  std::lock_guard<std::mutex> lk(g_mu);
  close_locked();
}

NcClass::~NcClass() {
  try {
    close();
  } catch (...) {
  }
}

// ---------------------------------------------------------------- canonical class files
namespace {
constexpr int kFillInt = -2147483647; // NC_FILL_INT

void nc_attr_f64(Api &a, hid_t obj, const char *name, double v) {
  const hsize_t one = 1;
  Id sp(chk(a.screate_simple(1, &one, nullptr), "H5Screate_simple"), a.sclose);
  Id at(chk(a.acreate(obj, name, *a.native_double, sp.id, P_DEFAULT, P_DEFAULT),
            std::string("attribute ") + name),
        a.aclose);
  if (a.awrite(at.id, *a.native_double, &v) < 0) throw Error(std::string("HDF5: writing ") + name);
}

void nc_attr_i32v(Api &a, hid_t obj, const char *name, const std::vector<int> &v) {
  const hsize_t len = v.size();
  Id sp(chk(a.screate_simple(1, &len, nullptr), "H5Screate_simple"), a.sclose);
  Id at(chk(a.acreate(obj, name, *a.native_int, sp.id, P_DEFAULT, P_DEFAULT),
            std::string("attribute ") + name),
        a.aclose);
  if (a.awrite(at.id, *a.native_int, v.data()) < 0) throw Error(std::string("HDF5: writing ") + name);
}

// A coordinate variable: the dimension's own dataset holds its values, is a dimension scale
// named after itself and carries the dimension id (NetCDF-4's layout of a variable that
// shares its dimension's name).
hid_t nc_coordinate(Api &a, hid_t f, const char *name, bool is_int, const void *values,
                    hsize_t len, int dimid) {
  Id sp(chk(a.screate_simple(1, &len, nullptr), "H5Screate_simple"), a.sclose);
  Id pl(chk(a.pcreate(*a.dataset_create), "H5Pcreate"), a.pclose);
  const hid_t ty = is_int ? *a.native_int : *a.native_float;
  if (is_int) a.pset_fill_value(pl.id, ty, &kFillInt);
  else a.pset_fill_value(pl.id, ty, &kFillFloat);
  const hid_t d = chk(a.dcreate(f, name, ty, sp.id, P_DEFAULT, pl.id, P_DEFAULT),
                      std::string("coordinate ") + name);
  if (a.dwrite(d, ty, sp.id, sp.id, P_DEFAULT, values) < 0)
    throw Error(std::string("HDF5: writing ") + name);
  if (a.ds_set_scale(d, name) < 0) throw Error(std::string("HDF5: dimension scale ") + name);
  nc_attr_i32(a, d, "_Netcdf4Dimid", dimid);
  return d;
}

// A float variable over the dimensions `dims` (their datasets, in order) with the extents
// `ext`, chunked as `chunk` when given, contiguous otherwise.
hid_t nc_variable_nd(Api &a, hid_t f, const char *name, const std::vector<hid_t> &dims,
                     const std::vector<int> &dimids, const std::vector<hsize_t> &ext,
                     const std::vector<hsize_t> &chunk) {
  Id sp(chk(a.screate_simple((int)ext.size(), ext.data(), nullptr), "H5Screate_simple"),
        a.sclose);
  Id pl(chk(a.pcreate(*a.dataset_create), "H5Pcreate"), a.pclose);
  if (!chunk.empty()) a.pset_chunk(pl.id, (int)chunk.size(), chunk.data());
  a.pset_fill_value(pl.id, *a.native_float, &kFillFloat);
  const hid_t d = chk(a.dcreate(f, name, *a.native_float, sp.id, P_DEFAULT, pl.id, P_DEFAULT),
                      std::string("variable ") + name);
  for (size_t k = 0; k < dims.size(); ++k)
    if (a.ds_attach_scale(d, dims[k], (unsigned)k) < 0)
      throw Error(std::string("HDF5: attaching ") + name);
  nc_attr_i32v(a, d, "_Netcdf4Coordinates", dimids);
  return d;
}

// rows [row, row + 1) of a variable whose first extent is the trajectory: count floats
void nc_put_row(Api &a, hid_t d, int rank, hsize_t row, const std::vector<hsize_t> &ext,
                const float *x) {
  if (!x) return;
  std::vector<hsize_t> off(rank, 0), cnt(ext.begin(), ext.end());
  off[0] = row;
  cnt[0] = 1;
  Id fs(chk(a.dget_space(d), "H5Dget_space"), a.sclose);
  a.sselect_hyperslab(fs.id, S_SELECT_SET, off.data(), nullptr, cnt.data(), nullptr);
  Id ms(chk(a.screate_simple(rank, cnt.data(), nullptr), "H5Screate_simple"), a.sclose);
  if (a.dwrite(d, *a.native_float, ms.id, fs.id, P_DEFAULT, x) < 0)
    throw Error("HDF5: writing a trajectory");
}
} // namespace

NcStandard::NcStandard(const std::string &path, int dims, const std::vector<int> &shape, int ns,
                       int ntraj, bool is_complex, const Attrs &at)
    : dims_(dims), ns_(ns), ntraj_(ntraj), shape_(shape) {
  NLS_REQUIRE(dims == 2 || dims == 3, "NetCDF standard file: dims must be 2 or 3");
  NLS_REQUIRE((int)shape.size() == dims, "NetCDF standard file: shape must have dims entries");
  NLS_REQUIRE(ns > 0 && ntraj > 0, "NetCDF standard file: need snapshots and trajectories");
  spat_ = 1;
  for (int x : shape) {
    NLS_REQUIRE(x > 0, "NetCDF standard file: bad shape");
    spat_ *= (hsize_t)x;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  Api &a = need();
  NLS_REQUIRE(a.ds_set_scale && a.ds_attach_scale,
              "NetCDF files need libhdf5_hl (dimension scales); set NLS_HDF5_HL_LIB");
  file_ = chk(a.fcreate(path.c_str(), ACC_TRUNC, P_DEFAULT, P_DEFAULT), "create " + path);
  try {
    // coordinates (canonicalize_data.py:218-223)
    std::vector<int> member(ntraj);
    for (int t = 0; t < ntraj; ++t) member[t] = t;
    auto linspace = [](double lo, double hi, int n) {
      std::vector<float> v(n);
      for (int k = 0; k < n; ++k)
        v[k] = (float)(n == 1 ? lo : lo + (hi - lo) * (double)k / (double)(n - 1));
      return v;
    };
    const std::vector<float> tv = linspace(0.0, at.T, ns);
    const double Ls[3] = {at.Lx, at.Ly, at.Lz};
    const char *axes[3] = {"x", "y", "z"};
    dim_[0] = nc_coordinate(a, file_, "member", true, member.data(), (hsize_t)ntraj, 0);
    dim_[1] = nc_coordinate(a, file_, "time", false, tv.data(), (hsize_t)ns, 1);
    for (int k = 0; k < dims; ++k) {
      const std::vector<float> xv = linspace(-Ls[k], Ls[k], shape[k]);
      dim_[2 + k] = nc_coordinate(a, file_, axes[k], false, xv.data(), (hsize_t)shape[k], 2 + k);
    }
    // fields (member, time, x, y[, z]), one trajectory per chunk; m, c (member, x, y[, z])
    std::vector<hid_t> fd = {dim_[0], dim_[1]}, pd = {dim_[0]};
    std::vector<int> fid = {0, 1}, pid = {0};
    std::vector<hsize_t> fext = {(hsize_t)ntraj, (hsize_t)ns}, pext = {(hsize_t)ntraj};
    for (int k = 0; k < dims; ++k) {
      fd.push_back(dim_[2 + k]);
      pd.push_back(dim_[2 + k]);
      fid.push_back(2 + k);
      pid.push_back(2 + k);
      fext.push_back((hsize_t)shape[k]);
      pext.push_back((hsize_t)shape[k]);
    }
    std::vector<hsize_t> fch(fext);
    fch[0] = 1;
    const char *names[2] = {is_complex ? "u_real" : "u", is_complex ? "u_imag" : "v"};
    for (int k = 0; k < 2; ++k) vu_[k] = nc_variable_nd(a, file_, names[k], fd, fid, fext, fch);
    vm_ = nc_variable_nd(a, file_, "m", pd, pid, pext, {});
    vc_ = nc_variable_nd(a, file_, "c", pd, pid, pext, {});
    // global attributes (canonicalize_data.py:264-279)
    nc_attr_text(a, file_, "pde_type", at.pde_type);
    nc_attr_text(a, file_, "boundary_condition", at.boundary_condition);
    nc_attr_i32(a, file_, "num_trajectories", ntraj);
    nc_attr_i32(a, file_, "num_snapshots", ns);
    nc_attr_i32(a, file_, "nt", at.nt);
    std::string sh = "(";
    for (int k = 0; k < dims; ++k) sh += std::to_string(shape[k]) + (k + 1 < dims ? ", " : "");
    sh += ")";
    nc_attr_text(a, file_, "spatial_shape", sh);
    nc_attr_i32(a, file_, "dims", dims);
    nc_attr_i32(a, file_, "is_complex", is_complex ? 1 : 0);
    nc_attr_f64(a, file_, "Lx", at.Lx);
    nc_attr_f64(a, file_, "Ly", at.Ly);
    if (dims == 3 && at.Lz != 0.0) nc_attr_f64(a, file_, "Lz", at.Lz);
    nc_attr_f64(a, file_, "T", at.T);
    if (!at.m_classes.empty()) nc_attr_text(a, file_, "m_classes", at.m_classes);
    if (!at.c_classes.empty()) nc_attr_text(a, file_, "c_classes", at.c_classes);
    a.fflush(file_, 0);
  } catch (...) {
    close_locked();
    throw;
  }
}

void NcStandard::put(int traj, const float *u, const float *v, const float *m, const float *c) {
  std::lock_guard<std::mutex> lk(g_mu);
  NLS_REQUIRE(file_ >= 0, "NetCDF standard file: closed");
  NLS_REQUIRE(traj >= 0 && traj < ntraj_, "NetCDF standard file: trajectory index out of range");
  Api &a = need();
  std::vector<hsize_t> fext = {(hsize_t)ntraj_, (hsize_t)ns_}, pext = {(hsize_t)ntraj_};
  for (int x : shape_) {
    fext.push_back((hsize_t)x);
    pext.push_back((hsize_t)x);
  }
  nc_put_row(a, vu_[0], dims_ + 2, (hsize_t)traj, fext, u);
  nc_put_row(a, vu_[1], dims_ + 2, (hsize_t)traj, fext, v);
  nc_put_row(a, vm_, dims_ + 1, (hsize_t)traj, pext, m);
  nc_put_row(a, vc_, dims_ + 1, (hsize_t)traj, pext, c);
}

void NcStandard::close_locked() {
  if (file_ < 0) return;
  Api &a = need();
  for (int64_t *d : {&vu_[0], &vu_[1], &vm_, &vc_, &dim_[0], &dim_[1], &dim_[2], &dim_[3], &dim_[4]})
    if (*d >= 0) {
      a.dclose(*d);
      *d = -1;
    }
  a.fclose(file_);
  file_ = -1;
}

void NcStandard::close() {
  std::lock_guard<std::mutex> lk(g_mu);
  close_locked();
}

NcStandard::~NcStandard() {
  try {
    close();
  } catch (...) {
  }
}

} // namespace h5
} // namespace nls

// ---------------------------------------------------------------- asynchronous writer
struct nls_h5_writer {
  size_t max_bytes = 0, queued = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_space, cv_idle;
  std::deque<std::unique_ptr<nls::h5::Run>> q;
  bool stop = false, busy = false;
  int64_t written = 0;
  std::string first_error;
  std::thread worker;

  void loop() {
    for (;;) {
      std::unique_ptr<nls::h5::Run> r;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !q.empty(); });
        if (q.empty()) return;
        r = std::move(q.front());
        q.pop_front();
        busy = true;
      }
      std::string err;
      try {
        nls::h5::write_run(*r);
      } catch (const std::exception &e) {
        err = e.what();
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        queued -= r->bytes();
        busy = false;
        if (err.empty()) ++written;
        else if (first_error.empty()) first_error = r->path + ": " + err;
      }
      cv_space.notify_all();
      cv_idle.notify_all();
    }
  }

  ~nls_h5_writer() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    if (worker.joinable()) worker.join();
  }
};

extern "C" {

int nls_h5_available(char *version, int len) {
  return nls::guard([&] {
    std::string v;
    NLS_REQUIRE(nls::h5::available(&v), "libhdf5 could not be loaded (set NLS_HDF5_LIB)");
    if (version && len > 0) {
      std::strncpy(version, v.c_str(), (size_t)len - 1);
      version[len - 1] = '\0';
    }
  });
}

int nls_h5_write_run(const nls_h5_run *r) {
  return nls::guard([&] {
    NLS_REQUIRE(r, "nls_h5_write_run: null run");
    nls::h5::write_run(nls::h5::from_abi(*r));
  });
}

int nls_h5_writer_create(size_t max_queued_bytes, nls_h5_writer **out) {
  return nls::guard([&] {
    NLS_REQUIRE(out, "nls_h5_writer_create: null out");
    NLS_REQUIRE(nls::h5::available(), "libhdf5 could not be loaded (set NLS_HDF5_LIB)");
    auto w = std::make_unique<nls_h5_writer>();
    w->max_bytes = max_queued_bytes > 0 ? max_queued_bytes : ((size_t)1 << 30);
    w->worker = std::thread([raw = w.get()] { raw->loop(); });
    *out = w.release();
  });
}

int nls_h5_writer_submit(nls_h5_writer *w, const nls_h5_run *r) {
  return nls::guard([&] {
    NLS_REQUIRE(w && r, "nls_h5_writer_submit: null argument");
    auto run = std::make_unique<nls::h5::Run>(nls::h5::from_abi(*r));
    const size_t b = run->bytes();
    std::unique_lock<std::mutex> lk(w->mu);
    // wait for room unless the queue is empty (a single large file always fits)
    w->cv_space.wait(lk, [&] { return w->queued == 0 || w->queued + b <= w->max_bytes; });
    w->queued += b;
    w->q.push_back(std::move(run));
    lk.unlock();
    w->cv_work.notify_one();
  });
}

int nls_h5_writer_drain(nls_h5_writer *w, int64_t *files_written) {
  return nls::guard([&] {
    NLS_REQUIRE(w, "nls_h5_writer_drain: null writer");
    std::unique_lock<std::mutex> lk(w->mu);
    w->cv_idle.wait(lk, [&] { return w->q.empty() && !w->busy; });
    if (files_written) *files_written = w->written;
    NLS_REQUIRE(w->first_error.empty(), "HDF5 writer: " + w->first_error);
  });
}

int nls_h5_writer_destroy(nls_h5_writer *w) { return nls::guard([&] { delete w; }); }

int nls_h5_read_dataset(const char *path, const char *name, double *buf, int64_t cap,
                        int64_t *dims, int *ndim, int *is_complex) {
  using namespace nls::h5;
  return nls::guard([&] {
    NLS_REQUIRE(path && name, "nls_h5_read_dataset: null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    Api &a = need();
    Id f(chk(a.fopen(path, ACC_RDONLY, P_DEFAULT), std::string("open ") + path), a.fclose);
    Id d(chk(a.dopen(f.id, name, P_DEFAULT), std::string("dataset ") + name), a.dclose);
    Id sp(chk(a.dget_space(d.id), "H5Dget_space"), a.sclose);
    const int nd = a.sget_ndims(sp.id);
    NLS_REQUIRE(nd >= 0 && nd <= 4, "nls_h5_read_dataset: rank above 4");
    hsize_t dd[4] = {0, 0, 0, 0};
    a.sget_dims(sp.id, dd, nullptr);
    size_t count = 1;
    for (int k = 0; k < nd; ++k) count *= dd[k];
    if (ndim) *ndim = nd;
    if (dims)
      for (int k = 0; k < 4; ++k) dims[k] = k < nd ? (int64_t)dd[k] : 0;
    Id ft(chk(a.dget_type(d.id), "H5Dget_type"), a.tclose);
    const bool cplx = a.tget_class(ft.id) == T_COMPOUND;
    if (is_complex) *is_complex = cplx ? 1 : 0;
    if (!buf) return;
    const size_t need_d = count * (cplx ? 2 : 1);
    NLS_REQUIRE((size_t)cap >= need_d, "nls_h5_read_dataset: buffer too small");
    Id mt(cplx ? complex_type(a) : a.tcopy(*a.native_double), a.tclose);
    if (a.dread(d.id, mt.id, S_ALL, S_ALL, P_DEFAULT, buf) < 0)
      throw nls::Error(std::string("HDF5: reading ") + name);
  });
}

int nls_h5_read_range_f32(const char *path, const char *name, int64_t offset, int64_t count,
                          float *buf) {
  using namespace nls::h5;
  return nls::guard([&] {
    NLS_REQUIRE(path && name && buf, "nls_h5_read_range_f32: null argument");
    NLS_REQUIRE(offset >= 0 && count >= 0, "nls_h5_read_range_f32: negative range");
    if (count == 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    Api &a = need();
    Id f(chk(a.fopen(path, ACC_RDONLY, P_DEFAULT), std::string("open ") + path), a.fclose);
    Id d(chk(a.dopen(f.id, name, P_DEFAULT), std::string("dataset ") + name), a.dclose);
    Id fs(chk(a.dget_space(d.id), "H5Dget_space"), a.sclose);
    NLS_REQUIRE(a.sget_ndims(fs.id) == 1, "nls_h5_read_range_f32: the dataset must be 1-D");
    hsize_t len = 0;
    a.sget_dims(fs.id, &len, nullptr);
    NLS_REQUIRE((hsize_t)offset + (hsize_t)count <= len,
                "nls_h5_read_range_f32: range beyond the dataset");
    Id ft(chk(a.dget_type(d.id), "H5Dget_type"), a.tclose);
    NLS_REQUIRE(a.tget_class(ft.id) != T_COMPOUND, "nls_h5_read_range_f32: compound dataset");
    const hsize_t off = (hsize_t)offset, cnt = (hsize_t)count;
    a.sselect_hyperslab(fs.id, S_SELECT_SET, &off, nullptr, &cnt, nullptr);
    Id ms(chk(a.screate_simple(1, &cnt, nullptr), "H5Screate_simple"), a.sclose);
    if (a.dread(d.id, *a.native_float, ms.id, fs.id, P_DEFAULT, buf) < 0)
      throw nls::Error(std::string("HDF5: reading ") + name);
  });
}

struct nls_nc_class {
  std::unique_ptr<nls::h5::NcClass> w;
};

int nls_nc_class_create(const char *path, int dims, const int *shape, int num_snapshots,
                        int num_trajectories, int is_complex, double Lx, double Ly, double Lz,
                        nls_nc_class **out) {
  return nls::guard([&] {
    NLS_REQUIRE(path && shape && out, "nls_nc_class_create: null argument");
    NLS_REQUIRE(dims == 2 || dims == 3, "nls_nc_class_create: dims must be 2 or 3");
    auto h = std::make_unique<nls_nc_class>();
    h->w = std::make_unique<nls::h5::NcClass>(path, dims, std::vector<int>(shape, shape + dims),
                                              num_snapshots, num_trajectories, is_complex != 0,
                                              Lx, Ly, Lz);
    *out = h.release();
  });
}

int nls_nc_class_put(nls_nc_class *h, int traj, const float *u, const float *v, const float *u0,
                     const float *v0, const float *m, const float *c) {
  return nls::guard([&] {
    NLS_REQUIRE(h && h->w, "nls_nc_class_put: null handle");
    h->w->put(traj, u, v, u0, v0, m, c);
  });
}

int nls_nc_class_close(nls_nc_class *h) {
  return nls::guard([&] {
    if (!h) return;
    std::unique_ptr<nls_nc_class> own(h);
    if (own->w) own->w->close();
  });
}

struct nls_nc_std {
  std::unique_ptr<nls::h5::NcStandard> w;
};

int nls_nc_std_create(const char *path, int dims, const int *shape, int num_snapshots,
                      int num_trajectories, int is_complex, double Lx, double Ly, double Lz,
                      double T, int nt, const char *pde_type, const char *boundary_condition,
                      const char *m_classes, const char *c_classes, nls_nc_std **out) {
  return nls::guard([&] {
    NLS_REQUIRE(path && shape && out && pde_type && boundary_condition,
                "nls_nc_std_create: null argument");
    NLS_REQUIRE(dims == 2 || dims == 3, "nls_nc_std_create: dims must be 2 or 3");
    nls::h5::NcStandard::Attrs at;
    at.pde_type = pde_type;
    at.boundary_condition = boundary_condition;
    at.m_classes = m_classes ? m_classes : "";
    at.c_classes = c_classes ? c_classes : "";
    at.Lx = Lx;
    at.Ly = Ly;
    at.Lz = Lz;
    at.T = T;
    at.nt = nt;
    auto h = std::make_unique<nls_nc_std>();
    h->w = std::make_unique<nls::h5::NcStandard>(path, dims, std::vector<int>(shape, shape + dims),
                                                 num_snapshots, num_trajectories, is_complex != 0,
                                                 at);
    *out = h.release();
  });
}

int nls_nc_std_put(nls_nc_std *h, int traj, const float *u, const float *v, const float *m,
                   const float *c) {
  return nls::guard([&] {
    NLS_REQUIRE(h && h->w, "nls_nc_std_put: null handle");
    h->w->put(traj, u, v, m, c);
  });
}

int nls_nc_std_close(nls_nc_std *h) {
  return nls::guard([&] {
    if (!h) return;
    std::unique_ptr<nls_nc_std> own(h);
    if (own->w) own->w->close();
  });
}

int nls_h5_read_attr(const char *path, const char *obj, const char *name, double *num, char *str,
                     int len, int *is_string) {
  using namespace nls::h5;
  return nls::guard([&] {
    NLS_REQUIRE(path && obj && name, "nls_h5_read_attr: null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    Api &a = need();
    Id f(chk(a.fopen(path, ACC_RDONLY, P_DEFAULT), std::string("open ") + path), a.fclose);
    Id at(chk(a.aopen_by_name(f.id, obj, name, P_DEFAULT, P_DEFAULT),
              std::string("attribute ") + obj + "/" + name),
          a.aclose);
    Id ty(chk(a.aget_type(at.id), "H5Aget_type"), a.tclose);
    const bool s = a.tis_variable_str(ty.id) > 0;
    const bool fixed = !s && a.tget_class(ty.id) == T_STRING; // NetCDF-4 text attributes
    if (is_string) *is_string = (s || fixed) ? 1 : 0;
    if (fixed) {
      const size_t sz = a.tget_size(ty.id);
      std::string buf(sz + 1, '\0');
      Id mt(chk(a.tcopy(*a.c_s1), "H5Tcopy"), a.tclose);
      a.tset_size(mt.id, sz);
      if (a.aread(at.id, mt.id, &buf[0]) < 0) throw nls::Error("HDF5: reading attribute");
      if (str && len > 0) {
        std::strncpy(str, buf.c_str(), (size_t)len - 1);
        str[len - 1] = '\0';
      }
    } else if (s) {
      Id mt(chk(a.tcopy(*a.c_s1), "H5Tcopy"), a.tclose);
      a.tset_size(mt.id, T_VARIABLE);
      a.tset_cset(mt.id, T_CSET_UTF8);
      char *p = nullptr;
      if (a.aread(at.id, mt.id, &p) < 0) throw nls::Error("HDF5: reading attribute");
      if (str && len > 0) {
        std::strncpy(str, p ? p : "", (size_t)len - 1);
        str[len - 1] = '\0';
      }
      if (p) a.free_memory(p);
    } else {
      double v = 0;
      if (a.aread(at.id, *a.native_double, &v) < 0) throw nls::Error("HDF5: reading attribute");
      if (num) *num = v;
    }
  });
}

} // extern "C"
