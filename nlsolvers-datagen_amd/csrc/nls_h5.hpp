// HDF5 per-trajectory files in the launchers' layout (scripts/real_launcher_2d.py:201-245,
// scripts/complex_launcher_2d.py:193-236), with libhdf5 (1.10 C API) loaded at run time.
//
// The library keeps working where libhdf5 is absent: only these calls fail.  Every
// HDF5 call goes through one process-wide mutex, so files may be written from the
// writer threads of several objects at once.
#pragma once

#include "nls.h"

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace nls {
namespace h5 {

// True when libhdf5 could be loaded; `version` receives "maj.min.rel".
bool available(std::string *version = nullptr);

// One trajectory file, held in host memory (owned copies of every array).
struct Run {
  std::string path;
  bool complex_field = false; // NLSE: u0 and u complex128, no v0 / v
  int dim = 2, nx = 0, ny = 0, nz = 1;
  double Lx = 0, Ly = 0, Lz = 0, T = 0;
  int nt = 0, num_snapshots = 0;
  std::string problem_type, run_id, phenomenon, m_type, c_type, timestamp;
  int run_index = 0;
  double elapsed_time = 0;
  std::vector<std::pair<std::string, std::string>> params; // metadata phenomenon_<key>
  std::vector<double> u0, v0, m, c;                         // grid arrays (u0: 2 doubles per cell if complex)
  int ns_out = 0, ox = 0, oy = 0, oz = 1;                   // stored trajectory shape
  std::vector<double> u, v;
  size_t bytes() const {
    return 8 * (u0.size() + v0.size() + m.size() + c.size() + u.size() + v.size());
  }
};

Run from_abi(const nls_h5_run &r); // copies the caller's arrays
void write_run(const Run &r);      // throws nls::Error on failure

// A file whose trajectory grows one snapshot at a time (serving loop): the
// metadata, grid arrays and an empty extendible u (and v) are written at open,
// snapshots are appended, elapsed_time is set at close.
class Stream {
 public:
  explicit Stream(const Run &header); // header.u / header.v ignored; ns_out = 0
  ~Stream();
  void append(const double *u, const double *v); // one snapshot of ox*oy*oz values each
  void close(double elapsed_time);
  int count() const { return count_; }

 private:
  int64_t file_ = -1, du_ = -1, dv_ = -1;
  int count_ = 0;
  bool complex_ = false, has_v_ = false;
  std::vector<uint64_t> dims_;
};

// A per-class NetCDF-4 file (aggregate_single_core.py:167-206): dimensions total_temporal
// (trajectories x snapshots x cells) and total_spatial (trajectories x cells), float
// variables u, v (re_u, im_u for complex classes) over total_temporal and u0, v0 (re_u0,
// im_u0), m, c over total_spatial, trajectory t at offset t x its per-trajectory size,
// and the global attributes num_snapshots, num_trajectories, spatial_shape, Lx, Ly[, Lz],
// is_complex, dims.  Written through libhdf5 with the NetCDF-4 conventions (dimension
// scales from libhdf5_hl).
class NcClass {
 public:
  NcClass(const std::string &path, int dims, const std::vector<int> &shape, int ns, int ntraj,
          bool is_complex, double Lx, double Ly, double Lz);
  ~NcClass();
  // one trajectory's arrays (any may be null: that part stays at the fill value)
  void put(int traj, const float *u, const float *v, const float *u0, const float *v0,
           const float *m, const float *c);
  void close();

 private:
  void close_locked();
  int64_t file_ = -1, dt_ = -1, ds_ = -1, vt_[2] = {-1, -1}, vs_[2] = {-1, -1}, vm_ = -1, vc_ = -1;
  uint64_t spat_ = 0;
  int ns_ = 0, ntraj_ = 0;
  bool complex_ = false;
};

// The canonical per-class NetCDF-4 file (canonicalize_data.py:174-262): dimensions member,
// time, x, y[, z] with coordinate variables (member int 0..; time linspace(0, T, ns); x, y,
// z linspace(-L, L, n)), float variables u, v (complex classes: u_real, u_imag) over
// (member, time, x, y[, z]) in chunks of one trajectory, m and c over (member, x, y[, z]),
// and the global attributes pde_type, boundary_condition, num_trajectories,
// num_snapshots, nt, spatial_shape, dims, is_complex, Lx, Ly[, Lz], T[, m_classes,
// c_classes].  NetCDF-4 conventions as NcClass.
class NcStandard {
 public:
  struct Attrs {
    std::string pde_type, boundary_condition, m_classes, c_classes;
    double Lx = 0, Ly = 0, Lz = 0, T = 0;
    int nt = 0;
  };
  NcStandard(const std::string &path, int dims, const std::vector<int> &shape, int ns,
             int ntraj, bool is_complex, const Attrs &at);
  ~NcStandard();
  // one trajectory: u, v of ns x cells (complex: real and imaginary parts), m, c of cells
  // (any may be null: that part stays at the fill value)
  void put(int traj, const float *u, const float *v, const float *m, const float *c);
  void close();

 private:
  void close_locked();
  int64_t file_ = -1, dim_[5] = {-1, -1, -1, -1, -1}, vu_[2] = {-1, -1}, vm_ = -1, vc_ = -1;
  uint64_t spat_ = 0;
  int dims_ = 0, ns_ = 0, ntraj_ = 0;
  std::vector<int> shape_;
};

} // namespace h5
} // namespace nls
