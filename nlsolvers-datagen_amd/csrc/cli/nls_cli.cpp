// Command-line drivers with the reference executables' names and argv, built on
// the libnls C ABI.  One binary; the mode comes from the name it is invoked as
// (install names: kg_gautschi_2d_dev, kg_sv_2d_dev, kg_gautschi_3d_dev,
// kg_sv_3d_dev, nlse_2d_dev, nlse_sewi_2d_dev, nlse_3d_dev, nlse_sewi_3d_dev,
// sp4_sv_2d_dev, sp4_sv_3d_dev -- device/drivers/CMakeLists.txt:15-31).
//
// Each mode reads the .npy inputs, runs the driver loop of the reference
// (nt-1 steps, ghost copy after each, snapshot every nt/num_snapshots) and
// writes the .npy outputs:
//   KG/SP4 2D: nx ny Lx Ly u0.npy v0.npy traj.npy vel.npy T nt ns m.npy c.npy
//              (kg_driver_dev_2d.cpp:19-49)
//   KG/SP4 3D: nx ny nz Lx Ly Lz u0 v0 traj vel T nt ns m c  (kg_driver_dev_3d.cpp:19-49)
//   NLSE 2D:   nx ny Lx Ly u0.npy traj.npy T nt ns m.npy c.npy (nlse_cubic_driver_2d.cpp:19-41)
//   NLSE 3D:   nx ny nz Lx Ly Lz u0 traj T nt ns m c           (nlse_cubic_driver_3d.cpp:19-44)
// Krylov dimensions as in the reference drivers: KG 10; NLSE 2D 20, "sewi" 2D 10
// (which runs Strang, nlse_cubic_sewi_driver_2d.cpp:104-108); NLSE 3D 25; SEWI 3D 15.
// Wrong-shape u0 returns 1; wrong-shape m or c throws (kg_driver_dev_2d.cpp:61-99).
// The stochastic phi^4 noise is Philox keyed by NLS_SP4_SEED (default 0) instead
// of the wall-clock seed of stochastic_phi4.cuh:25.
#include "nls.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// ---- .npy codec (NumPy format version 1.0/2.0, little-endian, C order)
struct Npy {
  std::string descr;
  std::vector<uint64_t> shape;
  std::vector<char> data;
};

Npy npy_read(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  char magic[8];
  f.read(magic, 8);
  if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error("not a .npy file: " + path);
  const int major = (unsigned char)magic[6];
  uint32_t hlen = 0;
  if (major == 1) {
    uint16_t h16;
    f.read(reinterpret_cast<char *>(&h16), 2);
    hlen = h16;
  } else {
    f.read(reinterpret_cast<char *>(&hlen), 4);
  }
  std::string header(hlen, '\0');
  f.read(&header[0], hlen);
  Npy out;
  auto field = [&](const char *key) {
    const size_t k = header.find(key);
    if (k == std::string::npos) throw std::runtime_error("bad .npy header in " + path);
    return header.substr(k + std::strlen(key));
  };
  std::string d = field("'descr':");
  const size_t q1 = d.find('\''), q2 = d.find('\'', q1 + 1);
  out.descr = d.substr(q1 + 1, q2 - q1 - 1);
  if (field("'fortran_order':").find("True") < field("'fortran_order':").find(','))
    throw std::runtime_error("Fortran-ordered .npy not supported: " + path);
  std::string sh = field("'shape':");
  sh = sh.substr(sh.find('(') + 1, sh.find(')') - sh.find('(') - 1);
  size_t pos = 0;
  while (pos < sh.size()) {
    const size_t c = sh.find(',', pos);
    const std::string tok = sh.substr(pos, c == std::string::npos ? std::string::npos : c - pos);
    if (tok.find_first_of("0123456789") != std::string::npos) out.shape.push_back(std::stoull(tok));
    if (c == std::string::npos) break;
    pos = c + 1;
  }
  uint64_t count = 1;
  for (auto s : out.shape) count *= s;
  const size_t item = out.descr == "<c16" ? 16 : out.descr == "<f8" ? 8 : 0;
  if (!item) throw std::runtime_error("unsupported dtype " + out.descr + " in " + path);
  out.data.resize(count * item);
  f.read(out.data.data(), (std::streamsize)out.data.size());
  if (!f) throw std::runtime_error("truncated .npy: " + path);
  return out;
}

void npy_write(const std::string &path, const std::string &descr, const std::vector<uint64_t> &shape,
               const void *data, size_t bytes) {
  std::string sh = "(";
  for (size_t i = 0; i < shape.size(); ++i) sh += std::to_string(shape[i]) + (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
  sh += ")";
  std::string header = "{'descr': '" + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
  const size_t total = 10 + header.size() + 1;
  header += std::string((64 - total % 64) % 64, ' ') + "\n";
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
  f.write(magic, 8);
  const uint16_t hl = (uint16_t)header.size();
  f.write(reinterpret_cast<const char *>(&hl), 2);
  f.write(header.data(), (std::streamsize)header.size());
  f.write(static_cast<const char *>(data), (std::streamsize)bytes);
}

void check(int rc) {
  if (rc != 0) throw std::runtime_error(nls_last_error());
}

std::vector<double> as_f64(const Npy &a) {
  if (a.descr != "<f8") throw std::runtime_error("expected float64 input");
  std::vector<double> v(a.data.size() / 8);
  std::memcpy(v.data(), a.data.data(), a.data.size());
  return v;
}

int run(const std::string &mode, int argc, char **argv) {
  const bool is3d = mode.find("_3d") != std::string::npos;
  const bool nlse = mode.rfind("nlse", 0) == 0;
  const int expect = nlse ? (is3d ? 14 : 12) : (is3d ? 16 : 14);
  if (argc != expect) {
    if (nlse)
      std::cerr << "Usage: " << argv[0]
                << (is3d ? " nx ny nz Lx Ly Lz input_u0.npy output_traj.npy T nt num_snapshots input_m.npy input_c.npy\n"
                         : " nx ny Lx Ly input_u0.npy output_traj.npy T nt num_snapshots input_m.npy input_c.npy\n");
    else
      std::cerr << "Usage: " << argv[0]
                << (is3d ? " nx ny nz Lx Ly Lz input_u0.npy input_v0.npy output_traj.npy output_vel.npy T nt num_snapshots input_m.npy input_c.npy\n"
                         : " nx ny Lx Ly input_u0.npy input_v0.npy output_traj.npy output_vel.npy T nt num_snapshots input_m.npy input_c.npy\n");
    return 1;
  }
  int k = 1;
  const uint32_t nx = std::stoul(argv[k++]);
  const uint32_t ny = std::stoul(argv[k++]);
  const uint32_t nz = is3d ? std::stoul(argv[k++]) : 1;
  const double Lx = std::stod(argv[k++]);
  k++; // Ly (square domains: dx from Lx, as the reference drivers do)
  if (is3d) k++; // Lz
  const std::string u0f = argv[k++];
  const std::string v0f = nlse ? "" : argv[k++];
  const std::string trajf = argv[k++];
  const std::string velf = nlse ? "" : argv[k++];
  const double T = std::stod(argv[k++]);
  const uint32_t nt = std::stoul(argv[k++]);
  const uint32_t ns = std::stoul(argv[k++]);
  const std::string mf = argv[k++], cf = argv[k++];
  if (nx != ny || (is3d && ny != nz)) throw std::runtime_error("grids must be square/cubic");
  const int n = (int)nx;
  const uint64_t N = is3d ? (uint64_t)n * n * n : (uint64_t)n * n;
  std::vector<uint64_t> fshape = is3d ? std::vector<uint64_t>{nz, ny, nx} : std::vector<uint64_t>{nx, ny};

  Npy u0 = npy_read(u0f);
  if (u0.shape != fshape) {
    std::cerr << "Error: Input array dimensions mismatch\n";
    return 1;
  }
  Npy m = npy_read(mf), c = npy_read(cf);
  if (m.shape != fshape) throw std::runtime_error("Faulty m (1)");
  if (c.shape != fshape) throw std::runtime_error("Faulty c (1)");
  const std::vector<double> mv = as_f64(m), cv = as_f64(c);
  std::vector<uint64_t> oshape = {ns};
  for (auto s : fshape) oshape.push_back(s);

  if (nlse) {
    if (u0.descr != "<c16") throw std::runtime_error("NLSE input must be complex128");
    nls_nlse_params p{};
    p.dim = is3d ? 3 : 2;
    p.n = n;
    p.L = Lx;
    p.T = T;
    p.nt = (int)nt;
    p.num_snapshots = (int)ns;
    const bool sewi = mode.find("sewi") != std::string::npos;
    p.krylov_dim = is3d ? (sewi ? 15 : 25) : (sewi ? 10 : 20);
    p.method = (sewi && is3d) ? NLS_NLSE_SEWI : NLS_NLSE_STRANG; // Q11: 2D "sewi" runs Strang
    p.q6 = 1;
    nls_nlse *h = nullptr;
    check(nls_nlse_create(&p, mv.data(), cv.data(), reinterpret_cast<const double *>(u0.data.data()), &h));
    check(nls_nlse_run(h));
    std::vector<double> out(2 * N * ns);
    check(nls_nlse_transfer_snapshots(h, out.data()));
    nls_nlse_destroy(h);
    npy_write(trajf, "<c16", oshape, out.data(), out.size() * 8);
    return 0;
  }
  Npy v0 = npy_read(v0f);
  if (v0.shape != fshape) {
    std::cerr << "Error: Input array dimensions mismatch\n";
    return 1;
  }
  const std::vector<double> u0v = as_f64(u0), v0v = as_f64(v0);
  std::vector<double> U(N * ns), V(N * ns);
  if (mode.rfind("sp4", 0) == 0) {
    nls_sp4_params p{};
    p.dim = is3d ? 3 : 2;
    p.n = n;
    p.L = Lx;
    p.T = T;
    p.nt = (int)nt;
    p.num_snapshots = (int)ns;
    p.mode = NLS_SP4_REFERENCE;
    p.q6 = 1;
    const char *seed = std::getenv("NLS_SP4_SEED");
    p.seed = seed ? std::strtoull(seed, nullptr, 10) : 0;
    nls_sp4 *h = nullptr;
    check(nls_sp4_create(&p, mv.data(), cv.data(), u0v.data(), v0v.data(), &h));
    check(nls_sp4_run(h));
    check(nls_sp4_transfer_snapshots(h, U.data(), V.data()));
    nls_sp4_destroy(h);
  } else {
    nls_kge_params p{};
    p.dim = is3d ? 3 : 2;
    p.n = n;
    p.L = Lx;
    p.T = T;
    p.nt = (int)nt;
    p.num_snapshots = (int)ns;
    p.krylov_dim = 10; // kg_driver_dev_2d.cpp:138
    p.method = mode.rfind("kg_sv", 0) == 0 ? NLS_KGE_SV : NLS_KGE_GAUTSCHI;
    p.mu2 = 0.0;
    p.q6 = 1;
    nls_kge *h = nullptr;
    check(nls_kge_create(&p, mv.data(), cv.data(), u0v.data(), v0v.data(), &h));
    check(nls_kge_run(h));
    check(nls_kge_transfer_snapshots(h, U.data(), V.data()));
    nls_kge_destroy(h);
  }
  npy_write(trajf, "<f8", oshape, U.data(), U.size() * 8);
  npy_write(velf, "<f8", oshape, V.data(), V.size() * 8);
  return 0;
}

} // namespace

int main(int argc, char **argv) {
  std::string mode = argv[0];
  const size_t sl = mode.find_last_of('/');
  if (sl != std::string::npos) mode = mode.substr(sl + 1);
  if (mode == "nls_driver") {
    if (argc < 2) {
      std::cerr << "usage: nls_driver <kg_gautschi_2d_dev|...> args...\n";
      return 1;
    }
    mode = argv[1];
    return run(mode, argc - 1, argv + 1);
  }
  return run(mode, argc, argv);
}
