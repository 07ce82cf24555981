// Seeded initial-condition samplers of the benchmark configurations, evaluated on the GPU.
//
// Field shapes (reference forms; the parameter laws are the configurations'):
//   gaussians 2D  u0 = sum_k A_k exp(-|x - x_k|^2 / w_k^2)            (util.hpp:94-125)
//   kink ring 2D  u0 = tanh((R - r) / w)                               (real_sampler.py:464-467)
//   packets 2D    psi0 = sum_k A_k exp(-|x - x_k|^2) exp(i(v_k.(x - x_k) + phi_k))
//                                                                      (nlse_sampler.py:43-100)
//   bubbles 3D    u0 = sum_k A_k exp(-|x - x_k|^2 / w_k^2)             (the Gaussian of util.hpp:94-125)
// on the grid linspace(-L, L, n) with "ij" indexing (real_sampler.py:17-25); v0 = 0.
// Centres and ring radii scale with the half-width L (centres U(-0.6 L, 0.6 L) in 2D and
// U(-0.5 L, 0.5 L) in 3D, R ~ U(0.2 L, 0.6 L)), so L = 10 (2D) and L = 8 (3D) give the
// configurations' stated laws.
//
// Parameters are drawn on the host by ParamStream: draw s of trajectory t is the first
// word (or the first words) of Philox4x32-10 at counter (s, t, tag, 0) under the 64-bit
// seed, u = (x + 0.5) 2^-32 (the same stream as nls/samplers.py, so a trajectory's data
// depend only on (seed, trajectory index)).  The field evaluation repeats numpy's
// operation order in FP64, so GPU and host fields differ only by the last bits of exp /
// tanh / sin / cos.
#pragma once

#include "nls_philox.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace nls {
namespace ic {

enum Kind { GAUSSIANS_2D = 0, KINK_RING_2D = 1, PACKETS_2D = 2, BUBBLES_3D = 3 };
constexpr int kMaxParams = 32;

struct ParamStream {
  uint64_t seed;
  uint32_t traj, tag;
  uint32_t slot = 0;
  u32x4 next() {
    return philox4x32_10(u32x4{slot++, traj, tag, 0u}, (uint32_t)seed, (uint32_t)(seed >> 32));
  }
  static double unit(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }
  double uniform(double lo, double hi) { return lo + (hi - lo) * unit(next().x); }
  int integer(int lo, int hi) {
    return lo + std::min((int)(uniform(0.0, 1.0) * (double)(hi - lo + 1)), hi - lo);
  }
  double normal() {
    const double u1 = uniform(0.0, 1.0), u2 = uniform(0.0, 1.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.141592653589793 * u2);
  }
};

inline int dim_of(int kind) { return kind == BUBBLES_3D ? 3 : 2; }

// Draws of one trajectory (domain [-L, L]^dim) in draw order; returns the count
// (<= kMaxParams), -1 for an unknown kind.  Layouts: gaussians [A, w, cx, cy] x 3; kink
// [R, w]; packets [K, (A, phi, vx, vy, cx, cy) x K]; bubbles [K, (A, w, cx, cy, cz) x K].
inline int draw(int kind, uint64_t seed, uint32_t traj, double L, double *p) {
  int c = 0;
  const double box2 = 6.0 * L / 10.0; // 2D centres U(-box2, box2)
  switch (kind) {
  case GAUSSIANS_2D: {
    ParamStream ps{seed, traj, 0x6A55u};
    for (int k = 0; k < 3; ++k) {
      p[c++] = ps.uniform(0.5, 2.0);
      p[c++] = ps.uniform(0.5, 2.0);
      p[c++] = ps.uniform(-box2, box2);
      p[c++] = ps.uniform(-box2, box2);
    }
    return c;
  }
  case KINK_RING_2D: {
    ParamStream ps{seed, traj, 0x4B12u};
    p[c++] = ps.uniform(2.0 * L / 10.0, 6.0 * L / 10.0);
    p[c++] = ps.uniform(0.5, 1.5);
    return c;
  }
  case PACKETS_2D: {
    ParamStream ps{seed, traj, 0x4E15u};
    const int K = ps.integer(2, 4);
    p[c++] = K;
    for (int k = 0; k < K; ++k) {
      p[c++] = ps.uniform(0.5, 1.5);
      p[c++] = ps.uniform(0.0, 2.0 * 3.141592653589793);
      p[c++] = ps.normal();
      p[c++] = ps.normal();
      p[c++] = ps.uniform(-box2, box2);
      p[c++] = ps.uniform(-box2, box2);
    }
    return c;
  }
  case BUBBLES_3D: {
    ParamStream ps{seed, traj, 0x3DB0u};
    const double b = 4.0 * L / 8.0;
    const int K = ps.integer(2, 5);
    p[c++] = K;
    for (int k = 0; k < K; ++k) {
      p[c++] = ps.uniform(0.5, 2.0);
      p[c++] = ps.uniform(0.8, 2.0);
      const u32x4 r = ps.next(); // uniform(-b, b, size=3): one counter, three words
      p[c++] = -b + (b - (-b)) * ParamStream::unit(r.x);
      p[c++] = -b + (b - (-b)) * ParamStream::unit(r.y);
      p[c++] = -b + (b - (-b)) * ParamStream::unit(r.z);
    }
    return c;
  }
  }
  return -1;
}

struct Params {
  int kind, n;
  double L;
  double p[kMaxParams];
};

// numpy.linspace(-L, L, n)[i]: i * step + start, the last point exactly stop
__device__ inline double grid_x(int i, int n, double L) {
  if (i == n - 1) return L;
  return (double)i * ((L - (-L)) / (double)(n - 1)) + (-L);
}

// Field value at cell g of an n^dim grid (complex for the packets: re, im).
__device__ inline void eval(const Params &P, int64_t g, double &re, double &im) {
#pragma clang fp contract(off)
  const int n = P.n;
  re = 0.0;
  im = 0.0;
  if (P.kind == BUBBLES_3D) {
    const int64_t n2 = (int64_t)n * n;
    const int i = (int)(g / n2), j = (int)((g / n) % n), k = (int)(g % n);
    const double X = grid_x(i, n, P.L), Y = grid_x(j, n, P.L), Z = grid_x(k, n, P.L);
    const int K = (int)P.p[0];
    for (int b = 0; b < K; ++b) {
      const double *q = P.p + 1 + 5 * b;
      const double dx = X - q[2], dy = Y - q[3], dz = Z - q[4];
      re += q[0] * exp(-(dx * dx + dy * dy + dz * dz) / (q[1] * q[1]));
    }
    return;
  }
  const int i = (int)(g / n), j = (int)(g % n);
  const double X = grid_x(i, n, P.L), Y = grid_x(j, n, P.L);
  if (P.kind == GAUSSIANS_2D) {
    for (int b = 0; b < 3; ++b) {
      const double *q = P.p + 4 * b;
      const double dx = X - q[2], dy = Y - q[3];
      re += q[0] * exp(-(dx * dx + dy * dy) / (q[1] * q[1]));
    }
  } else if (P.kind == KINK_RING_2D) {
    re = tanh((P.p[0] - sqrt(X * X + Y * Y)) / P.p[1]);
  } else {
    const int K = (int)P.p[0];
    for (int b = 0; b < K; ++b) {
      const double *q = P.p + 1 + 6 * b;
      const double dx = X - q[4], dy = Y - q[5];
      const double env = q[0] * exp(-(dx * dx + dy * dy) / 1.0);
      const double ph = q[2] * dx + q[3] * dy + q[1];
      double s, c;
      sincos(ph, &s, &c);
      re += env * c;
      im += env * s;
    }
  }
}

} // namespace ic
} // namespace nls
