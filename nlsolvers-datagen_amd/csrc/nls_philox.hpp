// Counter-based noise for the stochastic phi^4 stepper.
//
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11): round multipliers
// 0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85.
// The reference draws its noise from thrust::default_random_engine seeded with
// time(nullptr) every step (device/include/stochastic_phi4.cuh:25,47-48), which
// is not reproducible; here every normal is a pure function of
// (seed, stream, step, cell):  counter = (cell/4, step, stream, 0), key = seed.
//
// Normals come from Box-Muller whose log and sin/cos are evaluated with a fixed
// sequence of IEEE +,-,*,/,sqrt,fma (no libm), so the same bits come out on any
// device and on the host.  Contraction is disabled for these functions.
#pragma once

#include <cstdint>
#include <cstring>

#ifndef NLS_HD
#define NLS_HD __host__ __device__
#endif

namespace nls {

struct u32x4 {
  uint32_t x, y, z, w;
};

NLS_HD inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = u32x4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
              (uint32_t)p0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// ln(u), u in (0,1]: u = 2^e x with x in [1/sqrt2, sqrt2), ln x = 2 atanh(s),
// s = (x-1)/(x+1), odd series through s^23 (|s| <= 0.1716, truncation < 1e-17).
NLS_HD inline double det_log(double u) {
#pragma clang fp contract(off)
  uint64_t bits;
  memcpy(&bits, &u, 8);
  int e = (int)((bits >> 52) & 0x7ff) - 1023;
  bits = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double x;
  memcpy(&x, &bits, 8);
  if (x > 1.4142135623730951) { x = x * 0.5; e += 1; }
  const double s = (x - 1.0) / (x + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 23.0;
  p = fma(p, s2, 1.0 / 21.0); p = fma(p, s2, 1.0 / 19.0); p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0); p = fma(p, s2, 1.0 / 13.0); p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);  p = fma(p, s2, 1.0 / 7.0);  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  const double lx = 2.0 * fma(s * s2, p, s);
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  return fma((double)e, ln2_hi, fma((double)e, ln2_lo, lx));
}

// (cos 2 pi w, sin 2 pi w), w in [0,1): the quadrant comes from 4w exactly, the
// reduced angle phi in [0, pi/4] goes through Taylor series to phi^17 / phi^18.
NLS_HD inline void det_sincos2pi(double w, double &cs, double &sn) {
#pragma clang fp contract(off)
  const double t = 4.0 * w;
  const int q = (int)t;
  double r = t - (double)q;
  bool swap = false;
  if (r > 0.5) { r = 1.0 - r; swap = true; }
  const double phi = r * 1.5707963267948966;
  const double p2 = phi * phi;
  double s_ = 1.0 / 355687428096000.0;
  s_ = fma(s_, -p2, 1.0 / 1307674368000.0);
  s_ = fma(s_, -p2, 1.0 / 6227020800.0);
  s_ = fma(s_, -p2, 1.0 / 39916800.0);
  s_ = fma(s_, -p2, 1.0 / 362880.0);
  s_ = fma(s_, -p2, 1.0 / 5040.0);
  s_ = fma(s_, -p2, 1.0 / 120.0);
  s_ = fma(s_, -p2, 1.0 / 6.0);
  s_ = fma(s_, -p2, 1.0);
  const double sphi = phi * s_;
  double c_ = 1.0 / 6402373705728000.0;
  c_ = fma(c_, -p2, 1.0 / 20922789888000.0);
  c_ = fma(c_, -p2, 1.0 / 87178291200.0);
  c_ = fma(c_, -p2, 1.0 / 479001600.0);
  c_ = fma(c_, -p2, 1.0 / 3628800.0);
  c_ = fma(c_, -p2, 1.0 / 40320.0);
  c_ = fma(c_, -p2, 1.0 / 720.0);
  c_ = fma(c_, -p2, 1.0 / 24.0);
  c_ = fma(c_, -p2, 1.0 / 2.0);
  c_ = fma(c_, -p2, 1.0);
  const double cr = swap ? sphi : c_, sr = swap ? c_ : sphi;
  switch (q & 3) {
  case 0: cs = cr; sn = sr; break;
  case 1: cs = -sr; sn = cr; break;
  case 2: cs = -cr; sn = -sr; break;
  default: cs = sr; sn = -cr; break;
  }
}

// Four N(0,1) draws for cells 4*group .. 4*group+3.
NLS_HD inline void noise4(uint64_t seed, uint32_t stream, uint32_t step, uint32_t group,
                          double z[4]) {
#pragma clang fp contract(off)
  const u32x4 x = philox4x32_10(u32x4{group, step, stream, 0u}, (uint32_t)seed,
                                (uint32_t)(seed >> 32));
  const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const double u1 = ((double)xs[2 * h] + 0.5) * 2.3283064365386963e-10;
    const double u2 = (double)xs[2 * h + 1] * 2.3283064365386963e-10;
    const double r = sqrt(-2.0 * det_log(u1));
    double cs, sn;
    det_sincos2pi(u2, cs, sn);
    z[2 * h] = r * cs;
    z[2 * h + 1] = r * sn;
  }
}

// One N(0,1) draw for a single cell, used where a thread owns one cell: the Philox block of
// its group, then only the Box-Muller pair the cell belongs to (one logarithm, one square
// root and one sine/cosine instead of two each); the same expressions as noise4, so the
// same bits as noise4's z[cell & 3].
NLS_HD inline double noise1(uint64_t seed, uint32_t stream, uint32_t step, uint64_t cell) {
#pragma clang fp contract(off)
  const u32x4 x = philox4x32_10(u32x4{(uint32_t)(cell >> 2), step, stream, 0u}, (uint32_t)seed,
                                (uint32_t)(seed >> 32));
  const bool second = (cell & 2) != 0;
  const uint32_t a = second ? x.z : x.x, b = second ? x.w : x.y;
  const double u1 = ((double)a + 0.5) * 2.3283064365386963e-10;
  const double u2 = (double)b * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * det_log(u1));
  double cs, sn;
  det_sincos2pi(u2, cs, sn);
  return (cell & 1) ? r * sn : r * cs;
}

} // namespace nls
