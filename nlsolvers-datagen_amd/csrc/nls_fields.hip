// Coefficient fields m(x), c(x) and the reference's kink_field / multi_soliton initial
// data, evaluated on the GPU behind the C ABI.
//
// Classes and forms: scripts/c_fields_2d.py:12-168 (class dispatch :169-248) and
// scripts/m_fields_2d.py:14-146 (dispatch :147-238); kink_field real_sampler.py:223-271
// with the anisotropic Gaussian random field of :48-65; multi_soliton_state
// nlse_sampler.py:102-219 with fundamental_soliton :29-100; the parameter spaces the
// launchers draw phenomena from, valid_spaces_real.py:33-41 and valid_spaces_complex.py:11-21.
//
// Every scalar parameter is drawn on the host from the trajectory's Philox stream
// (ParamStream: draw s is word 0 of Philox4x32-10 at counter (s, traj, tag, 0)), in the
// order in which the reference generators make their np.random calls; arrays of draws
// (the white noise of the turbulent and multiscale classes and of the random field) are
// generated on the GPU from the same counters (element e: word e mod 4 of counter
// slot0 + e / 4).  A field therefore depends only on (seed, trajectory), and equals the
// host generators of the Python package (nls/fields.py, nls/samplers.py) up to the last
// bits of exp / sin / cos / log and the summation order of the Gaussian filters, the
// discrete Fourier transforms and the reductions.  The evaluation repeats numpy's
// operation order in FP64 with contraction off.
//
// Grid: linspace(-L, L, n) in both directions, "ij" indexing, row-major n x n doubles.
// Workspace (noise, filter and transform buffers) is allocated per call: sampling runs
// once per trajectory, outside the time stepping.
#include "nls.h"
#include "nls_common.hpp"
#include "nls_ic.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace nls {
namespace fields {

constexpr int BS = 256;
constexpr int kMaxP = 160;
constexpr double kPi = 3.141592653589793;
constexpr double k2Pi = 6.283185307179586; // 2 * np.pi

using ic::grid_x;
using ic::ParamStream;

// ------------------------------------------------------------------ device buffers
struct Buf {
  double *p = nullptr;
  explicit Buf(size_t count) { NLS_HIP(hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(double))); }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
};

// ------------------------------------------------------------------ kernels
// Class codes of the closed-form evaluator.
enum Closed {
  C_PERIODIC = 1,
  C_PIECEWISE = 2,
  C_SIGN_CHECKER = 3,
  C_SIGN_HALF = 4,
  C_LAYERED = 5,
  C_QUASI = 6,
  M_PIECEWISE = 11,
  M_DEFECTS = 15,
  M_QUASI = 16,
  CONSTANT = 20,
};

struct Args {
  int kind, n, count;
  double L, base;
  double p[kMaxP];
};

// Fields that are explicit functions of the drawn scalars (c_fields_2d.py:12-52, 53-66,
// 117-140; m_fields_2d.py:14-37, 83-126 and the dispatch's np.maximum floors).
__global__ void k_closed(Args a, double *out) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int n = a.n;
  if (g >= (int64_t)n * n) return;
  const int i = (int)(g / n), j = (int)(g % n);
  const double X = grid_x(i, n, a.L), Y = grid_x(j, n, a.L);
  const double *p = a.p;
  double v = 0.0;
  switch (a.kind) {
  case CONSTANT:
    v = a.base;
    break;
  case C_PERIODIC: // base (1 + amp sin(kX) sin(kY))
    v = a.base * (1.0 + p[0] * sin(p[1] * X) * sin(p[1] * Y));
    break;
  case C_PIECEWISE: // layers [lo, hi) along x at base * contrast
    v = a.base;
    for (int k = 0; k < a.count; ++k)
      if (X >= p[1 + 2 * k] && X < p[2 + 2 * k]) v = a.base * p[0];
    break;
  case C_SIGN_CHECKER: // base tanh(sharp sin(pi X / cell) sin(pi Y / cell))
    v = a.base * tanh(p[0] * sin(kPi * X / p[1]) * sin(kPi * Y / p[1]));
    break;
  case C_SIGN_HALF: // base tanh(sharp X / L)
    v = a.base * tanh(p[0] * X / a.L);
    break;
  case C_LAYERED: // base + sum amp sin(freq (cos th X + sin th Y) + ph), before the rescale
    v = a.base;
    for (int k = 0; k < a.count; ++k) {
      const double *q = p + 5 * k; // cos th, sin th, amp, freq, ph
      const double d = q[0] * X + q[1] * Y;
      v += q[2] * sin(q[3] * d + q[4]);
    }
    break;
  case C_QUASI:
  case M_QUASI: // base + sum a cos(kx X + ky Y + ph)
    v = a.base;
    for (int k = 0; k < a.count; ++k) {
      const double *q = p + 4 * k; // kx, ky, a, ph
      v += q[2] * cos(q[0] * X + q[1] * Y + q[3]);
    }
    if (a.kind == M_QUASI) v = fmax(v, 0.1 * a.base);
    break;
  case M_PIECEWISE: { // m0 + (m2 - m0) 0.5 (1 + tanh(b / (sw L)))
    const int bt = (int)p[1];
    double b;
    if (bt == 0) b = sqrt(X * X + Y * Y) - p[2] * a.L;
    else if (bt == 1) b = fmax(fabs(X), fabs(Y)) - p[2] * a.L;
    else if (bt == 2) b = Y;
    else if (bt == 3) b = X;
    else b = X + Y;
    v = a.base + p[0] * (1.0 + tanh(b / (p[3] * a.L)));
    break;
  }
  case M_DEFECTS: // m0 + sum st exp(-|x - x0|^2 / (2 w^2)), floored at 0.1 m0
    v = a.base;
    for (int k = 0; k < a.count; ++k) {
      const double *q = p + 4 * k; // x0, y0, st, w
      const double dx = X - q[0], dy = Y - q[1];
      v += q[2] * exp(-(dx * dx + dy * dy) / (2.0 * (q[3] * q[3])));
    }
    v = fmax(v, 0.1 * a.base);
    break;
  }
  out[g] = v;
}

// Wave-guiding structures (c_fields_2d.py:67-116): p = max(p, amp exp(-d^2 / (2 w^2))) per
// guide, d the distance to the guide's 100 sample points (curves: guide-major x then y).
__global__ void k_waveguide(int n, double L, double base, double amp, int guides,
                            const double *widths, const double *curves, double *out) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (g >= (int64_t)n * n) return;
  const int i = (int)(g / n), j = (int)(g % n);
  const double X = grid_x(i, n, L), Y = grid_x(j, n, L);
  double v = base;
  for (int k = 0; k < guides; ++k) {
    const double *cx = curves + 200 * k, *cy = cx + 100;
    double d = INFINITY;
    for (int q = 0; q < 100; ++q) {
      const double dx = X - cx[q], dy = Y - cy[q];
      d = fmin(d, sqrt(dx * dx + dy * dy));
    }
    const double w = widths[k];
    v = fmax(v, amp * exp(-(d * d) / (2.0 * (w * w))));
  }
  out[g] = v;
}

// Standard normals by Box-Muller from two blocks of uniform draws (ParamStream.normal with
// a size: u1 from slots [slot0, slot0 + G), u2 from [slot0 + G, slot0 + 2G), G = ceil(cnt/4)).
__global__ void k_noise(double *z, int64_t cnt, uint64_t seed, uint32_t traj, uint32_t tag,
                        uint32_t slot0) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (e >= cnt) return;
  const uint32_t G = (uint32_t)((cnt + 3) / 4);
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const u32x4 a = philox4x32_10(u32x4{slot0 + (uint32_t)(e / 4), traj, tag, 0u}, k0, k1);
  const u32x4 b = philox4x32_10(u32x4{slot0 + G + (uint32_t)(e / 4), traj, tag, 0u}, k0, k1);
  const int w = (int)(e & 3);
  const uint32_t xa = w == 0 ? a.x : w == 1 ? a.y : w == 2 ? a.z : a.w;
  const uint32_t xb = w == 0 ? b.x : w == 1 ? b.y : w == 2 ? b.z : b.w;
  const double u1 = ((double)xa + 0.5) * 2.3283064365386963e-10;
  const double u2 = ((double)xb + 0.5) * 2.3283064365386963e-10;
  z[e] = sqrt(-2.0 * log(u1)) * cos(k2Pi * u2);
}

// numpy/scipy "reflect" extension (d c b a | a b c d | d c b a): period 2n.
__device__ inline int reflect(int p, int n) {
  const int period = 2 * n;
  int q = p % period;
  if (q < 0) q += period;
  return q < n ? q : period - 1 - q;
}

// One axis of scipy.ndimage.gaussian_filter (correlate1d with the normalised Gaussian
// weights w[0..2r], mode "reflect").
__global__ void k_gauss_axis(const double *in, double *out, int n, int axis, const double *w,
                             int r) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (g >= (int64_t)n * n) return;
  const int i = (int)(g / n), j = (int)(g % n);
  double s = 0.0;
  if (axis == 0) {
    for (int k = -r; k <= r; ++k) s += w[k + r] * in[(int64_t)reflect(i + k, n) * n + j];
  } else {
    for (int k = -r; k <= r; ++k) s += w[k + r] * in[(int64_t)i * n + reflect(j + k, n)];
  }
  out[g] = s;
}

// Reductions over cnt values (one workgroup, fixed order): 0 min, 1 max, 2 max |x|, 3 sum,
// 4 sum (x - shift)^2, 5 max |z| of interleaved complex values.
__global__ void __launch_bounds__(1024) k_reduce(const double *x, int64_t cnt, int op,
                                                 double shift, double *res) {
#pragma clang fp contract(off)
  __shared__ double sh[1024];
  const int t = threadIdx.x;
  double a = op == 0 ? INFINITY : (op <= 2 || op == 5) ? -INFINITY : 0.0;
  for (int64_t e = t; e < cnt; e += 1024) {
    double v;
    if (op == 5) v = hypot(x[2 * e], x[2 * e + 1]);
    else v = x[e];
    if (op == 0) a = fmin(a, v);
    else if (op == 1) a = fmax(a, v);
    else if (op == 2) a = fmax(a, fabs(v));
    else if (op == 5) a = fmax(a, v);
    else if (op == 3) a += v;
    else {
      const double d = v - shift;
      a += d * d;
    }
  }
  sh[t] = a;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (t < s) {
      const double b = sh[t + s];
      sh[t] = op == 0 ? fmin(sh[t], b) : (op == 3 || op == 4) ? sh[t] + b : fmax(sh[t], b);
    }
    __syncthreads();
  }
  if (t == 0) *res = sh[0];
}

// Element-wise finishing steps.
enum Ew {
  EW_RESCALE = 0,   // y = base (y - lo) / (hi - lo)            (c_fields_2d.py _rescale forms)
  EW_AXPY = 1,      // y += a x                                  (turbulent octaves)
  EW_AXPY_DIV = 2,  // y += a x / d                              (multiscale scales)
  EW_SUB = 3,       // y -= a
  EW_DIV = 4,       // y /= a
  EW_EXP = 5,       // y = base exp(a (y - 0.5))                 (turbulent)
  EW_FLOOR = 6,     // y = max(y, a)
  EW_SCALE_DIV = 7, // y = (x / d) a                              (random field scaling)
  EW_MREL = 8,      // y = base (1 + a (x / d))                  (phase, topological)
};

__global__ void k_ew(int op, double *y, const double *x, int64_t cnt, double a, double d,
                     double base, double lo, double hi) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (e >= cnt) return;
  double v = y ? y[e] : 0.0;
  switch (op) {
  case EW_RESCALE: v = base * (v - lo) / (hi - lo); break;
  case EW_AXPY: v += a * x[e]; break;
  case EW_AXPY_DIV: v += a * x[e] / d; break;
  case EW_SUB: v -= a; break;
  case EW_DIV: v /= a; break;
  case EW_EXP: v = base * exp(a * (v - 0.5)); break;
  case EW_FLOOR: v = fmax(v, a); break;
  case EW_SCALE_DIV: v = x[e] / d * a; break;
  case EW_MREL: v = base * (1.0 + a * (x[e] / d)); break;
  }
  y[e] = v;
}

// m fields derived from c (m_fields_2d.py:38-82): gradient (central differences inside,
// zero on the first and last row / column), the phase magnitude from rolls by k, and the
// sign of laplace(c) - lambda c with the "reflect" boundary.
__global__ void k_from_c(int what, const double *c, double *out, int n, double p0, double p1,
                         int k, double m0) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (g >= (int64_t)n * n) return;
  const int i = (int)(g / n), j = (int)(g % n);
  auto C = [&](int a, int b) { return c[(int64_t)a * n + b]; };
  if (what == 0) { // m0 (1 + gam g2 / (g2 + eps^2)), p0 = gam, p1 = eps^2
    const double gx = (i > 0 && i < n - 1) ? (C(i + 1, j) - C(i - 1, j)) / 2.0 : 0.0;
    const double gy = (j > 0 && j < n - 1) ? (C(i, j + 1) - C(i, j - 1)) / 2.0 : 0.0;
    const double g2 = gx * gx + gy * gy;
    out[g] = m0 * (1.0 + p0 * g2 / (g2 + p1));
  } else if (what == 1) { // |(roll k - roll -k)| magnitude
    const int ip = ((i - k) % n + n) % n, im = (i + k) % n;
    const int jp = ((j - k) % n + n) % n, jm = (j + k) % n;
    const double dx = C(ip, j) - C(im, j);
    const double dy = C(i, jp) - C(i, jm);
    out[g] = sqrt(dx * dx + dy * dy);
  } else { // sign(laplace(c) - lam c), p0 = lam
    const double cc = C(i, j);
    const double lx = C(reflect(i - 1, n), j) - 2.0 * cc + C(reflect(i + 1, n), j);
    const double ly = C(i, reflect(j - 1, n)) - 2.0 * cc + C(i, reflect(j + 1, n));
    const double s = (lx + ly) - p0 * cc;
    out[g] = s > 0.0 ? 1.0 : s < 0.0 ? -1.0 : 0.0;
  }
}

// ---- phenomena
struct Kinks {
  int count, n;
  double L;
  double p[3 * 16]; // per kink: sign * 4, x0, width
  int axis[16];     // 0: along x, 1: along y
};

// u0 = sum sign 4 atan(exp((G - x0) / w)) over the kinks along x, then along y
// (real_sampler.py:243-262).
__global__ void k_kinks(Kinks a, double *u) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int n = a.n;
  if (g >= (int64_t)n * n) return;
  const int i = (int)(g / n), j = (int)(g % n);
  const double X = grid_x(i, n, a.L), Y = grid_x(j, n, a.L);
  double v = 0.0;
  for (int k = 0; k < a.count; ++k) {
    const double G = a.axis[k] == 0 ? X : Y;
    v += a.p[3 * k] * atan(exp((G - a.p[3 * k + 1]) / a.p[3 * k + 2]));
  }
  u[g] = v;
}

// complex noise (re = first normal block, im = second) -> interleaved z
__global__ void k_pack_complex(const double *re, const double *im, double *z, int64_t cnt) {
  const int64_t e = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (e >= cnt) return;
  z[2 * e] = re[e];
  z[2 * e + 1] = im[e];
}

// One axis of a 2D DFT of interleaved complex n x n data: out[k] = sum_j in[j] tw[(j k) mod n]
// (tw = exp(-+2 pi i m / n)), then scaled by `scale`.
__global__ void k_dft_axis(const double *in, double *out, int n, int axis, const double *tw,
                           double scale) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (g >= (int64_t)n * n) return;
  const int i = (int)(g / n), j = (int)(g % n);
  const int kk = axis == 0 ? i : j;
  double sr = 0.0, si = 0.0;
  int m = 0; // (q kk) mod n, advanced incrementally
  for (int q = 0; q < n; ++q) {
    const int64_t src = axis == 0 ? (int64_t)q * n + j : (int64_t)i * n + q;
    const double xr = in[2 * src], xi = in[2 * src + 1];
    const double wr = tw[2 * m], wi = tw[2 * m + 1];
    sr += xr * wr - xi * wi;
    si += xr * wi + xi * wr;
    m += kk;
    if (m >= n) m -= n;
  }
  out[2 * g] = sr * scale;
  out[2 * g + 1] = si * scale;
}

// z *= sqrt(spectrum) of the rotated anisotropic spectrum exp(-((kx'/lx)^2 + (ky'/ly)^2))
// (real_sampler.py:48-65 with power 2); k = 2 pi fftfreq(n, 2L/n).
__global__ void k_spectrum(double *z, const double *kv, int n, double ct, double st, double lx,
                           double ly) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (g >= (int64_t)n * n) return;
  const int i = (int)(g / n), j = (int)(g % n);
  const double KX = kv[i], KY = kv[j];
  const double xr = KX * ct - KY * st, yr = KX * st + KY * ct;
  const double a = xr / lx, b = yr / ly;
  const double s = sqrt(exp(-(a * a + b * b)));
  z[2 * g] *= s;
  z[2 * g + 1] *= s;
}

__global__ void k_real_part(const double *z, double *x, int64_t cnt) {
  const int64_t e = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (e < cnt) x[e] = z[2 * e];
}

struct Solitons {
  int count, n, system, order;
  double L, interaction;
  // per soliton: x0, y0, phase, amp, width, Lambda, chirp, aspect, cos o, sin o, vx, vy,
  // beta (cubic-quintic), sqrt Lambda
  double p[14 * 11];
};

__device__ inline double sech_pow(double x, int order) {
  const double s = 1.0 / cosh(x);
  return order == 1 ? s : pow(s, (double)order);
}

// sum of fundamental_soliton components (nlse_sampler.py:29-100 without the envelope),
// components after the first weighted by the interaction strength (:190-196).
__global__ void k_solitons(Solitons a, double *psi) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int n = a.n;
  if (g >= (int64_t)n * n) return;
  const int i = (int)(g / n), j = (int)(g % n);
  const double X = grid_x(i, n, a.L), Y = grid_x(j, n, a.L);
  double ur = 0.0, ui = 0.0;
  for (int k = 0; k < a.count; ++k) {
    const double *q = a.p + 14 * k;
    const double dx = X - q[0], dy = Y - q[1];
    const double xr = dx * q[8] + dy * q[9];
    const double yr = -dx * q[9] + dy * q[8];
    const double xa = xr / q[7];
    const double r = sqrt(xa * xa + yr * yr);
    const double mom = q[10] * dx + q[11] * dy;
    const double total = mom + q[2] + q[6] * (r * r);
    const double amp = q[3], w = q[4];
    double prof;
    if (a.system == 0) { // cubic
      prof = amp * sech_pow(r / w, a.order);
    } else if (a.system == 1) { // cubic-quintic, beta = -sigma2 amp^2 / sigma1 > 0
      const double s = sech_pow(r / w, a.order);
      const double s2 = a.order == 1 ? s * s : pow(1.0 / cosh(r / w), 2.0 * a.order);
      prof = amp * s / sqrt(1.0 + q[12] * s2);
    } else if (a.system == 2) { // saturable
      const double s = sech_pow(r / w, a.order);
      const double den = sqrt(1.0 + (amp * amp) * (s * s));
      prof = amp * s / den;
    } else { // Glasner-Allen-Flowers
      const double s = sech_pow(q[13] * r, a.order);
      const double s2 = a.order == 1 ? s * s : pow(s, 2.0 / a.order);
      prof = amp * s / sqrt(9.0 - 48.0 * q[5] * s2 + 31.0);
    }
    double sn, cs;
    sincos(total, &sn, &cs);
    double cr = prof * cs, ci = prof * sn;
    if (k > 0 && a.interaction < 1.0) {
      cr *= a.interaction;
      ci *= a.interaction;
    }
    ur += cr;
    ui += ci;
  }
  psi[2 * g] = ur;
  psi[2 * g + 1] = ui;
}

// ------------------------------------------------------------------ host helpers
struct Ctx {
  int n;
  int64_t cnt;
  hipStream_t s;
  unsigned grid() const { return blocks_for(cnt, BS); }
};

double reduce(const Ctx &c, const double *x, int op, double shift = 0.0, int64_t cnt = -1) {
  Buf r(1);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c.s, x, cnt < 0 ? c.cnt : cnt, op, shift,
                     r.p);
  check_launch("k_reduce");
  double h = 0.0;
  NLS_HIP(hipMemcpyAsync(&h, r.p, sizeof(double), hipMemcpyDeviceToHost, c.s));
  NLS_HIP(hipStreamSynchronize(c.s));
  return h;
}

void ew(const Ctx &c, int op, double *y, const double *x, double a, double d = 1.0,
        double base = 1.0, double lo = 0.0, double hi = 1.0) {
  hipLaunchKernelGGL(k_ew, dim3(c.grid()), dim3(BS), 0, c.s, op, y, x, c.cnt, a, d, base, lo,
                     hi);
  check_launch("k_ew");
}

// numpy.linspace(start, stop, num)[k]
double linspace_at(double start, double stop, int num, int k) {
  if (num > 1 && k == num - 1) return stop;
  const double step = (stop - start) / (double)(num - 1);
  return (double)k * step + start;
}

// scipy.ndimage gaussian_filter1d weights: radius int(4 sigma + 0.5),
// exp(-0.5 / sigma^2 x^2) normalised by their sum.
std::vector<double> gauss_weights(double sigma) {
  const int r = (int)(4.0 * sigma + 0.5);
  std::vector<double> w(2 * r + 1);
  const double c = -0.5 / (sigma * sigma);
  double sum = 0.0;
  for (int x = -r; x <= r; ++x) w[x + r] = std::exp(c * (double)(x * x));
  for (double v : w) sum += v;
  for (double &v : w) v /= sum;
  return w;
}

// y = gaussian_filter(x, sigma) (axis 0, then axis 1); tmp is n*n
void gaussian_filter(const Ctx &c, const double *x, double *y, double *tmp, double sigma) {
  const std::vector<double> w = gauss_weights(sigma);
  Buf dw(w.size());
  NLS_HIP(hipMemcpyAsync(dw.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
  const int r = (int)(w.size() / 2);
  hipLaunchKernelGGL(k_gauss_axis, dim3(c.grid()), dim3(BS), 0, c.s, x, tmp, c.n, 0, dw.p, r);
  check_launch("k_gauss_axis");
  hipLaunchKernelGGL(k_gauss_axis, dim3(c.grid()), dim3(BS), 0, c.s, (const double *)tmp, y, c.n,
                     1, dw.p, r);
  check_launch("k_gauss_axis");
  NLS_HIP(hipStreamSynchronize(c.s)); // dw goes out of scope
}

// ParamStream.normal(size=cnt) on the device; advances the stream's slot.
void noise(const Ctx &c, ParamStream &ps, double *z, int64_t cnt) {
  const uint32_t G = (uint32_t)((cnt + 3) / 4);
  hipLaunchKernelGGL(k_noise, dim3(blocks_for(cnt, BS)), dim3(BS), 0, c.s, z, cnt, ps.seed,
                     ps.traj, ps.tag, ps.slot);
  check_launch("k_noise");
  ps.slot += 2 * G;
}

void closed(const Ctx &c, const Args &a, double *out) {
  hipLaunchKernelGGL(k_closed, dim3(c.grid()), dim3(BS), 0, c.s, a, out);
  check_launch("k_closed");
}

Args args(int kind, int n, double L, double base) {
  Args a{};
  a.kind = kind;
  a.n = n;
  a.L = L;
  a.base = base;
  return a;
}

// _quasi_waves of nls/fields.py (c_fields_2d.py:117-140, m_fields_2d.py:101-126)
void quasi_waves(Args &a, ParamStream &ps, int waves, double amp_lo, double amp_hi,
                 double scale) {
  const double golden = (1.0 + std::sqrt(5.0)) / 2.0;
  a.count = waves;
  for (int i = 0; i < waves; ++i) {
    const double ang = (double)i * kPi / (double)waves;
    const double s = std::pow(golden, (double)i);
    double *q = a.p + 4 * i;
    q[0] = std::cos(ang) * s;
    q[1] = std::sin(ang) * s;
    q[2] = ps.uniform(amp_lo, amp_hi) * scale;
    q[3] = ps.uniform(0.0, k2Pi);
  }
}

void rescale(const Ctx &c, double *out, double base) {
  const double lo = reduce(c, out, 0), hi = reduce(c, out, 1);
  if (hi > lo) ew(c, EW_RESCALE, out, nullptr, 0.0, 1.0, base, lo, hi);
  else {
    Args a = args(CONSTANT, c.n, 1.0, base);
    closed(c, a, out);
  }
}

// ------------------------------------------------------------------ c classes
void c_field(const Ctx &c, const std::string &kind, double L, ParamStream &ps, double base,
             double *out) {
  const int n = c.n;
  if (kind == "constant") { // np.ones: 1 whatever the base value
    closed(c, args(CONSTANT, n, L, 1.0), out);
  } else if (kind == "periodic_structure") {
    Args a = args(C_PERIODIC, n, L, base);
    const double amp = ps.uniform(0.2, 0.5);
    const int freq = ps.integer(1, 2);
    a.p[0] = amp;
    a.p[1] = kPi * (double)freq / L;
    closed(c, a, out);
  } else if (kind == "piecewise_constant") {
    Args a = args(C_PIECEWISE, n, L, base);
    const int layers = ps.integer(2, 4);
    a.p[0] = ps.uniform(1.5, 2.5);
    const double w = 2.0 * L / (double)layers;
    int k = 0;
    for (int i = 1; i < layers; i += 2, ++k) {
      a.p[1 + 2 * k] = -L + (double)i * w;
      a.p[2 + 2 * k] = -L + (double)(i + 1) * w;
    }
    a.count = k;
    closed(c, a, out);
  } else if (kind == "sign_changing_mass") {
    const bool checker = ps.uniform(0.0, 1.0) < 0.5;
    const double sharp = ps.uniform(3.0, 6.0);
    Args a = args(checker ? C_SIGN_CHECKER : C_SIGN_HALF, n, L, base);
    a.p[0] = sharp;
    a.p[1] = L / 2.0; // cell = L / scale, scale = 2
    closed(c, a, out);
  } else if (kind == "layered") {
    Args a = args(C_LAYERED, n, L, base);
    const int layers = ps.integer(2, 5);
    const double a_lo = ps.uniform(0.1, 0.3), a_hi = ps.uniform(0.4, 0.8);
    const double f_lo = ps.uniform(1.0, 3.0), f_hi = ps.uniform(5.0, 15.0);
    a.count = layers;
    for (int k = 0; k < layers; ++k) {
      const double th = ps.uniform(0.0, kPi);
      double *q = a.p + 5 * k;
      q[0] = std::cos(th);
      q[1] = std::sin(th);
      q[2] = ps.uniform(a_lo, a_hi);
      q[3] = ps.uniform(f_lo, f_hi);
      q[4] = ps.uniform(0.0, k2Pi);
    }
    closed(c, a, out);
    rescale(c, out, base);
  } else if (kind == "waveguide") {
    const double w_lo = ps.uniform(0.1, 0.3), w_hi = ps.uniform(0.4, 0.8);
    const double gamp = ps.uniform(0.5, 2.0);
    const int guides = ps.integer(3, 11);
    std::vector<double> widths(guides), curves(200 * (size_t)guides);
    for (int k = 0; k < guides; ++k) {
      const bool line = ps.uniform(0.0, 1.0) < 0.5;
      widths[k] = ps.uniform(w_lo, w_hi);
      double *cx = curves.data() + 200 * k, *cy = cx + 100;
      if (line) { // straight guide through a random point
        const double x0 = ps.uniform(-L, L), y0 = ps.uniform(-L, L), th = ps.uniform(0.0, k2Pi);
        const double ct = std::cos(th), st = std::sin(th);
        for (int q = 0; q < 100; ++q) {
          const double t = linspace_at(-1.5 * L, 1.5 * L, 100, q);
          cx[q] = x0 + t * ct;
          cy[q] = y0 + t * st;
        }
      } else { // ellipse
        const double ea = ps.uniform(0.5, 2.0), eb = ps.uniform(0.5, 2.0),
                     ph = ps.uniform(0.0, k2Pi);
        for (int q = 0; q < 100; ++q) {
          const double t = linspace_at(0.0, k2Pi, 100, q);
          cx[q] = ea * std::cos(t + ph);
          cy[q] = eb * std::sin(t);
        }
      }
    }
    Buf dw(guides), dc(curves.size());
    NLS_HIP(hipMemcpyAsync(dw.p, widths.data(), guides * sizeof(double), hipMemcpyHostToDevice, c.s));
    NLS_HIP(hipMemcpyAsync(dc.p, curves.data(), curves.size() * sizeof(double),
                           hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(k_waveguide, dim3(c.grid()), dim3(BS), 0, c.s, n, L, base, gamp, guides,
                       (const double *)dw.p, (const double *)dc.p, out);
    check_launch("k_waveguide");
    NLS_HIP(hipStreamSynchronize(c.s));
  } else if (kind == "quasiperiodic") {
    Args a = args(C_QUASI, n, L, base);
    const int waves = ps.integer(3, 7);
    const double lo = ps.uniform(0.1, 0.3), hi = ps.uniform(0.4, 0.8);
    quasi_waves(a, ps, waves, lo, hi, 1.0);
    closed(c, a, out);
    rescale(c, out, base);
  } else if (kind == "turbulent") {
    const double inten = ps.uniform(0.3, 0.8), s_lo = ps.uniform(1.0, 3.0);
    const double s_hi = ps.uniform(10.0, 30.0), beta = ps.uniform(1.0, 3.0);
    const int octaves = ps.integer(3, 7);
    Buf z(c.cnt), sm(c.cnt), tmp(c.cnt);
    closed(c, args(CONSTANT, n, L, 0.0), out); // f = 0
    for (int o = 0; o < octaves; ++o) {
      const double sc = s_hi / std::pow(2.0, (double)o);
      if (sc < s_lo) break;
      noise(c, ps, z.p, c.cnt);
      gaussian_filter(c, z.p, sm.p, tmp.p, sc);
      ew(c, EW_AXPY, out, sm.p, std::pow(sc, beta));
    }
    ew(c, EW_SUB, out, nullptr, reduce(c, out, 0));
    ew(c, EW_DIV, out, nullptr, reduce(c, out, 1));
    ew(c, EW_EXP, out, nullptr, inten, 1.0, base);
    NLS_HIP(hipStreamSynchronize(c.s));
  } else {
    throw Error("nls_field_sample: unknown c class '" + kind +
                "' (constant, periodic_structure, piecewise_constant, sign_changing_mass, "
                "layered, waveguide, quasiperiodic, turbulent)");
  }
}

// ------------------------------------------------------------------ m classes
void m_field(const Ctx &c, std::string kind, double L, ParamStream &ps, double m0,
             const double *cf, double *out) {
  const int n = c.n;
  if ((kind == "gradient" || kind == "phase" || kind == "topological") && !cf)
    kind = "constant"; // as the reference's dispatch does without a c field
  if (kind == "constant") {
    closed(c, args(CONSTANT, n, L, m0), out);
  } else if (kind == "piecewise") {
    Args a = args(M_PIECEWISE, n, L, m0);
    const double m2 = ps.uniform(1.5, 3.0) * m0;
    const int bt = ps.integer(0, 4);
    const double bp = ps.uniform(0.3, 0.7), sw = ps.uniform(0.01, 0.1);
    a.p[0] = (m2 - m0) * 0.5;
    a.p[1] = bt;
    a.p[2] = bp;
    a.p[3] = sw;
    closed(c, a, out);
  } else if (kind == "gradient") {
    const double gam = ps.uniform(0.5, 2.0), eps = ps.uniform(1e-3, 1e-1);
    hipLaunchKernelGGL(k_from_c, dim3(c.grid()), dim3(BS), 0, c.s, 0, cf, out, n, gam, eps * eps,
                       0, m0);
    check_launch("k_from_c");
  } else if (kind == "phase") {
    const double delta = ps.uniform(0.3, 1.0), frac = ps.uniform(0.02, 0.1);
    const int k = std::max(1, (int)(frac * (double)n));
    Buf mag(c.cnt);
    hipLaunchKernelGGL(k_from_c, dim3(c.grid()), dim3(BS), 0, c.s, 1, cf, mag.p, n, 0.0, 0.0, k, m0);
    check_launch("k_from_c");
    const double mx = reduce(c, mag.p, 2);
    if (mx > 0.0) ew(c, EW_MREL, out, mag.p, delta, mx, m0);
    else ew(c, EW_MREL, out, mag.p, delta, 1.0, m0);
    NLS_HIP(hipStreamSynchronize(c.s));
  } else if (kind == "topological") {
    const double eta = ps.uniform(0.5, 1.0), lam = ps.uniform(0.3, 0.7);
    Buf sg(c.cnt), t(c.cnt), tmp(c.cnt);
    hipLaunchKernelGGL(k_from_c, dim3(c.grid()), dim3(BS), 0, c.s, 2, cf, sg.p, n, lam, 0.0, 0, m0);
    check_launch("k_from_c");
    gaussian_filter(c, sg.p, t.p, tmp.p, 1.0);
    const double mx = reduce(c, t.p, 2);
    ew(c, EW_MREL, out, t.p, eta, mx > 0.0 ? mx : 1.0, m0);
    NLS_HIP(hipStreamSynchronize(c.s));
  } else if (kind == "defects") {
    Args a = args(M_DEFECTS, n, L, m0);
    const int num = ps.integer(5, 19);
    const double s_lo = ps.uniform(-0.5, -0.1), s_hi = ps.uniform(0.5, 1.0);
    const double w_lo = ps.uniform(0.03, 0.08), w_hi = ps.uniform(0.1, 0.3);
    NLS_REQUIRE(4 * num <= kMaxP, "defect count over the parameter capacity");
    a.count = num;
    for (int k = 0; k < num; ++k) {
      double *q = a.p + 4 * k;
      q[0] = ps.uniform(-L, L);
      q[1] = ps.uniform(-L, L);
      q[2] = ps.uniform(s_lo, s_hi) * m0;
      q[3] = ps.uniform(w_lo, w_hi) * L;
    }
    closed(c, a, out);
  } else if (kind == "quasiperiodic") {
    Args a = args(M_QUASI, n, L, m0);
    const int waves = ps.integer(3, 7);
    const double lo = ps.uniform(0.1, 0.3), hi = ps.uniform(0.4, 0.8);
    quasi_waves(a, ps, waves, lo, hi, m0);
    closed(c, a, out);
  } else if (kind == "multiscale") {
    const int scales = ps.integer(3, 5);
    const double s_lo = ps.uniform(1.0, 3.0), s_hi = ps.uniform(8.0, 20.0);
    const double a_lo = ps.uniform(0.1, 0.3), a_hi = ps.uniform(0.4, 0.8);
    Buf z(c.cnt), sm(c.cnt), tmp(c.cnt);
    closed(c, args(CONSTANT, n, L, m0), out);
    const double l0 = std::log10(s_lo), l1 = std::log10(s_hi);
    for (int k = 0; k < scales; ++k) {
      const double sc = std::pow(10.0, linspace_at(l0, l1, scales, k));
      noise(c, ps, z.p, c.cnt);
      gaussian_filter(c, z.p, sm.p, tmp.p, sc);
      const double mx = reduce(c, sm.p, 2);
      ew(c, EW_AXPY_DIV, out, sm.p, ps.uniform(a_lo, a_hi) * m0, mx);
    }
    ew(c, EW_FLOOR, out, nullptr, 0.1 * m0);
    NLS_HIP(hipStreamSynchronize(c.s));
  } else {
    throw Error("nls_field_sample: unknown m class '" + kind +
                "' (constant, piecewise, gradient, phase, topological, defects, "
                "quasiperiodic, multiscale)");
  }
}

} // namespace fields
} // namespace nls

namespace nls {
namespace fields {

// ------------------------------------------------------------------ phenomena
// Parameter spaces (valid_spaces_real.py:33-41, valid_spaces_complex.py:11-21): the
// number of choices per key, in key order.
constexpr int kKinkKeys = 5;
constexpr int kKinkSizes[kKinkKeys] = {5, 9, 9, 9, 2}; // system, winding x, y, widths, randomize
constexpr double kKinkWidthLo[3] = {0.3, 0.5, 0.7}, kKinkWidthHi[3] = {1.5, 2.0, 3.0};
constexpr int kSolKeys = 9;
// system, width range, amplitude range, phase pattern, arrangement, coherence,
// velocity scale, chirp range, aspect-ratio range
constexpr int kSolSizes[kSolKeys] = {4, 3, 3, 4, 4, 3, 3, 2, 2};
constexpr double kSolRanges[3][2] = {{0.5, 1.0}, {1.0, 1.5}, {1.5, 2.0}};
constexpr double kSolVelocity[3] = {0.0, 0.5, 1.0};
constexpr double kSolChirp[2][2] = {{-0.5, 0.0}, {0.0, 0.5}};
constexpr double kSolAspect[2][2] = {{1.0, 1.0}, {1.0, 1.5}};

// ParamStream.normal(loc, scale) for one value
double normal1(ParamStream &ps, double loc, double scale) {
  const double z = ps.normal();
  return (loc == 0.0 && scale == 1.0) ? z : loc + scale * z;
}

// numpy's add.reduce of a short vector: the first element plus the pairwise sum of the
// rest (eight partial sums from eight elements on)
double np_sum(const std::vector<double> &a) {
  if (a.empty()) return 0.0;
  const size_t m = a.size() - 1;
  const double *r0 = a.data() + 1;
  double rest;
  if (m < 8) {
    rest = -0.0;
    for (size_t i = 0; i < m; ++i) rest += r0[i];
  } else {
    double r[8];
    for (int k = 0; k < 8; ++k) r[k] = r0[k];
    size_t i = 8;
    for (; i < m - (m % 8); i += 8)
      for (int k = 0; k < 8; ++k) r[k] += r0[i + k];
    rest = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < m; ++i) rest += r0[i];
  }
  return a[0] + rest;
}

// anisotropic Gaussian random field (real_sampler.py:48-65): theta 30 degrees, anisotropy
// ratio 2, power 2; written to v scaled to standard deviation `amplitude`.
void anisotropic_grf(const Ctx &c, double L, ParamStream &rng, double length_scale,
                     double amplitude, double *v) {
  const int n = c.n;
  const int64_t cnt = c.cnt;
  std::vector<double> kv(n), tw(2 * (size_t)n);
  const double d = 2.0 * L / (double)n, val = 1.0 / ((double)n * d);
  const int N = (n - 1) / 2 + 1;
  for (int i = 0; i < n; ++i) kv[i] = k2Pi * ((double)(i < N ? i : i - n) * val);
  const double th = 30.0 * (kPi / 180.0);
  const double lx = length_scale * std::sqrt(2.0), ly = length_scale / std::sqrt(2.0);
  Buf re(cnt), im(cnt), z(2 * cnt), w(2 * cnt), dk(n), dtw(2 * (size_t)n);
  noise(c, rng, re.p, cnt);
  noise(c, rng, im.p, cnt);
  hipLaunchKernelGGL(k_pack_complex, dim3(c.grid()), dim3(BS), 0, c.s, (const double *)re.p,
                     (const double *)im.p, z.p, cnt);
  check_launch("k_pack_complex");
  NLS_HIP(hipMemcpyAsync(dk.p, kv.data(), n * sizeof(double), hipMemcpyHostToDevice, c.s));
  auto dft = [&](double sign, double scale) {
    for (int m = 0; m < n; ++m) {
      const double a = k2Pi * (double)m / (double)n;
      tw[2 * m] = std::cos(a);
      tw[2 * m + 1] = sign * std::sin(a);
    }
    NLS_HIP(hipMemcpyAsync(dtw.p, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice,
                           c.s));
    hipLaunchKernelGGL(k_dft_axis, dim3(c.grid()), dim3(BS), 0, c.s, (const double *)z.p, w.p, n,
                       0, (const double *)dtw.p, scale);
    check_launch("k_dft_axis");
    hipLaunchKernelGGL(k_dft_axis, dim3(c.grid()), dim3(BS), 0, c.s, (const double *)w.p, z.p, n,
                       1, (const double *)dtw.p, scale);
    check_launch("k_dft_axis");
    NLS_HIP(hipStreamSynchronize(c.s)); // tw is rewritten by the next transform
  };
  dft(-1.0, 1.0); // fft2
  hipLaunchKernelGGL(k_spectrum, dim3(c.grid()), dim3(BS), 0, c.s, z.p, (const double *)dk.p, n,
                     std::cos(th), std::sin(th), lx, ly);
  check_launch("k_spectrum");
  dft(1.0, 1.0 / (double)n); // ifft2
  hipLaunchKernelGGL(k_real_part, dim3(c.grid()), dim3(BS), 0, c.s, (const double *)z.p, re.p,
                     cnt);
  check_launch("k_real_part");
  const double mean = reduce(c, re.p, 3) / (double)cnt;
  const double sd = std::sqrt(reduce(c, re.p, 4, mean) / (double)cnt);
  ew(c, EW_SCALE_DIV, v, re.p, amplitude, sd);
  NLS_HIP(hipStreamSynchronize(c.s));
}

// kink_field (real_sampler.py:223-271) with the parameters of the reference's space
// (space stream tag 0x4B1E, sampler stream tag 0x4B1F)
void kink_field(const Ctx &c, double L, uint64_t seed, uint32_t traj, double *u, double *v,
                int *ch) {
  ParamStream sp{seed, traj, 0x4B1Eu};
  for (int k = 0; k < kKinkKeys; ++k) ch[k] = sp.integer(0, kKinkSizes[k] - 1);
  const int wind[2] = {ch[1] - 4, ch[2] - 4};
  const double wr0 = kKinkWidthLo[ch[3] / 3], wr1 = kKinkWidthHi[ch[3] % 3];
  const bool randomize = ch[4] == 0; // [True, False]
  ParamStream rng{seed, traj, 0x4B1Fu};
  Kinks K{};
  K.n = c.n;
  K.L = L;
  for (int axis = 0; axis < 2; ++axis) {
    const int w = wind[axis];
    if (w == 0) continue;
    const int aw = std::abs(w);
    const double width = wr0 + (wr1 - wr0) * rng.uniform(0.0, 1.0);
    double pos[4];
    for (int i = 0; i < aw; ++i)
      pos[i] = randomize ? L * (2.0 * rng.uniform(0.0, 1.0) - 1.0)
                         : L * (-0.8 + 1.6 * (double)i / (double)aw);
    for (int i = 0; i < aw; ++i) {
      const double kw = width * (0.8 + 0.4 * rng.uniform(0.0, 1.0));
      K.p[3 * K.count] = w > 0 ? 4.0 : -4.0;
      K.p[3 * K.count + 1] = pos[i];
      K.p[3 * K.count + 2] = kw;
      K.axis[K.count++] = axis;
    }
  }
  hipLaunchKernelGGL(k_kinks, dim3(c.grid()), dim3(BS), 0, c.s, K, u);
  check_launch("k_kinks");
  const double amp = reduce(c, u, 2) * 0.1;
  anisotropic_grf(c, L, rng, (wr0 + wr1) / 2.0 * 2.0, amp, v);
}

// multi_soliton_state (nlse_sampler.py:102-219) with the parameters of the reference's
// space (space stream tag 0x50A1, sampler stream tag 0x50A2); psi interleaved complex
void multi_soliton(const Ctx &c, double L, uint64_t seed, uint32_t traj, bool normalize,
                   double *psi, int *ch) {
  ParamStream sp{seed, traj, 0x50A1u};
  for (int k = 0; k < kSolKeys; ++k) ch[k] = sp.integer(0, kSolSizes[k] - 1);
  const int system = ch[0], pattern = ch[3], arrangement = ch[4];
  const double *wr = kSolRanges[ch[1]], *ar = kSolRanges[ch[2]];
  const double vs = kSolVelocity[ch[6]];
  const double *cr = kSolChirp[ch[7]], *asr = kSolAspect[ch[8]];
  const double separation = 5.0, interaction = 0.5;
  ParamStream rng{seed, traj, 0x50A2u};
  const int ns = rng.integer(3, 11); // randint(3, 12)
  ch[kSolKeys] = ns;
  std::vector<double> px(ns), py(ns), phase(ns);
  if (arrangement == 0) { // linear
    for (int i = 0; i < ns; ++i) {
      px[i] = ((double)i - (double)(ns - 1) / 2.0) * separation;
      py[i] = 0.0;
    }
  } else if (arrangement == 1) { // circular
    for (int i = 0; i < ns; ++i) {
      const double ang = k2Pi * (double)i / (double)ns;
      px[i] = separation * std::cos(ang);
      py[i] = separation * std::sin(ang);
    }
  } else if (arrangement == 2) { // random, N(0, L/4)
    for (int i = 0; i < ns; ++i) {
      px[i] = normal1(rng, 0.0, 1.0 * L / 4.0);
      py[i] = normal1(rng, 0.0, 1.0 * L / 4.0);
    }
  } else { // lattice
    const int side = (int)std::ceil(std::sqrt((double)ns));
    int k = 0;
    for (int i = 0; i < side; ++i)
      for (int j = 0; j < side; ++j)
        if (k < ns) {
          px[k] = ((double)i - (double)(side - 1) / 2.0) * separation;
          py[k++] = ((double)j - (double)(side - 1) / 2.0) * separation;
        }
  }
  if (pattern == 0) { // random: uniform(0, 2 pi, size=ns), one block of counters
    const int groups = (ns + 3) / 4;
    for (int gi = 0; gi < groups; ++gi) {
      const u32x4 r = rng.next();
      const uint32_t wv[4] = {r.x, r.y, r.z, r.w};
      for (int q = 0; q < 4 && 4 * gi + q < ns; ++q)
        phase[4 * gi + q] = 0.0 + (k2Pi - 0.0) * ParamStream::unit(wv[q]);
    }
  } else if (pattern == 1) { // alternating
    for (int i = 0; i < ns; ++i) phase[i] = (double)i * kPi;
  } else if (pattern == 2) { // synchronized
    for (int i = 0; i < ns; ++i) phase[i] = 0.0;
  } else { // vortex
    const double cx = np_sum(px) / (double)ns, cy = np_sum(py) / (double)ns;
    for (int i = 0; i < ns; ++i) phase[i] = std::atan2(py[i] - cy, px[i] - cx);
  }
  Solitons S{};
  S.count = ns;
  S.n = c.n;
  S.system = system;
  S.order = 1;
  S.L = L;
  S.interaction = interaction;
  for (int i = 0; i < ns; ++i) {
    double *q = S.p + 14 * i;
    const double amp = rng.uniform(ar[0], ar[1]);
    const double width = rng.uniform(wr[0], wr[1]);
    const double Lambda = rng.uniform(0.04, 0.14);
    const int order = rng.integer(1, 1); // randint(1, 2)
    const double chirp = rng.uniform(cr[0], cr[1]);
    const double aspect = rng.uniform(asr[0], asr[1]);
    const double orient = rng.uniform(0.0, k2Pi);
    double vx = 0.0, vy = 0.0;
    if (vs > 0.0) {
      if (arrangement == 1) {
        const double ang = k2Pi * (double)i / (double)ns;
        vx = -vs * std::cos(ang);
        vy = -vs * std::sin(ang);
      } else {
        vx = normal1(rng, 0.0, vs);
        vy = normal1(rng, 0.0, vs);
      }
    }
    (void)order;
    q[0] = px[i];
    q[1] = py[i];
    q[2] = phase[i];
    q[3] = amp;
    q[4] = width;
    q[5] = Lambda;
    q[6] = chirp;
    q[7] = aspect;
    q[8] = std::cos(orient);
    q[9] = std::sin(orient);
    q[10] = vx;
    q[11] = vy;
    q[12] = 0.1 * (amp * amp) / 1.0; // -sigma2 amp^2 / sigma1
    q[13] = std::sqrt(Lambda);
  }
  hipLaunchKernelGGL(k_solitons, dim3(c.grid()), dim3(BS), 0, c.s, S, psi);
  check_launch("k_solitons");
  if (normalize) {
    const double mx = reduce(c, psi, 5);
    ew(Ctx{c.n, 2 * c.cnt, c.s}, EW_DIV, psi, nullptr, mx);
  }
  NLS_HIP(hipStreamSynchronize(c.s));
}

} // namespace fields
} // namespace nls

using namespace nls;

extern "C" {

int nls_field_sample(const char *family, const char *type, int n, double L, uint64_t seed,
                     int traj, double base, const double *c, double *out, void *stream) {
  return guard([&] {
    NLS_REQUIRE(family && type && out, "nls_field_sample: null argument");
    NLS_REQUIRE(n >= 2 && n <= 8192, "nls_field_sample: n must lie in [2, 8192]");
    NLS_REQUIRE(L > 0.0 && traj >= 0, "nls_field_sample: bad domain or trajectory index");
    const fields::Ctx ctx{n, (int64_t)n * n, (hipStream_t)stream};
    const std::string fam(family);
    if (fam == "c") {
      ic::ParamStream ps{seed, (uint32_t)traj, 0xC0C0u};
      fields::c_field(ctx, type, L, ps, base, out);
    } else if (fam == "m") {
      ic::ParamStream ps{seed, (uint32_t)traj, 0x3A55u};
      fields::m_field(ctx, type, L, ps, base, c, out);
    } else {
      throw Error("nls_field_sample: family must be \"m\" or \"c\"");
    }
    NLS_HIP(hipStreamSynchronize(ctx.s));
  });
}

int nls_phenomenon_sample(const char *name, int n, double L, uint64_t seed, int traj,
                          int normalize, double *u, double *v, int *choices, int cap,
                          int *count, void *stream) {
  return guard([&] {
    NLS_REQUIRE(name && u, "nls_phenomenon_sample: null argument");
    NLS_REQUIRE(n >= 2 && n <= 8192, "nls_phenomenon_sample: n must lie in [2, 8192]");
    NLS_REQUIRE(L > 0.0 && traj >= 0, "nls_phenomenon_sample: bad domain or trajectory index");
    const fields::Ctx ctx{n, (int64_t)n * n, (hipStream_t)stream};
    const std::string ph(name);
    int ch[16] = {0};
    int nch;
    if (ph == "kink_field") {
      NLS_REQUIRE(v, "nls_phenomenon_sample: kink_field needs the v0 buffer");
      fields::kink_field(ctx, L, seed, (uint32_t)traj, u, v, ch);
      nch = fields::kKinkKeys;
    } else if (ph == "multi_soliton") {
      fields::multi_soliton(ctx, L, seed, (uint32_t)traj, normalize != 0, u, ch);
      nch = fields::kSolKeys + 1;
    } else {
      throw Error("nls_phenomenon_sample: unknown phenomenon '" + ph +
                  "' (kink_field, multi_soliton)");
    }
    if (choices) {
      NLS_REQUIRE(cap >= nch, "nls_phenomenon_sample: choice buffer too small");
      std::copy(ch, ch + nch, choices);
    }
    if (count) *count = nch;
  });
}

} // extern "C"
