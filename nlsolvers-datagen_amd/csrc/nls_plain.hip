// Plain path: the drop-in for the reference's per-trajectory GPU code.
//
// CSR SpMV (device/include/spmv.hpp) and the Lanczos matrix function at the
// reference's kernel granularity (device/include/lanczos.hpp:208-267,
// lanczos_complex.hpp:320-519, matfunc_real.hpp:190-313, matfunc_complex.hpp:254-375):
// one SpMV, two dot+axpy pairs, a two-GEMV classical Gram-Schmidt pass, a norm
// and a scale per iteration, all in FP64 / complex128.  Differences from the
// reference are mechanical only: scalars stay on the device (no host round trip
// per dot), reductions are deterministic two-stage sums, and workspaces are
// allocated once instead of per call.
#include "nls.h"
#include "nls_common.hpp"
#include "nls_philox.hpp"
#include "nls_plain.hpp"
#include "nls_small.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace nls {

// ------------------------------------------------------------------ scalar helpers
__host__ __device__ inline c128 operator+(c128 a, c128 b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline c128 operator-(c128 a, c128 b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline c128 operator*(c128 a, c128 b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline c128 operator*(double s, c128 a) { return {s * a.re, s * a.im}; }
__host__ __device__ inline c128 operator*(c128 a, double s) { return {a.re * s, a.im * s}; }
__host__ __device__ inline c128 operator-(c128 a) { return {-a.re, -a.im}; }
__host__ __device__ inline c128 conj(c128 a) { return {a.re, -a.im}; }
__host__ __device__ inline double conj(double a) { return a; }
__host__ __device__ inline double abs2(double a) { return a * a; }
__host__ __device__ inline double abs2(c128 a) { return a.re * a.re + a.im * a.im; }

template <class S> struct Acc;
template <> struct Acc<double> {
  double v = 0.0;
  __device__ void add_dotc(double x, double y) { v += x * y; }
  __device__ void sum(double *sh) { v = block_sum<256>(v, sh); }
  __device__ void store(double *p) const { p[0] = v; }
  __device__ void load_add(const double *p) { v += p[0]; }
  __device__ double get() const { return v; }
  static constexpr int W = 1;
};
template <> struct Acc<c128> {
  double r = 0.0, i = 0.0;
  __device__ void add_dotc(c128 x, c128 y) { // conj(x) * y
    r += x.re * y.re + x.im * y.im;
    i += x.re * y.im - x.im * y.re;
  }
  __device__ void sum(double *sh) {
    r = block_sum<256>(r, sh);
    i = block_sum<256>(i, sh);
  }
  __device__ void store(double *p) const { p[0] = r; p[1] = i; }
  __device__ void load_add(const double *p) { r += p[0]; i += p[1]; }
  __device__ c128 get() const { return {r, i}; }
  static constexpr int W = 2;
};

template <class S> __device__ inline S from_parts(const double *p);
template <> __device__ inline double from_parts<double>(const double *p) { return p[0]; }
template <> __device__ inline c128 from_parts<c128>(const double *p) { return {p[0], p[1]}; }
template <class S> __device__ inline S real_scalar(double v);
template <> __device__ inline double real_scalar<double>(double v) { return v; }
template <> __device__ inline c128 real_scalar<c128>(double v) { return {v, 0.0}; }

constexpr int BS = 256;

// ------------------------------------------------------------------ SpMV kernels
__device__ inline void madd(double &acc, double v, double x) { acc += v * x; }
__device__ inline void madd(float &acc, float v, float x) { acc += v * x; }
__device__ inline void madd(c128 &acc, double v, c128 x) {
  acc.re += v * x.re;
  acc.im += v * x.im;
}
__device__ inline void madd(c128 &acc, c128 v, c128 x) {
  acc.re += v.re * x.re - v.im * x.im;
  acc.im += v.re * x.im + v.im * x.re;
}
__device__ inline void madd(c64 &acc, c64 v, c64 x) {
  acc.re += v.re * x.re - v.im * x.im;
  acc.im += v.re * x.im + v.im * x.re;
}
template <class X> __device__ inline X zero_of();
template <> __device__ inline double zero_of<double>() { return 0.0; }
template <> __device__ inline float zero_of<float>() { return 0.0f; }
template <> __device__ inline c128 zero_of<c128>() { return {0.0, 0.0}; }
template <> __device__ inline c64 zero_of<c64>() { return {0.0f, 0.0f}; }

// One row per thread: the rows hold 3..7 entries, so a row's entries are one
// short contiguous run and neighbouring lanes read neighbouring runs.
// Row-per-thread CSR with the workgroup's entries staged through LDS: the column indices
// and values of the workgroup's BS rows are read once, contiguously and coalesced, then
// each thread sums its row from LDS in CSR order, acc += val[k] * x[ci[k]] for k = rp[r]
// .. rp[r+1]-1 (the order of a one-thread-per-row CSR kernel).  Workgroups whose rows hold
// more entries than the LDS tile read them directly.  For the 5- and 7-point operators (5-7 entries per row) this replaces strided
// per-thread reads of ci/val (each wave load touching ~7x the cache lines it uses).
template <class V> constexpr int spmv_tile() { return 24576 / (int)(sizeof(int32_t) + sizeof(V)); }

template <class V, class X>
__global__ __launch_bounds__(BS) void k_spmv_lds(int64_t n, const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ ci,
                                                 const V *__restrict__ val,
                                                 const X *__restrict__ x, X *__restrict__ y) {
  constexpr int TILE = spmv_tile<V>();
  __shared__ int32_t sci[TILE];
  __shared__ V sval[TILE];
  const int64_t r0 = (int64_t)blockIdx.x * BS;
  const int64_t r1 = r0 + BS < n ? r0 + BS : n;
  const int32_t b0 = rp[r0], b1 = rp[r1];
  const int64_t r = r0 + threadIdx.x;
  const int32_t b = r < n ? rp[r] : 0, e = r < n ? rp[r + 1] : 0;
  if (b1 - b0 <= TILE) { // uniform per workgroup
    for (int32_t k = threadIdx.x; k < b1 - b0; k += BS) {
      sci[k] = ci[b0 + k];
      sval[k] = val[b0 + k];
    }
    __syncthreads();
    if (r >= n) return;
    X acc = zero_of<X>();
    for (int32_t k = b - b0; k < e - b0; ++k) madd(acc, sval[k], x[sci[k]]);
    y[r] = acc;
  } else {
    if (r >= n) return;
    X acc = zero_of<X>();
    for (int32_t k = b; k < e; ++k) madd(acc, val[k], x[ci[k]]);
    y[r] = acc;
  }
}

template <class V, class X>
void launch_spmv(int64_t n, const int32_t *rp, const int32_t *ci, const V *val, const X *x, X *y,
                 hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_spmv_lds<V, X>), dim3(blocks_for(n, BS)), dim3(BS), 0, s, n, rp, ci, val,
                     x, y);
  check_launch("spmv");
}

// ------------------------------------------------------------------ BLAS-1 style kernels
template <class S>
__global__ __launch_bounds__(BS) void k_dot_s1(int64_t n, const S *__restrict__ x,
                                               const S *__restrict__ y, double *__restrict__ part) {
  __shared__ double sh[BS / kWave];
  Acc<S> a;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    a.add_dotc(x[i], y[i]);
  a.sum(sh);
  if (threadIdx.x == 0) a.store(part + (size_t)blockIdx.x * Acc<S>::W);
}

// Reduces G partials in fixed order; writes `value` to t1 (if non-null), -value to neg;
// real_only keeps the real part (the textbook complex Lanczos' alpha).
__device__ inline double real_part(double v) { return v; }
__device__ inline c128 real_part(c128 v) { return c128{v.re, 0.0}; }
template <class S>
__global__ __launch_bounds__(BS) void k_dot_s2(int G, const double *__restrict__ part, S *t1,
                                               S *neg, int real_only = 0) {
  __shared__ double sh[BS / kWave];
  Acc<S> a;
  for (int g = threadIdx.x; g < G; g += BS) a.load_add(part + (size_t)g * Acc<S>::W);
  a.sum(sh);
  if (threadIdx.x == 0) {
    const S v = real_only ? real_part(a.get()) : a.get();
    if (t1) *t1 = v;
    if (neg) *neg = -v;
  }
}

// Norm: v = sqrt(sum); writes v to t1 and t2 (T entries, may be null), 1/v (or
// `inv_if_zero` when v == 0) to inv, and v to vout.
template <class S>
__global__ __launch_bounds__(BS) void k_norm_s2(int G, const double *__restrict__ part, S *t1,
                                                S *t2, double *inv, double *vout,
                                                double inv_if_zero) {
  __shared__ double sh[BS / kWave];
  double a = 0.0;
  for (int g = threadIdx.x; g < G; g += BS) a += part[(size_t)g * Acc<S>::W];
  a = block_sum<BS>(a, sh);
  if (threadIdx.x == 0) {
    const double v = sqrt(a);
    if (t1) *t1 = real_scalar<S>(v);
    if (t2) *t2 = real_scalar<S>(v);
    if (inv) *inv = v > 0.0 ? 1.0 / v : inv_if_zero;
    if (vout) *vout = v;
  }
}

// dst = -src (one scalar on the device)
template <class S> __global__ void k_negate(const S *__restrict__ src, S *__restrict__ dst) {
  if (threadIdx.x == 0) *dst = -*src;
}

template <class S>
__global__ __launch_bounds__(BS) void k_axpy_dev(int64_t n, const S *__restrict__ a,
                                                 const S *__restrict__ x, S *__restrict__ y) {
  const S av = *a;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    y[i] = y[i] + av * x[i];
}

template <class S>
__global__ __launch_bounds__(BS) void k_scale_copy(int64_t n, const double *__restrict__ s,
                                                   const S *__restrict__ x, S *__restrict__ y) {
  const double sv = *s;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    y[i] = x[i] * sv;
}

// h_c = sum_i conj(V_c[i]) w[i], c < k: per-block partials part[blk*k + c].
template <class S>
__global__ __launch_bounds__(BS) void k_gemvT_s1(int64_t n, int k, const S *__restrict__ V,
                                                 const S *__restrict__ w,
                                                 double *__restrict__ part) {
  __shared__ double sh[BS / kWave];
  for (int c = 0; c < k; ++c) {
    Acc<S> a;
    const S *vc = V + (size_t)c * n;
    for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
      a.add_dotc(vc[i], w[i]);
    a.sum(sh);
    if (threadIdx.x == 0) a.store(part + ((size_t)blockIdx.x * k + c) * Acc<S>::W);
  }
}

template <class S>
__global__ __launch_bounds__(BS) void k_gemvT_s2(int G, int k, const double *__restrict__ part,
                                                 S *__restrict__ h) {
  __shared__ double sh[BS / kWave];
  for (int c = 0; c < k; ++c) {
    Acc<S> a;
    for (int g = threadIdx.x; g < G; g += BS) a.load_add(part + ((size_t)g * k + c) * Acc<S>::W);
    a.sum(sh);
    if (threadIdx.x == 0) h[c] = a.get();
  }
}

// w -= V[:, 0..k) h   (cublas gemv N with alpha = -1, beta = 1)
template <class S>
__global__ __launch_bounds__(BS) void k_gemvN(int64_t n, int k, const S *__restrict__ V,
                                              const S *__restrict__ h, S *__restrict__ w) {
  __shared__ S hs[kMaxKrylov];
  if (threadIdx.x < k) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    S acc = V[i] * hs[0];
    for (int c = 1; c < k; ++c) acc = acc + V[(size_t)c * n + i] * hs[c];
    w[i] = w[i] - acc;
  }
}

// out = V[:, 0..m) y
template <class S>
__global__ __launch_bounds__(BS) void k_combine(int64_t n, int m, const S *__restrict__ V,
                                                const S *__restrict__ y, S *__restrict__ out) {
  __shared__ S ys[kMaxKrylov];
  if (threadIdx.x < m) ys[threadIdx.x] = y[threadIdx.x];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    S acc = V[i] * ys[0];
    for (int c = 1; c < m; ++c) acc = acc + V[(size_t)c * n + i] * ys[c];
    out[i] = acc;
  }
}

// Small stage on the column-major Lanczos T (lower triangle read, as syevd/heevd
// with CUBLAS_FILL_MODE_LOWER).  One wavefront.
template <class S> __device__ inline double re_of(S v);
template <> __device__ inline double re_of<double>(double v) { return v; }
template <> __device__ inline double re_of<c128>(c128 v) { return v.re; }

template <class S>
__global__ __launch_bounds__(kWave) void k_small_plain(int m, const S *__restrict__ T, int kind,
                                                       double t_re, double t_im,
                                                       const double *__restrict__ beta0,
                                                       S *__restrict__ y, int *status) {
  __shared__ SmallWS ws;
  const int lane = threadIdx.x;
  if (lane == 0) {
    ws.phr[0] = 1.0;
    ws.phi[0] = 0.0;
    for (int i = 0; i < m; ++i) ws.d[i] = re_of<S>(T[(size_t)m * i + i]);
    for (int i = 0; i < m - 1; ++i) {
      const S t = T[(size_t)m * i + i + 1];
      double tr, ti;
      if constexpr (sizeof(S) == sizeof(double)) { tr = re_of<S>(t); ti = 0.0; }
      else { const c128 tc = *reinterpret_cast<const c128 *>(&t); tr = tc.re; ti = tc.im; }
      const double a = sqrt(tr * tr + ti * ti);
      ws.e[i] = a;
      if (a > 0.0) {
        const double ur = tr / a, ui = ti / a;
        ws.phr[i + 1] = ws.phr[i] * ur - ws.phi[i] * ui;
        ws.phi[i + 1] = ws.phr[i] * ui + ws.phi[i] * ur;
      } else {
        ws.phr[i + 1] = ws.phr[i];
        ws.phi[i + 1] = ws.phi[i];
      }
    }
  }
  __syncthreads();
  const int rc = small_ql(m, ws);
  if (rc) {
    if (lane == 0) *status = 1;
    return;
  }
  __shared__ double yr[kMaxKrylov], yi[kMaxKrylov];
  small_apply_f(m, ws, kind, t_re, t_im, *beta0, yr, yi);
  __syncthreads();
  if (lane < m) {
    if constexpr (sizeof(S) == sizeof(double)) reinterpret_cast<double *>(y)[lane] = yr[lane];
    else reinterpret_cast<c128 *>(y)[lane] = c128{yr[lane], yi[lane]};
  }
}

// Host-callable variant on a (d, e) pair.
__global__ __launch_bounds__(kWave) void k_small_de(int m, const double *__restrict__ d,
                                                    const double *__restrict__ e, int kind,
                                                    double t_re, double t_im, double beta0,
                                                    double *yr, double *yi, int *status) {
  __shared__ SmallWS ws;
  if (kind == F_EXP_SERIES) { // exp(t T) e1 without the eigen-decomposition
    const int lane = threadIdx.x;
    double r, i;
    small_exp_series_wave(m, lane < m ? d[lane] : 0.0, lane < m - 1 ? e[lane] : 0.0, t_re, t_im,
                          beta0, r, i);
    if (lane < m) {
      yr[lane] = r;
      yi[lane] = i;
    }
    return;
  }
  if (kind & F_KGE_SERIES) { // the fused batch's power series (T with no negative eigenvalue)
    const int lane = threadIdx.x;
    const double y = small_kge_series_wave(m, lane < m ? d[lane] : 0.0,
                                           lane < m - 1 ? e[lane] : 0.0, kind & ~F_KGE_SERIES,
                                           t_re, beta0);
    if (lane < m) {
      yr[lane] = y;
      yi[lane] = 0.0;
    }
    return;
  }
  if (kind & F_WAVE_QL) { // the fused batch's register-resident QL (real symmetric T)
    const int lane = threadIdx.x;
    double dk = lane < m ? d[lane] : 0.0, ek = lane < m - 1 ? e[lane] : 0.0;
    if (small_ql_wave(m, dk, ek, ws.Z)) {
      if (lane == 0) *status = 1;
      return;
    }
    double r, i;
    small_apply_f_wave(m, dk, ws.Z, kind & ~F_WAVE_QL, t_re, t_im, beta0, 1.0, 0.0, r, i);
    if (lane < m) {
      yr[lane] = r;
      yi[lane] = i;
    }
    return;
  }
  if (threadIdx.x == 0) {
    for (int i = 0; i < m; ++i) { ws.d[i] = d[i]; ws.phr[i] = 1.0; ws.phi[i] = 0.0; }
    for (int i = 0; i < m - 1; ++i) ws.e[i] = e[i];
  }
  __syncthreads();
  if (small_ql(m, ws)) {
    if (threadIdx.x == 0) *status = 1;
    return;
  }
  small_apply_f(m, ws, kind, t_re, t_im, beta0, yr, yi);
}

// ------------------------------------------------------------------ pointwise kernels
// Ghost ring <- interior: with the reference's copy order (rows, then columns;
// 3D: slow-axis faces, middle-axis faces, fast-axis faces; boundaries.cuh:9-81)
// every ghost cell ends up equal to the cell at its coordinates clamped to
// [1, n-2], so a single gather over the faces reproduces it.
template <class S>
__global__ __launch_bounds__(BS) void k_bc(S *__restrict__ u, int n, int dim) {
  const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
  auto cl = [n](int64_t a) { return a < 1 ? (int64_t)1 : (a > n - 2 ? (int64_t)n - 2 : a); };
  if (dim == 2) {
    if (t >= 4LL * n) return;
    const int side = (int)(t / n);
    const int64_t p = t % n;
    int64_t i, j;
    if (side == 0) { i = 0; j = p; }
    else if (side == 1) { i = n - 1; j = p; }
    else if (side == 2) { i = p; j = 0; }
    else { i = p; j = n - 1; }
    u[i * n + j] = u[cl(i) * n + cl(j)];
  } else {
    const int64_t n2 = (int64_t)n * n;
    if (t >= 6 * n2) return;
    const int face = (int)(t / n2);
    const int64_t p = t % n2, q0 = p / n, q1 = p % n;
    int64_t a, b, k;
    const int64_t fix = (face & 1) ? n - 1 : 0;
    if (face < 2) { a = fix; b = q0; k = q1; }
    else if (face < 4) { a = q0; b = fix; k = q1; }
    else { a = q0; b = q1; k = fix; }
    u[(a * n + b) * n + k] = u[(cl(a) * n + cl(b)) * n + cl(k)];
  }
}

// -m u^3 with the reference's operation order ((-m*u)*u)*u (kg_single.cuh:106-109).
__global__ __launch_bounds__(BS) void k_neg_mu3(int64_t n, const double *__restrict__ m,
                                                const double *__restrict__ u,
                                                double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) out[i] = -m[i] * u[i] * u[i] * u[i];
}

// Gautschi combine (kg_single.cuh:103-126): b2 = 2*cos-part, b3 = tau^2 * sinc-part,
// u+ = (b2 - u_past) + b3; u_past = old u; v = (u+ - u_past)/tau.
__global__ __launch_bounds__(BS) void k_gautschi_update(int64_t n, double *__restrict__ u,
                                                        double *__restrict__ up,
                                                        double *__restrict__ v,
                                                        const double *__restrict__ bcos,
                                                        const double *__restrict__ bsinc,
                                                        double tau) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double uo = u[i];
  const double b2 = bcos[i] * 2.0;
  const double b3 = bsinc[i] * (tau * tau);
  const double un = (b2 - up[i]) + b3;
  u[i] = un;
  up[i] = uo;
  v[i] = (un - uo) / tau;
}

// Stormer-Verlet combine (kg_single.cuh:145-178).
__global__ __launch_bounds__(BS) void k_sv_update(int64_t n, double *__restrict__ u,
                                                  double *__restrict__ up, double *__restrict__ v,
                                                  const double *__restrict__ Lu,
                                                  const double *__restrict__ m, double tau) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double uo = u[i];
  const double t2 = tau * tau;
  const double nl = -m[i] * uo * uo * uo;
  const double a = t2 * Lu[i], b = t2 * nl;
  double x = 2.0 * uo - up[i];
  x = x + a;
  x = x + b;
  u[i] = x;
  up[i] = uo;
  v[i] = (x - uo) / tau;
}

// Stochastic phi^4 (stochastic_phi4.cuh:27-76) with Philox normals for step `it`.
__global__ __launch_bounds__(BS) void k_sp4_update(int64_t n, double *__restrict__ u,
                                                   double *__restrict__ up,
                                                   double *__restrict__ v,
                                                   const double *__restrict__ Lu,
                                                   const double *__restrict__ m, double tau,
                                                   int mode, double s_eff, double namp,
                                                   double gamma, uint64_t seed, uint32_t stream,
                                                   uint32_t it) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double z = noise1(seed, stream, it, (uint64_t)i);
  const double uo = u[i];
  const double t2 = tau * tau;
  double x;
  if (mode == NLS_SP4_REFERENCE) {
    const double nl = -m[i] * (uo - uo * uo * uo + s_eff * z);
    x = 2.0 * uo - up[i];
    x = x + t2 * Lu[i];
    x = x + t2 * nl;
  } else {
    const double gp = 1.0 + 0.5 * gamma * tau, gm = 1.0 - 0.5 * gamma * tau;
    const double f = Lu[i] + m[i] * (uo - uo * uo * uo);
    x = (2.0 * uo - gm * up[i] + t2 * f + namp * z) / gp;
  }
  u[i] = x;
  up[i] = uo;
  v[i] = (x - uo) / tau;
}

// NLSE pointwise (nlse_dev.hpp:20-64).  phase: out = in * exp(i * a * m|in|^2).
__global__ __launch_bounds__(BS) void k_nl_phase(int64_t n, c128 *__restrict__ out,
                                                 const c128 *__restrict__ in,
                                                 const double *__restrict__ m, double a) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const c128 x = in[i];
  const double rho = m[i] * (x.re * x.re + x.im * x.im);
  const double ang = a * rho;
  const c128 e{cos(ang), sin(ang)};
  out[i] = x * e;
}

// B(psi) = -m |psi|^2 psi  (nlse_dev.hpp:43-52)
__global__ __launch_bounds__(BS) void k_nlse_B(int64_t n, c128 *__restrict__ out,
                                               const c128 *__restrict__ in,
                                               const double *__restrict__ m) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const c128 x = in[i];
  const double s = -m[i] * (x.re * x.re + x.im * x.im);
  out[i] = s * x;
}

// apply_sewi(out, exp_prev, exp_psi_B, tau): out = exp_prev - 2 tau exp_psi_B with
// tau = i dt (nlse_dev.hpp:54-64).
__global__ __launch_bounds__(BS) void k_sewi_combine(int64_t n, c128 *__restrict__ out,
                                                     const c128 *__restrict__ a,
                                                     const c128 *__restrict__ b, double dt) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const c128 two_tau{0.0, 2.0 * dt};
  out[i] = a[i] - two_tau * b[i];
}

__global__ __launch_bounds__(BS) void k_philox_fill(int64_t n, uint64_t seed, uint32_t stream,
                                                    uint32_t step, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) out[i] = noise1(seed, stream, step, (uint64_t)i);
}

// ------------------------------------------------------------------ Krylov workspace
static int reduce_grid(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(512, (n + 4 * BS - 1) / (4 * BS)));
}

// Zeroes `words` 4-byte words (a kernel rather than hipMemsetAsync, so that a caller may
// capture the calls into a HIP graph).
__global__ __launch_bounds__(BS) void k_zero_words(uint32_t *__restrict__ p, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < words; i += (int64_t)gridDim.x * BS)
    p[i] = 0u;
}

static void zero_words(void *p, size_t bytes, hipStream_t s) {
  const int64_t words = (int64_t)(bytes / 4);
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (words + BS - 1) / BS));
  hipLaunchKernelGGL(k_zero_words, dim3(g), dim3(BS), 0, s, static_cast<uint32_t *>(p), words);
}

// Byte offsets of the buffers in the workspace block (each 256-byte aligned).
namespace {
struct KrylovCarve {
  size_t V, w, T, part, h, y, scal, dscal, status, total;
  template <class S> static KrylovCarve of(int64_t n, int m, int G) {
    KrylovCarve c{};
    size_t o = 0;
    auto take = [&o](size_t bytes) {
      const size_t at = o;
      o += (bytes + 255) / 256 * 256;
      return at;
    };
    c.V = take(sizeof(S) * (size_t)n * m);
    c.w = take(sizeof(S) * (size_t)n);
    c.T = take(sizeof(S) * (size_t)m * m);
    c.part = take(sizeof(double) * 2 * (size_t)G * (m + 1));
    c.h = take(sizeof(S) * kMaxKrylov);
    c.y = take(sizeof(S) * kMaxKrylov);
    c.scal = take(sizeof(S) * 8);
    c.dscal = take(sizeof(double) * 8);
    c.status = take(sizeof(int));
    c.total = o;
    return c;
  }
};
} // namespace

template <class S> size_t PlainKrylov<S>::workspace_bytes(int64_t n, int m) {
  return KrylovCarve::of<S>(n, m, reduce_grid(n)).total;
}

template <class S>
PlainKrylov<S>::PlainKrylov(int64_t n_, int m_, hipStream_t s_, char *mem)
    : n(n_), m(m_), stream(s_) {
  NLS_REQUIRE(m >= 2 && m <= kMaxKrylov, "Krylov dimension must be in [2, 32]");
  G = reduce_grid(n);
  const KrylovCarve c = KrylovCarve::of<S>(n, m, G);
  if (mem) {
    block = mem;
  } else {
    NLS_HIP(hipMalloc(&block, c.total));
    owns = true;
  }
  V = reinterpret_cast<S *>(block + c.V);
  w = reinterpret_cast<S *>(block + c.w);
  T = reinterpret_cast<S *>(block + c.T);
  part = reinterpret_cast<double *>(block + c.part);
  h = reinterpret_cast<S *>(block + c.h);
  y = reinterpret_cast<S *>(block + c.y);
  scal = reinterpret_cast<S *>(block + c.scal);
  dscal = reinterpret_cast<double *>(block + c.dscal);
  status = reinterpret_cast<int *>(block + c.status);
  zero_words(status, sizeof(int), stream);
}

template <class S> PlainKrylov<S>::~PlainKrylov() {
  if (owns) (void)hipFree(block);
}

// Reference Lanczos (lanczos.hpp:208-267 / lanczos_complex.hpp:320-519).
template <class S> void PlainKrylov<S>::lanczos(const CsrView &A, const S *in) {
  const dim3 g(G), b(BS);
  hipStream_t s = stream;
  zero_words(T, sizeof(S) * m * m, s);
  double *beta0 = dscal + 0, *inv0 = dscal + 1, *invb = dscal + 2;
  S *negh = scal + 0, *nega = scal + 1;
  hipLaunchKernelGGL((k_dot_s1<S>), g, b, 0, s, n, in, in, part);
  // beta0 = ||in||; inv0 = 1/beta0, or 0 for a zero input (Q14 guard: result 0).
  hipLaunchKernelGGL((k_norm_s2<S>), dim3(1), b, 0, s, G, part, (S *)nullptr, (S *)nullptr, inv0,
                     beta0, 0.0);
  hipLaunchKernelGGL((k_scale_copy<S>), g, b, 0, s, n, inv0, in, V);
  for (int j = 0; j < m - 1; ++j) {
    S *vj = V + (size_t)j * n;
    launch_spmv(n, A.rp, A.ci, A.val, vj, w, s);
    if (j > 0) {
      S *vp = V + (size_t)(j - 1) * n;
      if (variant == 0) { // Q2: the lower subdiagonal recomputed as <A v_j, v_{j-1}>
        hipLaunchKernelGGL((k_dot_s1<S>), g, b, 0, s, n, w, vp, part);
        hipLaunchKernelGGL((k_dot_s2<S>), dim3(1), b, 0, s, G, part,
                           T + (size_t)m * (j - 1) + j, negh, 0);
      } else { // textbook: w -= b_j v_{j-1} with b_j from the previous norm
        hipLaunchKernelGGL((k_negate<S>), dim3(1), dim3(kWave), 0, s,
                           T + (size_t)m * (j - 1) + j, negh);
      }
      hipLaunchKernelGGL((k_axpy_dev<S>), g, b, 0, s, n, negh, vp, w);
    }
    hipLaunchKernelGGL((k_dot_s1<S>), g, b, 0, s, n, w, vj, part);
    hipLaunchKernelGGL((k_dot_s2<S>), dim3(1), b, 0, s, G, part, T + (size_t)m * j + j, nega,
                       variant);
    hipLaunchKernelGGL((k_axpy_dev<S>), g, b, 0, s, n, nega, vj, w);
    if (variant == 0) { // Q3: one classical Gram-Schmidt pass over v_0 .. v_j
      hipLaunchKernelGGL((k_gemvT_s1<S>), g, b, 0, s, n, j + 1, V, w, part);
      hipLaunchKernelGGL((k_gemvT_s2<S>), dim3(1), b, 0, s, G, j + 1, part, h);
      hipLaunchKernelGGL((k_gemvN<S>), g, b, 0, s, n, j + 1, V, h, w);
    }
    hipLaunchKernelGGL((k_dot_s1<S>), g, b, 0, s, n, w, w, part);
    hipLaunchKernelGGL((k_norm_s2<S>), dim3(1), b, 0, s, G, part, T + (size_t)m * j + j + 1,
                       T + (size_t)m * (j + 1) + j, invb, (double *)nullptr, 1.0);
    hipLaunchKernelGGL((k_scale_copy<S>), g, b, 0, s, n, invb, w, V + (size_t)(j + 1) * n);
  }
  if (variant == 1) { // alpha_{m-1} = <A v_{m-1}, v_{m-1}> (real part for complex vectors)
    S *vl = V + (size_t)(m - 1) * n;
    launch_spmv(n, A.rp, A.ci, A.val, vl, w, s);
    hipLaunchKernelGGL((k_dot_s1<S>), g, b, 0, s, n, w, vl, part);
    hipLaunchKernelGGL((k_dot_s2<S>), dim3(1), b, 0, s, G, part,
                       T + (size_t)m * (m - 1) + m - 1, (S *)nullptr, 1);
  }
  check_launch("lanczos");
}

template <class S>
void PlainKrylov<S>::apply(const CsrView &A, S *out, const S *in, double t_re, double t_im,
                           int kind) {
  lanczos(A, in);
  hipLaunchKernelGGL((k_small_plain<S>), dim3(1), dim3(kWave), 0, stream, m, T, kind, t_re, t_im,
                     dscal + 0, y, status);
  hipLaunchKernelGGL((k_combine<S>), dim3(G), dim3(BS), 0, stream, n, m, V, y, out);
  check_launch("matfunc");
}

template <class S> void PlainKrylov<S>::check_status() {
  int st = 0;
  NLS_HIP(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, stream));
  NLS_HIP(hipStreamSynchronize(stream));
  NLS_REQUIRE(st == 0, "tridiagonal eigensolver did not converge");
}

template struct PlainKrylov<double>;
template struct PlainKrylov<c128>;

// ------------------------------------------------------------------ CSR upload
CsrDev::CsrDev(int64_t n_, int64_t nnz_, const int32_t *rp_h, const int32_t *ci_h,
               const double *val_h)
    : n(n_), nnz(nnz_) {
  NLS_HIP(hipMalloc(&rp, sizeof(int32_t) * (n + 1)));
  NLS_HIP(hipMalloc(&ci, sizeof(int32_t) * nnz));
  NLS_HIP(hipMalloc(&val, sizeof(double) * nnz));
  NLS_HIP(hipMemcpy(rp, rp_h, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
  NLS_HIP(hipMemcpy(ci, ci_h, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  NLS_HIP(hipMemcpy(val, val_h, sizeof(double) * nnz, hipMemcpyHostToDevice));
}
CsrDev::~CsrDev() {
  (void)hipFree(rp); (void)hipFree(ci); (void)hipFree(val);
}

std::unique_ptr<CsrDev> build_operator(int dim, int n, double L, const double *c, int negate,
                                       double shift, int q6) {
  const double dx = 2.0 * L / (n - 1);
  const int64_t N = dim == 3 ? (int64_t)n * n * n : (int64_t)n * n;
  int64_t nnz = nls_laplacian_nnz(dim, n, q6);
  std::vector<int32_t> rp(N + 1), ci(nnz);
  std::vector<double> val(nnz);
  int rc = dim == 3 ? nls_build_laplacian_3d(n, dx, c, negate, shift, q6, rp.data(), ci.data(),
                                             val.data(), &nnz)
                    : nls_build_laplacian_2d(n, dx, dx, c, negate, shift, rp.data(), ci.data(),
                                             val.data(), &nnz);
  NLS_REQUIRE(rc == 0, std::string("operator assembly failed: ") + nls_last_error());
  return std::make_unique<CsrDev>(N, nnz, rp.data(), ci.data(), val.data());
}

void launch_bc_f64(double *u, int n, int dim, hipStream_t s) {
  const int64_t cnt = dim == 3 ? 6LL * n * n : 4LL * n;
  hipLaunchKernelGGL((k_bc<double>), dim3(blocks_for(cnt, BS)), dim3(BS), 0, s, u, n, dim);
  check_launch("bc");
}
void launch_bc_c128(c128 *u, int n, int dim, hipStream_t s) {
  const int64_t cnt = dim == 3 ? 6LL * n * n : 4LL * n;
  hipLaunchKernelGGL((k_bc<c128>), dim3(blocks_for(cnt, BS)), dim3(BS), 0, s, u, n, dim);
  check_launch("bc");
}

// ------------------------------------------------------------------ solvers
static int64_t cells(int dim, int n) { return dim == 3 ? (int64_t)n * n * n : (int64_t)n * n; }

struct DevBuf {
  void *p = nullptr;
  explicit DevBuf(size_t bytes) { NLS_HIP(hipMalloc(&p, bytes)); }
  ~DevBuf() { (void)hipFree(p); }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

} // namespace nls

using namespace nls;

// KGESolverDevice (device/include/kg_dev.hpp:19-157).
struct nls_kge {
  nls_kge_params p;
  int64_t N;
  double dt;
  int freq;
  hipStream_t s = nullptr;
  std::unique_ptr<CsrDev> A;
  std::unique_ptr<DevBuf> u, v, up, b2, b3, m, snapU, snapV;
  std::unique_ptr<PlainKrylov<double>> kry;
  int current_snapshot = 0;

  nls_kge(const nls_kge_params &p_, std::unique_ptr<CsrDev> op, const double *m_h,
          const double *u0, const double *v0)
      : p(p_), A(std::move(op)) {
    N = cells(p.dim, p.n);
    NLS_REQUIRE(A->n == N, "operator size does not match the grid");
    dt = p.T / p.nt;
    freq = p.nt / p.num_snapshots;
    NLS_HIP(hipStreamCreate(&s));
    const size_t B = sizeof(double) * N;
    u = std::make_unique<DevBuf>(B); v = std::make_unique<DevBuf>(B);
    up = std::make_unique<DevBuf>(B); b2 = std::make_unique<DevBuf>(B);
    b3 = std::make_unique<DevBuf>(B); m = std::make_unique<DevBuf>(B);
    snapU = std::make_unique<DevBuf>(B * p.num_snapshots);
    snapV = std::make_unique<DevBuf>(B * p.num_snapshots);
    std::vector<double> upast(N);
    for (int64_t i = 0; i < N; ++i) upast[i] = u0[i] - dt * v0[i]; // kg_dev.hpp:55-58
    NLS_HIP(hipMemcpy(u->p, u0, B, hipMemcpyHostToDevice));
    NLS_HIP(hipMemcpy(v->p, v0, B, hipMemcpyHostToDevice));
    NLS_HIP(hipMemcpy(up->p, upast.data(), B, hipMemcpyHostToDevice));
    NLS_HIP(hipMemcpy(m->p, m_h, B, hipMemcpyHostToDevice));
    NLS_REQUIRE(p.variant == 0 || p.variant == 1, "variant must be 0 or 1");
    kry = std::make_unique<PlainKrylov<double>>(N, p.krylov_dim, s);
    kry->variant = p.variant;
    store_snapshot(0);
  }
  ~nls_kge() {
    if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  }
  void step() { // kg_single.cuh:94-126
    const unsigned g = blocks_for(N, BS);
    kry->apply(A->view(), b2->as<double>(), u->as<double>(), dt, 0.0, F_COS_SQRT);
    hipLaunchKernelGGL(k_neg_mu3, dim3(g), dim3(BS), 0, s, N, m->as<double>(), u->as<double>(),
                       b3->as<double>());
    kry->apply(A->view(), b3->as<double>(), b3->as<double>(), dt, 0.0,
               p.method == NLS_KGE_SEWI ? F_SINC2_HALF : F_SINC2_SQRT);
    hipLaunchKernelGGL(k_gautschi_update, dim3(g), dim3(BS), 0, s, N, u->as<double>(),
                       up->as<double>(), v->as<double>(), b2->as<double>(), b3->as<double>(), dt);
    check_launch("kge step");
  }
  void step_sv() { // kg_single.cuh:145-178
    const unsigned g = blocks_for(N, BS);
    launch_spmv(N, A->rp, A->ci, A->val, u->as<double>(), b2->as<double>(), s);
    hipLaunchKernelGGL(k_sv_update, dim3(g), dim3(BS), 0, s, N, u->as<double>(), up->as<double>(),
                       v->as<double>(), b2->as<double>(), m->as<double>(), dt);
    check_launch("kge step_sv");
  }
  void apply_bc() { launch_bc_f64(u->as<double>(), p.n, p.dim, s); }
  void store_snapshot(int idx) { // kg_dev.hpp:123-131
    if (idx < 0 || idx >= p.num_snapshots) return;
    NLS_HIP(hipMemcpyAsync(snapU->as<double>() + (size_t)idx * N, u->p, sizeof(double) * N,
                           hipMemcpyDeviceToDevice, s));
    NLS_HIP(hipMemcpyAsync(snapV->as<double>() + (size_t)idx * N, v->p, sizeof(double) * N,
                           hipMemcpyDeviceToDevice, s));
    current_snapshot = idx + 1;
  }
  void run() { // kg_driver_dev_2d.cpp:144-153
    for (int i = 1; i < p.nt; ++i) {
      if (p.method == NLS_KGE_SV) step_sv(); else step();
      apply_bc();
      if (freq > 0 && i % freq == 0 && i / freq < p.num_snapshots) store_snapshot(i / freq);
    }
    kry->check_status();
  }
};

// NLSESolverDevice (device/include/nlse_dev.hpp:66-362).
struct nls_nlse {
  nls_nlse_params p;
  int64_t N;
  double dt;
  int freq;
  hipStream_t s = nullptr;
  std::unique_ptr<CsrDev> A;
  std::unique_ptr<DevBuf> u, buf, b2, b3, uprev, m, snap;
  std::unique_ptr<PlainKrylov<c128>> kry;
  int current_snapshot = 0;

  nls_nlse(const nls_nlse_params &p_, std::unique_ptr<CsrDev> op, const double *m_h,
           const double *u0)
      : p(p_), A(std::move(op)) {
    N = cells(p.dim, p.n);
    dt = p.T / p.nt;
    freq = p.nt / p.num_snapshots;
    NLS_HIP(hipStreamCreate(&s));
    const size_t B = sizeof(c128) * N;
    u = std::make_unique<DevBuf>(B); buf = std::make_unique<DevBuf>(B);
    b2 = std::make_unique<DevBuf>(B); b3 = std::make_unique<DevBuf>(B);
    uprev = std::make_unique<DevBuf>(B);
    m = std::make_unique<DevBuf>(sizeof(double) * N);
    snap = std::make_unique<DevBuf>(B * p.num_snapshots);
    NLS_HIP(hipMemcpy(u->p, u0, B, hipMemcpyHostToDevice));
    NLS_HIP(hipMemcpy(m->p, m_h, sizeof(double) * N, hipMemcpyHostToDevice));
    NLS_REQUIRE(p.variant == 0 || p.variant == 1, "variant must be 0 or 1");
    kry = std::make_unique<PlainKrylov<c128>>(N, p.krylov_dim, s);
    kry->variant = p.variant;
    store_snapshot_online(); // nlse_cubic_driver_2d.cpp:106
  }
  ~nls_nlse() {
    if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  }
  void store_snapshot_online() { // nlse_dev.hpp:324-335
    if (current_snapshot >= p.num_snapshots) return;
    NLS_HIP(hipMemcpyAsync(snap->as<c128>() + (size_t)current_snapshot * N, u->p,
                           sizeof(c128) * N, hipMemcpyDeviceToDevice, s));
    current_snapshot++;
  }
  void strang(int step_number) { // nlse_dev.hpp:187-197
    const unsigned g = blocks_for(N, BS);
    hipLaunchKernelGGL(k_nl_phase, dim3(g), dim3(BS), 0, s, N, buf->as<c128>(), u->as<c128>(),
                       m->as<double>(), 0.5 * dt);
    kry->apply(A->view(), u->as<c128>(), buf->as<c128>(), 0.0, dt, F_EXP);
    hipLaunchKernelGGL(k_nl_phase, dim3(g), dim3(BS), 0, s, N, u->as<c128>(), u->as<c128>(),
                       m->as<double>(), 0.5 * dt);
    check_launch("nlse step");
    if (step_number % freq == 0) store_snapshot_online();
  }
  void step_sewi(int step_number) { // nlse_dev.hpp:205-239
    if (step_number == 1) {
      NLS_HIP(hipMemcpyAsync(uprev->p, u->p, sizeof(c128) * N, hipMemcpyDeviceToDevice, s));
      strang(1);
      return;
    }
    const unsigned g = blocks_for(N, BS);
    NLS_HIP(hipMemcpyAsync(buf->p, u->p, sizeof(c128) * N, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_nlse_B, dim3(g), dim3(BS), 0, s, N, b2->as<c128>(), u->as<c128>(),
                       m->as<double>());
    kry->apply(A->view(), b3->as<c128>(), b2->as<c128>(), dt, 0.0, F_SINC);
    kry->apply(A->view(), b2->as<c128>(), b3->as<c128>(), 0.0, dt, F_EXP);
    kry->apply(A->view(), b3->as<c128>(), uprev->as<c128>(), 0.0, 2.0 * dt, F_EXP);
    // Q20: the reference passes (e^{tau L} sinc B, e^{2 tau L} psi-) to apply_sewi in swapped
    // order (nlse_dev.hpp:54-64, 219-225); variant 1 forms the intended
    // psi+ = e^{2 tau L} psi- - 2 tau e^{tau L} sinc(tau L) B(psi)
    if (p.variant == 0)
      hipLaunchKernelGGL(k_sewi_combine, dim3(g), dim3(BS), 0, s, N, u->as<c128>(),
                         b2->as<c128>(), b3->as<c128>(), dt);
    else
      hipLaunchKernelGGL(k_sewi_combine, dim3(g), dim3(BS), 0, s, N, u->as<c128>(),
                         b3->as<c128>(), b2->as<c128>(), dt);
    NLS_HIP(hipMemcpyAsync(uprev->p, buf->p, sizeof(c128) * N, hipMemcpyDeviceToDevice, s));
    check_launch("nlse sewi");
    if (step_number > 1 && step_number % freq == 0) store_snapshot_online();
  }
  void apply_bc() { launch_bc_c128(u->as<c128>(), p.n, p.dim, s); }
  void run() {
    for (int i = 1; i < p.nt; ++i) {
      if (p.method == NLS_NLSE_SEWI) step_sewi(i); else strang(i);
      apply_bc();
    }
    kry->check_status();
  }
};

// SP4SolverDevice (device/include/stochastic_phi4_dev.hpp:17-160).
struct nls_sp4 {
  nls_sp4_params p;
  int64_t N;
  double dt, dx;
  int freq;
  int it = 0;
  hipStream_t s = nullptr;
  std::unique_ptr<CsrDev> A;
  std::unique_ptr<DevBuf> u, v, up, Lu, m, snapU, snapV;

  nls_sp4(const nls_sp4_params &p_, std::unique_ptr<CsrDev> op, const double *m_h,
          const double *u0, const double *v0)
      : p(p_), A(std::move(op)) {
    N = cells(p.dim, p.n);
    dt = p.T / p.nt;
    dx = 2.0 * p.L / (p.n - 1);
    freq = p.nt / p.num_snapshots;
    NLS_HIP(hipStreamCreate(&s));
    const size_t B = sizeof(double) * N;
    u = std::make_unique<DevBuf>(B); v = std::make_unique<DevBuf>(B);
    up = std::make_unique<DevBuf>(B); Lu = std::make_unique<DevBuf>(B);
    m = std::make_unique<DevBuf>(B);
    snapU = std::make_unique<DevBuf>(B * p.num_snapshots);
    snapV = std::make_unique<DevBuf>(B * p.num_snapshots);
    std::vector<double> upast(N);
    for (int64_t i = 0; i < N; ++i) upast[i] = u0[i] - dt * v0[i];
    NLS_HIP(hipMemcpy(u->p, u0, B, hipMemcpyHostToDevice));
    NLS_HIP(hipMemcpy(v->p, v0, B, hipMemcpyHostToDevice));
    NLS_HIP(hipMemcpy(up->p, upast.data(), B, hipMemcpyHostToDevice));
    NLS_HIP(hipMemcpy(m->p, m_h, B, hipMemcpyHostToDevice));
    store_snapshot(0);
  }
  ~nls_sp4() {
    if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  }
  void step() {
    ++it;
    const unsigned g = blocks_for(N, BS);
    launch_spmv(N, A->rp, A->ci, A->val, u->as<double>(), Lu->as<double>(), s);
    const double s_eff = 0.01 * std::sqrt(2.0 * dx) / dt; // stochastic_phi4_dev.hpp:47-51
    const double dxd = p.dim == 3 ? dx * std::sqrt(dx) : dx;
    const double namp = p.sigma * dt * std::sqrt(dt) / dxd;
    hipLaunchKernelGGL(k_sp4_update, dim3(g), dim3(BS), 0, s, N, u->as<double>(),
                       up->as<double>(), v->as<double>(), Lu->as<double>(), m->as<double>(), dt,
                       p.mode, s_eff, namp, p.gamma, p.seed, p.stream_id, (uint32_t)it);
    check_launch("sp4 step");
  }
  void apply_bc() { launch_bc_f64(u->as<double>(), p.n, p.dim, s); }
  void store_snapshot(int idx) {
    if (idx < 0 || idx >= p.num_snapshots) return;
    NLS_HIP(hipMemcpyAsync(snapU->as<double>() + (size_t)idx * N, u->p, sizeof(double) * N,
                           hipMemcpyDeviceToDevice, s));
    NLS_HIP(hipMemcpyAsync(snapV->as<double>() + (size_t)idx * N, v->p, sizeof(double) * N,
                           hipMemcpyDeviceToDevice, s));
  }
  void run() {
    for (int i = 1; i < p.nt; ++i) {
      step();
      apply_bc();
      if (freq > 0 && i % freq == 0 && i / freq < p.num_snapshots) store_snapshot(i / freq);
    }
  }
};

struct nls_matfunc {
  std::unique_ptr<CsrDev> A;
  hipStream_t s = nullptr;
  std::unique_ptr<PlainKrylov<double>> kr;
  std::unique_ptr<PlainKrylov<c128>> kc;
  ~nls_matfunc() {
    if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  }
};

// ------------------------------------------------------------------ fixed scratch
// Krylov workspace of nls_matfunc_apply_real_noalloc: a file-scope device array, so that
// entry point allocates nothing.  Every GPU the library runs on holds its own copy (at
// its own address) until the process ends.  On the host `scratch` names a stand-in, not
// device memory: only the per-device address from hipGetSymbolAddress is used.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

static char *scratch_for_current_device() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static char *addr[kMaxDevices] = {};
  int dev = 0;
  NLS_HIP(hipGetDevice(&dev));
  NLS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    NLS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NLS_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    addr[dev] = static_cast<char *>(p);
  }
  return addr[dev];
}

// ------------------------------------------------------------------ C ABI
static void check_params_common(int dim, int n, double L, double T, int nt, int ns) {
  NLS_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  NLS_REQUIRE(n >= 3, "grid side must be >= 3");
  NLS_REQUIRE(L > 0 && T > 0, "L and T must be positive");
  NLS_REQUIRE(nt >= 1 && ns >= 1, "nt and num_snapshots must be >= 1");
  NLS_REQUIRE(ns <= nt, "num_snapshots > nt gives a zero snapshot interval "
                        "(the reference then takes i % 0, kg_driver_dev_2d.cpp:55,147)");
}

extern "C" {

int nls_csr_spmv_f64(int64_t n, const int32_t *rp, const int32_t *ci, const double *val,
                     const double *x, double *y, void *stream) {
  return guard([&] { launch_spmv(n, rp, ci, val, x, y, (hipStream_t)stream); });
}
int nls_csr_spmv_f32(int64_t n, const int32_t *rp, const int32_t *ci, const float *val,
                     const float *x, float *y, void *stream) {
  return guard([&] { launch_spmv(n, rp, ci, val, x, y, (hipStream_t)stream); });
}
int nls_csr_spmv_c128(int64_t n, const int32_t *rp, const int32_t *ci, const double *val,
                      const double *x, double *y, void *stream) {
  return guard([&] {
    launch_spmv(n, rp, ci, reinterpret_cast<const c128 *>(val), reinterpret_cast<const c128 *>(x),
                reinterpret_cast<c128 *>(y), (hipStream_t)stream);
  });
}
int nls_csr_spmv_c64(int64_t n, const int32_t *rp, const int32_t *ci, const float *val,
                     const float *x, float *y, void *stream) {
  return guard([&] {
    launch_spmv(n, rp, ci, reinterpret_cast<const c64 *>(val), reinterpret_cast<const c64 *>(x),
                reinterpret_cast<c64 *>(y), (hipStream_t)stream);
  });
}

int nls_matfunc_create(int64_t n, int64_t nnz, const int32_t *rp, const int32_t *ci,
                       const double *val, int m, int complex_vectors, nls_matfunc **out) {
  return guard([&] {
    NLS_REQUIRE(out && rp && ci && val, "nls_matfunc_create: null argument");
    auto h = std::make_unique<nls_matfunc>();
    h->A = std::make_unique<CsrDev>(n, nnz, rp, ci, val);
    NLS_HIP(hipStreamCreate(&h->s));
    if (complex_vectors) h->kc = std::make_unique<PlainKrylov<c128>>(n, m, h->s);
    else h->kr = std::make_unique<PlainKrylov<double>>(n, m, h->s);
    *out = h.release();
  });
}
int nls_matfunc_destroy(nls_matfunc *h) {
  return guard([&] { delete h; });
}
int nls_matfunc_apply_real(nls_matfunc *h, double *out, const double *in, double t, int kind) {
  return guard([&] {
    NLS_REQUIRE(h && h->kr, "nls_matfunc_apply_real: handle has no real workspace");
    NLS_REQUIRE(kind >= 0 && kind <= 4, "unknown real function kind");
    h->kr->apply(h->A->view(), out, in, t, 0.0, kind);
    h->kr->check_status();
  });
}
int nls_matfunc_apply_real_noalloc(int64_t n, const int32_t *row_ptr, const int32_t *col,
                                   const double *val, double *out, const double *in, double t,
                                   int kind, int m, void *stream) {
  return guard([&] {
    NLS_REQUIRE(row_ptr && col && val && out && in, "nls_matfunc_apply_real_noalloc: null argument");
    NLS_REQUIRE(n >= 1, "nls_matfunc_apply_real_noalloc: n must be >= 1");
    NLS_REQUIRE(m >= 2 && m <= kMaxKrylov, "nls_matfunc_apply_real_noalloc: m must be in [2, 32]");
    NLS_REQUIRE(kind >= 0 && kind <= 4, "unknown real function kind");
    const size_t need = PlainKrylov<double>::workspace_bytes(n, m);
    NLS_REQUIRE(need <= kScratchBytes,
                "nls_matfunc_apply_real_noalloc: n = " + std::to_string(n) + ", m = " +
                    std::to_string(m) + " needs " + std::to_string(need) +
                    " bytes of Krylov workspace, more than the fixed scratch of " +
                    std::to_string(kScratchBytes) + " bytes (use nls_matfunc_create)");
    char *ws = scratch_for_current_device();
    hipStream_t s = static_cast<hipStream_t>(stream);
    PlainKrylov<double> kr(n, m, s, ws); // a view of the scratch: allocates nothing
    kr.apply(CsrView{n, row_ptr, col, val}, out, in, t, 0.0, kind);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    NLS_HIP(hipStreamIsCapturing(s, &cap));
    if (cap == hipStreamCaptureStatusNone) kr.check_status(); // synchronises the stream
  });
}
int nls_matfunc_apply_complex(nls_matfunc *h, double *out, const double *in, double dt_re,
                              double dt_im, int kind) {
  return guard([&] {
    NLS_REQUIRE(h && h->kc, "nls_matfunc_apply_complex: handle has no complex workspace");
    NLS_REQUIRE(kind == F_EXP || kind == F_SINC, "unknown complex function kind");
    h->kc->apply(h->A->view(), reinterpret_cast<c128 *>(out), reinterpret_cast<const c128 *>(in), dt_re,
                 dt_im, kind);
    h->kc->check_status();
  });
}
int nls_matfunc_last_krylov(nls_matfunc *h, double *T_host, double *beta0) {
  return guard([&] {
    NLS_REQUIRE(h, "null handle");
    hipStream_t s = h->s;
    if (h->kr) {
      NLS_HIP(hipMemcpyAsync(T_host, h->kr->T, sizeof(double) * h->kr->m * h->kr->m,
                             hipMemcpyDeviceToHost, s));
      NLS_HIP(hipMemcpyAsync(beta0, h->kr->dscal, sizeof(double), hipMemcpyDeviceToHost, s));
    } else {
      NLS_HIP(hipMemcpyAsync(T_host, h->kc->T, sizeof(c128) * h->kc->m * h->kc->m,
                             hipMemcpyDeviceToHost, s));
      NLS_HIP(hipMemcpyAsync(beta0, h->kc->dscal, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    NLS_HIP(hipStreamSynchronize(s));
  });
}

int nls_small_matfunc(int m, const double *d, const double *e, int kind, double t_re,
                      double t_im, double beta0, double *yr, double *yi) {
  return guard([&] {
    NLS_REQUIRE(m >= 1 && m <= kMaxKrylov, "m must be in [1, 32]");
    DevBuf dd(sizeof(double) * kMaxKrylov), de(sizeof(double) * kMaxKrylov);
    DevBuf dyr(sizeof(double) * kMaxKrylov), dyi(sizeof(double) * kMaxKrylov), st(sizeof(int));
    NLS_HIP(hipMemcpy(dd.p, d, sizeof(double) * m, hipMemcpyHostToDevice));
    if (m > 1) NLS_HIP(hipMemcpy(de.p, e, sizeof(double) * (m - 1), hipMemcpyHostToDevice));
    NLS_HIP(hipMemset(st.p, 0, sizeof(int)));
    hipLaunchKernelGGL(k_small_de, dim3(1), dim3(kWave), 0, 0, m, dd.as<double>(),
                       de.as<double>(), kind, t_re, t_im, beta0, dyr.as<double>(),
                       dyi.as<double>(), st.as<int>());
    check_launch("small matfunc");
    int status = 0;
    NLS_HIP(hipMemcpy(&status, st.p, sizeof(int), hipMemcpyDeviceToHost));
    NLS_REQUIRE(status == 0, "tridiagonal eigensolver did not converge");
    NLS_HIP(hipMemcpy(yr, dyr.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (yi) NLS_HIP(hipMemcpy(yi, dyi.p, sizeof(double) * m, hipMemcpyDeviceToHost));
  });
}

int nls_kge_create(const nls_kge_params *p, const double *m, const double *c, const double *u0,
                   const double *v0, nls_kge **out) {
  return guard([&] {
    NLS_REQUIRE(p && m && c && u0 && v0 && out, "nls_kge_create: null argument");
    check_params_common(p->dim, p->n, p->L, p->T, p->nt, p->num_snapshots);
    const bool sv = p->method == NLS_KGE_SV;
    // Gautschi/sEWI drivers negate the operator; SV keeps L (kg_sv_driver_dev_2d.cpp:103).
    auto A = build_operator(p->dim, p->n, p->L, c, sv ? 0 : 1, sv ? -p->mu2 : p->mu2, p->q6);
    *out = new nls_kge(*p, std::move(A), m, u0, v0);
  });
}
int nls_kge_create_csr(const nls_kge_params *p, int64_t nnz, const int32_t *rp, const int32_t *ci,
                       const double *val, const double *m, const double *u0, const double *v0,
                       nls_kge **out) {
  return guard([&] {
    NLS_REQUIRE(p && rp && ci && val && m && u0 && v0 && out, "nls_kge_create_csr: null argument");
    check_params_common(p->dim, p->n, p->L, p->T, p->nt, p->num_snapshots);
    auto A = std::make_unique<CsrDev>(cells(p->dim, p->n), nnz, rp, ci, val);
    *out = new nls_kge(*p, std::move(A), m, u0, v0);
  });
}
int nls_kge_destroy(nls_kge *h) { return guard([&] { delete h; }); }
int nls_kge_step(nls_kge *h) { return guard([&] { h->step(); }); }
int nls_kge_step_sv(nls_kge *h) { return guard([&] { h->step_sv(); }); }
int nls_kge_apply_bc(nls_kge *h) { return guard([&] { h->apply_bc(); }); }
int nls_kge_store_snapshot(nls_kge *h, int idx) { return guard([&] { h->store_snapshot(idx); }); }
int nls_kge_transfer_snapshots(nls_kge *h, double *u, double *v) {
  return guard([&] {
    const size_t B = sizeof(double) * h->N * h->p.num_snapshots;
    NLS_HIP(hipMemcpyAsync(u, h->snapU->p, B, hipMemcpyDeviceToHost, h->s));
    if (v) NLS_HIP(hipMemcpyAsync(v, h->snapV->p, B, hipMemcpyDeviceToHost, h->s));
    NLS_HIP(hipStreamSynchronize(h->s));
  });
}
int nls_kge_run(nls_kge *h) { return guard([&] { h->run(); NLS_HIP(hipStreamSynchronize(h->s)); }); }
int nls_kge_get_state(nls_kge *h, double *u, double *v) {
  return guard([&] {
    NLS_HIP(hipMemcpyAsync(u, h->u->p, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->s));
    if (v) NLS_HIP(hipMemcpyAsync(v, h->v->p, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->s));
    NLS_HIP(hipStreamSynchronize(h->s));
    h->kry->check_status();
  });
}

int nls_nlse_create(const nls_nlse_params *p, const double *m, const double *c, const double *u0,
                    nls_nlse **out) {
  return guard([&] {
    NLS_REQUIRE(p && m && c && u0 && out, "nls_nlse_create: null argument");
    check_params_common(p->dim, p->n, p->L, p->T, p->nt, p->num_snapshots);
    auto A = build_operator(p->dim, p->n, p->L, c, 0, 0.0, p->q6);
    *out = new nls_nlse(*p, std::move(A), m, u0);
  });
}
int nls_nlse_destroy(nls_nlse *h) { return guard([&] { delete h; }); }
int nls_nlse_step(nls_nlse *h, int k) { return guard([&] { h->strang(k); }); }
int nls_nlse_step_sewi(nls_nlse *h, int k) { return guard([&] { h->step_sewi(k); }); }
int nls_nlse_apply_bc(nls_nlse *h) { return guard([&] { h->apply_bc(); }); }
int nls_nlse_run(nls_nlse *h) { return guard([&] { h->run(); NLS_HIP(hipStreamSynchronize(h->s)); }); }
int nls_nlse_transfer_snapshots(nls_nlse *h, double *u) {
  return guard([&] {
    NLS_HIP(hipMemcpyAsync(u, h->snap->p, sizeof(c128) * h->N * h->p.num_snapshots,
                           hipMemcpyDeviceToHost, h->s));
    NLS_HIP(hipStreamSynchronize(h->s));
  });
}
int nls_nlse_store_snapshot_online(nls_nlse *h, double *host_dst) {
  return guard([&] {
    NLS_REQUIRE(h, "nls_nlse_store_snapshot_online: null handle");
    const int slot = h->current_snapshot;
    if (slot >= h->p.num_snapshots) return; // nlse_dev.hpp:325-327
    h->store_snapshot_online();             // the library's own snapshot array
    if (host_dst) {
      NLS_HIP(hipMemcpyAsync(host_dst + (size_t)2 * slot * h->N, h->u->p, sizeof(c128) * h->N,
                             hipMemcpyDeviceToHost, h->s));
      NLS_HIP(hipStreamSynchronize(h->s));
    }
  });
}
int nls_nlse_get_state(nls_nlse *h, double *u) {
  return guard([&] {
    NLS_HIP(hipMemcpyAsync(u, h->u->p, sizeof(c128) * h->N, hipMemcpyDeviceToHost, h->s));
    NLS_HIP(hipStreamSynchronize(h->s));
    h->kry->check_status();
  });
}

int nls_sp4_create(const nls_sp4_params *p, const double *m, const double *c, const double *u0,
                   const double *v0, nls_sp4 **out) {
  return guard([&] {
    NLS_REQUIRE(p && m && c && u0 && v0 && out, "nls_sp4_create: null argument");
    check_params_common(p->dim, p->n, p->L, p->T, p->nt, p->num_snapshots);
    NLS_REQUIRE(p->mode == NLS_SP4_REFERENCE || p->mode == NLS_SP4_DAMPED, "unknown sp4 mode");
    auto A = build_operator(p->dim, p->n, p->L, c, 0, 0.0, p->q6);
    *out = new nls_sp4(*p, std::move(A), m, u0, v0);
  });
}
int nls_sp4_destroy(nls_sp4 *h) { return guard([&] { delete h; }); }
int nls_sp4_step(nls_sp4 *h) { return guard([&] { h->step(); }); }
int nls_sp4_apply_bc(nls_sp4 *h) { return guard([&] { h->apply_bc(); }); }
int nls_sp4_store_snapshot(nls_sp4 *h, int idx) { return guard([&] { h->store_snapshot(idx); }); }
int nls_sp4_transfer_snapshots(nls_sp4 *h, double *u, double *v) {
  return guard([&] {
    const size_t B = sizeof(double) * h->N * h->p.num_snapshots;
    NLS_HIP(hipMemcpyAsync(u, h->snapU->p, B, hipMemcpyDeviceToHost, h->s));
    if (v) NLS_HIP(hipMemcpyAsync(v, h->snapV->p, B, hipMemcpyDeviceToHost, h->s));
    NLS_HIP(hipStreamSynchronize(h->s));
  });
}
int nls_sp4_run(nls_sp4 *h) { return guard([&] { h->run(); NLS_HIP(hipStreamSynchronize(h->s)); }); }

int nls_philox_normals(uint64_t seed, uint32_t stream_id, uint32_t step, int64_t count,
                       double *out_host) {
  return guard([&] {
    NLS_REQUIRE(count >= 0 && out_host, "nls_philox_normals: bad arguments");
    if (count == 0) return;
    DevBuf d(sizeof(double) * count);
    hipLaunchKernelGGL(k_philox_fill, dim3(blocks_for(count, BS)), dim3(BS), 0, 0, count, seed,
                       stream_id, step, d.as<double>());
    check_launch("philox");
    NLS_HIP(hipMemcpy(out_host, d.p, sizeof(double) * count, hipMemcpyDeviceToHost));
  });
}

// STREAM-style copy used as the achievable-bandwidth reference next to the roofline figures
// and to calibrate the HBM counters: one float4 per thread and a grid that covers the buffer.
// On MI355X this reaches about 6.2 TB/s on a 1 GiB buffer; a grid-stride loop over the same
// copy (each wave waits for its previous store before its next load) reaches 4.6-4.8 TB/s.
__global__ __launch_bounds__(BS) void k_stream_copy(int64_t n4, const float4 *__restrict__ x,
                                                    float4 *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n4) y[i] = x[i];
}

int nls_stream_copy_bench(size_t bytes, int reps, double *gbps, double *avg_ms) {
  return guard([&] {
    NLS_REQUIRE(bytes >= 1024 && reps >= 1 && gbps, "nls_stream_copy_bench: bad arguments");
    const int64_t n4 = (int64_t)(bytes / 16);
    DevBuf a(16 * n4), b(16 * n4);
    NLS_HIP(hipMemset(a.p, 0, 16 * n4));
    hipStream_t st;
    NLS_HIP(hipStreamCreate(&st));
    NLS_REQUIRE((n4 + BS - 1) / BS < ((int64_t)1 << 31), "nls_stream_copy_bench: buffer too large");
    const unsigned grid = (unsigned)blocks_for(n4, BS);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(BS), 0, st, n4, a.as<float4>(), b.as<float4>());
    hipEvent_t e0, e1;
    NLS_HIP(hipEventCreate(&e0));
    NLS_HIP(hipEventCreate(&e1));
    NLS_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(BS), 0, st, n4, a.as<float4>(), b.as<float4>());
    NLS_HIP(hipEventRecord(e1, st));
    NLS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    NLS_HIP(hipEventElapsedTime(&ms, e0, e1));
    check_launch("stream copy");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    *gbps = 2.0 * 16.0 * n4 * reps / (ms / 1e3) / 1e9;
    if (avg_ms) *avg_ms = ms / reps;
  });
}

} // extern "C"
