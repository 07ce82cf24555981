// FP64 small-matrix stage of the Krylov matrix function: eigen-decomposition of
// the m x m Lanczos tridiagonal T and y = beta0 * Q f(Lambda) Q^T e1.
//
// Replaces cusolverDnDsyevd/Zheevd + the f(lambda) Thrust lambdas + two GEMMs and
// a GEMV of the reference (device/include/matfunc_real.hpp:190-313,385-392;
// device/include/matfunc_complex.hpp:254-375,382-390).
//
// One 64-lane wavefront solves one problem: implicit QL with a Wilkinson-type
// shift (Golub & Van Loan sec. 8.3) on the symmetric tridiagonal; lane k owns row
// k of the eigenvector matrix Z (in LDS) and applies every Givens rotation to it,
// lane 0 owns the diagonal/subdiagonal updates.  A Hermitian T with complex
// subdiagonal t_i (the reference's complex Lanczos) is first reduced to a real
// symmetric S = D^H T D with D = diag(ph), ph_0 = 1, ph_{i+1} = ph_i t_i/|t_i|,
// so f(T) e1 = D f(S) e1.
#pragma once

#include "nls_common.hpp"

namespace nls {

constexpr int kMaxKrylov = 32;

enum FuncKind : int {
  F_COS_SQRT = 0,   // cos(t sqrt|l|)                (matfunc_real.hpp:221-228)
  F_SINC2_SQRT = 1, // (sin x / x)^2, x = t sqrt|l|  (matfunc_real.hpp:212-219)
  F_ID_SQRT = 2,    // t sqrt|l|                     (matfunc_real.hpp:230-237)
  F_SINC2_HALF = 3, // (sin(x/2)/(x/2))^2: sEWI filter for Klein-Gordon
  F_COSM1_SQRT = 4, // cos(x) - 1 = -2 sin^2(x/2)    (increment form used by the fused path)
  F_EXP = 8,        // exp(dt l), dt complex         (matfunc_complex.hpp:281-287)
  F_SINC = 9,       // sin(dt l)/(dt l), 1 if |.| < 1e-8  (matfunc_complex.hpp:293-300)
  F_EXP_SERIES = 10, // exp(dt T) e1 by a scaled Taylor series (no eigen-decomposition)
};
// Added to a real kind: the register-resident single-wavefront QL of the fused batch
// (small_ql_wave + small_apply_f_wave) instead of the workgroup version.
constexpr int F_WAVE_QL = 0x100;
// Added to F_COSM1_SQRT, F_SINC2_SQRT or F_SINC2_HALF: the fused batch's power-series
// route (small_kge_series_wave), for a T with no negative eigenvalue.
constexpr int F_KGE_SERIES = 0x200;

struct cplx {
  double re, im;
};

__device__ inline cplx eval_f(int kind, double t_re, double t_im, double lam) {
  if (kind < 8) {
    const double x = t_re * sqrt(fabs(lam));
    double f;
    switch (kind) {
    case F_COS_SQRT: f = cos(x); break;
    case F_SINC2_SQRT: {
      if (fabs(x) < 1e-8) f = 1.0;
      else { double q = sin(x) / x; f = q * q; }
      break;
    }
    case F_ID_SQRT: f = x; break;
    case F_SINC2_HALF: {
      const double y = 0.5 * x;
      if (fabs(y) < 1e-8) f = 1.0;
      else { double q = sin(y) / y; f = q * q; }
      break;
    }
    default: { double s = sin(0.5 * x); f = -2.0 * s * s; break; }
    }
    return {f, 0.0};
  }
  const double a = t_re * lam, b = t_im * lam;
  if (kind == F_EXP) {
    const double ea = exp(a);
    return {ea * cos(b), ea * sin(b)};
  }
  // sin(z)/z for z = a + i b
  if (sqrt(a * a + b * b) < 1e-8) return {1.0, 0.0};
  const double sr = sin(a) * cosh(b), si = cos(a) * sinh(b);
  const double den = a * a + b * b;
  return {(sr * a + si * b) / den, (si * a - sr * b) / den};
}

// Shared-memory workspace of one problem.
struct SmallWS {
  double d[kMaxKrylov], e[kMaxKrylov];
  double Z[kMaxKrylov][kMaxKrylov + 1];
  double phr[kMaxKrylov], phi[kMaxKrylov];
  int status;
};

// QL on ws.d/ws.e (size m) with eigenvectors accumulated in ws.Z.  Must be
// called by all 64 lanes of a single-wavefront workgroup.  Returns 0 on success.
__device__ inline int small_ql(int m, SmallWS &ws) {
  const int lane = threadIdx.x;
  for (int k = lane; k < m * m; k += kWave) ws.Z[k / m][k % m] = (k / m == k % m) ? 1.0 : 0.0;
  if (lane == 0) ws.e[m - 1] = 0.0;
  __syncthreads();
  for (int l = 0; l < m; ++l) {
    int iter = 0;
    for (;;) {
      int mm;
      for (mm = l; mm < m - 1; ++mm) {
        const double dd = fabs(ws.d[mm]) + fabs(ws.d[mm + 1]);
        if (fabs(ws.e[mm]) <= 2.220446049250313e-16 * dd) break;
      }
      if (mm == l) break;
      if (++iter > 60) return 1;
      double g = (ws.d[l + 1] - ws.d[l]) / (2.0 * ws.e[l]);
      double r = hypot(g, 1.0);
      g = ws.d[mm] - ws.d[l] + ws.e[l] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      for (int i = mm - 1; i >= l; --i) {
        const double f = s * ws.e[i], b = c * ws.e[i];
        r = hypot(f, g);
        __syncthreads();
        if (lane == 0) ws.e[i + 1] = r;
        if (r == 0.0) {
          if (lane == 0) { ws.d[i + 1] -= p; ws.e[mm] = 0.0; }
          __syncthreads();
          deflated = true;
          break;
        }
        const double ir = 1.0 / r;
        s = f * ir;
        c = g * ir;
        g = ws.d[i + 1] - p;
        r = (ws.d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        const double dnew = g + p;
        g = c * r - b;
        if (lane < m) {
          const double zf = ws.Z[lane][i + 1], zi = ws.Z[lane][i];
          ws.Z[lane][i + 1] = s * zi + c * zf;
          ws.Z[lane][i] = c * zi - s * zf;
        }
        __syncthreads();
        if (lane == 0) ws.d[i + 1] = dnew;
        __syncthreads();
      }
      if (deflated) continue;
      __syncthreads();
      if (lane == 0) { ws.d[l] -= p; ws.e[l] = g; ws.e[mm] = 0.0; }
      __syncthreads();
    }
  }
  __syncthreads();
  return 0;
}

// After small_ql: y_k = beta0 * ph_k * sum_i Z[k][i] f(l_i) Z[0][i], lanes k < m.
__device__ inline void small_apply_f(int m, SmallWS &ws, int kind, double t_re, double t_im,
                                     double beta0, double *yr, double *yi) {
  const int lane = threadIdx.x;
  if (lane < m) {
    double ar = 0.0, ai = 0.0;
    for (int i = 0; i < m; ++i) {
      const cplx f = eval_f(kind, t_re, t_im, ws.d[i]);
      const double w = ws.Z[lane][i] * ws.Z[0][i];
      ar += w * f.re;
      ai += w * f.im;
    }
    const double pr = ws.phr[lane], pi = ws.phi[lane];
    yr[lane] = beta0 * (pr * ar - pi * ai);
    if (yi) yi[lane] = beta0 * (pr * ai + pi * ar);
  }
}

// ---------------------------------------------------------------- register-resident variant
// Same implicit QL and the same f evaluation as small_ql / small_apply_f, organised
// for a single wavefront with no barriers: lane k < m keeps d_k and e_k in registers,
// every lane follows the (uniform) rotation recurrence reading d_i, e_i with
// readlane, lane k applies each rotation to its own row of Z (LDS, rows never
// read by other lanes during the sweep), and lane i evaluates f(lambda_i) for its
// own eigenvalue.  r = sqrt(f^2 + g^2) replaces hypot: the entries of a Lanczos T
// are far from the overflow range.  Used by the fused batch, where this stage is
// on the critical path of every step.
__device__ inline double lane_read(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// dk, ek: d_lane, e_lane on entry (lanes >= m ignored); eigenvalues in dk on exit.
// Z: row stride kMaxKrylov + 1.  Returns 0 on success (uniform).
__device__ inline int small_ql_wave(int m, double &dk, double &ek, double (*Z)[kMaxKrylov + 1]) {
  const int lane = threadIdx.x & (kWave - 1);
  if (lane < m)
    for (int i = 0; i < m; ++i) Z[lane][i] = lane == i ? 1.0 : 0.0;
  if (lane == m - 1) ek = 0.0;
  for (int l = 0; l < m; ++l) {
    int iter = 0;
    for (;;) {
      // first mm >= l with a negligible e_mm (or m - 1): every lane tests its own e_k
      // against |d_k| + |d_{k+1}| at once, and the ballot gives the first one
      const double dnext = __shfl_down(dk, 1);
      const bool neg = lane < m - 1 && fabs(ek) <= 2.220446049250313e-16 * (fabs(dk) + fabs(dnext));
      const unsigned long long stop =
          (__ballot(neg) | (1ull << (m - 1))) & ~((1ull << l) - 1ull);
      const int mm = __ffsll((long long)stop) - 1;
      if (mm == l) break;
      if (++iter > 60) return 1;
      const double dl = lane_read(dk, l), el = lane_read(ek, l);
      double g = (lane_read(dk, l + 1) - dl) / (2.0 * el);
      double r = sqrt(g * g + 1.0);
      g = lane_read(dk, mm) - dl + el / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      // Z[lane][i + 1] of the rotation in flight stays in a register: each rotation reads
      // one element Z[lane][i] no rotation has written yet and writes one finished element
      double zc = lane < m ? Z[lane][mm] : 0.0;
      for (int i = mm - 1; i >= l; --i) {
        const double ei = lane_read(ek, i);
        const double f = s * ei, b = c * ei;
        // r = sqrt(q), 1 / r = rsq(q) with two Newton steps: one long-latency operation
        // per rotation instead of a square root and a division
        const double q = f * f + g * g;
        double rs = __builtin_amdgcn_rsq(q);
        rs = fma(0.5 * rs, fma(-q * rs, rs, 1.0), rs);
        rs = fma(0.5 * rs, fma(-q * rs, rs, 1.0), rs);
        r = q > 0.0 ? q * rs : 0.0;
        if (lane == i + 1) ek = r;
        if (r == 0.0) {
          const double di1 = lane_read(dk, i + 1);
          if (lane == i + 1) dk = di1 - p;
          if (lane == mm) ek = 0.0;
          if (lane < m) Z[lane][i + 1] = zc;
          deflated = true;
          break;
        }
        s = f * rs;
        c = g * rs;
        g = lane_read(dk, i + 1) - p;
        r = (lane_read(dk, i) - g) * s + 2.0 * c * b;
        p = s * r;
        const double dnew = g + p;
        g = c * r - b;
        if (lane < m) {
          const double zi = Z[lane][i];
          Z[lane][i + 1] = s * zi + c * zc;
          zc = c * zi - s * zc;
        }
        if (lane == i + 1) dk = dnew;
      }
      if (deflated) continue;
      if (lane < m) Z[lane][l] = zc;
      if (lane == l) {
        dk -= p;
        ek = g;
      }
      if (lane == mm) ek = 0.0;
    }
  }
  return 0;
}

// After small_ql_wave: y_k = beta0 * ph_k * sum_i Z[k][i] f(lambda_i) Z[0][i] for lanes
// k < m (ph_k = phr + i phi, 1 for a real symmetric T).
__device__ inline void small_apply_f_wave(int m, double lam, double (*Z)[kMaxKrylov + 1],
                                          int kind, double t_re, double t_im, double beta0,
                                          double phr, double phi, double &yr, double &yi) {
  const int lane = threadIdx.x & (kWave - 1);
  cplx f = {0.0, 0.0};
  if (lane < m) f = eval_f(kind, t_re, t_im, lam);
  // rows of Z written by every lane of this wavefront are read below: a wavefront-scope
  // barrier (LDS operations of one wavefront complete in order), so a workgroup may run
  // this stage on some of its waves only
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double ar = 0.0, ai = 0.0;
  for (int i = 0; i < m; ++i) {
    const double fr = lane_read(f.re, i), fi = lane_read(f.im, i);
    if (lane < m) {
      const double w = Z[lane][i] * Z[0][i];
      ar += w * fr;
      ai += w * fi;
    }
  }
  yr = beta0 * (phr * ar - phi * ai);
  yi = beta0 * (phr * ai + phi * ar);
}

// T * x for the tridiagonal held one row per lane (lane k: d = T[k][k], e = T[k][k+1],
// eu = T[k][k-1]; lanes >= m hold zeros): the neighbours' x come from DPP row shifts when
// every row lies in the first 16 lanes (m <= 16; lane 15's right neighbour and lane 0's
// left one are not needed there), from cross-lane shuffles otherwise.
__device__ inline double tridiag_apply_wave(int m, double d, double e, double eu, double x) {
  const int lane = threadIdx.x & (kWave - 1);
  double up, dn;
  if (m <= 16) {
    up = dpp_move<0x101>(x); // row_shl:1, lane k + 1
    dn = dpp_move<0x111>(x); // row_shr:1, lane k - 1
  } else {
    up = __shfl_down(x, 1, kWave);
    dn = __shfl_up(x, 1, kWave);
  }
  double r = d * x + e * (lane + 1 < m ? up : 0.0);
  if (lane > 0) r += eu * dn;
  return r;
}

// Largest row sum of |T| for the wavefront layout of small_ql_wave (lane k: d_k, e_k =
// T[k][k+1]; lanes >= m ignored), in every lane.
__device__ inline double tridiag_norm_inf_wave(int m, double dk, double ek) {
  const int lane = threadIdx.x & (kWave - 1);
  const double e = lane < m - 1 ? ek : 0.0;
  const double el = __shfl_up(e, 1, kWave);
  double nu = lane < m ? fabs(dk) + fabs(lane > 0 ? el : 0.0) + fabs(e) : 0.0;
  for (int off = 32; off >= 1; off >>= 1) nu = fmax(nu, __shfl_xor(nu, off, kWave));
  return nu;
}

// Whether the symmetric tridiagonal T (wavefront layout of small_ql_wave) has no negative
// eigenvalue, from the signs of the pivots of T = L D L^T (Sylvester's law of inertia: the
// number of negative pivots is the number of negative eigenvalues).  A zero pivot answers
// false.  Uniform across the wavefront.
__device__ inline bool tridiag_nonneg_wave(int m, double dk, double ek) {
  double p = lane_read(dk, 0);
  bool ok = p > 0.0;
  for (int k = 1; k < m && ok; ++k) {
    const double e = lane_read(ek, k - 1);
    p = lane_read(dk, k) - e * e / p;
    ok = p > 0.0;
  }
  return ok;
}

// y = beta0 f(T) e1 for the Klein-Gordon kinds F_COSM1_SQRT, F_SINC2_SQRT and
// F_SINC2_HALF on a symmetric tridiagonal T with nonnegative spectrum (the full
// Lanczos T of a shifted Laplacian), lane k < m receiving y_k.  These f are entire
// functions of lambda (power series in t^2 lambda), so on such a T the series is the
// same function of T the eigen-decomposition gives:
//   cos(t sqrt l) - 1    = sum_{k>=1} (-t^2 l)^k / (2k)!
//   (sin x / x)^2, x^2 = t^2 l = sum_{k>=0} (-1)^k 2^(2k+1) (t^2 l)^k / (2k+2)!
// (F_SINC2_HALF: t/2 in place of t).  Term k is ratio_k * T * term_{k-1}.  With
// rho = tt ||T||_inf, ||term_k|| is at most b_k = b_{k-1} rho / ((2k)(2k - 1)) (cos - 1,
// b_0 = 1, the sum starting at term 1) or b_{k-1} 4 rho / ((2k + 2)(2k + 1)) (sinc^2, the
// sum starting at term 0 = e1); the sum takes K terms, K the first with b_{K+1} <= 1e-17
// of the first summed term's bound, fixed before the loop, so no term needs a reduction
// across the wavefront (at most 60 terms).  Meant for t^2 ||T||_inf <= 4, where no term
// exceeds the sum by more than a factor of a few.
__device__ inline double small_kge_series_wave(int m, double dk, double ek, int kind, double t,
                                               double beta0) {
  const int lane = threadIdx.x & (kWave - 1);
  const bool in = lane < m;
  const double d = in ? dk : 0.0;
  const double e = lane < m - 1 ? ek : 0.0;  // T[lane][lane+1]
  const double el = __shfl_up(e, 1, kWave);  // T[lane-1][lane]
  const double eu = lane > 0 ? el : 0.0;
  const double tt = kind == F_SINC2_HALF ? 0.25 * t * t : t * t;
  const bool cosm1 = kind == F_COSM1_SQRT;
  const double rho = tt * tridiag_norm_inf_wave(m, dk, ek);
  auto ratio_bound = [&](int k) {
    return cosm1 ? rho / ((2.0 * k) * (2.0 * k - 1.0))
                 : 4.0 * rho / ((2.0 * k + 2.0) * (2.0 * k + 1.0));
  };
  double b = ratio_bound(1); // bound of term 1
  const double first = cosm1 ? b : 1.0;
  int K = 1;
  while (K < 60) {
    const double next = b * ratio_bound(K + 1);
    if (next <= 1e-17 * first) break;
    b = next;
    ++K;
  }
  double term = lane == 0 ? 1.0 : 0.0; // e1
  double acc = cosm1 ? 0.0 : term;
  for (int k = 1; k <= K; ++k) {
    const double x = tridiag_apply_wave(m, d, e, eu, term);
    const double ratio = cosm1 ? -tt / ((2.0 * k) * (2.0 * k - 1.0))
                               : -4.0 * tt / ((2.0 * k + 2.0) * (2.0 * k + 1.0));
    term = in ? ratio * x : 0.0;
    acc += term;
  }
  return beta0 * acc;
}

// y = beta0 exp(t T) e1 (t = t_re + i t_im) for the symmetric tridiagonal T of the
// wavefront layout of small_ql_wave (lane k: d_k, e_k = T[k][k+1]; lanes >= m ignored),
// lane k < m receiving y_k.  exp is analytic, so the series gives what the
// eigen-decomposition gives (the same function of T) without the QL iteration: the step
// t is cut into s pieces with rho = |t| ||T||_inf / s <= 2 and each piece applies
// sum_{k <= K} (t/s)^k T^k / k! to the running vector v.  Term k is bounded by
// rho^k / k! ||v||_inf, so K is the smallest k with rho^k / k! <= 1e-17 (at most 40;
// 25 at rho = 2, 19 at rho = 1), fixed before the loop: no per-term reductions across the
// wavefront.  No term exceeds 2 ||v||_inf, so the sum loses at most a bit to cancellation;
// rho <= 2 takes fewer products than rho <= 1 (2 x 25 against 4 x 19 at |t| ||T|| = 4).
// 1/k comes from a lane table (lane k - 1 holds 1/k) instead of a division per term.
__device__ inline void small_exp_series_wave(int m, double dk, double ek, double t_re,
                                             double t_im, double beta0, double &yr,
                                             double &yi) {
  const int lane = threadIdx.x & (kWave - 1);
  const bool in = lane < m;
  const double d = in ? dk : 0.0;
  const double e = lane < m - 1 ? ek : 0.0;      // T[lane][lane+1]
  const double el = __shfl_up(e, 1, kWave);      // T[lane-1][lane]
  const double eu = lane > 0 ? el : 0.0;
  double nu = in ? fabs(d) + fabs(eu) + fabs(e) : 0.0;
  for (int off = 32; off >= 1; off >>= 1) nu = fmax(nu, __shfl_xor(nu, off, kWave));
  const double tabs = sqrt(t_re * t_re + t_im * t_im);
  const int s = (int)fmax(1.0, ceil(0.5 * tabs * nu));
  const double cr = t_re / s, ci = t_im / s;
  const double rho = tabs * nu / s;
  int K = 1;
  for (double bnd = rho; K < 40 && bnd > 1e-17;) {
    ++K;
    bnd *= rho / K;
  }
  const double rk = 1.0 / (double)(lane + 1); // lane k - 1: 1 / k
  double vr = lane == 0 ? 1.0 : 0.0, vi = 0.0; // e1
  for (int sub = 0; sub < s; ++sub) {
    double ar = vr, ai = vi, tr = vr, ti = vi;
    for (int k = 1; k <= K; ++k) {
      // T * term
      const double xr = tridiag_apply_wave(m, d, e, eu, tr);
      const double xi = tridiag_apply_wave(m, d, e, eu, ti);
      // term <- (c / k) * (T term)
      const double inv = lane_read(rk, k - 1);
      tr = in ? (cr * xr - ci * xi) * inv : 0.0;
      ti = in ? (cr * xi + ci * xr) * inv : 0.0;
      ar += tr;
      ai += ti;
    }
    vr = ar;
    vi = ai;
  }
  yr = beta0 * vr;
  yi = beta0 * vi;
}

} // namespace nls
