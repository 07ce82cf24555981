// Fused batched path: many trajectories (uneven grids allowed) advanced together.
//
// Klein-Gordon Gautschi / sEWI step (reference: device/include/kg_single.cuh:57-127)
//   u+ = 2 cos(tau W) u - u- + tau^2 psi(tau W) (-m u^3),   W = sqrt(A), A = -L_c + mu2 I
// is evaluated in increment form
//   u+ = 2u - u- + 2 [(cos(tau W) - I) u] + tau^2 [psi(tau W) g],  g = -m u^3,
// so the two Krylov results are small corrections and FP32 Krylov vectors cost
// no accuracy in the FP64 state (the two-step recurrence amplifies an absolute
// error e in u+ by ~1/(tau w_min), which FP32 state would not survive over
// 10^4 steps).
//
// Both Krylov sequences (seq 0 on u, seq 1 on g) advance together.  One pass j
// per Lanczos iteration does, for every cell of every trajectory:
//   v_j   = (w'_{j-1} - alpha_{j-1} v_{j-1}) / b_j      at the cell and its neighbours
//   A v_j  by the matrix-free 5-point / 7-point stencil of div(c grad) (+ mu2)
//   w'_j  = A v_j - b_j v_{j-1}
// stores v_j and w'_j and emits per-chunk partial sums <w',v>, <w',w'>, <v,v>.
// The next pass reduces the partials of its own trajectory in a fixed order
// (alpha_j = <w',v>/<v,v>, b_{j+1}^2 = <w',w'> - <w',v>^2/<v,v>), so each pass
// needs exactly one global reduction, results do not depend on which other
// trajectories share the launch, and the stencil, axpy, dot and norm of the
// reference's iteration (lanczos.hpp:234-263) are one sweep over memory.
// As in the reference, alpha_{m-1} is not computed (T[m-1,m-1] = 0, Q1).
// No Gram-Schmidt pass is made (the reference makes one per iteration, Q3).
//
// Then one wavefront per (trajectory, sequence) eigen-decomposes T in FP64 and
// forms y = b0 Q f(L) Q^T e1 (nls_small.hpp), and the combine pass forms
// sum_k y_k v_k, the FP64 update, the velocity at snapshot steps, and the
// Neumann ghost copy (boundaries.cuh:9-81: ghost = value at clamped coordinates).
#include "nls.h"
#include "nls_common.hpp"
#include "nls_philox.hpp"
#include "nls_plain.hpp"
#include "nls_small.hpp"
#include "nls_ic.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <type_traits>
#include <vector>

namespace nls {
namespace fused {

constexpr int FBS = 256; // threads per workgroup
constexpr int NSEQ = 2;  // Krylov sequences per KGE step
constexpr int PN = 3;    // partial sums per sequence
constexpr int NXCD = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin

struct TrajDev {
  int n, dim, nchunks, chunk0;
  int range; // cells per chunk (one workgroup's contiguous range)
  // 3D plane march: chunk c = in-plane range c % npen (prange cells per plane) over
  // the planes [zseg * (c / npen), zseg * (c / npen + 1))
  int npen, prange, zseg;
  int64_t off, N;
  double tau;
  // The Krylov passes apply the unscaled operator B (weights (c_i + c_nb)/2; KGE: the
  // negated div(c grad), NLSE: div(c grad)) and the small stages form T of the operator
  // sd B + mu2d I in FP64 (the Krylov space of B is that of A = sd B + mu2d I, and
  // T_A = sd T_B + mu2d I), so the FP32 rounding of 1/dx^2 does not shift the spectrum.
  double sd;   // FP64 1 / dx^2
  double mu2d; // KGE mass term mu2
  // spatially uniform coefficient fields (bit 0: c, bit 1: m) and their FP32 values; a
  // batch whose trajectories all have both runs the UNI kernel variants, which use
  // cval / mval instead of streaming the fields (the same numbers)
  int uni;
  float cval, mval;
};
constexpr int kUniC = 1, kUniM = 2;

struct TrajScal {
  double b0[NSEQ];
  double alpha[NSEQ][kMaxKrylov];
  double beta[NSEQ][kMaxKrylov + 1]; // beta[s][j] = b_j, j >= 1
  double y[NSEQ][kMaxKrylov];
  int status;
  // Krylov monitor, per sequence: the largest weight |y_{m-1}| / |y| of the last basis
  // vector in y = f(T) e1 seen so far, and the number of steps where it exceeded
  // kStallTol.  With T[m-1][m-1] = 0 (the reference's Lanczos variant) this weight
  // falls as dt^2 as dt shrinks, not as dt^(2(m-1)): it reads the accuracy floor of
  // the variant as well as an unconverged Krylov space.
  int stalls[NSEQ];
  double max_tail[NSEQ];
  // Orthogonality of the FP32 basis (reorthogonalisation modes 1 and 2): |<v_{m-2}, v_0>|
  // of the previous step and its largest value; reorth_on: full reorthogonalisation is
  // active for this trajectory (mode 2 sets it once the loss exceeds the threshold).
  double orth[NSEQ];
  double max_orth[NSEQ];
  int reorth_on;
  double ralpha[NSEQ]; // alpha_j of the pass being reorthogonalised
};
constexpr double kStallTol = 1e-6;
// Default loss of orthogonality that switches reorthogonalisation on (mode 2): about the
// square root of the FP32 unit roundoff (semi-orthogonality).
constexpr double kReorthTol = 2.5e-4;

// Records the tail weight of y (lane k < m holds y_k, complex if yi != 0) for one
// sequence; y_0 is taken as yr + y0_add (the fused KGE step evaluates cos - 1, the
// weight is that of cos).  All lanes of the wavefront call it.
__device__ inline void monitor_tail(TrajScal &S, int s, int m, double yr, double yi,
                                    double y0_add) {
  const int lane = threadIdx.x & (kWave - 1);
  const double r = lane == 0 ? yr + y0_add : yr;
  const double v2 = lane < m ? r * r + yi * yi : 0.0;
  const double tot = wave_sum(v2);
  const double last = lane_read(v2, m - 1);
  if (lane == 0 && tot > 0.0) {
    const double tail = sqrt(last / tot);
    if (tail > S.max_tail[s]) S.max_tail[s] = tail;
    if (tail > kStallTol) S.stalls[s] += 1;
  }
}

struct KgeArgs {
  const TrajDev *traj;
  const int2 *cmap;
  const int *tlist; // small stage: block -> trajectory (nullptr: identity)
  TrajScal *scal;
  double *part; // [2 parity][total_chunks][NSEQ*PN]
  int total_chunks;
  const float *mf, *cf; // coefficient fields m (cubic) and c (conductivity)
  const double *u, *up;
  double *un;
  float *V; // [NSEQ][m-2][cells]
  float *W; // [NSEQ][2][cells]
  int64_t cells;
  int m;
  int q6;
  int kind1; // f of sequence 1
  // full reorthogonalisation: 0 off, 1 every iteration, 2 once the basis has lost
  // semi-orthogonality (|<v_{m-2}, v_0>| > reorth_tol)
  int reorth;
  double reorth_tol;
  double *part2; // [total_chunks][NSEQ][2][kMaxKrylov]: <w'_j, v_k>, <v_j, v_k>
  double *part3; // [total_chunks][NSEQ]: <v_{m-2}, v_0> of the last combine
  int *any_reorth; // mode 2: set once any trajectory has switched reorthogonalisation on
};

__device__ inline bool reorth_active(const KgeArgs &a, const TrajScal &S) {
  return a.reorth == 1 || (a.reorth == 2 && S.reorth_on);
}

// v_0 of both sequences at cell g (x / b0, x the FP32 seed)
__device__ inline void basis0(const KgeArgs &a, const TrajScal &S, int64_t g, double &x0,
                              double &x1) {
  const double uu = a.u[g];
  x0 = S.b0[0] > 0.0 ? (double)(float)uu / S.b0[0] : 0.0;
  x1 = S.b0[1] > 0.0 ? (double)(float)(-(double)a.mf[g] * uu * uu * uu) / S.b0[1] : 0.0;
}

// Orthogonality monitor: per-chunk <v_{m-2}, x> of both sequences at cell g, x the FP32
// seed (v_0 = x / b0; the combine kernels scale the chunk sums by 1 / b0).
__device__ inline void orth_partial(const KgeArgs &a, int64_t g, double *o) {
  const size_t m2 = (size_t)(a.m - 2);
  const double uu = a.u[g];
  o[0] += (double)(float)uu * (double)a.V[(0 * m2 + m2 - 1) * a.cells + g];
  o[1] += (double)(float)(-(double)a.mf[g] * uu * uu * uu) *
          (double)a.V[(1 * m2 + m2 - 1) * a.cells + g];
}

__device__ inline void store_orth(const KgeArgs &a, const TrajDev &T, const TrajScal &S, int chunk,
                                  const double *o) {
#pragma unroll
  for (int s = 0; s < NSEQ; ++s)
    a.part3[(size_t)(T.chunk0 + chunk) * NSEQ + s] = S.b0[s] > 0.0 ? o[s] / S.b0[s] : 0.0;
}

// ---------------------------------------------------------------- neighbour helpers
struct Nbr {
  int64_t off[6];
  bool has[6];
  int cnt;
};

template <int DIM>
__device__ inline void neighbours(int64_t i, int n, int64_t N, int q6, Nbr &nb) {
  if (DIM == 2) {
    const int64_t r = i / n, col = i - r * n;
    nb.cnt = 4;
    nb.off[0] = -n; nb.has[0] = r > 0;
    nb.off[1] = -1; nb.has[1] = col > 0;
    nb.off[2] = 1;  nb.has[2] = col < n - 1;
    nb.off[3] = n;  nb.has[3] = r < n - 1;
  } else {
    const int64_t n2 = (int64_t)n * n;
    const int64_t col = i % n, inplane = i % n2;
    nb.cnt = 6;
    nb.off[0] = -n2; nb.has[0] = i >= n2;
    nb.off[1] = -n;  nb.has[1] = q6 ? (i >= n) : (inplane >= n);
    nb.off[2] = -1;  nb.has[2] = col > 0;
    nb.off[3] = 1;   nb.has[3] = col < n - 1;
    nb.off[4] = n;   nb.has[4] = q6 ? (i + n < N) : (inplane < n2 - n);
    nb.off[5] = n2;  nb.has[5] = i + n2 < N;
  }
}

template <int DIM> __device__ inline int64_t clamp_src(int64_t i, int n) {
  auto cl = [n](int64_t a) { return a < 1 ? (int64_t)1 : (a > n - 2 ? (int64_t)n - 2 : a); };
  if (DIM == 2) {
    const int64_t r = i / n, col = i - r * n;
    return cl(r) * n + cl(col);
  } else {
    const int64_t n2 = (int64_t)n * n;
    const int64_t a = i / n2, rem = i - a * n2, b = rem / n, k = rem - b * n;
    return (cl(a) * n + cl(b)) * n + cl(k);
  }
}

// Seed vectors of the two sequences: x0 = u, x1 = g = -m u^3 (FP64 -> FP32).
__device__ inline float seed_x(int s, const KgeArgs &a, int64_t g) {
  const double uu = a.u[g];
  if (s == 0) return (float)uu;
  return (float)(-(double)a.mf[g] * uu * uu * uu);
}

// ---------------------------------------------------------------- LDS-only barrier
// A workgroup barrier that orders LDS accesses only.  __syncthreads() also waits for every
// outstanding global memory operation of the wave (s_waitcnt vmcnt(0)), so in a march the
// previous sub-chunk's stores and the prefetched loads would complete before each barrier.
// The passes exchange nothing through global memory inside a launch (each workgroup reads
// what earlier launches wrote and stores its own cells), so their barriers need only LDS
// ordering (s_waitcnt lgkmcnt(0); s_barrier) and the stores drain while the march goes on.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------- per-trajectory reduction
// Sums the partials of this trajectory's chunks (fixed order) into red[0..NSEQ*PN).
__device__ inline void reduce_partials(const KgeArgs &a, const TrajDev &T, int parity,
                                       double *red) {
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    double acc[NSEQ * PN];
#pragma unroll
    for (int k = 0; k < NSEQ * PN; ++k) acc[k] = 0.0;
    const double *p = a.part + ((size_t)parity * a.total_chunks + T.chunk0) * (NSEQ * PN);
    for (int g = lane; g < T.nchunks; g += kWave) {
#pragma unroll
      for (int k = 0; k < NSEQ * PN; ++k) acc[k] += p[(size_t)g * (NSEQ * PN) + k];
    }
#pragma unroll
    for (int k = 0; k < NSEQ * PN; ++k) {
      const double v = wave_sum_dpp(acc[k]);
      if (lane == 0) red[k] = v;
    }
  }
  lds_barrier();
}

struct Coef {
  float A, B, C; // v = (w - A*vprev) * B ; w' = Av - C * vprev(center)
  double alpha, b;
};

// From <w',v>, <w',w'>, <v,v> of pass j-1: alpha_{j-1}, b_j.
__device__ inline void lanczos_scalars(double q0, double q1, double q2, double &alpha,
                                       double &b) {
  alpha = q2 > 0.0 ? q0 / q2 : 0.0;
  double b2 = q2 > 0.0 ? q1 - q0 * q0 / q2 : 0.0;
  // Breakdown (invariant subspace, or cancellation down to FP32 noise): b = 0.
  b = (b2 > 1e-14 * q1 && b2 > 0.0) ? sqrt(b2) : 0.0;
}

// ---------------------------------------------------------------- Lanczos pass
// STAGE 0: pass j = 0 from the seeds; STAGE 1: j = 1; STAGE 2: j >= 2.
// LAST: pass j = m - 1 of the full-T Lanczos variant (partials only, nothing stored).
template <int DIM, int STAGE, int CPT, bool LAST = false>
__global__ __launch_bounds__(FBS) void k_kge_pass(KgeArgs a, int j) {
  static_assert(!(LAST && STAGE == 0), "the last pass follows pass 0");
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double red[NSEQ * PN];
  __shared__ double sh[FBS / kWave];
  Coef cf[NSEQ];
  double inv_b0[NSEQ];
  if (STAGE >= 1) {
    reduce_partials(a, T, (j - 1) & 1, red);
#pragma unroll
    for (int s = 0; s < NSEQ; ++s) {
      const double q0 = red[s * PN], q1 = red[s * PN + 1], q2 = red[s * PN + 2];
      if (STAGE == 1) {
        // pass-0 partials are <x,x>, <x,z>, <z,z> with z = A x (x unnormalised)
        const double b0 = sqrt(q0);
        const double al = q0 > 0.0 ? q1 / q0 : 0.0;
        double b1sq = q0 > 0.0 ? (q2 - q1 * q1 / q0) / q0 : 0.0;
        const double b1 = (b1sq > 1e-14 * (q0 > 0.0 ? q2 / q0 : 0.0) && b1sq > 0.0) ? sqrt(b1sq) : 0.0;
        inv_b0[s] = b0 > 0.0 ? 1.0 / b0 : 0.0;
        cf[s].alpha = al;
        cf[s].b = b1;
        cf[s].A = (float)al;
        cf[s].B = (b0 > 0.0 && b1 > 0.0) ? (float)(1.0 / (b0 * b1)) : 0.0f;
        cf[s].C = (float)(b1 * inv_b0[s]);
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].b0[s] = b0;
          a.scal[cm.x].alpha[s][0] = al;
          a.scal[cm.x].beta[s][1] = b1;
        }
      } else {
        double al, b;
        lanczos_scalars(q0, q1, q2, al, b);
        cf[s].alpha = al;
        cf[s].b = b;
        cf[s].A = (float)al;
        cf[s].B = b > 0.0 ? (float)(1.0 / b) : 0.0f;
        cf[s].C = (float)b;
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].alpha[s][j - 1] = al;
          a.scal[cm.x].beta[s][j] = b;
        }
      }
    }
  }
  const int n = T.n;
  const int64_t cells = a.cells;
  double acc[NSEQ][PN];
#pragma unroll
  for (int s = 0; s < NSEQ; ++s)
#pragma unroll
    for (int k = 0; k < PN; ++k) acc[s][k] = 0.0;

  const int64_t base = (int64_t)cm.y * (FBS * CPT);
#pragma unroll
  for (int kk = 0; kk < CPT; ++kk) {
    const int64_t i = base + threadIdx.x + (int64_t)kk * FBS;
    if (i >= T.N) break;
    const int64_t gi = T.off + i;
    Nbr nb;
    neighbours<DIM>(i, n, T.N, a.q6, nb);
    const float ci = a.cf[gi];
    float w[6];
#pragma unroll
    for (int q = 0; q < 2 * DIM; ++q) w[q] = nb.has[q] ? 0.5f * (ci + a.cf[gi + nb.off[q]]) : 0.0f;
#pragma unroll
    for (int s = 0; s < NSEQ; ++s) {
      auto val = [&](int64_t g) -> float {
        if (STAGE == 0) return seed_x(s, a, g);
        if (STAGE == 1) {
          const float z = a.W[((size_t)s * 2 + 0) * cells + g];
          return (z - cf[s].A * seed_x(s, a, g)) * cf[s].B;
        }
        const float wv = a.W[((size_t)s * 2 + ((j - 1) & 1)) * cells + g];
        const float vp = a.V[((size_t)s * (a.m - 2) + (j - 2)) * cells + g];
        return (wv - cf[s].A * vp) * cf[s].B;
      };
      const float vi = val(gi);
      float Av = 0.0f; // flux form sum_nb w (v_i - v_nb), as apply5
#pragma unroll
      for (int q = 0; q < 2 * DIM; ++q)
        if (nb.has[q]) Av += w[q] * (vi - val(gi + nb.off[q]));
      if (STAGE == 0) {
        a.W[((size_t)s * 2 + 0) * cells + gi] = Av;
        acc[s][0] += (double)vi * vi;
        acc[s][1] += (double)vi * Av;
        acc[s][2] += (double)Av * Av;
      } else {
        float vprev;
        if (STAGE == 1) vprev = seed_x(s, a, gi);
        else vprev = a.V[((size_t)s * (a.m - 2) + (j - 2)) * cells + gi];
        const float wp = Av - cf[s].C * vprev;
        if (!LAST) {
          a.V[((size_t)s * (a.m - 2) + (j - 1)) * cells + gi] = vi;
          a.W[((size_t)s * 2 + (j & 1)) * cells + gi] = wp;
        }
        acc[s][0] += (double)wp * vi;
        acc[s][1] += (double)wp * wp;
        acc[s][2] += (double)vi * vi;
      }
    }
  }
  // block partials -> part[parity j&1][global chunk]
  double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * (NSEQ * PN);
#pragma unroll
  for (int s = 0; s < NSEQ; ++s)
#pragma unroll
    for (int k = 0; k < PN; ++k) {
      const double t = block_sum<FBS>(acc[s][k], sh);
      if (threadIdx.x == 0) dst[s * PN + k] = t;
    }
}

// ---------------------------------------------------------------- small stage
// One wavefront per (trajectory, sequence): the scalars of the last pass, T of
// A = sd B + mu2 I (B the operator the passes apply) and y = b0 Q f(L) Q^T e1.
// last_pass = m - 2: the reference's variant, alpha_{m-2} and b_{m-1} from the last
// pass and T = tridiag(b, [alpha_0..alpha_{m-2}, 0], b) (lanczos.hpp:234-263, Q1).
// last_pass = m - 1: full T, alpha_{m-1} from the extra pass (b_{m-1} was set by it).
// One wavefront's small stage for (trajectory t, sequence s) with Z in LDS; no workgroup
// barrier, so the persistent step kernel can run the two sequences on two waves.
__device__ __forceinline__ void kge_small_wave(const KgeArgs &a, int t, int s, int last_pass,
                                               double (*Z)[kMaxKrylov + 1]) {
  const TrajDev T = a.traj[t];
  const int m = a.m;
  const int lane = threadIdx.x & (kWave - 1);
  // reduce last-pass partials (the wave sum leaves the same total in every lane)
  double red[PN] = {0.0, 0.0, 0.0};
  {
    const double *p =
        a.part + ((size_t)(last_pass & 1) * a.total_chunks + T.chunk0) * (NSEQ * PN) + s * PN;
    for (int g = lane; g < T.nchunks; g += kWave)
      for (int k = 0; k < PN; ++k) red[k] += p[(size_t)g * (NSEQ * PN) + k];
    for (int k = 0; k < PN; ++k) red[k] = wave_sum(red[k]);
  }
  TrajScal &S = a.scal[t];
  if (a.reorth != 0 && m >= 3) {
    // orthogonality of the previous step's basis (combine partials, fixed order)
    double o = 0.0;
    for (int g = lane; g < T.nchunks; g += kWave) o += a.part3[(size_t)(T.chunk0 + g) * NSEQ + s];
    o = fabs(wave_sum(o));
    if (lane == 0) {
      S.orth[s] = o;
      if (o > S.max_orth[s]) S.max_orth[s] = o;
      if (a.reorth == 2 && o > a.reorth_tol) {
        S.reorth_on = 1;
        *a.any_reorth = 1;
      }
    }
  }
  double al, b;
  lanczos_scalars(red[0], red[1], red[2], al, b);
  const bool full = last_pass == m - 1;
  if (lane == 0) {
    S.alpha[s][last_pass] = al;
    if (!full) S.beta[s][m - 1] = b;
  }
  // lane k: d_k = sd alpha_k + mu2 (reference variant: d_{m-1} = 0, Q1), e_k = sd b_{k+1}
  const double alk = lane < last_pass ? S.alpha[s][lane] : (lane == last_pass ? al : 0.0);
  const double bk = lane < m - 2 ? S.beta[s][lane + 1]
                                 : (lane == m - 2 ? (full ? S.beta[s][m - 1] : b) : 0.0);
  double dk = lane <= last_pass ? T.sd * alk + T.mu2d : 0.0;
  double ek = T.sd * bk;
  const double b0 = S.b0[s];
  const int kind = s == 0 ? F_COSM1_SQRT : a.kind1;
  double y;
  // Full T with no negative eigenvalue (the shifted Laplacian's): f is entire in lambda
  // and f(|lambda|) = f(lambda), so the power series gives y without the QL (7 terms at
  // tau^2 ||T|| = 0.05) when tau^2 ||T||_inf <= 4.  The reference variant's Q1 matrix can
  // have a negative eigenvalue, where the reference takes f at |lambda|: the QL route.
  if (full && T.tau * T.tau * tridiag_norm_inf_wave(m, dk, ek) <= 4.0 &&
      tridiag_nonneg_wave(m, dk, ek)) {
    y = small_kge_series_wave(m, dk, ek, kind, T.tau, b0);
  } else {
    if (small_ql_wave(m, dk, ek, Z)) {
      if (lane == 0) S.status = 1;
      return;
    }
    double yi;
    small_apply_f_wave(m, dk, Z, kind, T.tau, 0.0, b0, 1.0, 0.0, y, yi);
  }
  if (lane < m) S.y[s][lane] = y;
  monitor_tail(S, s, m, y, 0.0, s == 0 ? b0 : 0.0);
}

__global__ __launch_bounds__(kWave) void k_kge_small(KgeArgs a, int last_pass) {
  const int t = a.tlist ? a.tlist[blockIdx.x / NSEQ] : (int)(blockIdx.x / NSEQ);
  const int s = blockIdx.x % NSEQ;
  __shared__ SmallWS ws;
  kge_small_wave(a, t, s, last_pass, ws.Z);
}

// ---------------------------------------------------------------- combine
struct SnapArgs {
  double *u; // snapshot slot base (pool, cells)
  double *v;
};

// ---------------------------------------------------------------- 2D LDS-tiled Lanczos pass
// Same arithmetic as k_kge_pass<2, STAGE, CPT>, restructured for memory-level
// parallelism: the block first streams its chunk plus one grid row of halo on
// each side ([base-n, base+CH+n)) with independent coalesced loads, forms
// v_j (both sequences) once per element into LDS, then evaluates the stencil
// from LDS.  Halo rows are shared with the neighbouring chunks, which run on
// the same XCD (block map), so they come from L2.
template <int STAGE, int CPT>
__global__ __launch_bounds__(FBS) void k_kge_pass2d(KgeArgs a, int j, int emax) {
  extern __shared__ float lds[];
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double red[NSEQ * PN];
  __shared__ double sh[FBS / kWave];
  float cA[NSEQ] = {0.f, 0.f}, cB[NSEQ] = {0.f, 0.f}, cC[NSEQ] = {0.f, 0.f};
  if (STAGE >= 1) {
    reduce_partials(a, T, (j - 1) & 1, red);
#pragma unroll
    for (int s = 0; s < NSEQ; ++s) {
      const double q0 = red[s * PN], q1 = red[s * PN + 1], q2 = red[s * PN + 2];
      if (STAGE == 1) {
        const double b0 = sqrt(q0);
        const double al = q0 > 0.0 ? q1 / q0 : 0.0;
        const double b1sq = q0 > 0.0 ? (q2 - q1 * q1 / q0) / q0 : 0.0;
        const double b1 = (b1sq > 1e-14 * (q0 > 0.0 ? q2 / q0 : 0.0) && b1sq > 0.0) ? sqrt(b1sq) : 0.0;
        cA[s] = (float)al;
        cB[s] = (b0 > 0.0 && b1 > 0.0) ? (float)(1.0 / (b0 * b1)) : 0.0f;
        cC[s] = b0 > 0.0 ? (float)(b1 / b0) : 0.0f;
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].b0[s] = b0;
          a.scal[cm.x].alpha[s][0] = al;
          a.scal[cm.x].beta[s][1] = b1;
        }
      } else {
        double al, b;
        lanczos_scalars(q0, q1, q2, al, b);
        cA[s] = (float)al;
        cB[s] = b > 0.0 ? (float)(1.0 / b) : 0.0f;
        cC[s] = (float)b;
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].alpha[s][j - 1] = al;
          a.scal[cm.x].beta[s][j] = b;
        }
      }
    }
  }
  const int n = T.n;
  const int64_t N = T.N, cells = a.cells;
  constexpr int CH = FBS * CPT;
  const int64_t base = (int64_t)cm.y * CH;
  const int64_t cend = base + CH < N ? base + CH : N;
  const int64_t lo = base - n > 0 ? base - n : 0;
  const int64_t hi = cend + n < N ? cend + n : N;
  const int E = (int)(hi - lo);
  float *sv0 = lds, *sv1 = lds + emax, *sc = lds + 2 * emax;
  const float *W0 = a.W + (size_t)(0 * 2 + ((j - 1) & 1)) * cells;
  const float *W1 = a.W + (size_t)(1 * 2 + ((j - 1) & 1)) * cells;
  const float *V0 = a.V + (size_t)(0 * (a.m - 2) + (j - 2)) * cells;
  const float *V1 = a.V + (size_t)(1 * (a.m - 2) + (j - 2)) * cells;
  const float *Z0 = a.W + (size_t)(0 * 2) * cells, *Z1 = a.W + (size_t)(1 * 2) * cells;
  // phase 1: stream the extended range, UNR elements per thread in flight
  constexpr int UNR = 4;
  const int64_t gbase = T.off + lo;
  for (int e0 = threadIdx.x; e0 < E; e0 += UNR * FBS) {
    float x0[UNR], x1[UNR], y0[UNR], y1[UNR], cc[UNR];
#pragma unroll
    for (int q = 0; q < UNR; ++q) {
      const int e = e0 + q * FBS;
      const int64_t g = gbase + (e < E ? e : 0);
      cc[q] = a.cf[g];
      if (STAGE == 0) {
        x0[q] = seed_x(0, a, g);
        x1[q] = seed_x(1, a, g);
      } else if (STAGE == 1) {
        x0[q] = Z0[g];
        x1[q] = Z1[g];
        y0[q] = seed_x(0, a, g);
        y1[q] = seed_x(1, a, g);
      } else {
        x0[q] = W0[g];
        x1[q] = W1[g];
        y0[q] = V0[g];
        y1[q] = V1[g];
      }
    }
#pragma unroll
    for (int q = 0; q < UNR; ++q) {
      const int e = e0 + q * FBS;
      if (e < E) {
        sc[e] = cc[q];
        if (STAGE == 0) {
          sv0[e] = x0[q];
          sv1[e] = x1[q];
        } else {
          sv0[e] = (x0[q] - cA[0] * y0[q]) * cB[0];
          sv1[e] = (x1[q] - cA[1] * y1[q]) * cB[1];
        }
      }
    }
  }
  __syncthreads();
  double acc[NSEQ][PN];
#pragma unroll
  for (int s = 0; s < NSEQ; ++s)
#pragma unroll
    for (int k = 0; k < PN; ++k) acc[s][k] = 0.0;
#pragma unroll
  for (int kk = 0; kk < CPT; ++kk) {
    const int64_t i = base + threadIdx.x + (int64_t)kk * FBS;
    if (i < N) {
      const int e = (int)(i - lo);
      const int64_t r = i / n, col = i - r * n;
      const float ci = sc[e];
      const float wu = r > 0 ? 0.5f * (ci + sc[e - n]) : 0.0f;
      const float wl = col > 0 ? 0.5f * (ci + sc[e - 1]) : 0.0f;
      const float wr = col < n - 1 ? 0.5f * (ci + sc[e + 1]) : 0.0f;
      const float wd = r < n - 1 ? 0.5f * (ci + sc[e + n]) : 0.0f;
      const int64_t gi = T.off + i;
#pragma unroll
      for (int s = 0; s < NSEQ; ++s) {
        const float *sv = s == 0 ? sv0 : sv1;
        const float vi = sv[e];
        float Av = 0.0f; // flux form, as apply5
        if (r > 0) Av += wu * (vi - sv[e - n]);
        if (col > 0) Av += wl * (vi - sv[e - 1]);
        if (col < n - 1) Av += wr * (vi - sv[e + 1]);
        if (r < n - 1) Av += wd * (vi - sv[e + n]);
        if (STAGE == 0) {
          a.W[(size_t)(s * 2) * cells + gi] = Av;
          acc[s][0] += (double)vi * vi;
          acc[s][1] += (double)vi * Av;
          acc[s][2] += (double)Av * Av;
        } else {
          // v_{j-1} at the cell (just streamed by this block: an L2 hit)
          const float vprev = STAGE == 1 ? seed_x(s, a, gi) : (s == 0 ? V0 : V1)[gi];
          const float wp = Av - cC[s] * vprev;
          a.V[((size_t)s * (a.m - 2) + (j - 1)) * cells + gi] = vi;
          a.W[((size_t)s * 2 + (j & 1)) * cells + gi] = wp;
          acc[s][0] += (double)wp * vi;
          acc[s][1] += (double)wp * wp;
          acc[s][2] += (double)vi * vi;
        }
      }
    }
  }
  double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * (NSEQ * PN);
#pragma unroll
  for (int s = 0; s < NSEQ; ++s)
#pragma unroll
    for (int k = 0; k < PN; ++k) {
      const double t = block_sum<FBS>(acc[s][k], sh);
      if (threadIdx.x == 0) dst[s * PN + k] = t;
    }
}

inline size_t pass2d_lds_bytes(int cpt, int nmax) {
  const int CH = FBS * cpt, emax = CH + 2 * nmax;
  return sizeof(float) * 3 * (size_t)emax;
}

// ---------------------------------------------------------------- 2D row-march Lanczos pass
// Each block owns one contiguous range of a trajectory's cells (a few grid rows
// to tens of rows) and walks it in sub-chunks of SUB cells.  The LDS window
// holds [cb - n, cb + SUB + n) of v_j (both sequences), c and the raw v_{j-1}
// (or seed); moving to the next sub-chunk, the 2n cells both windows share are
// shifted inside LDS and only SUB new cells are read from memory, so every
// input element is read about once per pass.  The per-trajectory Lanczos
// scalars are formed once per block and the block writes one set of partial
// sums for its range.
template <int NV>
__device__ inline void block_sum_n(double (&v)[NV], double *sh /* NV * (FBS/64) */) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum_dpp(v[k]);
  lds_barrier();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k * (FBS / kWave) + w] = v[k];
  lds_barrier();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = sh[k * (FBS / kWave)];
#pragma unroll
    for (int q = 1; q < FBS / kWave; ++q) t += sh[k * (FBS / kWave) + q];
    v[k] = t;
  }
}

constexpr int MSPT = 4;               // cells per thread per sub-chunk
constexpr int MSUB = FBS * MSPT;      // sub-chunk cells

// c, v_j (two arrays) and the previous vector (two arrays); without c for uniform fields
inline size_t march2d_lds_bytes(int nmax, bool uni = false) {
  return sizeof(float) * (uni ? 4 : 5) * (size_t)(MSUB + 2 * nmax);
}

// ---------------------------------------------------------------- 2D row march, recomputed w'
// Variant of the row march that stores only the basis vectors.  Pass j re-reads
// v_{j-1} (two grid rows of halo), v_{j-2} (one row) and c (two rows), recomputes
//   w'_{j-1} = A v_{j-1} - b_{j-1} v_{j-2}
// at the cells it needs, forms v_j = (w'_{j-1} - alpha_{j-1} v_{j-1}) / b_j in LDS
// (same float expressions as the stored-w' passes), applies A and emits the
// partials of w'_j = A v_j - b_j v_{j-1}.  Per cell, both sequences: read
// v_{j-1}, v_{j-2}, c (20 B), write v_j (8 B) = 28 B, against 36 B when w' is
// stored.  Pass 0 stores nothing (pass 1 recomputes its A x); the last pass also
// stores w'_{m-2} for the combine.
// Blocks with an odd range index march downward (alt != 0): the halo rows that
// two neighbouring ranges share are then read by both blocks at about the same
// time, so the second read is served by the XCD's L2 instead of HBM.
// STAGE 0: j = 0 on the seeds x (partials <x,x>, <x,Ax>, <Ax,Ax>);
// STAGE 1: j = 1 (P = x, no Q); STAGE 2: j = 2 (P = v_1, Q = x);
// STAGE 3: j >= 3 (P = v_{j-1}, Q = v_{j-2}).
struct Sten5 {
  float wu, wl, wr, wd;
  bool hu, hl, hr, hd;
};

// Row of cell i (0 <= i < 2^24) from a float reciprocal of n, corrected to the
// exact quotient.
__device__ inline int row_of(int i, int n, float rn) {
  int r = (int)((float)i * rn);
  if (r * n > i) --r;
  else if ((r + 1) * n <= i) ++r;
  return r;
}

// Weights (c_i + c_nb)/2 of cell i (grid column from i; N = n * n cells) read from an LDS
// window at e, neighbour offsets (up, left, right, down) in window coordinates; a missing
// neighbour has weight 0.
__device__ inline Sten5 sten5(const float *sc, int e, int oU, int oL, int oR, int oD, int i,
                              int n, int N, float rn) {
  const int col = i - row_of(i, n, rn) * n;
  Sten5 w;
  w.hu = i >= n;
  w.hl = col > 0;
  w.hr = col < n - 1;
  w.hd = i < N - n;
  const float ci = sc[e];
  w.wu = w.hu ? 0.5f * (ci + sc[e + oU]) : 0.0f;
  w.wl = w.hl ? 0.5f * (ci + sc[e + oL]) : 0.0f;
  w.wr = w.hr ? 0.5f * (ci + sc[e + oR]) : 0.0f;
  w.wd = w.hd ? 0.5f * (ci + sc[e + oD]) : 0.0f;
  return w;
}

// sten5 for a uniform c: every weight (c + c) / 2 is c, formed as the field path forms it
__device__ inline Sten5 sten5_uniform(float cu, int i, int n, int N, float rn) {
  const int col = i - row_of(i, n, rn) * n;
  Sten5 w;
  w.hu = i >= n;
  w.hl = col > 0;
  w.hr = col < n - 1;
  w.hd = i < N - n;
  const float wc = 0.5f * (cu + cu);
  w.wu = w.hu ? wc : 0.0f;
  w.wl = w.hl ? wc : 0.0f;
  w.wr = w.hr ? wc : 0.0f;
  w.wd = w.hd ? wc : 0.0f;
  return w;
}

// B v = sum_nb w (v_i - v_nb) (the negated div(c grad) without 1/dx^2) in flux form: each
// difference is exact for neighbouring values within a factor 2, and the implicit diagonal
// is exactly the sum of the rounded off-diagonal weights, so B 1 = 0 holds in FP32 as in
// the exact operator (a diagonal formed as a rounded sum of the weights would add a
// random potential of ~1/dx^2 * 6e-8 per cell, which shifts the low frequencies and
// accumulates over long runs).  apply5: neighbours read only where present.
__device__ inline float apply5(const Sten5 &w, const float *v, int e, int oU, int oL, int oR,
                               int oD) {
  const float vi = v[e];
  float r = 0.0f;
  if (w.hu) r += w.wu * (vi - v[e + oU]);
  if (w.hl) r += w.wl * (vi - v[e + oL]);
  if (w.hr) r += w.wr * (vi - v[e + oR]);
  if (w.hd) r += w.wd * (vi - v[e + oD]);
  return r;
}

// apply5 on a window whose cells outside the grid hold finite values (zeros): every
// neighbour is read and a missing one contributes w = 0 (the same sum, no branches).
__device__ inline float apply5z(const Sten5 &w, const float *v, int e, int oU, int oL, int oR,
                                int oD) {
  const float vi = v[e];
  float r = w.wu * (vi - v[e + oU]);
  r += w.wl * (vi - v[e + oL]);
  r += w.wr * (vi - v[e + oR]);
  r += w.wd * (vi - v[e + oD]);
  return r;
}

constexpr int RKS = 6; // window shift registers: supports n <= RKS * FBS / 3 = 512

template <int STAGE>
__global__ __launch_bounds__(FBS) void k_kge_rmarch2d(KgeArgs a, int j, int emaxP, int emaxV, int alt) {
  extern __shared__ float lds[];
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double red[NSEQ * PN];
  __shared__ double shn[NSEQ * PN * (FBS / kWave)];
  // v_j = ((A P - qc Q) - pa P) * pb ;  w'_j = A v_j - pc P
  float pa[NSEQ] = {0.f, 0.f}, pb[NSEQ] = {0.f, 0.f}, qc[NSEQ] = {0.f, 0.f}, pc[NSEQ] = {0.f, 0.f};
  if (STAGE >= 1) {
    reduce_partials(a, T, (j - 1) & 1, red);
#pragma unroll
    for (int s = 0; s < NSEQ; ++s) {
      const double q0 = red[s * PN], q1 = red[s * PN + 1], q2 = red[s * PN + 2];
      if (STAGE == 1) {
        const double b0 = sqrt(q0);
        const double al = q0 > 0.0 ? q1 / q0 : 0.0;
        const double b1sq = q0 > 0.0 ? (q2 - q1 * q1 / q0) / q0 : 0.0;
        const double b1 = (b1sq > 1e-14 * (q0 > 0.0 ? q2 / q0 : 0.0) && b1sq > 0.0) ? sqrt(b1sq) : 0.0;
        pa[s] = (float)al;
        pb[s] = (b0 > 0.0 && b1 > 0.0) ? (float)(1.0 / (b0 * b1)) : 0.0f;
        pc[s] = b0 > 0.0 ? (float)(b1 / b0) : 0.0f;
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].b0[s] = b0;
          a.scal[cm.x].alpha[s][0] = al;
          a.scal[cm.x].beta[s][1] = b1;
        }
      } else {
        double al, b;
        lanczos_scalars(q0, q1, q2, al, b);
        pa[s] = (float)al;
        pb[s] = b > 0.0 ? (float)(1.0 / b) : 0.0f;
        pc[s] = (float)b;
        if (STAGE == 2) {
          const double b0 = a.scal[cm.x].b0[s], b1 = a.scal[cm.x].beta[s][1];
          qc[s] = b0 > 0.0 ? (float)(b1 / b0) : 0.0f;
        } else {
          qc[s] = (float)a.scal[cm.x].beta[s][j - 1];
        }
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].alpha[s][j - 1] = al;
          a.scal[cm.x].beta[s][j] = b;
        }
      }
    }
  }
  const int n = T.n;
  const int N = (int)T.N; // 2D grids: N < 2^31
  const int64_t cells = a.cells;
  // LDS: c, P (both sequences) over local [cb - 2n, cb + SUB + 2n); v_j over [cb - n, cb + SUB + n)
  float *sc = lds, *sp0 = lds + emaxP, *sp1 = lds + 2 * emaxP;
  float *sv0 = lds + 3 * emaxP, *sv1 = sv0 + emaxV;
  // per-trajectory bases (uniform), 32-bit cell offsets
  const size_t m2 = (size_t)(a.m - 2);
  const float *cT = a.cf + T.off, *mT = a.mf + T.off;
  const double *uT = a.u + T.off;
  const float *P0 = a.V + (STAGE >= 2 ? (0 * m2 + (size_t)(j - 2)) * cells : 0) + T.off;
  const float *P1 = a.V + (STAGE >= 2 ? (1 * m2 + (size_t)(j - 2)) * cells : 0) + T.off;
  const float *Q0 = a.V + (STAGE >= 3 ? (0 * m2 + (size_t)(j - 3)) * cells : 0) + T.off;
  const float *Q1 = a.V + (STAGE >= 3 ? (1 * m2 + (size_t)(j - 3)) * cells : 0) + T.off;
  float *Vout0 = a.V + (STAGE >= 1 ? (0 * m2 + (size_t)(j - 1)) * cells : 0) + T.off;
  float *Vout1 = a.V + (STAGE >= 1 ? (1 * m2 + (size_t)(j - 1)) * cells : 0) + T.off;
  float *Wout0 = a.W + ((size_t)0 * 2 + (j & 1)) * cells + T.off;
  float *Wout1 = a.W + ((size_t)1 * 2 + (j & 1)) * cells + T.off;
  const bool last = j == a.m - 2;
  double acc[NSEQ * PN];
#pragma unroll
  for (int k = 0; k < NSEQ * PN; ++k) acc[k] = 0.0;
  const int rbeg = cm.y * T.range;
  const int rend = rbeg + T.range < N ? rbeg + T.range : N;
  const int len = rend - rbeg;
  // local coordinate l -> cell org + dir * l (valid while 0 <= cell < N)
  const bool down = alt != 0 && (cm.y & 1);
  const int org = down ? rend - 1 : rbeg;
  const int dir = down ? -1 : 1;
  const int oU = -dir * n, oL = -dir, oR = dir, oD = dir * n; // grid neighbours in local offsets
  const float rn = 1.0f / (float)n;
  auto cell = [&](int l) { return org + dir * l; };
  auto seed = [&](int s, int i) -> float {
    const double uu = uT[i];
    if (s == 0) return (float)uu;
    return (float)(-(double)mT[i] * uu * uu * uu);
  };
  auto loadP = [&](int i, float &x0, float &x1) {
    if (STAGE <= 1) {
      x0 = seed(0, i);
      x1 = seed(1, i);
    } else {
      x0 = P0[i];
      x1 = P1[i];
    }
  };
  auto loadQ = [&](int i, float &x0, float &x1) {
    if (STAGE == 2) {
      x0 = seed(0, i);
      x1 = seed(1, i);
    } else if (STAGE == 3) {
      x0 = Q0[i];
      x1 = Q1[i];
    } else {
      x0 = 0.0f;
      x1 = 0.0f;
    }
  };
  // v_j at cell i: P window index eP, v window index eV
  auto phaseA = [&](int i, int eP, int eV, float q0, float q1) {
    const Sten5 w = sten5(sc, eP, oU, oL, oR, oD, i, n, N, rn);
    const float a0 = apply5(w, sp0, eP, oU, oL, oR, oD);
    const float a1 = apply5(w, sp1, eP, oU, oL, oR, oD);
    const float w0 = a0 - qc[0] * q0;
    const float w1 = a1 - qc[1] * q1;
    sv0[eV] = (w0 - pa[0] * sp0[eP]) * pb[0];
    sv1[eV] = (w1 - pa[1] * sp1[eP]) * pb[1];
  };
  // first window: c, P on [-2n, SUB + 2n), Q parked in the v_j window on [-n, SUB + n);
  // loads are issued MSPT at a time per thread before their LDS stores
  for (int e0 = 0; e0 < MSUB + 4 * n; e0 += MSPT * FBS) {
    float rc[MSPT], r0[MSPT], r1[MSPT], s0[MSPT], s1[MSPT];
#pragma unroll
    for (int q = 0; q < MSPT; ++q) {
      const int e = e0 + threadIdx.x + q * FBS;
      const int i = cell(e - 2 * n);
      if (e < MSUB + 4 * n && i >= 0 && i < N) {
        rc[q] = cT[i];
        loadP(i, r0[q], r1[q]);
      }
      const int iq = cell(e - n);
      if (STAGE >= 2 && e < MSUB + 2 * n && iq >= 0 && iq < N) loadQ(iq, s0[q], s1[q]);
    }
#pragma unroll
    for (int q = 0; q < MSPT; ++q) {
      const int e = e0 + threadIdx.x + q * FBS;
      const int i = cell(e - 2 * n);
      if (e < MSUB + 4 * n && i >= 0 && i < N) {
        sc[e] = rc[q];
        sp0[e] = r0[q];
        sp1[e] = r1[q];
      }
      const int iq = cell(e - n);
      if (STAGE >= 2 && e < MSUB + 2 * n && iq >= 0 && iq < N) {
        sv0[e] = s0[q];
        sv1[e] = s1[q];
      }
    }
  }
  // Register prefetch of the new cells of sub-chunk cb: c, P at [cb + 2n, cb + SUB + 2n),
  // Q at [cb + n, cb + SUB + n).  Issued before the compute phases that precede
  // their deposit, so the loads overlap two phases of work.
  float nc[MSPT], np0[MSPT], np1[MSPT], nq0[MSPT], nq1[MSPT];
  auto prefetch = [&](int cb) {
#pragma unroll
    for (int q = 0; q < MSPT; ++q) {
      const int l = cb + threadIdx.x + q * FBS;
      const int ip = cell(l + 2 * n);
      if (ip >= 0 && ip < N) {
        nc[q] = cT[ip];
        loadP(ip, np0[q], np1[q]);
      }
      const int iq = cell(l + n);
      if (STAGE >= 2 && iq >= 0 && iq < N) loadQ(iq, nq0[q], nq1[q]);
    }
  };
  if (MSUB < len) prefetch(MSUB);
  __syncthreads();
  if (STAGE >= 1) {
    // v_j on [-n, SUB + n); each element's Q is read and overwritten by the same thread
    for (int e = threadIdx.x; e < MSUB + 2 * n; e += FBS) {
      const int i = cell(e - n);
      if (i >= 0 && i < N) phaseA(i, e + n, e, STAGE >= 2 ? sv0[e] : 0.0f, STAGE >= 2 ? sv1[e] : 0.0f);
    }
    __syncthreads();
  }
  for (int cb = 0; cb < len; cb += MSUB) {
    const int ce = cb + MSUB < len ? cb + MSUB : len;
#pragma unroll
    for (int kk = 0; kk < MSPT; ++kk) {
      const int l = cb + threadIdx.x + kk * FBS;
      if (l < ce) {
        const int i = cell(l);
        const int eP = l - cb + 2 * n, eV = l - cb + n;
        const Sten5 w = sten5(sc, eP, oU, oL, oR, oD, i, n, N, rn);
        if (STAGE == 0) {
#pragma unroll
          for (int s = 0; s < NSEQ; ++s) {
            const float *sp = s == 0 ? sp0 : sp1;
            const float x = sp[eP];
            const float Av = apply5(w, sp, eP, oU, oL, oR, oD);
            acc[s * PN + 0] += (double)x * x;
            acc[s * PN + 1] += (double)x * Av;
            acc[s * PN + 2] += (double)Av * Av;
          }
        } else {
#pragma unroll
          for (int s = 0; s < NSEQ; ++s) {
            const float *sv = s == 0 ? sv0 : sv1;
            const float vi = sv[eV];
            const float Av = apply5(w, sv, eV, oU, oL, oR, oD);
            const float wp = Av - pc[s] * (s == 0 ? sp0 : sp1)[eP];
            (s == 0 ? Vout0 : Vout1)[i] = vi;
            if (last) (s == 0 ? Wout0 : Wout1)[i] = wp;
            acc[s * PN + 0] += (double)wp * vi;
            acc[s * PN + 1] += (double)wp * wp;
            acc[s * PN + 2] += (double)vi * vi;
          }
        }
      }
    }
    if (cb + MSUB >= len) break;
    __syncthreads(); // every read of this window is done
    // keep c on [cb' - n, cb' + 2n), v_j on [cb' - n, cb' + n) and P on [cb', cb' + 2n)
    // (STAGE 0 applies A to P itself: P on [cb' - n, cb' + 2n))
    {
      const int pk = STAGE == 0 ? 3 * n : 2 * n, p0 = 4 * n - pk;
      if (3 * n <= MSUB) {
        // source and destination ranges are disjoint: copy in place
        for (int e = threadIdx.x; e < 3 * n; e += FBS) sc[n + e] = sc[MSUB + n + e];
        for (int e = threadIdx.x; e < pk; e += FBS) {
          sp0[p0 + e] = sp0[MSUB + p0 + e];
          sp1[p0 + e] = sp1[MSUB + p0 + e];
        }
        if (STAGE >= 1)
          for (int e = threadIdx.x; e < 2 * n; e += FBS) {
            sv0[e] = sv0[MSUB + e];
            sv1[e] = sv1[MSUB + e];
          }
        __syncthreads(); // copies done before the deposit overwrites their sources
      } else {
        // ranges of length 3n overlap (n > SUB / 3): stage them through registers
        float tc[RKS], t0[RKS], t1[RKS];
#pragma unroll
        for (int k = 0; k < RKS; ++k) {
          const int e = threadIdx.x + k * FBS;
          if (e < 3 * n) {
            tc[k] = sc[MSUB + n + e];
            if (STAGE == 0) {
              t0[k] = sp0[MSUB + n + e];
              t1[k] = sp1[MSUB + n + e];
            }
          }
        }
        if (STAGE >= 1)
          for (int e = threadIdx.x; e < 2 * n; e += FBS) { // 2n <= SUB: disjoint
            sp0[2 * n + e] = sp0[MSUB + 2 * n + e];
            sp1[2 * n + e] = sp1[MSUB + 2 * n + e];
            sv0[e] = sv0[MSUB + e];
            sv1[e] = sv1[MSUB + e];
          }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RKS; ++k) {
          const int e = threadIdx.x + k * FBS;
          if (e < 3 * n) {
            sc[n + e] = tc[k];
            if (STAGE == 0) {
              sp0[n + e] = t0[k];
              sp1[n + e] = t1[k];
            }
          }
        }
      }
    }
    // deposit the prefetched cells: c, P at [4n, SUB + 4n), Q parked at [2n, SUB + 2n)
#pragma unroll
    for (int q = 0; q < MSPT; ++q) {
      const int l = cb + MSUB + threadIdx.x + q * FBS;
      const int e = 4 * n + threadIdx.x + q * FBS;
      const int ip = cell(l + 2 * n);
      if (ip >= 0 && ip < N) {
        sc[e] = nc[q];
        sp0[e] = np0[q];
        sp1[e] = np1[q];
      }
      const int iq = cell(l + n);
      if (STAGE >= 2 && iq >= 0 && iq < N) {
        sv0[e - 2 * n] = nq0[q];
        sv1[e - 2 * n] = nq1[q];
      }
    }
    if (cb + 2 * MSUB < len) prefetch(cb + 2 * MSUB);
    __syncthreads();
    if (STAGE >= 1) {
      // v_j on the next window's new cells [cb' + n, cb' + SUB + n)
#pragma unroll
      for (int q = 0; q < MSPT; ++q) {
        const int eV = 2 * n + threadIdx.x + q * FBS;
        const int i = cell(cb + MSUB + n + threadIdx.x + q * FBS);
        if (i >= 0 && i < N) phaseA(i, eV + n, eV, STAGE >= 2 ? sv0[eV] : 0.0f, STAGE >= 2 ? sv1[eV] : 0.0f);
      }
      __syncthreads();
    }
  }
  block_sum_n<NSEQ * PN>(acc, shn);
  if (threadIdx.x == 0) {
    double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * (NSEQ * PN);
#pragma unroll
    for (int k = 0; k < NSEQ * PN; ++k) dst[k] = acc[k];
  }
}

inline size_t rmarch2d_lds_bytes(int nmax) {
  return sizeof(float) * (3 * (size_t)(MSUB + 4 * nmax) + 2 * (size_t)(MSUB + 2 * nmax));
}

// Element idx of a per-trajectory array through a 32-bit byte offset from its (uniform)
// base, so the access uses the scalar-base addressing form instead of a 64-bit vector
// address per lane.  Needs idx * sizeof(T) < 2^32 (2D trajectories: fewer than 2^29 cells).
template <class T> __device__ inline T &at32(T *base, uint32_t idx) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T));
}
template <class T> __device__ inline const T &at32(const T *base, uint32_t idx) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                      idx * (uint32_t)sizeof(T));
}

// ---------------------------------------------------------------- 2D row march, stored w'
// Pass j reads w'_{j-1} and v_{j-1} (one grid row of halo), forms
// v_j = (w'_{j-1} - alpha_{j-1} v_{j-1}) / b_j as it deposits each cell into the LDS
// window [cb - n, cb + SUB + n), applies A from LDS and stores v_j and
// w'_j = A v_j - b_j v_{j-1}.  Same local coordinates as k_kge_rmarch2d, but the
// direction alternates with the pass as well (alt != 0): a block marches downward when its
// range index plus j is odd.  Neighbouring ranges still meet at their shared halo rows at
// the same time (the second read comes from L2), and each pass starts at the end of its
// range that the previous pass wrote last; the outputs are ordinary stores, so the first
// windows of the next pass find them in cache (62 against 66 us per pass, measured).
// STAGE 0: j = 0 on the seeds x (stores z = A x); 1: j = 1 (from z and x); 2: j >= 2.
// UNI: every trajectory of the batch has uniform m and c (values from the descriptor).
// LAST: pass j = m - 1 of the full-T Lanczos variant: v_{m-1} is formed and A applied for
// the partials that give alpha_{m-1}; nothing is stored (the combine folds v_{m-1}).
// The pass body for the chunk cm (trajectory, range) with the dynamic LDS window lds;
// shared by the pass kernel and the persistent step kernel.
template <int STAGE, bool UNI, bool LAST>
__device__ __forceinline__ void march2d_body(const KgeArgs a, const int2 cm, int j, int emax,
                                             int alt, float *lds) {
  static_assert(!(LAST && STAGE == 0), "the last pass follows pass 0");
  const TrajDev T = a.traj[cm.x];
  __shared__ double red[NSEQ * PN];
  __shared__ double shn[NSEQ * PN * (FBS / kWave)];
  // v = (X - cA P) * cB at deposit;  w' = A v - cC P
  float cA[NSEQ] = {0.f, 0.f}, cB[NSEQ] = {0.f, 0.f}, cC[NSEQ] = {0.f, 0.f};
  const int n = T.n;
  const int N = (int)T.N;
  const int64_t cells = a.cells;
  // no c window for uniform fields; v_j and the previous vector of the two sequences are
  // interleaved (one 8-byte LDS access per cell for both)
  float *sc = lds;
  float2 *sv = reinterpret_cast<float2 *>(lds + (UNI ? 0 : emax)), *sp = sv + emax;
  const size_t m2 = (size_t)(a.m - 2);
  const float *cT = a.cf + T.off, *mT = a.mf + T.off;
  const double *uT = a.u + T.off;
  // X source: STAGE 1 z = W[s][0]; STAGE 2 w'_{j-1}.  P source: STAGE 2 v_{j-1}.
  const float *X0 = a.W + ((size_t)0 * 2 + (STAGE == 1 ? 0 : ((j - 1) & 1))) * cells + T.off;
  const float *X1 = a.W + ((size_t)1 * 2 + (STAGE == 1 ? 0 : ((j - 1) & 1))) * cells + T.off;
  const float *P0 = a.V + (STAGE >= 2 ? (0 * m2 + (size_t)(j - 2)) * cells : 0) + T.off;
  const float *P1 = a.V + (STAGE >= 2 ? (1 * m2 + (size_t)(j - 2)) * cells : 0) + T.off;
  float *Vout0 = a.V + (STAGE >= 1 ? (0 * m2 + (size_t)(j - 1)) * cells : 0) + T.off;
  float *Vout1 = a.V + (STAGE >= 1 ? (1 * m2 + (size_t)(j - 1)) * cells : 0) + T.off;
  float *Wout0 = a.W + ((size_t)0 * 2 + (j & 1)) * cells + T.off;
  float *Wout1 = a.W + ((size_t)1 * 2 + (j & 1)) * cells + T.off;
  double acc[NSEQ * PN];
#pragma unroll
  for (int k = 0; k < NSEQ * PN; ++k) acc[k] = 0.0;
  const int rbeg = cm.y * T.range;
  const int rend = rbeg + T.range < N ? rbeg + T.range : N;
  const int len = rend - rbeg;
  // ranges whose index plus the pass index is odd march downward (see the comment above)
  const bool down = alt != 0 && ((cm.y ^ j) & 1);
  const int org = down ? rend - 1 : rbeg;
  const int dir = down ? -1 : 1;
  const int oU = -dir * n, oL = -dir, oR = dir, oD = dir * n;
  const float rn = 1.0f / (float)n;
  auto cell = [&](int l) { return org + dir * l; };
  // raw inputs of cell i: c, X (both sequences), P (both sequences), and u, m for
  // the seeds; nothing is computed on them here, so the loads stay in flight until
  // deposit() uses them
  auto fetch = [&](int i, float &c, float &x0, float &x1, float &p0, float &p1, double &uu,
                   float &mm) {
    const uint32_t u = (uint32_t)i;
    c = UNI ? T.cval : at32(cT, u);
    if (STAGE <= 1) {
      uu = at32(uT, u);
      mm = UNI ? T.mval : at32(mT, u);
    }
    if (STAGE >= 1) {
      x0 = at32(X0, u);
      x1 = at32(X1, u);
    }
    if (STAGE >= 2) {
      p0 = at32(P0, u);
      p1 = at32(P1, u);
    }
  };
  auto deposit = [&](int e, float c, float x0, float x1, float p0, float p1, double uu,
                     float mm) {
    if (!UNI) sc[e] = c;
    if (STAGE <= 1) { // seeds x0 = u, x1 = -m u^3 (FP64 -> FP32)
      const float s0 = (float)uu, s1 = (float)(-(double)mm * uu * uu * uu);
      if (STAGE == 0) {
        x0 = s0;
        x1 = s1;
      } else {
        p0 = s0;
        p1 = s1;
      }
    }
    if (STAGE == 0) {
      sv[e] = make_float2(x0, x1);
    } else {
      sv[e] = make_float2((x0 - cA[0] * p0) * cB[0], (x1 - cA[1] * p1) * cB[1]);
      sp[e] = make_float2(p0, p1);
    }
  };
  // Lanczos scalars of the previous pass; formed after the first window loads are
  // issued, so the reduction's latency overlaps theirs
  auto coefs = [&]() {
    if (STAGE >= 1) {
      reduce_partials(a, T, (j - 1) & 1, red);
#pragma unroll
      for (int s = 0; s < NSEQ; ++s) {
        const double q0 = red[s * PN], q1 = red[s * PN + 1], q2 = red[s * PN + 2];
        if (STAGE == 1) {
          const double b0 = sqrt(q0);
          const double al = q0 > 0.0 ? q1 / q0 : 0.0;
          const double b1sq = q0 > 0.0 ? (q2 - q1 * q1 / q0) / q0 : 0.0;
          const double b1 = (b1sq > 1e-14 * (q0 > 0.0 ? q2 / q0 : 0.0) && b1sq > 0.0) ? sqrt(b1sq) : 0.0;
          cA[s] = (float)al;
          cB[s] = (b0 > 0.0 && b1 > 0.0) ? (float)(1.0 / (b0 * b1)) : 0.0f;
          cC[s] = b0 > 0.0 ? (float)(b1 / b0) : 0.0f;
          if (cm.y == 0 && threadIdx.x == 0) {
            a.scal[cm.x].b0[s] = b0;
            a.scal[cm.x].alpha[s][0] = al;
            a.scal[cm.x].beta[s][1] = b1;
          }
        } else {
          double al, b;
          lanczos_scalars(q0, q1, q2, al, b);
          cA[s] = (float)al;
          cB[s] = b > 0.0 ? (float)(1.0 / b) : 0.0f;
          cC[s] = (float)b;
          if (cm.y == 0 && threadIdx.x == 0) {
            a.scal[cm.x].alpha[s][j - 1] = al;
            a.scal[cm.x].beta[s][j] = b;
          }
        }
      }
    }
  };
  // Window cells outside the trajectory (i < 0 or i >= N) are deposited as zeros, so the
  // stencil reads all four window neighbours unconditionally (a missing one has weight 0
  // and value 0).  Loads use clamped indices and are all issued: their number, and the
  // number of stores of a full sub-chunk, are fixed, so waiting for the prefetched window
  // does not have to wait for the stores issued after it.
  auto clampi = [N](int i) { return i < 0 ? 0 : (i >= N ? N - 1 : i); };
  // first window [-n, SUB + n), at most 2 SUB cells (2n <= SUB): the loads of both halves
  // are in flight before any deposit
  {
    float rc[2][MSPT], x0[2][MSPT], x1[2][MSPT], p0[2][MSPT], p1[2][MSPT], rm[2][MSPT];
    double ru[2][MSPT];
    bool ok[2][MSPT];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int q = 0; q < MSPT; ++q) {
        const int e = r * MSUB + threadIdx.x + q * FBS, i = cell(e - n);
        ok[r][q] = i >= 0 && i < N;
        if (e < MSUB + 2 * n)
          fetch(clampi(i), rc[r][q], x0[r][q], x1[r][q], p0[r][q], p1[r][q], ru[r][q], rm[r][q]);
      }
    coefs();
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int q = 0; q < MSPT; ++q) {
        const int e = r * MSUB + threadIdx.x + q * FBS;
        if (e < MSUB + 2 * n) {
          if (!ok[r][q])
            rc[r][q] = x0[r][q] = x1[r][q] = p0[r][q] = p1[r][q] = rm[r][q] = 0.0f, ru[r][q] = 0.0;
          deposit(e, rc[r][q], x0[r][q], x1[r][q], p0[r][q], p1[r][q], ru[r][q], rm[r][q]);
        }
      }
  }
  // Grid column of this thread's cell at local coordinate l = cb + tid + kk * FBS, from
  // the column of l = tid and the offset (cb + kk * FBS) mod n, which is the same for the
  // whole workgroup (no per-cell division).
  const int col_t = cell(threadIdx.x) - row_of(cell(threadIdx.x), n, rn) * n;
  auto column = [&](int off_mod_n) {
    int c = col_t + dir * off_mod_n;
    c = c >= n ? c - n : c;
    return c < 0 ? c + n : c;
  };
  // stencil, stores and partials of the cells [cb, ce); FULL: ce = cb + SUB
  auto compute = [&](auto full_t, int cb, int ce) {
    constexpr bool FULL = decltype(full_t)::value;
    const int cbm = cb % n;
#pragma unroll
    for (int kk = 0; kk < MSPT; ++kk) {
      const int l = cb + threadIdx.x + kk * FBS;
      if (FULL || l < ce) {
        const int i = cell(l), e = l - cb + n;
        int om = cbm + (kk * FBS) % n;
        om = om >= n ? om - n : om;
        const int col = column(om);
        Sten5 w;
        w.hu = i >= n;
        w.hl = col > 0;
        w.hr = col < n - 1;
        w.hd = i < N - n;
        if (UNI) {
          const float wc = 0.5f * (T.cval + T.cval); // (c + c) / 2, as the field path
          w.wu = w.hu ? wc : 0.0f;
          w.wl = w.hl ? wc : 0.0f;
          w.wr = w.hr ? wc : 0.0f;
          w.wd = w.hd ? wc : 0.0f;
        } else {
          const float ci = sc[e];
          w.wu = w.hu ? 0.5f * (ci + sc[e + oU]) : 0.0f;
          w.wl = w.hl ? 0.5f * (ci + sc[e + oL]) : 0.0f;
          w.wr = w.hr ? 0.5f * (ci + sc[e + oR]) : 0.0f;
          w.wd = w.hd ? 0.5f * (ci + sc[e + oD]) : 0.0f;
        }
        const float2 vc = sv[e], vU = sv[e + oU], vL = sv[e + oL], vR = sv[e + oR],
                     vD = sv[e + oD];
        const float2 pc = STAGE == 0 ? make_float2(0.0f, 0.0f) : sp[e];
#pragma unroll
        for (int s = 0; s < NSEQ; ++s) {
          // apply5z's sum, term by term in the same order
          const float vi = s == 0 ? vc.x : vc.y;
          float Av = w.wu * (vi - (s == 0 ? vU.x : vU.y));
          Av += w.wl * (vi - (s == 0 ? vL.x : vL.y));
          Av += w.wr * (vi - (s == 0 ? vR.x : vR.y));
          Av += w.wd * (vi - (s == 0 ? vD.x : vD.y));
          if (STAGE == 0) {
            at32(s == 0 ? Wout0 : Wout1, (uint32_t)i) = Av;
            acc[s * PN + 0] += (double)vi * vi;
            acc[s * PN + 1] += (double)vi * Av;
            acc[s * PN + 2] += (double)Av * Av;
          } else {
            const float wp = Av - cC[s] * (s == 0 ? pc.x : pc.y);
            if (!LAST) {
              at32(s == 0 ? Vout0 : Vout1, (uint32_t)i) = vi;
              at32(s == 0 ? Wout0 : Wout1, (uint32_t)i) = wp;
            }
            acc[s * PN + 0] += (double)wp * vi;
            acc[s * PN + 1] += (double)wp * wp;
            acc[s * PN + 2] += (double)vi * vi;
          }
        }
      }
    }
  };
  float nc[MSPT], nx0[MSPT], nx1[MSPT], np0[MSPT], np1[MSPT], nm[MSPT];
  double nu[MSPT];
  bool nok[MSPT];
  for (int cb = 0; cb < len; cb += MSUB) {
    const bool more = cb + MSUB < len;
    if (more) {
      // the next window's new cells [cb' + n, cb' + SUB + n)
#pragma unroll
      for (int q = 0; q < MSPT; ++q) {
        const int i = cell(cb + MSUB + n + threadIdx.x + q * FBS);
        nok[q] = i >= 0 && i < N;
        fetch(clampi(i), nc[q], nx0[q], nx1[q], np0[q], np1[q], nu[q], nm[q]);
      }
    }
    lds_barrier();
    if (!more) {
      compute(std::false_type{}, cb, len);
      break;
    }
    compute(std::true_type{}, cb, cb + MSUB);
    lds_barrier(); // every read of this window is done
    // keep [cb' - n, cb' + n) (disjoint source and destination: 2n <= SUB)
    for (int e = threadIdx.x; e < 2 * n; e += FBS) {
      if (!UNI) sc[e] = sc[MSUB + e];
      sv[e] = sv[MSUB + e];
      if (STAGE >= 1 && e >= n) // the previous vector is read at own cells only
        sp[e] = sp[MSUB + e];
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < MSPT; ++q) {
      if (!nok[q]) nc[q] = nx0[q] = nx1[q] = np0[q] = np1[q] = nm[q] = 0.0f, nu[q] = 0.0;
      deposit(2 * n + threadIdx.x + q * FBS, nc[q], nx0[q], nx1[q], np0[q], np1[q], nu[q], nm[q]);
    }
  }
  block_sum_n<NSEQ * PN>(acc, shn);
  if (threadIdx.x == 0) {
    double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * (NSEQ * PN);
#pragma unroll
    for (int k = 0; k < NSEQ * PN; ++k) dst[k] = acc[k];
  }
}

template <int STAGE, bool UNI = false, bool LAST = false>
__global__ __launch_bounds__(FBS) void k_kge_march2d(KgeArgs a, int j, int emax, int alt) {
  extern __shared__ float lds[];
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  march2d_body<STAGE, UNI, LAST>(a, cm, j, emax, alt, lds);
}

// ---------------------------------------------------------------- 3D plane march, stored w'
// A workgroup owns an in-plane range [pa, pb) of one trajectory (flat offsets within a
// plane, at most 256 * CPT cells) and marches through the planes [z0, z1).  Plane z
// needs v_j on the flat window [z n^2 + pa - n, z n^2 + pb + n) for the x and y
// neighbours, and v_j of planes z - 1 and z + 1 at its own cells for the z neighbours.
// In flat coordinates the window's ends hold the previous plane's last row and the next
// plane's first row, which are exactly the cross-plane y couplings of the reference's
// 3D operator (laplacians.hpp:158-218, q6 set); with q6 clear the flags drop them.
// The window lives in LDS (c and v_j of both sequences); each thread keeps, for its own
// cells, v_j and c of planes z - 1, z, z + 1 and the previous vector of planes z, z + 1
// in registers.  The raw inputs of plane z + 2 (own cells) and of the window ends of
// plane z + 1 are in flight while plane z is computed.  Per cell and pass: the same
// 36 B as the 2D row march, plus the window ends (2n cells per plane, shared with the
// neighbouring ranges, which the block map places on the same XCD) and, per z segment,
// the own cells of planes z0 - 1 and z1.
// STAGE 0: j = 0 on the seeds x (stores z = A x); 1: j = 1 (from z and x); 2: j >= 2.
constexpr int M3_HPT = 2; // window-end cells per thread: 2n <= 256 * M3_HPT

// c and v_j (two arrays); without c for uniform fields
inline size_t march3d_lds_bytes(int cpt, int nmax, bool uni = false) {
  return sizeof(float) * (uni ? 2 : 3) * (size_t)(FBS * cpt + 2 * nmax);
}

// Register caps, measured on one box on 40 x 220^3 (tile 4, uniform fields): the general
// stage-2 pass runs uncapped (132 VGPRs, 3 waves per SIMD, no spill): 2.44 ms, against 2.90 ms
// held to 4 waves (128 VGPRs, a 12-byte spill) and 4.54 ms held to 5 (a 120-byte spill).  The
// last pass of the full-T variant is held to 4 waves: 1.46 ms, against 2.25 ms uncapped at
// 126 VGPRs (the same occupancy, scheduled for fewer waves).
// LAST: pass j = m - 1 of the full-T Lanczos variant (partials only, nothing stored).
template <int STAGE, int CPT, bool UNI = false, bool LAST = false>
__global__ __launch_bounds__(FBS, (CPT == 4 && UNI && STAGE == 2 && LAST) ? 4 : 1) void k_kge_march3d(KgeArgs a, int j, int emax) {
  static_assert(!(LAST && STAGE == 0), "the last pass follows pass 0");
  extern __shared__ float lds[];
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double red[NSEQ * PN];
  __shared__ double shn[NSEQ * PN * (FBS / kWave)];
  // v = (X - cA P) * cB when formed;  w' = A v - cC P
  float cA[NSEQ] = {0.f, 0.f}, cB[NSEQ] = {0.f, 0.f}, cC[NSEQ] = {0.f, 0.f};
  if (STAGE >= 1) {
    reduce_partials(a, T, (j - 1) & 1, red);
#pragma unroll
    for (int s = 0; s < NSEQ; ++s) {
      const double q0 = red[s * PN], q1 = red[s * PN + 1], q2 = red[s * PN + 2];
      if (STAGE == 1) {
        const double b0 = sqrt(q0);
        const double al = q0 > 0.0 ? q1 / q0 : 0.0;
        const double b1sq = q0 > 0.0 ? (q2 - q1 * q1 / q0) / q0 : 0.0;
        const double b1 = (b1sq > 1e-14 * (q0 > 0.0 ? q2 / q0 : 0.0) && b1sq > 0.0) ? sqrt(b1sq) : 0.0;
        cA[s] = (float)al;
        cB[s] = (b0 > 0.0 && b1 > 0.0) ? (float)(1.0 / (b0 * b1)) : 0.0f;
        cC[s] = b0 > 0.0 ? (float)(b1 / b0) : 0.0f;
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].b0[s] = b0;
          a.scal[cm.x].alpha[s][0] = al;
          a.scal[cm.x].beta[s][1] = b1;
        }
      } else {
        double al, b;
        lanczos_scalars(q0, q1, q2, al, b);
        cA[s] = (float)al;
        cB[s] = b > 0.0 ? (float)(1.0 / b) : 0.0f;
        cC[s] = (float)b;
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].alpha[s][j - 1] = al;
          a.scal[cm.x].beta[s][j] = b;
        }
      }
    }
  }
  const int n = T.n, n2 = n * n, N = (int)T.N;
  const int64_t cells = a.cells;
  const size_t m2 = (size_t)(a.m - 2);
  // two windows (planes of even and odd z), each c and v_j of both sequences (no c if UNI)
  const int wstride = (UNI ? 2 : 3) * emax;
  auto win = [&](int z, float *&c_, float *&v0_, float *&v1_) {
    float *b_ = lds + (z & 1) * wstride;
    c_ = b_;
    v0_ = b_ + (UNI ? 0 : emax);
    v1_ = v0_ + emax;
  };
  const float *cT = a.cf + T.off, *mT = a.mf + T.off;
  const double *uT = a.u + T.off;
  const float *X0 = a.W + ((size_t)0 * 2 + (STAGE == 1 ? 0 : ((j - 1) & 1))) * cells + T.off;
  const float *X1 = a.W + ((size_t)1 * 2 + (STAGE == 1 ? 0 : ((j - 1) & 1))) * cells + T.off;
  const float *P0 = a.V + (STAGE >= 2 ? (0 * m2 + (size_t)(j - 2)) * cells : 0) + T.off;
  const float *P1 = a.V + (STAGE >= 2 ? (1 * m2 + (size_t)(j - 2)) * cells : 0) + T.off;
  float *Vout0 = a.V + (STAGE >= 1 ? (0 * m2 + (size_t)(j - 1)) * cells : 0) + T.off;
  float *Vout1 = a.V + (STAGE >= 1 ? (1 * m2 + (size_t)(j - 1)) * cells : 0) + T.off;
  float *Wout0 = a.W + ((size_t)0 * 2 + (j & 1)) * cells + T.off;
  float *Wout1 = a.W + ((size_t)1 * 2 + (j & 1)) * cells + T.off;
  const int q6 = a.q6;
  double acc[NSEQ * PN];
#pragma unroll
  for (int k = 0; k < NSEQ * PN; ++k) acc[k] = 0.0;

  const int pen = cm.y % T.npen, zs = cm.y / T.npen;
  const int pa = pen * T.prange;
  const int pb = pa + T.prange < n2 ? pa + T.prange : n2;
  const int z0 = zs * T.zseg;
  const int z1 = z0 + T.zseg < n ? z0 + T.zseg : n;
  const int S = pb - pa;

  // raw inputs of one cell (loads only: they stay in flight until form() uses them)
  struct Raw {
    float c, x0, x1, p0, p1, mm;
    double uu;
  };
  auto fetch = [&](int g, Raw &r) {
    r.c = UNI ? T.cval : cT[g];
    if (STAGE <= 1) {
      r.uu = uT[g];
      r.mm = UNI ? T.mval : mT[g];
    }
    if (STAGE >= 1) {
      r.x0 = X0[g];
      r.x1 = X1[g];
    }
    if (STAGE >= 2) {
      r.p0 = P0[g];
      r.p1 = P1[g];
    }
  };
  // v_j of both sequences and the previous vector P from the raw inputs
  auto form = [&](const Raw &r, float &v0, float &v1, float &q0, float &q1) {
    float x0 = r.x0, x1 = r.x1, p0 = r.p0, p1 = r.p1;
    if (STAGE <= 1) { // seeds x0 = u, x1 = -m u^3 (FP64 -> FP32)
      const float s0 = (float)r.uu, s1 = (float)(-(double)r.mm * r.uu * r.uu * r.uu);
      if (STAGE == 0) {
        x0 = s0;
        x1 = s1;
      } else {
        p0 = s0;
        p1 = s1;
      }
    }
    if (STAGE == 0) {
      v0 = x0;
      v1 = x1;
      q0 = q1 = 0.0f;
    } else {
      v0 = (x0 - cA[0] * p0) * cB[0];
      v1 = (x1 - cA[1] * p1) * cB[1];
      q0 = p0;
      q1 = p1;
    }
  };

  // own cells: in-plane offset, column flags
  int po[CPT];
  bool own[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    po[q] = pa + threadIdx.x + q * FBS;
    own[q] = threadIdx.x + q * FBS < S;
  }
  // planes z - 1, z, z + 1 of the own cells
  float vm0[CPT], vm1[CPT], vc0[CPT], vc1[CPT], vn0[CPT], vn1[CPT];
  float cmz[CPT], ccz[CPT], cnz[CPT], pc0[CPT], pc1[CPT], pn0[CPT], pn1[CPT];
  Raw ro[CPT];        // own cells of plane z + 2 (in flight)
  Raw rh[M3_HPT];     // window ends of plane z + 2 (in flight)
  // window end h in [0, 2n): window index e and flat cell of plane z
  auto hwin = [&](int h) { return h < n ? h : S + h; };
  auto hcell = [&](int h, int z) { return z * n2 + pa - n + hwin(h); };
  auto fetch_own = [&](int z) {
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (own[q]) fetch(z * n2 + po[q], ro[q]);
  };
  auto fetch_halo = [&](int z) {
#pragma unroll
    for (int k = 0; k < M3_HPT; ++k) {
      const int h = threadIdx.x + k * FBS;
      if (h < 2 * n) {
        const int g = hcell(h, z);
        if (g >= 0 && g < N) fetch(g, rh[k]);
      }
    }
  };
  // window ends of plane z from rh into plane z's window
  auto put_halo = [&](int z) {
    float *hc, *h0, *h1;
    win(z, hc, h0, h1);
#pragma unroll
    for (int k = 0; k < M3_HPT; ++k) {
      const int h = threadIdx.x + k * FBS;
      if (h < 2 * n) {
        const int g = hcell(h, z);
        if (g >= 0 && g < N) {
          const int e = hwin(h);
          float t0, t1;
          if (!UNI) hc[e] = rh[k].c;
          form(rh[k], h0[e], h1[e], t0, t1);
        }
      }
    }
  };

  if (z0 < z1) {
    // prologue: plane z0 - 1 and plane z0 formed, plane z0 + 1 and the ends of z0 in flight
    if (z0 > 0) {
      fetch_own(z0 - 1);
#pragma unroll
      for (int q = 0; q < CPT; ++q)
        if (own[q]) {
          float t0, t1;
          form(ro[q], vm0[q], vm1[q], t0, t1);
          cmz[q] = ro[q].c;
        }
    }
    fetch_own(z0);
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (own[q]) {
        form(ro[q], vc0[q], vc1[q], pc0[q], pc1[q]);
        ccz[q] = ro[q].c;
      }
    if (z0 + 1 < n) fetch_own(z0 + 1);
    fetch_halo(z0);
    put_halo(z0);
    if (z0 + 1 < z1) fetch_halo(z0 + 1);
  }
  // The window ends of plane z + 2 are loaded with the own cells of plane z + 2, two planes
  // ahead: the neighbouring ranges load the same cells as their own cells at about the same
  // time, so one of the two reads is a cache hit.  They are deposited one plane ahead into
  // the other window, which plane z + 1 then completes with its own cells.
  for (int z = z0; z < z1; ++z) {
    lds_barrier(); // plane z - 1's window reads are done: its window takes plane z + 1
    float *sc, *sv0, *sv1;
    win(z, sc, sv0, sv1);
    // window of plane z: own cells from registers (its ends are in place)
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (own[q]) {
        const int e = po[q] - pa + n;
        if (!UNI) sc[e] = ccz[q];
        sv0[e] = vc0[q];
        sv1[e] = vc1[q];
      }
    // ends of plane z + 1 (loaded one plane ago) into the other window
    if (z + 1 < z1) put_halo(z + 1);
    // plane z + 1 of the own cells
    const bool up = z + 1 < n;
    if (up) {
#pragma unroll
      for (int q = 0; q < CPT; ++q)
        if (own[q]) {
          form(ro[q], vn0[q], vn1[q], pn0[q], pn1[q]);
          cnz[q] = ro[q].c;
        }
    }
    // next loads: own cells and window ends of plane z + 2
    if (z + 1 < z1) {
      if (z + 2 < n) fetch_own(z + 2);
      if (z + 2 < z1) fetch_halo(z + 2);
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      if (!own[q]) continue;
      const int p = po[q], e = p - pa + n, g = z * n2 + p;
      const int col = p % n;
      const bool hzm = z > 0, hzp = up;
      const bool hym = q6 ? g >= n : p >= n;
      const bool hyp = q6 ? g + n < N : p < n2 - n;
      const bool hxm = col > 0, hxp = col < n - 1;
      const float ci = ccz[q];
      const float wzm = hzm ? 0.5f * (ci + cmz[q]) : 0.0f;
      const float wym = hym ? 0.5f * (ci + (UNI ? ci : sc[e - n])) : 0.0f;
      const float wxm = hxm ? 0.5f * (ci + (UNI ? ci : sc[e - 1])) : 0.0f;
      const float wxp = hxp ? 0.5f * (ci + (UNI ? ci : sc[e + 1])) : 0.0f;
      const float wyp = hyp ? 0.5f * (ci + (UNI ? ci : sc[e + n])) : 0.0f;
      const float wzp = hzp ? 0.5f * (ci + cnz[q]) : 0.0f;
#pragma unroll
      for (int s = 0; s < NSEQ; ++s) {
        const float *sv = s == 0 ? sv0 : sv1;
        const float vi = s == 0 ? vc0[q] : vc1[q];
        float Av = 0.0f; // flux form sum_nb w (v_i - v_nb), as apply5
        if (hzm) Av += wzm * (vi - (s == 0 ? vm0[q] : vm1[q]));
        if (hym) Av += wym * (vi - sv[e - n]);
        if (hxm) Av += wxm * (vi - sv[e - 1]);
        if (hxp) Av += wxp * (vi - sv[e + 1]);
        if (hyp) Av += wyp * (vi - sv[e + n]);
        if (hzp) Av += wzp * (vi - (s == 0 ? vn0[q] : vn1[q]));
        if (STAGE == 0) {
          __builtin_nontemporal_store(Av, &(s == 0 ? Wout0 : Wout1)[g]);
          acc[s * PN + 0] += (double)vi * vi;
          acc[s * PN + 1] += (double)vi * Av;
          acc[s * PN + 2] += (double)Av * Av;
        } else {
          const float wp = Av - cC[s] * (s == 0 ? pc0[q] : pc1[q]);
          if (!LAST) {
            __builtin_nontemporal_store(vi, &(s == 0 ? Vout0 : Vout1)[g]);
            __builtin_nontemporal_store(wp, &(s == 0 ? Wout0 : Wout1)[g]);
          }
          acc[s * PN + 0] += (double)wp * vi;
          acc[s * PN + 1] += (double)wp * wp;
          acc[s * PN + 2] += (double)vi * vi;
        }
      }
    }
    // roll the planes
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      vm0[q] = vc0[q];
      vm1[q] = vc1[q];
      cmz[q] = ccz[q];
      vc0[q] = vn0[q];
      vc1[q] = vn1[q];
      ccz[q] = cnz[q];
      pc0[q] = pn0[q];
      pc1[q] = pn1[q];
    }
  }
  block_sum_n<NSEQ * PN>(acc, shn);
  if (threadIdx.x == 0) {
    double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * (NSEQ * PN);
#pragma unroll
    for (int k = 0; k < NSEQ * PN; ++k) dst[k] = acc[k];
  }
}

// Combine over the workgroup's range in groups of G cells per thread (independent loads).
// v_{m-1} = (w'_{m-2} - alpha_{m-2} v_{m-2}) / b_{m-1} is folded into the
// coefficients: d = y0 x/b0 + sum_{k<m-2} y_k v_k + (y_{m-2} - y_{m-1} alpha/b) v_{m-2}
// + (y_{m-1}/b) w'_{m-2}, so each stored vector is read exactly once.
// ORTH: also the orthogonality partials <v_{m-2}, v_0> (reorthogonalisation modes 1, 2).
template <int DIM, bool ORTH, bool UNI>
__device__ __forceinline__ void combine_rt_body(const KgeArgs a, const int2 cm) {
  const TrajDev T = a.traj[cm.x];
  const TrajScal &S = a.scal[cm.x];
  __shared__ double cv[NSEQ][kMaxKrylov]; // coefficient of V slot k-1 (k = 1..m-2); [0]: seed
  __shared__ double cw[NSEQ];             // coefficient of w'_{m-2}
  __shared__ double sho[NSEQ * (FBS / kWave)];
  const int m = a.m;
  const bool orth = ORTH && m >= 3;
  double o[NSEQ] = {0.0, 0.0};
  if (threadIdx.x < NSEQ) {
    const int s = threadIdx.x;
    const double ib0 = S.b0[s] > 0.0 ? 1.0 / S.b0[s] : 0.0;
    const double ibl = S.beta[s][m - 1] > 0.0 ? 1.0 / S.beta[s][m - 1] : 0.0;
    const double al = (double)(float)S.alpha[s][m - 2];
    cv[s][0] = S.y[s][0] * ib0;
    for (int k = 1; k <= m - 2; ++k) cv[s][k] = S.y[s][k];
    cv[s][m - 2] -= S.y[s][m - 1] * al * ibl;
    cw[s] = S.y[s][m - 1] * ibl;
  }
  __syncthreads();
  // 2 cells per thread at a time: 118 us per launch at 100 x 300^2, m = 8, against 130 us
  // at 1, 142 us at 4 and 251 us at 8 (more loads in flight per wave run slower); at
  // 40 x 220^3, m = 12, the group size makes no difference (9.54 / 9.56 / 9.48 ms at 2 / 1 / 4)
  constexpr int G = 2;
  const int n = T.n;
  const int64_t cells = a.cells;
  const int64_t cbeg = (int64_t)cm.y * T.range;
  const int64_t cend = cbeg + T.range < T.N ? cbeg + T.range : T.N;
  const double t2 = T.tau * T.tau;
  const float *Wl0 = a.W + ((size_t)0 * 2 + ((m - 2) & 1)) * cells;
  const float *Wl1 = a.W + ((size_t)1 * 2 + ((m - 2) & 1)) * cells;
  // every stream except u is read once per step: non-temporal loads (u is read twice)
  for (int64_t gb = cbeg; gb < cend; gb += G * FBS) {
    int64_t gs[G];
    bool ok[G];
    double d0[G], d1[G];
    float xs0[G], xs1[G]; // seeds at the source cell (orthogonality monitor)
#pragma unroll
    for (int kk = 0; kk < G; ++kk) {
      const int64_t i = gb + threadIdx.x + (int64_t)kk * FBS;
      ok[kk] = i < cend;
      gs[kk] = T.off + (ok[kk] ? clamp_src<DIM>(i, n) : 0);
      const double uu = a.u[gs[kk]];
      const double mm = UNI ? (double)T.mval : (double)__builtin_nontemporal_load(a.mf + gs[kk]);
      const double g1 = -mm * uu * uu * uu;
      d0[kk] = cv[0][0] * (double)(float)uu + cw[0] * (double)__builtin_nontemporal_load(Wl0 + gs[kk]);
      d1[kk] = cv[1][0] * (double)(float)g1 + cw[1] * (double)__builtin_nontemporal_load(Wl1 + gs[kk]);
      xs0[kk] = (float)uu;
      xs1[kk] = (float)g1;
    }
    // with the monitor, v_{m-2} is read after the loop and used twice
#pragma unroll 2
    for (int k = 1; k <= (orth ? m - 3 : m - 2); ++k) {
      const float *Vk0 = a.V + ((size_t)0 * (m - 2) + (k - 1)) * cells;
      const float *Vk1 = a.V + ((size_t)1 * (m - 2) + (k - 1)) * cells;
      const double y0 = cv[0][k], y1 = cv[1][k];
#pragma unroll
      for (int kk = 0; kk < G; ++kk) {
        d0[kk] += y0 * (double)__builtin_nontemporal_load(Vk0 + gs[kk]);
        d1[kk] += y1 * (double)__builtin_nontemporal_load(Vk1 + gs[kk]);
      }
    }
    if (orth) {
      const float *VL0 = a.V + ((size_t)0 * (m - 2) + (m - 3)) * cells;
      const float *VL1 = a.V + ((size_t)1 * (m - 2) + (m - 3)) * cells;
      const double y0 = cv[0][m - 2], y1 = cv[1][m - 2];
#pragma unroll
      for (int kk = 0; kk < G; ++kk) {
        const float l0 = __builtin_nontemporal_load(VL0 + gs[kk]),
                    l1 = __builtin_nontemporal_load(VL1 + gs[kk]);
        d0[kk] += y0 * (double)l0;
        d1[kk] += y1 * (double)l1;
        const int64_t gi = T.off + gb + threadIdx.x + (int64_t)kk * FBS;
        if (ok[kk]) {
          if (gs[kk] == gi) { // interior cell: its own values are the ones just read
            o[0] += (double)xs0[kk] * (double)l0;
            o[1] += (double)xs1[kk] * (double)l1;
          } else {
            orth_partial(a, gi, o);
          }
        }
      }
    }
#pragma unroll
    for (int kk = 0; kk < G; ++kk) {
      const int64_t i = gb + threadIdx.x + (int64_t)kk * FBS;
      if (ok[kk])
        __builtin_nontemporal_store(2.0 * a.u[gs[kk]] - __builtin_nontemporal_load(a.up + gs[kk]) +
                                        2.0 * d0[kk] + t2 * d1[kk],
                                    &a.un[T.off + i]);
    }
  }
  if (orth) {
    block_sum_n<NSEQ>(o, sho);
    if (threadIdx.x == 0) store_orth(a, T, S, cm.y, o);
  }
}

template <int DIM, bool ORTH, bool UNI = false>
__global__ __launch_bounds__(FBS) void k_kge_combine_rt(KgeArgs a) {
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  combine_rt_body<DIM, ORTH, UNI>(a, cm);
}

// ---------------------------------------------------------------- persistent step kernel
// One cooperative launch advances a 2D row-march batch (Gautschi / sEWI) by nsteps steps
// that take no snapshot.  Workgroup b owns chunk cmap[b] in every phase of every step
// (the grid is the chunk map; a cooperative launch keeps every workgroup resident).  Per
// trajectory, an arrival counter orders the phases: pass j (j = 0 .. np - 1) of a chunk
// starts once every chunk of its trajectory has arrived from the phase before; the chunk
// that arrives last from the last pass runs the small stage (waves 0 and 1, one Krylov
// sequence each, Z in the LDS of the pass window) and arrives once more, which releases
// the combine; the combine's arrivals release pass 0 of the next step.  Per step the
// counter rises by (np + 1) * nchunks + 1.  An arrival is an agent-scope release
// read-modify-write after a workgroup barrier; a wait polls with relaxed loads and ends
// with one agent-scope acquire fence.  Waits are bounded: one that gives up sets the
// error flag, and every other wait then ends too, so every workgroup exits.
constexpr int kCntStride = 64;          // ints between two trajectories' counters (256 B)
constexpr int64_t kSpinMax = 1ll << 21; // poll rounds before a wait gives up (seconds)
// Poll rounds between two reads of the host's abort flag in a wait: the flag is in host
// memory, and a read of it in every round by every waiting workgroup made the step 3.3x
// slower (1.05 -> 3.5 ms on 100 x 300^2).
constexpr int64_t kAbortPoll = 4096;
// Waves per SIMD the kernel is compiled for: every phase inlined needs 203-243 VGPRs
// uncapped (2 waves, 512 resident workgroups); capped at 128 VGPRs (4 waves, 1024
// workgroups) it spills 252-340 bytes per lane, nearly all outside the phases' loops.
constexpr int kPersistWaves = 4;

struct PersistArgs {
  double *state[3]; // the FP64 state ring
  int *cnt;         // [ntraj * kCntStride], zero at launch
  int *err;         // 1: a wait gave up; 2: the host set the abort flag
  const int *abort; // pinned host memory: nonzero ends every wait
  int64_t step0;    // steps taken before this launch (position in the ring)
  int nsteps;
  int emax;         // LDS window cells of the pass
  int last_pass;    // m - 2 (reference variant) or m - 1 (full T)
};

// step_start: also read the error flag before polling (once per step), so an abort ends
// every workgroup at its next step even where no counter falls short; watch: this
// workgroup reads the host's abort flag there too (one workgroup per launch does, and
// turns the abort into the error flag for the others).
__device__ inline bool persist_wait(const int *cnt, int target, const PersistArgs &q, int *flag,
                                    bool step_start = false, bool watch = false) {
  if (threadIdx.x == 0) {
    // code: 1 a wait gave up, 2 the host set the abort flag, 0 the error flag is already set
    int ok = 0, code = 1;
    bool stop = false;
    if (step_start) {
      if (__hip_atomic_load(q.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        stop = true;
        code = 0;
      } else if (watch && __hip_atomic_load(q.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) {
        stop = true;
        code = 2;
      }
    }
    for (int64_t it = 0; it < kSpinMax && !stop; ++it) {
      if (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) {
        ok = 1;
        break;
      }
      // the flags are read only once the counter has not reached the target; the host's
      // flag once every kAbortPoll rounds
      if (__hip_atomic_load(q.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        code = 0;
        break;
      }
      if ((it & (kAbortPoll - 1)) == kAbortPoll - 1 &&
          __hip_atomic_load(q.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) {
        code = 2;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    else if (code != 0) __hip_atomic_store(q.err, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *flag = ok;
  }
  __syncthreads();
  return *flag != 0;
}

// Arrival of a chunk at the end of a phase (after the workgroup barrier that completes
// every store of the phase): agent-scope release, or acquire-release for the last pass,
// whose last arrival also acquires the other chunks' partials.  Returns, in thread 0,
// whether this was the last arrival of the phase (target: the count it completes).
__device__ inline bool persist_arrive(int *cnt, bool acq, int target) {
  __syncthreads();
  bool last = false;
  if (threadIdx.x == 0) {
    const int old = acq ? __hip_atomic_fetch_add(cnt, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT)
                        : __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    last = old == target - 1;
  }
  return last;
}

// An opaque copy of the (uniform) chunk descriptor: what a phase derives from it cannot be
// hoisted out of the step loop, so the values of all phases are not live at once (203
// against 223 VGPRs uncapped; under the 4-wave cap 252-340 against 336-424 bytes of spill).
__device__ inline int2 opaque_chunk(int2 v) {
  asm volatile("" : "+s"(v.x), "+s"(v.y));
  return v;
}

template <bool UNI, bool ORTH>
__global__ __launch_bounds__(FBS, kPersistWaves) void k_kge_persist2d(KgeArgs a0, PersistArgs q) {
  extern __shared__ float lds[];
  __shared__ int wait_ok, last_arriver;
  const int2 cm = a0.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const int C = a0.traj[cm.x].nchunks;
  int *cnt = q.cnt + (size_t)cm.x * kCntStride;
  const int np = q.last_pass + 1; // passes per step (at least 3: m >= 3)
  const bool full = q.last_pass == a0.m - 1;
  const int per_step = (np + 1) * C + 1;
  const bool watch = cm.x == 0 && cm.y == 0; // reads the host's abort flag at each step start
  for (int st = 0; st < q.nsteps; ++st) {
    KgeArgs a = a0;
    const int cur = (int)((q.step0 + st) % 3);
    a.u = q.state[cur];
    a.up = q.state[(cur + 2) % 3];
    a.un = q.state[(cur + 1) % 3];
    const int base = st * per_step;
    // pass 0 on the seeds, pass 1, the general passes 2 .. np - 2
    if (!persist_wait(cnt, base, q, &wait_ok, true, watch)) return;
    march2d_body<0, UNI, false>(a, opaque_chunk(cm), 0, q.emax, 1, lds);
    persist_arrive(cnt, false, 0);
    if (!persist_wait(cnt, base + C, q, &wait_ok)) return;
    march2d_body<1, UNI, false>(a, opaque_chunk(cm), 1, q.emax, 1, lds);
    persist_arrive(cnt, false, 0);
    for (int j = 2; j < np - 1; ++j) {
      if (!persist_wait(cnt, base + j * C, q, &wait_ok)) return;
      march2d_body<2, UNI, false>(a, opaque_chunk(cm), j, q.emax, 1, lds);
      persist_arrive(cnt, false, 0);
    }
    // the last pass: j = m - 1 of the full T (partials only) or j = m - 2
    const int jl = np - 1;
    if (!persist_wait(cnt, base + jl * C, q, &wait_ok)) return;
    if (full) march2d_body<2, UNI, true>(a, opaque_chunk(cm), jl, q.emax, 1, lds);
    else march2d_body<2, UNI, false>(a, opaque_chunk(cm), jl, q.emax, 1, lds);
    const bool last = persist_arrive(cnt, true, base + np * C);
    if (threadIdx.x == 0) last_arriver = last;
    __syncthreads();
    if (last_arriver) {
      // small stage: wave w < NSEQ for sequence w, Z rows [kMaxKrylov w, kMaxKrylov (w + 1))
      // in the pass window's LDS
      const int w = threadIdx.x / kWave;
      double(*Z)[kMaxKrylov + 1] = reinterpret_cast<double(*)[kMaxKrylov + 1]>(lds);
      if (w < NSEQ) kge_small_wave(a, opaque_chunk(cm).x, w, q.last_pass, Z + w * kMaxKrylov);
      persist_arrive(cnt, false, 0);
    }
    if (!persist_wait(cnt, base + np * C + 1, q, &wait_ok)) return;
    combine_rt_body<2, ORTH, UNI>(a, opaque_chunk(cm));
    persist_arrive(cnt, false, 0);
  }
}

inline size_t persist2d_lds_bytes(int nmax, bool uni) {
  return std::max(march2d_lds_bytes(nmax, uni),
                  sizeof(double) * NSEQ * kMaxKrylov * (kMaxKrylov + 1));
}

// Snapshot combine over the workgroup's range: also the velocity (u+ - u)/tau before
// the ghost copy and both copies of u+ (same folded coefficients as above).
template <int DIM, bool ORTH, bool UNI = false>
__global__ __launch_bounds__(FBS) void k_kge_combine_snap(KgeArgs a, SnapArgs snap) {
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  const TrajScal &S = a.scal[cm.x];
  __shared__ double cv[NSEQ][kMaxKrylov];
  __shared__ double cw[NSEQ];
  __shared__ double sho[NSEQ * (FBS / kWave)];
  const int m = a.m;
  const bool orth = ORTH && m >= 3;
  double o[NSEQ] = {0.0, 0.0};
  if (threadIdx.x < NSEQ) {
    const int s = threadIdx.x;
    const double ib0 = S.b0[s] > 0.0 ? 1.0 / S.b0[s] : 0.0;
    const double ibl = S.beta[s][m - 1] > 0.0 ? 1.0 / S.beta[s][m - 1] : 0.0;
    const double al = (double)(float)S.alpha[s][m - 2];
    cv[s][0] = S.y[s][0] * ib0;
    for (int k = 1; k <= m - 2; ++k) cv[s][k] = S.y[s][k];
    cv[s][m - 2] -= S.y[s][m - 1] * al * ibl;
    cw[s] = S.y[s][m - 1] * ibl;
  }
  __syncthreads();
  const int n = T.n;
  const int64_t cells = a.cells;
  const double t2 = T.tau * T.tau;
  auto unew = [&](int64_t g) {
    const double uu = a.u[g];
    const double g1 = -(UNI ? (double)T.mval : (double)a.mf[g]) * uu * uu * uu;
    double d0 = cv[0][0] * (double)(float)uu + cw[0] * (double)a.W[((size_t)0 * 2 + ((m - 2) & 1)) * cells + g];
    double d1 = cv[1][0] * (double)(float)g1 + cw[1] * (double)a.W[((size_t)1 * 2 + ((m - 2) & 1)) * cells + g];
    for (int k = 1; k <= m - 2; ++k) {
      d0 += cv[0][k] * (double)a.V[((size_t)0 * (m - 2) + (k - 1)) * cells + g];
      d1 += cv[1][k] * (double)a.V[((size_t)1 * (m - 2) + (k - 1)) * cells + g];
    }
    return 2.0 * uu - a.up[g] + 2.0 * d0 + t2 * d1;
  };
  const int64_t cbeg = (int64_t)cm.y * T.range;
  const int64_t cend = cbeg + T.range < T.N ? cbeg + T.range : T.N;
  for (int64_t i = cbeg + threadIdx.x; i < cend; i += FBS) {
    const int64_t src = clamp_src<DIM>(i, n);
    const int64_t gi = T.off + i, gs = T.off + src;
    const double us = unew(gs);
    a.un[gi] = us;
    const double ui = src == i ? us : unew(gi);
    snap.u[gi] = us;
    snap.v[gi] = (ui - a.u[gi]) / T.tau;
    if (orth) orth_partial(a, gi, o);
  }
  if (orth) {
    block_sum_n<NSEQ>(o, sho);
    if (threadIdx.x == 0) store_orth(a, T, S, cm.y, o);
  }
}

// ---------------------------------------------------------------- Stormer-Verlet step
// kg_single.cuh:129-179 on the batch, FP64 throughout:
//   u+ = ((2u - u-) + tau^2 (L_c - mu2) u) + tau^2 (-m u^3),   v = (u+ - u) / tau,
// with the operator row formed as the reference assembles it (laplacians.hpp:54-103 in 2D,
// :158-218 in 3D incl. the Q6 couplings: weights (c_a + c_b) / 2, diagonal -(sum in the
// loop order) times 1/dx^2, then the shift) and summed in ascending column order as the CSR
// SpMV sums it, without contraction, so the result is the reference arithmetic's.  The
// ghost ring takes its source cell's u+ (the combine's clamp); the velocity snapshot is
// taken before the copy.  UNI: m and c from the descriptor.
template <int DIM, bool UNI>
__global__ __launch_bounds__(FBS) void k_kge_sv(KgeArgs a, SnapArgs snap) {
#pragma clang fp contract(off)
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  // 32-bit cell indices and quotients from a float reciprocal corrected to the exact
  // value (row_of; every trajectory of the batch holds fewer than 2^24 cells)
  const int n = T.n, n2 = n * n, N = (int)T.N;
  const float rn = 1.0f / (float)n, rn2 = 1.0f / (float)n2;
  const double *u = a.u + T.off;
  const float *cT = a.cf + T.off;
  const double sc = T.sd, shift = -T.mu2d, t2 = T.tau * T.tau;
  auto C = [&](int r) { return UNI ? (double)T.cval : (double)cT[r]; };
  auto unew = [&](int r) {
    const double cr = C(r);
    double acc = 0.0;
    if (DIM == 2) {
      const int i = row_of(r, n, rn), j = r - i * n;
      const bool hl = j > 0, hr = j < n - 1, hu = i > 0, hd = i < n - 1;
      double wl = 0.0, wr = 0.0, wu = 0.0, wd = 0.0, dsum = 0.0;
      if (hl) wl = (C(r - 1) + cr) / 2.0;
      if (hr) wr = (cr + C(r + 1)) / 2.0;
      if (hu) wu = (C(r - n) + cr) / 2.0;
      if (hd) wd = (cr + C(r + n)) / 2.0;
      if (hl) dsum += wl;
      if (hr) dsum += wr;
      if (hu) dsum += wu;
      if (hd) dsum += wd;
      if (hu) acc += (wu * sc) * u[r - n];
      if (hl) acc += (wl * sc) * u[r - 1];
      acc += ((-dsum * sc) + shift) * u[r];
      if (hr) acc += (wr * sc) * u[r + 1];
      if (hd) acc += (wd * sc) * u[r + n];
    } else {
      const int jp = r - row_of(r, n2, rn2) * n2, i = r - row_of(r, n, rn) * n;
      const bool hxm = i > 0, hxp = i < n - 1;
      const bool hym = a.q6 ? r >= n : jp >= n, hyp = a.q6 ? r + n < N : jp < n2 - n;
      const bool hzm = r >= n2, hzp = r + n2 < N;
      double wxm = 0.0, wxp = 0.0, wym = 0.0, wyp = 0.0, wzm = 0.0, wzp = 0.0, dsum = 0.0;
      if (hxm) wxm = (C(r - 1) + cr) / 2.0;
      if (hxp) wxp = (cr + C(r + 1)) / 2.0;
      if (hym) wym = (C(r - n) + cr) / 2.0;
      if (hyp) wyp = (cr + C(r + n)) / 2.0;
      if (hzm) wzm = (C(r - n2) + cr) / 2.0;
      if (hzp) wzp = (cr + C(r + n2)) / 2.0;
      if (hxm) dsum += wxm;
      if (hxp) dsum += wxp;
      if (hym) dsum += wym;
      if (hyp) dsum += wyp;
      if (hzm) dsum += wzm;
      if (hzp) dsum += wzp;
      if (hzm) acc += (wzm * sc) * u[r - n2];
      if (hym) acc += (wym * sc) * u[r - n];
      if (hxm) acc += (wxm * sc) * u[r - 1];
      acc += ((-dsum * sc) + shift) * u[r];
      if (hxp) acc += (wxp * sc) * u[r + 1];
      if (hyp) acc += (wyp * sc) * u[r + n];
      if (hzp) acc += (wzp * sc) * u[r + n2];
    }
    const double ur = u[r];
    const double mm = UNI ? (double)T.mval : (double)a.mf[T.off + r];
    const double g = -mm * ur * ur * ur;
    double x = 2.0 * ur - a.up[T.off + r];
    x = x + t2 * acc;
    return x + t2 * g;
  };
  const int cbeg = cm.y * T.range;
  const int cend = cbeg + T.range < N ? cbeg + T.range : N;
  auto cl = [n](int x) { return x < 1 ? 1 : (x > n - 2 ? n - 2 : x); };
  auto source = [&](int i) { // Neumann ghost copy (boundaries.cuh:9-81), as clamp_src
    if (DIM == 2) {
      const int r = row_of(i, n, rn);
      return cl(r) * n + cl(i - r * n);
    }
    const int z = row_of(i, n2, rn2), rem = i - z * n2, y = row_of(rem, n, rn);
    return (cl(z) * n + cl(y)) * n + cl(rem - y * n);
  };
  // G cells per thread at a time, every load before the stores (independent loads)
  constexpr int G = 4;
  for (int i0 = cbeg + (int)threadIdx.x; i0 < cend; i0 += G * FBS) {
    double us[G], ui[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const int i = i0 + k * FBS;
      if (i < cend) {
        const int src = source(i);
        us[k] = unew(src);
        if (snap.u) ui[k] = src == i ? us[k] : unew(i);
      }
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const int i = i0 + k * FBS;
      if (i < cend) {
        a.un[T.off + i] = us[k];
        if (snap.u) {
          snap.u[T.off + i] = us[k];
          snap.v[T.off + i] = (ui[k] - u[i]) / T.tau;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- full reorthogonalisation
// Classical Gram-Schmidt of the Lanczos residual against the earlier basis vectors, with
// FP64 dot products and coefficients (the reference makes one such pass in every
// iteration, lanczos.hpp:249-250).  Pass j has stored w'_j = A v_j - b_j v_{j-1}; the
// next pass forms v_{j+1} from r = w'_j - alpha_j v_j.  Here r is made orthogonal to
// v_0 .. v_{j-1}:
//   h_k = <r, v_k> = <w'_j, v_k> - alpha_j <v_j, v_k>,   w'_j <- w'_j - sum_{k<j} h_k v_k,
// and the pass-j partials are recomputed from the corrected w'_j, so the next pass's
// alpha_j and b_{j+1} (and the combine's folded v_{m-1}) follow it.  v_0 = x / b0 is formed
// from the seed x.  Both kernels run over the combine's flat ranges and return at once
// for trajectories where reorthogonalisation is not active.

// Per-chunk partials <w'_j, v_k> and <v_j, v_k> for k in [4 blockIdx.y, 4 blockIdx.y + 4),
// k < j.  The first workgroup of a trajectory also records alpha_j from the pass partials.
__global__ __launch_bounds__(FBS) void k_kge_reorth_dots(KgeArgs a, int j) {
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  TrajScal &S = a.scal[cm.x];
  if (!reorth_active(a, S)) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double sh[NSEQ * 8 * (FBS / kWave)];
  if (cm.y == 0 && blockIdx.y == 0 && threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    const double *p = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0) * (NSEQ * PN);
#pragma unroll
    for (int s = 0; s < NSEQ; ++s) {
      double q0 = 0.0, q2 = 0.0;
      for (int g = lane; g < T.nchunks; g += kWave) {
        q0 += p[(size_t)g * (NSEQ * PN) + s * PN];
        q2 += p[(size_t)g * (NSEQ * PN) + s * PN + 2];
      }
      q0 = wave_sum(q0);
      q2 = wave_sum(q2);
      if (lane == 0) S.ralpha[s] = q2 > 0.0 ? q0 / q2 : 0.0;
    }
  }
  const int k0 = blockIdx.y * 4;
  const int64_t cells = a.cells;
  const size_t m2 = (size_t)(a.m - 2);
  const float *W0 = a.W + ((size_t)0 * 2 + (j & 1)) * cells + T.off;
  const float *W1 = a.W + ((size_t)1 * 2 + (j & 1)) * cells + T.off;
  const float *Vj0 = a.V + (0 * m2 + (size_t)(j - 1)) * cells + T.off;
  const float *Vj1 = a.V + (1 * m2 + (size_t)(j - 1)) * cells + T.off;
  double acc[NSEQ * 8]; // [s][q][w' or v_j]
#pragma unroll
  for (int k = 0; k < NSEQ * 8; ++k) acc[k] = 0.0;
  const int64_t cbeg = (int64_t)cm.y * T.range;
  const int64_t cend = cbeg + T.range < T.N ? cbeg + T.range : T.N;
  for (int64_t i = cbeg + threadIdx.x; i < cend; i += FBS) {
    const double w0 = W0[i], w1 = W1[i], v0 = Vj0[i], v1 = Vj1[i];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + q;
      if (k < j) {
        double x0, x1;
        if (k == 0) {
          basis0(a, S, T.off + i, x0, x1);
        } else {
          x0 = a.V[(0 * m2 + (size_t)(k - 1)) * cells + T.off + i];
          x1 = a.V[(1 * m2 + (size_t)(k - 1)) * cells + T.off + i];
        }
        acc[(0 * 4 + q) * 2 + 0] += w0 * x0;
        acc[(0 * 4 + q) * 2 + 1] += v0 * x0;
        acc[(1 * 4 + q) * 2 + 0] += w1 * x1;
        acc[(1 * 4 + q) * 2 + 1] += v1 * x1;
      }
    }
  }
  block_sum_n<NSEQ * 8>(acc, sh);
  if (threadIdx.x == 0) {
    double *dst = a.part2 + (size_t)(T.chunk0 + cm.y) * NSEQ * 2 * kMaxKrylov;
#pragma unroll
    for (int s = 0; s < NSEQ; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (k0 + q < j) {
          dst[(s * 2 + 0) * kMaxKrylov + k0 + q] = acc[(s * 4 + q) * 2 + 0];
          dst[(s * 2 + 1) * kMaxKrylov + k0 + q] = acc[(s * 4 + q) * 2 + 1];
        }
  }
}

// w'_j <- w'_j - sum_{k<j} h_k v_k and the pass-j partials <w',v_j>, <w',w'>, <v_j,v_j>
// of the corrected w'_j (written over the pass's own).
__global__ __launch_bounds__(FBS) void k_kge_reorth_apply(KgeArgs a, int j) {
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajScal &S = a.scal[cm.x];
  if (!reorth_active(a, S)) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double hk[NSEQ][kMaxKrylov];
  __shared__ double shn[NSEQ * PN * (FBS / kWave)];
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    const double *p2 = a.part2 + (size_t)T.chunk0 * NSEQ * 2 * kMaxKrylov;
    for (int s = 0; s < NSEQ; ++s) {
      const double al = S.ralpha[s];
      for (int k = 0; k < j; ++k) {
        double d1 = 0.0, d2 = 0.0;
        for (int g = lane; g < T.nchunks; g += kWave) {
          d1 += p2[((size_t)g * NSEQ + s) * 2 * kMaxKrylov + k];
          d2 += p2[((size_t)g * NSEQ + s) * 2 * kMaxKrylov + kMaxKrylov + k];
        }
        d1 = wave_sum(d1);
        d2 = wave_sum(d2);
        if (lane == 0) hk[s][k] = d1 - al * d2;
      }
    }
  }
  __syncthreads();
  const int64_t cells = a.cells;
  const size_t m2 = (size_t)(a.m - 2);
  float *W0 = a.W + ((size_t)0 * 2 + (j & 1)) * cells + T.off;
  float *W1 = a.W + ((size_t)1 * 2 + (j & 1)) * cells + T.off;
  const float *Vj0 = a.V + (0 * m2 + (size_t)(j - 1)) * cells + T.off;
  const float *Vj1 = a.V + (1 * m2 + (size_t)(j - 1)) * cells + T.off;
  double acc[NSEQ * PN];
#pragma unroll
  for (int k = 0; k < NSEQ * PN; ++k) acc[k] = 0.0;
  const int64_t cbeg = (int64_t)cm.y * T.range;
  const int64_t cend = cbeg + T.range < T.N ? cbeg + T.range : T.N;
  for (int64_t i = cbeg + threadIdx.x; i < cend; i += FBS) {
    double x0, x1;
    basis0(a, S, T.off + i, x0, x1);
    double c0 = hk[0][0] * x0, c1 = hk[1][0] * x1;
    for (int k = 1; k < j; ++k) {
      c0 += hk[0][k] * (double)a.V[(0 * m2 + (size_t)(k - 1)) * cells + T.off + i];
      c1 += hk[1][k] * (double)a.V[(1 * m2 + (size_t)(k - 1)) * cells + T.off + i];
    }
    const float w0 = (float)((double)W0[i] - c0), w1 = (float)((double)W1[i] - c1);
    W0[i] = w0;
    W1[i] = w1;
    const float v0 = Vj0[i], v1 = Vj1[i];
    acc[0] += (double)w0 * v0;
    acc[1] += (double)w0 * w0;
    acc[2] += (double)v0 * v0;
    acc[PN + 0] += (double)w1 * v1;
    acc[PN + 1] += (double)w1 * w1;
    acc[PN + 2] += (double)v1 * v1;
  }
  block_sum_n<NSEQ * PN>(acc, shn);
  if (threadIdx.x == 0) {
    double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * (NSEQ * PN);
#pragma unroll
    for (int k = 0; k < NSEQ * PN; ++k) dst[k] = acc[k];
  }
}

// ---------------------------------------------------------------- sampled initial data
// Initial condition of one trajectory evaluated at every cell (nls_ic.hpp) and written
// where set_initial would put the host arrays: KGE: u0 into the state ring's current and
// previous slots (v0 = 0), NLSE: psi0 as complex64; m, c constant; snapshot 0 if kept.
__global__ __launch_bounds__(FBS) void k_ic_sample(ic::Params P, int64_t N, int nlse, float mval,
                                                   float cval, double *u, double *up, float *psi,
                                                   float *mf, float *cf, double *snapU,
                                                   double *snapV) {
  const int64_t g = (int64_t)blockIdx.x * FBS + threadIdx.x;
  if (g >= N) return;
  double re, im;
  ic::eval(P, g, re, im);
  mf[g] = mval;
  cf[g] = cval;
  if (nlse) {
    psi[2 * g] = (float)re;
    psi[2 * g + 1] = (float)im;
    if (snapU) {
      snapU[2 * g] = re;
      snapU[2 * g + 1] = im;
    }
  } else {
    u[g] = re;
    up[g] = re;
    if (snapU) {
      snapU[g] = re;
      snapV[g] = 0.0;
    }
  }
}

// ---------------------------------------------------------------- batch object
struct Layout {
  std::vector<TrajDev> traj;
  std::vector<int2> cmap;
  int total_chunks = 0;
  int64_t cells = 0;
};

// XCD-aware block map of the trajectories in `subset`: workgroup b runs on XCD
// b % 8.  The chunks, in trajectory order, are cut into 8 contiguous runs of equal
// length, one per XCD, so neighbouring ranges of a trajectory share an XCD (their halo
// rows meet in that XCD's L2; a trajectory is split between two XCDs at most at the 7
// cut points) and every XCD holds the same number of workgroups (a map that keeps
// trajectories whole leaves XCDs with one trajectory more than others, and those finish
// a full range later).
static std::vector<int2> xcd_cmap(const std::vector<TrajDev> &traj, const std::vector<int> &subset) {
  std::vector<int2> all;
  for (int t : subset)
    for (int c = 0; c < traj[t].nchunks; ++c) all.push_back(int2{t, c});
  const size_t per = (all.size() + NXCD - 1) / NXCD;
  std::vector<int2> cmap(per * NXCD, int2{-1, -1});
  for (size_t k = 0; k < all.size(); ++k) cmap[(k % per) * NXCD + k / per] = all[k];
  return cmap;
}

// Chunking: fixed ranges of 256 * cpt cells (the tiled pass kernels need exactly
// that), or, for the row-march kernels (balanced), the same number of chunks with
// the trajectory split into equal ranges (multiples of 64 cells), so the last
// workgroup of a trajectory is not a short one.  Either way a trajectory's
// chunking depends only on its own size and the tile.
// Plane march (3D): in-plane ranges of at most 256 * cpt cells (equal, multiples of 64),
// and the planes split into equal segments of at most kPlaneSeg planes (each segment
// re-reads the two planes beside it), more of them (down to kPlaneSegMin planes) when a
// trajectory has fewer than kPlaneChunks chunks.  Both depend only on the trajectory's
// own size.
constexpr int kPlaneSeg = 64, kPlaneSegMin = 16, kPlaneChunks = 96;

static Layout make_layout(const nls_traj_desc *td, int ntraj, int dim, int cpt, double dt,
                          double mu2, bool balanced = false, bool plane3d = false) {
  Layout L;
  const int CH = FBS * cpt;
  L.traj.resize(ntraj);
  for (int t = 0; t < ntraj; ++t) {
    TrajDev &T = L.traj[t];
    T.n = td[t].n;
    T.dim = dim;
    T.N = dim == 3 ? (int64_t)T.n * T.n * T.n : (int64_t)T.n * T.n;
    T.npen = 1;
    T.prange = 0;
    T.zseg = T.n;
    if (plane3d) {
      const int n2 = T.n * T.n;
      T.npen = (n2 + CH - 1) / CH;
      const int r = (n2 + T.npen - 1) / T.npen;
      T.prange = std::min(CH, (r + 63) / 64 * 64);
      T.npen = (n2 + T.prange - 1) / T.prange;
      const int nseg = std::max((T.n + kPlaneSeg - 1) / kPlaneSeg,
                                std::min((kPlaneChunks + T.npen - 1) / T.npen, T.n / kPlaneSegMin));
      T.zseg = (T.n + nseg - 1) / nseg;
    }
  }
  for (int t = 0; t < ntraj; ++t) {
    TrajDev &T = L.traj[t];
    T.off = L.cells;
    L.cells += (T.N + 63) / 64 * 64; // keep every trajectory 256-byte aligned
    T.nchunks = (int)((T.N + CH - 1) / CH);
    T.range = CH;
    if (plane3d) T.nchunks = T.npen * ((T.n + T.zseg - 1) / T.zseg);
    if (balanced || plane3d) {
      // the combine's flat ranges (for the plane march: any cover of [0, N))
      const int64_t r = (T.N + T.nchunks - 1) / T.nchunks;
      T.range = (int)((r + 63) / 64 * 64);
      if (!plane3d) T.nchunks = (int)((T.N + T.range - 1) / T.range);
    }
    T.chunk0 = L.total_chunks;
    L.total_chunks += T.nchunks;
    const double dx = 2.0 * td[t].L / (T.n - 1);
    T.sd = 1.0 / (dx * dx);
    T.mu2d = mu2;
    T.tau = dt;
  }
  std::vector<int> all(ntraj);
  std::iota(all.begin(), all.end(), 0);
  L.cmap = xcd_cmap(L.traj, all);
  return L;
}

// Krylov dimension from an a-priori bound: the error of the m-step Lanczos
// approximation of cos(tau sqrt(A)) / sinc^2 behaves like (tau^2 rho)^m / (2m)!
// with rho >= lambda_max(A) (Gershgorin: 4 dim max(c) / dx^2 + mu2).  Smallest m
// in [8, 16] whose bound is below 1e-14.
static int autotune_krylov(const std::vector<double> &rho_tau2) {
  double worst = 0.0;
  for (double r : rho_tau2) worst = std::max(worst, r);
  for (int m = 8; m <= 16; ++m) {
    double lg = m * std::log(std::max(worst, 1e-300)) - std::lgamma(2.0 * m + 1.0);
    if (lg < std::log(1e-14)) return m;
  }
  return 16;
}

// ================================================================ cubic NLSE (Strang)
// Reference step (nlse_dev.hpp:187-197):
//   psi1 = psi exp(i dt/2 m|psi|^2);  psi2 = exp(i dt L) psi1;  psi3 = psi2 exp(i dt/2 m|psi2|^2)
// with the snapshot taken before the ghost copy (nlse_dev.hpp:195-197).  One
// complex64 Krylov sequence; L is real symmetric, so alpha and b are real and
// T is real (the reference's Zdotc-based T is Hermitian with rounding-level
// imaginary parts).  The first pass forms psi1 at each cell and its neighbours
// and stores it (seed X); the combine pass forms psi2 = sum_k y_k v_k, the
// second half phase, the snapshot and the ghost copy.
struct c32 {
  float re, im;
};
__device__ inline c32 operator+(c32 a, c32 b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline c32 operator-(c32 a, c32 b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline c32 operator*(float s, c32 a) { return {s * a.re, s * a.im}; }

struct NlseArgs {
  const TrajDev *traj;
  const int2 *cmap;
  const int *tlist; // small stage: block -> trajectory (nullptr: identity)
  TrajScal *scal; // sequence 0 arrays; y[1] holds Im(y)
  double *part;   // [2 parity][total_chunks][PN]
  int total_chunks;
  const float *mf, *cf;
  const c32 *psi;
  c32 *psi_next;
  c32 *X; // seed psi1
  c32 *V; // [m-2][cells]
  c32 *W; // [2][cells]
  int64_t cells;
  int m, q6;
  double dt;
  int reorth; // as KgeArgs (complex classical Gram-Schmidt)
  double reorth_tol;
  double *part2;   // [total_chunks][4][kMaxKrylov]: <v_k, w'_j>, <v_k, v_j> (re, im)
  double *part3;   // [total_chunks][2]: <v_{m-2}, v_0> (re, im) of the last combine
  int *any_reorth;
};

// sin and cos of the half-step phase x = dt/2 m |psi|^2 (FP64).  The workloads' phases are
// small (|x| <= 0.05 at dt = 0.005), so |x| <= 1/2 takes the Taylor series in nested form,
// 8 terms for sin and cos (truncation below 2e-17 relative at |x| = 1/2, i.e. below FP64
// rounding); larger phases take the library's sincos, with its range reduction.
__device__ inline void phase_sincos(double x, double &sn, double &cs) {
  if (fabs(x) <= 0.5) {
    // term ratios 1 / (k (k + 1)) as constants: a division per term would cost more than
    // the library call
    constexpr double s1 = 1.0 / 6, s2 = 1.0 / 20, s3 = 1.0 / 42, s4 = 1.0 / 72, s5 = 1.0 / 110,
                     s6 = 1.0 / 156, s7 = 1.0 / 210;
    constexpr double c1 = 1.0 / 2, c2 = 1.0 / 12, c3 = 1.0 / 30, c4 = 1.0 / 56, c5 = 1.0 / 90,
                     c6 = 1.0 / 132, c7 = 1.0 / 182;
    const double t = x * x;
    sn = x * (1.0 - t * s1 * (1.0 - t * s2 * (1.0 - t * s3 * (1.0 - t * s4 *
             (1.0 - t * s5 * (1.0 - t * s6 * (1.0 - t * s7)))))));
    cs = 1.0 - t * c1 * (1.0 - t * c2 * (1.0 - t * c3 * (1.0 - t * c4 *
             (1.0 - t * c5 * (1.0 - t * c6 * (1.0 - t * c7))))));
  } else {
    sincos(x, &sn, &cs);
  }
}

__device__ inline c32 nl_half_phase(c32 p, float mval, double dt) {
  const double rho = (double)mval * ((double)p.re * p.re + (double)p.im * p.im);
  const double ang = 0.5 * dt * rho;
  double sn, cs;
  phase_sincos(ang, sn, cs);
  return {(float)(p.re * cs - p.im * sn), (float)(p.re * sn + p.im * cs)};
}

__device__ inline void reduce_partials3(const NlseArgs &a, const TrajDev &T, int parity,
                                        double *red) {
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    double acc[PN] = {0.0, 0.0, 0.0};
    const double *p = a.part + ((size_t)parity * a.total_chunks + T.chunk0) * PN;
    for (int g = lane; g < T.nchunks; g += kWave)
#pragma unroll
      for (int k = 0; k < PN; ++k) acc[k] += p[(size_t)g * PN + k];
#pragma unroll
    for (int k = 0; k < PN; ++k) {
      const double v = wave_sum(acc[k]);
      if (lane == 0) red[k] = v;
    }
  }
  __syncthreads();
}

// LAST: pass j = m - 1 of the full-T Lanczos variant (partials only, nothing stored).
template <int DIM, int STAGE, int CPT, bool LAST = false>
__global__ __launch_bounds__(FBS) void k_nlse_pass(NlseArgs a, int j) {
  static_assert(!(LAST && STAGE == 0), "the last pass follows pass 0");
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double red[PN];
  __shared__ double sh[FBS / kWave];
  float cA = 0.f, cB = 0.f, cC = 0.f;
  if (STAGE >= 1) {
    reduce_partials3(a, T, (j - 1) & 1, red);
    const double q0 = red[0], q1 = red[1], q2 = red[2];
    if (STAGE == 1) {
      const double b0 = sqrt(q0);
      const double al = q0 > 0.0 ? q1 / q0 : 0.0;
      const double b1sq = q0 > 0.0 ? (q2 - q1 * q1 / q0) / q0 : 0.0;
      const double b1 = (b1sq > 1e-14 * (q0 > 0.0 ? q2 / q0 : 0.0) && b1sq > 0.0) ? sqrt(b1sq) : 0.0;
      cA = (float)al;
      cB = (b0 > 0.0 && b1 > 0.0) ? (float)(1.0 / (b0 * b1)) : 0.0f;
      cC = b0 > 0.0 ? (float)(b1 / b0) : 0.0f;
      if (cm.y == 0 && threadIdx.x == 0) {
        a.scal[cm.x].b0[0] = b0;
        a.scal[cm.x].alpha[0][0] = al;
        a.scal[cm.x].beta[0][1] = b1;
      }
    } else {
      double al, b;
      lanczos_scalars(q0, q1, q2, al, b);
      cA = (float)al;
      cB = b > 0.0 ? (float)(1.0 / b) : 0.0f;
      cC = (float)b;
      if (cm.y == 0 && threadIdx.x == 0) {
        a.scal[cm.x].alpha[0][j - 1] = al;
        a.scal[cm.x].beta[0][j] = b;
      }
    }
  }
  const int n = T.n;
  const int64_t cells = a.cells;
  double acc[PN] = {0.0, 0.0, 0.0};
  const int64_t base = (int64_t)cm.y * (FBS * CPT);
#pragma unroll
  for (int kk = 0; kk < CPT; ++kk) {
    const int64_t i = base + threadIdx.x + (int64_t)kk * FBS;
    if (i >= T.N) break;
    const int64_t gi = T.off + i;
    Nbr nb;
    neighbours<DIM>(i, n, T.N, a.q6, nb);
    const float ci = a.cf[gi];
    float w[6];
#pragma unroll
    for (int q = 0; q < 2 * DIM; ++q) w[q] = nb.has[q] ? 0.5f * (ci + a.cf[gi + nb.off[q]]) : 0.0f;
    auto val = [&](int64_t g) -> c32 {
      if (STAGE == 0) return nl_half_phase(a.psi[g], a.mf[g], a.dt);
      if (STAGE == 1) return cB * (a.W[g] - cA * a.X[g]);
      return cB * (a.W[(size_t)((j - 1) & 1) * cells + g] - cA * a.V[(size_t)(j - 2) * cells + g]);
    };
    const c32 vi = val(gi);
    // B v = sum_nb w (v_nb - v_i): div(c grad) without 1/dx^2, not negated
    // (nlse_cubic_driver_2d.cpp:94-97), in flux form (as apply5)
    c32 Lv{0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 2 * DIM; ++q)
      if (nb.has[q]) Lv = Lv + w[q] * (val(gi + nb.off[q]) - vi);
    if (STAGE == 0) {
      a.X[gi] = vi;
      a.W[gi] = Lv;
      acc[0] += (double)vi.re * vi.re + (double)vi.im * vi.im;
      acc[1] += (double)vi.re * Lv.re + (double)vi.im * Lv.im;
      acc[2] += (double)Lv.re * Lv.re + (double)Lv.im * Lv.im;
    } else {
      const c32 vprev = STAGE == 1 ? a.X[gi] : a.V[(size_t)(j - 2) * cells + gi];
      const c32 wp = Lv - cC * vprev;
      if (!LAST) {
        a.V[(size_t)(j - 1) * cells + gi] = vi;
        a.W[(size_t)(j & 1) * cells + gi] = wp;
      }
      acc[0] += (double)wp.re * vi.re + (double)wp.im * vi.im;
      acc[1] += (double)wp.re * wp.re + (double)wp.im * wp.im;
      acc[2] += (double)vi.re * vi.re + (double)vi.im * vi.im;
    }
  }
  double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * PN;
#pragma unroll
  for (int k = 0; k < PN; ++k) {
    const double t = block_sum<FBS>(acc[k], sh);
    if (threadIdx.x == 0) dst[k] = t;
  }
}

__global__ __launch_bounds__(kWave) void k_nlse_small(NlseArgs a, int last_pass) {
  const int t = a.tlist ? a.tlist[blockIdx.x] : (int)blockIdx.x;
  const TrajDev T = a.traj[t];
  __shared__ double red[PN];
  const int m = a.m;
  {
    const int lane = threadIdx.x;
    double acc[PN] = {0.0, 0.0, 0.0};
    const double *p = a.part + ((size_t)(last_pass & 1) * a.total_chunks + T.chunk0) * PN;
    for (int g = lane; g < T.nchunks; g += kWave)
      for (int k = 0; k < PN; ++k) acc[k] += p[(size_t)g * PN + k];
    for (int k = 0; k < PN; ++k) {
      const double v = wave_sum(acc[k]);
      if (lane == 0) red[k] = v;
    }
  }
  __syncthreads();
  TrajScal &S = a.scal[t];
  const int lane = threadIdx.x;
  if (a.reorth != 0 && m >= 3) {
    // orthogonality of the previous step's basis, |<v_{m-2}, v_0>| (combine partials)
    double o0 = 0.0, o1 = 0.0;
    for (int g = lane; g < T.nchunks; g += kWave) {
      o0 += a.part3[(size_t)(T.chunk0 + g) * 2];
      o1 += a.part3[(size_t)(T.chunk0 + g) * 2 + 1];
    }
    o0 = wave_sum(o0);
    o1 = wave_sum(o1);
    if (lane == 0) {
      const double o = sqrt(o0 * o0 + o1 * o1);
      S.orth[0] = o;
      if (o > S.max_orth[0]) S.max_orth[0] = o;
      if (a.reorth == 2 && o > a.reorth_tol) {
        S.reorth_on = 1;
        *a.any_reorth = 1;
      }
    }
  }
  double al, b;
  lanczos_scalars(red[0], red[1], red[2], al, b);
  const bool full = last_pass == m - 1; // as k_kge_small: full T or d_{m-1} = 0 (Q1)
  if (lane == 0) {
    S.alpha[0][last_pass] = al;
    if (!full) S.beta[0][m - 1] = b;
  }
  // T of L = sd B (B the unscaled operator the passes apply)
  const double alk = lane < last_pass ? S.alpha[0][lane] : (lane == last_pass ? al : 0.0);
  const double bk = lane < m - 2 ? S.beta[0][lane + 1]
                                 : (lane == m - 2 ? (full ? S.beta[0][m - 1] : b) : 0.0);
  double dk = lane <= last_pass ? T.sd * alk : 0.0;
  double ek = T.sd * bk;
  // exp(i dt T) e1 is analytic in T: a scaled Taylor series gives it without the QL
  // iteration (the eigen-decomposition route, small_ql_wave + small_apply_f_wave, takes
  // ~84 us for m = 12 against a few us here)
  double yr, yi;
  small_exp_series_wave(m, dk, ek, 0.0, a.dt, S.b0[0], yr, yi);
  if (lane < m) {
    S.y[0][lane] = yr;
    S.y[1][lane] = yi;
  }
  monitor_tail(S, 0, m, yr, yi, 0.0);
}

// Non-temporal load of a complex64 value (the combine reads each basis vector once).
__device__ inline c32 ld_stream(const c32 *p) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 v = __builtin_nontemporal_load(reinterpret_cast<const f2 *>(p));
  return c32{v.x, v.y};
}

template <bool UNI>
__device__ inline void nlse_psi3(const NlseArgs &a, const TrajDev &T, const TrajScal &S, const double *yr,
                                 const double *yi, int64_t g, double &ore, double &oim) {
  const int m = a.m;
  const int64_t cells = a.cells;
  const double b0 = S.b0[0];
  double re = 0.0, im = 0.0;
  if (b0 > 0.0) {
    const c32 x = ld_stream(a.X + g);
    const double xr = x.re / b0, xi = x.im / b0;
    re += yr[0] * xr - yi[0] * xi;
    im += yr[0] * xi + yi[0] * xr;
  }
  for (int k = 1; k <= m - 2; ++k) {
    const c32 v = ld_stream(a.V + (size_t)(k - 1) * cells + g);
    re += yr[k] * v.re - yi[k] * v.im;
    im += yr[k] * v.im + yi[k] * v.re;
  }
  const double bl = S.beta[0][m - 1];
  if (bl > 0.0) {
    const c32 wl = ld_stream(a.W + (size_t)((m - 2) & 1) * cells + g);
    const c32 vl = ld_stream(a.V + (size_t)(m - 3) * cells + g);
    const float al = (float)S.alpha[0][m - 2];
    const double vr = (double)(wl.re - al * vl.re) / bl, vim = (double)(wl.im - al * vl.im) / bl;
    re += yr[m - 1] * vr - yi[m - 1] * vim;
    im += yr[m - 1] * vim + yi[m - 1] * vr;
  }
  // second half phase with |psi2|^2 (nlse_dev.hpp:192-193)
  const double mm = UNI ? (double)T.mval : (double)__builtin_nontemporal_load(a.mf + g);
  const double rho = mm * (re * re + im * im);
  double sn, cs;
  phase_sincos(0.5 * a.dt * rho, sn, cs);
  ore = re * cs - im * sn;
  oim = re * sn + im * cs;
}

struct SnapC {
  double *psi; // complex128 interleaved, slot base
};

// <v_{m-2}, x> at cell g (x the seed psi1 = b0 v_0), accumulated as (re, im)
__device__ inline void nlse_orth_partial(const NlseArgs &a, int64_t g, double *o) {
  const c32 v = a.V[(size_t)(a.m - 3) * a.cells + g], x = a.X[g];
  o[0] += (double)v.re * x.re + (double)v.im * x.im;
  o[1] += (double)v.re * x.im - (double)v.im * x.re;
}

__device__ inline void nlse_store_orth(const NlseArgs &a, const TrajDev &T, const TrajScal &S,
                                       int chunk, const double *o) {
  const double b0 = S.b0[0];
  a.part3[(size_t)(T.chunk0 + chunk) * 2] = b0 > 0.0 ? o[0] / b0 : 0.0;
  a.part3[(size_t)(T.chunk0 + chunk) * 2 + 1] = b0 > 0.0 ? o[1] / b0 : 0.0;
}

// ORTH: also the orthogonality partials (reorthogonalisation modes 1 and 2).
template <int DIM, int CPT, bool ORTH, bool UNI = false>
__global__ __launch_bounds__(FBS) void k_nlse_combine(NlseArgs a, SnapC snap) {
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  const TrajScal &S = a.scal[cm.x];
  __shared__ double ys[2][kMaxKrylov];
  __shared__ double sho[2 * (FBS / kWave)];
  const bool orth = ORTH && a.m >= 3;
  double o[2] = {0.0, 0.0};
  if (threadIdx.x < a.m) {
    ys[0][threadIdx.x] = S.y[0][threadIdx.x];
    ys[1][threadIdx.x] = S.y[1][threadIdx.x];
  }
  __syncthreads();
  const int n = T.n;
  const int64_t base = (int64_t)cm.y * (FBS * CPT);
#pragma unroll
  for (int kk = 0; kk < CPT; ++kk) {
    const int64_t i = base + threadIdx.x + (int64_t)kk * FBS;
    if (i >= T.N) break;
    const int64_t src = clamp_src<DIM>(i, n);
    const int64_t gi = T.off + i, gs = T.off + src;
    double sr, si;
    nlse_psi3<UNI>(a, T, S, ys[0], ys[1], gs, sr, si);
    a.psi_next[gi] = c32{(float)sr, (float)si};
    if (snap.psi) { // snapshot before the ghost copy (nlse_dev.hpp:195-197)
      double ir = sr, ii = si;
      if (src != i) nlse_psi3<UNI>(a, T, S, ys[0], ys[1], gi, ir, ii);
      snap.psi[2 * gi] = ir;
      snap.psi[2 * gi + 1] = ii;
    }
    if (orth) nlse_orth_partial(a, gi, o);
  }
  if (orth) {
    block_sum_n<2>(o, sho);
    if (threadIdx.x == 0) nlse_store_orth(a, T, S, cm.y, o);
  }
}


// ---------------------------------------------------------------- NLSE 2D row march
// The stored-w' row march of k_kge_march2d for the complex64 NLSE sequence: the real
// and imaginary parts of v_j are two real sequences sharing the (real) Lanczos
// scalars, B v = sum_nb w (v_nb - v) (not negated, as in k_nlse_pass), and the
// partials add both parts.  STAGE 0 forms the seed psi1 = psi exp(i dt/2 m|psi|^2) when
// a cell is deposited and stores it (X) with W[0] = L psi1; STAGE 1 forms v_1 from W[0]
// and X; STAGE 2 is the general pass.
typedef float f2v __attribute__((ext_vector_type(2)));

// LAST: pass j = m - 1 of the full-T Lanczos variant (partials only, nothing stored).
template <int STAGE, bool UNI = false, bool LAST = false>
__global__ __launch_bounds__(FBS) void k_nlse_march2d(NlseArgs a, int j, int emax, int alt) {
  static_assert(!(LAST && STAGE == 0), "the last pass follows pass 0");
  extern __shared__ float lds[];
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double red[PN];
  __shared__ double shn[PN * (FBS / kWave)];
  float cA = 0.f, cB = 0.f, cC = 0.f;
  const int n = T.n;
  const int N = (int)T.N;
  const int64_t cells = a.cells;
  // no c window for uniform fields
  float *sc = lds, *sv0 = lds + (UNI ? 0 : emax), *sv1 = sv0 + emax, *sp0 = sv1 + emax,
        *sp1 = sp0 + emax;
  const c32 *psiT = a.psi + T.off;
  const float *cT = a.cf + T.off, *mT = a.mf + T.off;
  const c32 *Xin = a.W + (STAGE == 1 ? 0 : (size_t)((j - 1) & 1) * cells) + T.off; // z or w'
  const c32 *Pin = (STAGE == 1 ? a.X : a.V + (STAGE >= 2 ? (size_t)(j - 2) * cells : 0)) + T.off;
  c32 *Vout = a.V + (STAGE >= 1 ? (size_t)(j - 1) * cells : 0) + T.off;
  c32 *Wout = a.W + (size_t)(j & 1) * cells + T.off;
  c32 *Xout = a.X + T.off;
  double acc[PN] = {0.0, 0.0, 0.0};
  const int rbeg = cm.y * T.range;
  const int rend = rbeg + T.range < N ? rbeg + T.range : N;
  const int len = rend - rbeg;
  // direction alternating with the range and the pass, ordinary stores: as k_kge_march2d
  const bool down = alt != 0 && ((cm.y ^ j) & 1);
  const int org = down ? rend - 1 : rbeg;
  const int dir = down ? -1 : 1;
  const int oU = -dir * n, oL = -dir, oR = dir, oD = dir * n;
  const float rn = 1.0f / (float)n;
  auto cell = [&](int l) { return org + dir * l; };
  auto store2 = [](c32 *p, float re, float im) { *reinterpret_cast<f2v *>(p) = f2v{re, im}; };
  // raw inputs of cell i (no arithmetic here, so the loads stay in flight)
  auto fetch = [&](int i, float &c, c32 &x, c32 &pv) {
    c = UNI ? T.cval : cT[i];
    if (STAGE == 0) {
      x = psiT[i];
      pv.re = UNI ? T.mval : mT[i];
    } else {
      x = Xin[i];
      pv = Pin[i];
    }
  };
  auto deposit = [&](int e, float c, c32 x, c32 pv) {
    if (!UNI) sc[e] = c;
    if (STAGE == 0) {
      const c32 v = nl_half_phase(x, pv.re, a.dt);
      sv0[e] = v.re;
      sv1[e] = v.im;
    } else {
      sv0[e] = cB * (x.re - cA * pv.re);
      sv1[e] = cB * (x.im - cA * pv.im);
      sp0[e] = pv.re;
      sp1[e] = pv.im;
    }
  };
  // B v at LDS index e: sum_nb w (v_nb - v) in flux form (the small stage scales T by
  // 1/dx^2): minus apply5
  auto lap = [&](const Sten5 &w, const float *v, int e) {
    return -apply5(w, v, e, oU, oL, oR, oD);
  };
  // Lanczos scalars of the previous pass; formed after the first window loads are
  // issued, so the reduction's latency overlaps theirs
  auto coefs = [&]() {
    if (STAGE >= 1) {
      reduce_partials3(a, T, (j - 1) & 1, red);
      const double q0 = red[0], q1 = red[1], q2 = red[2];
      if (STAGE == 1) {
        const double b0 = sqrt(q0);
        const double al = q0 > 0.0 ? q1 / q0 : 0.0;
        const double b1sq = q0 > 0.0 ? (q2 - q1 * q1 / q0) / q0 : 0.0;
        const double b1 = (b1sq > 1e-14 * (q0 > 0.0 ? q2 / q0 : 0.0) && b1sq > 0.0) ? sqrt(b1sq) : 0.0;
        cA = (float)al;
        cB = (b0 > 0.0 && b1 > 0.0) ? (float)(1.0 / (b0 * b1)) : 0.0f;
        cC = b0 > 0.0 ? (float)(b1 / b0) : 0.0f;
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].b0[0] = b0;
          a.scal[cm.x].alpha[0][0] = al;
          a.scal[cm.x].beta[0][1] = b1;
        }
      } else {
        double al, b;
        lanczos_scalars(q0, q1, q2, al, b);
        cA = (float)al;
        cB = b > 0.0 ? (float)(1.0 / b) : 0.0f;
        cC = (float)b;
        if (cm.y == 0 && threadIdx.x == 0) {
          a.scal[cm.x].alpha[0][j - 1] = al;
          a.scal[cm.x].beta[0][j] = b;
        }
      }
    }
  };
  for (int e0 = 0; e0 < MSUB + 2 * n; e0 += MSPT * FBS) {
    float rc[MSPT];
    c32 rx[MSPT], rp[MSPT];
#pragma unroll
    for (int q = 0; q < MSPT; ++q) {
      const int e = e0 + threadIdx.x + q * FBS, i = cell(e - n);
      if (e < MSUB + 2 * n && i >= 0 && i < N) fetch(i, rc[q], rx[q], rp[q]);
    }
    if (e0 == 0) coefs();
#pragma unroll
    for (int q = 0; q < MSPT; ++q) {
      const int e = e0 + threadIdx.x + q * FBS, i = cell(e - n);
      if (e < MSUB + 2 * n && i >= 0 && i < N) deposit(e, rc[q], rx[q], rp[q]);
    }
  }
  float nc[MSPT];
  c32 nx[MSPT], np[MSPT];
  for (int cb = 0; cb < len; cb += MSUB) {
    const int ce = cb + MSUB < len ? cb + MSUB : len;
    const bool more = cb + MSUB < len;
    if (more) {
#pragma unroll
      for (int q = 0; q < MSPT; ++q) {
        const int i = cell(cb + MSUB + n + threadIdx.x + q * FBS);
        if (i >= 0 && i < N) fetch(i, nc[q], nx[q], np[q]);
      }
    }
    lds_barrier();
#pragma unroll
    for (int kk = 0; kk < MSPT; ++kk) {
      const int l = cb + threadIdx.x + kk * FBS;
      if (l < ce) {
        const int i = cell(l), e = l - cb + n;
        const Sten5 w = UNI ? sten5_uniform(T.cval, i, n, N, rn) : sten5(sc, e, oU, oL, oR, oD, i, n, N, rn);
        const float vr = sv0[e], vi = sv1[e];
        const float lr = lap(w, sv0, e), li = lap(w, sv1, e);
        if (STAGE == 0) {
          store2(Xout + i, vr, vi);
          store2(Wout + i, lr, li);
          acc[0] += (double)vr * vr + (double)vi * vi;
          acc[1] += (double)vr * lr + (double)vi * li;
          acc[2] += (double)lr * lr + (double)li * li;
        } else {
          const float wr = lr - cC * sp0[e], wi = li - cC * sp1[e];
          if (!LAST) {
            store2(Vout + i, vr, vi);
            store2(Wout + i, wr, wi);
          }
          acc[0] += (double)wr * vr + (double)wi * vi;
          acc[1] += (double)wr * wr + (double)wi * wi;
          acc[2] += (double)vr * vr + (double)vi * vi;
        }
      }
    }
    if (!more) break;
    lds_barrier(); // every read of this window is done
    for (int e = threadIdx.x; e < 2 * n; e += FBS) { // keep [cb' - n, cb' + n); 2n <= SUB
      if (!UNI) sc[e] = sc[MSUB + e];
      sv0[e] = sv0[MSUB + e];
      sv1[e] = sv1[MSUB + e];
      if (STAGE >= 1 && e >= n) { // the previous vector is read at own cells only
        sp0[e] = sp0[MSUB + e];
        sp1[e] = sp1[MSUB + e];
      }
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < MSPT; ++q) {
      const int i = cell(cb + MSUB + n + threadIdx.x + q * FBS);
      if (i >= 0 && i < N) deposit(2 * n + threadIdx.x + q * FBS, nc[q], nx[q], np[q]);
    }
  }
  block_sum_n<PN>(acc, shn);
  if (threadIdx.x == 0) {
    double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * PN;
#pragma unroll
    for (int k = 0; k < PN; ++k) dst[k] = acc[k];
  }
}

// The plane march of k_kge_march3d for the complex64 NLSE sequence: real and imaginary
// parts of v_j share the (real) Lanczos scalars, B v = sum_nb w (v_nb - v), and the
// partials add both parts; STAGE 0 forms the seed psi1 = psi exp(i dt/2 m|psi|^2) at each
// cell it loads and stores it (X) with W[0] = L psi1, STAGE 1 forms v_1 from W[0] and X,
// STAGE 2 is the general pass.  Same window, registers and cross-plane couplings as the
// Klein-Gordon kernel.
// LAST: pass j = m - 1 of the full-T Lanczos variant (partials only, nothing stored).
template <int STAGE, int CPT, bool UNI = false, bool LAST = false>
__global__ __launch_bounds__(FBS) void k_nlse_march3d(NlseArgs a, int j, int emax) {
  static_assert(!(LAST && STAGE == 0), "the last pass follows pass 0");
  extern __shared__ float lds[];
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double red[PN];
  __shared__ double shn[PN * (FBS / kWave)];
  float cA = 0.f, cB = 0.f, cC = 0.f;
  if (STAGE >= 1) {
    reduce_partials3(a, T, (j - 1) & 1, red);
    const double q0 = red[0], q1 = red[1], q2 = red[2];
    if (STAGE == 1) {
      const double b0 = sqrt(q0);
      const double al = q0 > 0.0 ? q1 / q0 : 0.0;
      const double b1sq = q0 > 0.0 ? (q2 - q1 * q1 / q0) / q0 : 0.0;
      const double b1 = (b1sq > 1e-14 * (q0 > 0.0 ? q2 / q0 : 0.0) && b1sq > 0.0) ? sqrt(b1sq) : 0.0;
      cA = (float)al;
      cB = (b0 > 0.0 && b1 > 0.0) ? (float)(1.0 / (b0 * b1)) : 0.0f;
      cC = b0 > 0.0 ? (float)(b1 / b0) : 0.0f;
      if (cm.y == 0 && threadIdx.x == 0) {
        a.scal[cm.x].b0[0] = b0;
        a.scal[cm.x].alpha[0][0] = al;
        a.scal[cm.x].beta[0][1] = b1;
      }
    } else {
      double al, b;
      lanczos_scalars(q0, q1, q2, al, b);
      cA = (float)al;
      cB = b > 0.0 ? (float)(1.0 / b) : 0.0f;
      cC = (float)b;
      if (cm.y == 0 && threadIdx.x == 0) {
        a.scal[cm.x].alpha[0][j - 1] = al;
        a.scal[cm.x].beta[0][j] = b;
      }
    }
  }
  const int n = T.n, n2 = n * n, N = (int)T.N;
  const int64_t cells = a.cells;
  // two windows (planes of even and odd z), as k_kge_march3d; no c window if UNI
  const int wstride = (UNI ? 2 : 3) * emax;
  auto win = [&](int z, float *&c_, float *&v0_, float *&v1_) {
    float *b_ = lds + (z & 1) * wstride;
    c_ = b_;
    v0_ = b_ + (UNI ? 0 : emax);
    v1_ = v0_ + emax;
  };
  const c32 *psiT = a.psi + T.off;
  const float *cT = a.cf + T.off, *mT = a.mf + T.off;
  const c32 *Xin = a.W + (STAGE == 1 ? 0 : (size_t)((j - 1) & 1) * cells) + T.off; // z or w'
  const c32 *Pin = (STAGE == 1 ? a.X : a.V + (STAGE >= 2 ? (size_t)(j - 2) * cells : 0)) + T.off;
  c32 *Vout = a.V + (STAGE >= 1 ? (size_t)(j - 1) * cells : 0) + T.off;
  c32 *Wout = a.W + (size_t)(j & 1) * cells + T.off;
  c32 *Xout = a.X + T.off;
  const int q6 = a.q6;
  double acc[PN] = {0.0, 0.0, 0.0};
  auto store2 = [](c32 *p, float re, float im) {
    __builtin_nontemporal_store(f2v{re, im}, reinterpret_cast<f2v *>(p));
  };

  const int pen = cm.y % T.npen, zs = cm.y / T.npen;
  const int pa = pen * T.prange;
  const int pb = pa + T.prange < n2 ? pa + T.prange : n2;
  const int z0 = zs * T.zseg;
  const int z1 = z0 + T.zseg < n ? z0 + T.zseg : n;
  const int S = pb - pa;

  struct Raw {
    float c;
    c32 x, p; // STAGE 0: x = psi, p.re = m
  };
  auto fetch = [&](int g, Raw &r) {
    r.c = UNI ? T.cval : cT[g];
    if (STAGE == 0) {
      r.x = psiT[g];
      r.p.re = UNI ? T.mval : mT[g];
    } else {
      r.x = Xin[g];
      r.p = Pin[g];
    }
  };
  // v_j (re, im) and the previous vector P from the raw inputs
  auto form = [&](const Raw &r, float &v0, float &v1, float &q0, float &q1) {
    if (STAGE == 0) {
      const c32 v = nl_half_phase(r.x, r.p.re, a.dt);
      v0 = v.re;
      v1 = v.im;
      q0 = q1 = 0.0f;
    } else {
      v0 = cB * (r.x.re - cA * r.p.re);
      v1 = cB * (r.x.im - cA * r.p.im);
      q0 = r.p.re;
      q1 = r.p.im;
    }
  };

  int po[CPT];
  bool own[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    po[q] = pa + threadIdx.x + q * FBS;
    own[q] = threadIdx.x + q * FBS < S;
  }
  float vm0[CPT], vm1[CPT], vc0[CPT], vc1[CPT], vn0[CPT], vn1[CPT];
  float cmz[CPT], ccz[CPT], cnz[CPT], pc0[CPT], pc1[CPT], pn0[CPT], pn1[CPT];
  Raw ro[CPT];
  Raw rh[M3_HPT];
  auto hwin = [&](int h) { return h < n ? h : S + h; };
  auto hcell = [&](int h, int z) { return z * n2 + pa - n + hwin(h); };
  auto fetch_own = [&](int z) {
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (own[q]) fetch(z * n2 + po[q], ro[q]);
  };
  auto fetch_halo = [&](int z) {
#pragma unroll
    for (int k = 0; k < M3_HPT; ++k) {
      const int h = threadIdx.x + k * FBS;
      if (h < 2 * n) {
        const int g = hcell(h, z);
        if (g >= 0 && g < N) fetch(g, rh[k]);
      }
    }
  };
  auto put_halo = [&](int z) {
    float *hc, *h0, *h1;
    win(z, hc, h0, h1);
#pragma unroll
    for (int k = 0; k < M3_HPT; ++k) {
      const int h = threadIdx.x + k * FBS;
      if (h < 2 * n) {
        const int g = hcell(h, z);
        if (g >= 0 && g < N) {
          const int e = hwin(h);
          float t0, t1;
          if (!UNI) hc[e] = rh[k].c;
          form(rh[k], h0[e], h1[e], t0, t1);
        }
      }
    }
  };

  if (z0 < z1) {
    if (z0 > 0) {
      fetch_own(z0 - 1);
#pragma unroll
      for (int q = 0; q < CPT; ++q)
        if (own[q]) {
          float t0, t1;
          form(ro[q], vm0[q], vm1[q], t0, t1);
          cmz[q] = ro[q].c;
        }
    }
    fetch_own(z0);
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (own[q]) {
        form(ro[q], vc0[q], vc1[q], pc0[q], pc1[q]);
        ccz[q] = ro[q].c;
      }
    if (z0 + 1 < n) fetch_own(z0 + 1);
    fetch_halo(z0);
    put_halo(z0);
    if (z0 + 1 < z1) fetch_halo(z0 + 1);
  }
  // window ends loaded with the own cells two planes ahead, deposited one plane ahead
  for (int z = z0; z < z1; ++z) {
    lds_barrier(); // plane z - 1's window reads are done: its window takes plane z + 1
    float *sc, *sv0, *sv1;
    win(z, sc, sv0, sv1);
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (own[q]) {
        const int e = po[q] - pa + n;
        if (!UNI) sc[e] = ccz[q];
        sv0[e] = vc0[q];
        sv1[e] = vc1[q];
      }
    if (z + 1 < z1) put_halo(z + 1);
    const bool up = z + 1 < n;
    if (up) {
#pragma unroll
      for (int q = 0; q < CPT; ++q)
        if (own[q]) {
          form(ro[q], vn0[q], vn1[q], pn0[q], pn1[q]);
          cnz[q] = ro[q].c;
        }
    }
    if (z + 1 < z1) {
      if (z + 2 < n) fetch_own(z + 2);
      if (z + 2 < z1) fetch_halo(z + 2);
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      if (!own[q]) continue;
      const int p = po[q], e = p - pa + n, g = z * n2 + p;
      const int col = p % n;
      const bool hzm = z > 0, hzp = up;
      const bool hym = q6 ? g >= n : p >= n;
      const bool hyp = q6 ? g + n < N : p < n2 - n;
      const bool hxm = col > 0, hxp = col < n - 1;
      const float ci = ccz[q];
      const float wzm = hzm ? 0.5f * (ci + cmz[q]) : 0.0f;
      const float wym = hym ? 0.5f * (ci + (UNI ? ci : sc[e - n])) : 0.0f;
      const float wxm = hxm ? 0.5f * (ci + (UNI ? ci : sc[e - 1])) : 0.0f;
      const float wxp = hxp ? 0.5f * (ci + (UNI ? ci : sc[e + 1])) : 0.0f;
      const float wyp = hyp ? 0.5f * (ci + (UNI ? ci : sc[e + n])) : 0.0f;
      const float wzp = hzp ? 0.5f * (ci + cnz[q]) : 0.0f;
      float lv[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const float *sv = s == 0 ? sv0 : sv1;
        const float vi = s == 0 ? vc0[q] : vc1[q];
        float r = 0.0f; // flux form, minus the apply5 sum
        if (hzm) r += wzm * (vi - (s == 0 ? vm0[q] : vm1[q]));
        if (hym) r += wym * (vi - sv[e - n]);
        if (hxm) r += wxm * (vi - sv[e - 1]);
        if (hxp) r += wxp * (vi - sv[e + 1]);
        if (hyp) r += wyp * (vi - sv[e + n]);
        if (hzp) r += wzp * (vi - (s == 0 ? vn0[q] : vn1[q]));
        lv[s] = -r;
      }
      const float vr = vc0[q], vi = vc1[q];
      if (STAGE == 0) {
        store2(Xout + g, vr, vi);
        store2(Wout + g, lv[0], lv[1]);
        acc[0] += (double)vr * vr + (double)vi * vi;
        acc[1] += (double)vr * lv[0] + (double)vi * lv[1];
        acc[2] += (double)lv[0] * lv[0] + (double)lv[1] * lv[1];
      } else {
        const float wr = lv[0] - cC * pc0[q], wi = lv[1] - cC * pc1[q];
        if (!LAST) {
          store2(Vout + g, vr, vi);
          store2(Wout + g, wr, wi);
        }
        acc[0] += (double)wr * vr + (double)wi * vi;
        acc[1] += (double)wr * wr + (double)wi * wi;
        acc[2] += (double)vr * vr + (double)vi * vi;
      }
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      vm0[q] = vc0[q];
      vm1[q] = vc1[q];
      cmz[q] = ccz[q];
      vc0[q] = vn0[q];
      vc1[q] = vn1[q];
      ccz[q] = cnz[q];
      pc0[q] = pn0[q];
      pc1[q] = pn1[q];
    }
  }
  block_sum_n<PN>(acc, shn);
  if (threadIdx.x == 0) {
    double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * PN;
#pragma unroll
    for (int k = 0; k < PN; ++k) dst[k] = acc[k];
  }
}

// Combine over the workgroup's range (any layout): psi2 = sum_k y_k v_k, the second half
// phase, the pre-copy snapshot and the ghost copy.
template <int DIM, bool ORTH, bool UNI = false>
__global__ __launch_bounds__(FBS) void k_nlse_combine_rt(NlseArgs a, SnapC snap) {
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajDev T = a.traj[cm.x];
  const TrajScal &S = a.scal[cm.x];
  __shared__ double ys[2][kMaxKrylov];
  __shared__ double sho[2 * (FBS / kWave)];
  const bool orth = ORTH && a.m >= 3;
  double o[2] = {0.0, 0.0};
  if (threadIdx.x < a.m) {
    ys[0][threadIdx.x] = S.y[0][threadIdx.x];
    ys[1][threadIdx.x] = S.y[1][threadIdx.x];
  }
  __syncthreads();
  const int n = T.n;
  const int64_t cbeg = (int64_t)cm.y * T.range;
  const int64_t cend = cbeg + T.range < T.N ? cbeg + T.range : T.N;
  // 2 cells per thread at a time: both cells' loads are issued before either store
  constexpr int G = 2;
  for (int64_t i0 = cbeg + threadIdx.x; i0 < cend; i0 += G * FBS) {
    double sr[G], si[G];
    int64_t srcs[G];
#pragma unroll
    for (int kk = 0; kk < G; ++kk) {
      const int64_t i = i0 + (int64_t)kk * FBS;
      if (i < cend) {
        srcs[kk] = clamp_src<DIM>(i, n);
        nlse_psi3<UNI>(a, T, S, ys[0], ys[1], T.off + srcs[kk], sr[kk], si[kk]);
      }
    }
#pragma unroll
    for (int kk = 0; kk < G; ++kk) {
      const int64_t i = i0 + (int64_t)kk * FBS;
      if (i < cend) {
        const int64_t gi = T.off + i;
        a.psi_next[gi] = c32{(float)sr[kk], (float)si[kk]};
        if (snap.psi) { // snapshot before the ghost copy (nlse_dev.hpp:195-197)
          double ir = sr[kk], ii = si[kk];
          if (srcs[kk] != i) nlse_psi3<UNI>(a, T, S, ys[0], ys[1], gi, ir, ii);
          snap.psi[2 * gi] = ir;
          snap.psi[2 * gi + 1] = ii;
        }
        if (orth) nlse_orth_partial(a, gi, o);
      }
    }
  }
  if (orth) {
    block_sum_n<2>(o, sho);
    if (threadIdx.x == 0) nlse_store_orth(a, T, S, cm.y, o);
  }
}

// ---------------------------------------------------------------- NLSE reorthogonalisation
// Complex classical Gram-Schmidt of the residual r = w'_j - alpha_j v_j against
// v_0 .. v_{j-1} with FP64 dot products and coefficients (lanczos_complex.hpp's CGS pass):
// h_k = <v_k, r> = <v_k, w'_j> - alpha_j <v_k, v_j>, w'_j <- w'_j - sum_k h_k v_k, then the
// pass partials of the corrected w'_j.  v_0 = X / b0.  As the KGE kernels otherwise.
__global__ __launch_bounds__(FBS) void k_nlse_reorth_dots(NlseArgs a, int j) {
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  TrajScal &S = a.scal[cm.x];
  if (!(a.reorth == 1 || (a.reorth == 2 && S.reorth_on))) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double sh[16 * (FBS / kWave)];
  if (cm.y == 0 && blockIdx.y == 0 && threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    const double *p = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0) * PN;
    double q0 = 0.0, q2 = 0.0;
    for (int g = lane; g < T.nchunks; g += kWave) {
      q0 += p[(size_t)g * PN];
      q2 += p[(size_t)g * PN + 2];
    }
    q0 = wave_sum(q0);
    q2 = wave_sum(q2);
    if (lane == 0) S.ralpha[0] = q2 > 0.0 ? q0 / q2 : 0.0;
  }
  const int k0 = blockIdx.y * 4;
  const int64_t cells = a.cells;
  const double ib0 = S.b0[0] > 0.0 ? 1.0 / S.b0[0] : 0.0;
  double acc[16]; // [q][<v_k,w'> re, im, <v_k,v_j> re, im]
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.0;
  const int64_t cbeg = (int64_t)cm.y * T.range;
  const int64_t cend = cbeg + T.range < T.N ? cbeg + T.range : T.N;
  for (int64_t i = cbeg + threadIdx.x; i < cend; i += FBS) {
    const int64_t g = T.off + i;
    const c32 w = a.W[(size_t)(j & 1) * cells + g], v = a.V[(size_t)(j - 1) * cells + g];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + q;
      if (k < j) {
        double xr, xi;
        if (k == 0) {
          const c32 x = a.X[g];
          xr = x.re * ib0;
          xi = x.im * ib0;
        } else {
          const c32 x = a.V[(size_t)(k - 1) * cells + g];
          xr = x.re;
          xi = x.im;
        }
        acc[q * 4 + 0] += xr * w.re + xi * w.im;
        acc[q * 4 + 1] += xr * w.im - xi * w.re;
        acc[q * 4 + 2] += xr * v.re + xi * v.im;
        acc[q * 4 + 3] += xr * v.im - xi * v.re;
      }
    }
  }
  block_sum_n<16>(acc, sh);
  if (threadIdx.x == 0) {
    double *dst = a.part2 + (size_t)(T.chunk0 + cm.y) * 4 * kMaxKrylov;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (k0 + q < j)
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[r * kMaxKrylov + k0 + q] = acc[q * 4 + r];
  }
}

__global__ __launch_bounds__(FBS) void k_nlse_reorth_apply(NlseArgs a, int j) {
  const int2 cm = a.cmap[blockIdx.x];
  if (cm.x < 0) return;
  const TrajScal &S = a.scal[cm.x];
  if (!(a.reorth == 1 || (a.reorth == 2 && S.reorth_on))) return;
  const TrajDev T = a.traj[cm.x];
  __shared__ double hr[kMaxKrylov], hi[kMaxKrylov];
  __shared__ double shn[PN * (FBS / kWave)];
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    const double *p2 = a.part2 + (size_t)T.chunk0 * 4 * kMaxKrylov;
    const double al = S.ralpha[0];
    for (int k = 0; k < j; ++k) {
      double d[4] = {0.0, 0.0, 0.0, 0.0};
      for (int g = lane; g < T.nchunks; g += kWave)
        for (int r = 0; r < 4; ++r) d[r] += p2[((size_t)g * 4 + r) * kMaxKrylov + k];
      for (int r = 0; r < 4; ++r) d[r] = wave_sum(d[r]);
      if (lane == 0) {
        hr[k] = d[0] - al * d[2];
        hi[k] = d[1] - al * d[3];
      }
    }
  }
  __syncthreads();
  const int64_t cells = a.cells;
  const double ib0 = S.b0[0] > 0.0 ? 1.0 / S.b0[0] : 0.0;
  double acc[PN] = {0.0, 0.0, 0.0};
  const int64_t cbeg = (int64_t)cm.y * T.range;
  const int64_t cend = cbeg + T.range < T.N ? cbeg + T.range : T.N;
  for (int64_t i = cbeg + threadIdx.x; i < cend; i += FBS) {
    const int64_t g = T.off + i;
    const c32 x = a.X[g];
    double xr = x.re * ib0, xi = x.im * ib0;
    double cr = hr[0] * xr - hi[0] * xi, ci = hr[0] * xi + hi[0] * xr;
    for (int k = 1; k < j; ++k) {
      const c32 v = a.V[(size_t)(k - 1) * cells + g];
      cr += hr[k] * v.re - hi[k] * v.im;
      ci += hr[k] * v.im + hi[k] * v.re;
    }
    c32 &wr = a.W[(size_t)(j & 1) * cells + g];
    const c32 w = {(float)((double)wr.re - cr), (float)((double)wr.im - ci)};
    wr = w;
    const c32 v = a.V[(size_t)(j - 1) * cells + g];
    acc[0] += (double)w.re * v.re + (double)w.im * v.im;
    acc[1] += (double)w.re * w.re + (double)w.im * w.im;
    acc[2] += (double)v.re * v.re + (double)v.im * v.im;
  }
  block_sum_n<PN>(acc, shn);
  if (threadIdx.x == 0) {
    double *dst = a.part + ((size_t)(j & 1) * a.total_chunks + T.chunk0 + cm.y) * PN;
#pragma unroll
    for (int k = 0; k < PN; ++k) dst[k] = acc[k];
  }
}

} // namespace fused
} // namespace nls

using namespace nls;
using namespace nls::fused;

struct nls_batch {
  nls_batch_params p{};
  int ntraj = 0;
  std::vector<nls_traj_desc> desc;
  Layout lay;
  // per trajectory: uniform coefficient fields found at upload (TrajDev::uni, cval, mval)
  std::vector<int> uni;
  std::vector<float> uni_c, uni_m;
  bool uni_all = false;  // every trajectory has uniform m and c: the UNI kernel variants
  bool recapture = false; // the step graphs hold the other variants
  int m = 10, cpt = 8;
  bool tuned = false;
  hipStream_t s = nullptr;
  TrajDev *d_traj = nullptr;
  int2 *d_cmap = nullptr;
  size_t cmap_cap = 0;
  TrajScal *d_scal = nullptr;
  double *d_part = nullptr;
  double *d_part2 = nullptr, *d_part3 = nullptr; // reorthogonalisation partials
  int *d_any = nullptr;                            // mode 2: any trajectory switched on
  // The reorthogonalisation kernels are launched (mode 1 always; mode 2 once the device
  // reports a trajectory that needs them, checked every kReorthCheck steps).
  bool reorth_live = false;
  static constexpr int kReorthCheck = 16;
  float *d_mf = nullptr, *d_cf = nullptr, *d_V = nullptr, *d_W = nullptr, *d_X = nullptr;
  double *d_state[3] = {nullptr, nullptr, nullptr}; // KGE: FP64 ring of 3; NLSE: complex64 pair
  double *d_snapU = nullptr, *d_snapV = nullptr;
  size_t bytes = 0;
  int64_t step = 0;
  int nsnap = 0;
  hipGraphExec_t graphs[3] = {nullptr, nullptr, nullptr};
  bool profiling = false;
  // 0 direct, 1 LDS-tiled chunk, 2 row march, 3 row march with recomputed w', 4 plane march (3D)
  int pass_kind = 0;
  int nmax = 0;
  // Lanes: the batch split into up to MAXL parts (balanced by cells) whose steps run on
  // their own streams, so one part's small stage and kernel tails overlap the other
  // parts' passes.  Per-trajectory results do not depend on the lane split.
  static constexpr int MAXL = 4;
  int lanes = 1, max_lanes = 1;
  hipStream_t sl[MAXL] = {}; // sl[0] is s
  hipEvent_t ev_fork = nullptr, ev_join[MAXL] = {};
  int2 *d_cmapL[MAXL] = {};
  int *d_tlistL[MAXL] = {};
  size_t cmapL_size[MAXL] = {};
  int ntrajL[MAXL] = {};
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<int> ev_launches; // general passes between the events of each recorded pair
  double pass_ms = 0.0;
  int pass_launches = 0;
  // Persistent step kernel (p.persistent): arrival counters, error flag, the host-set abort
  // flag (pinned host memory), steps it has run.
  int *d_cnt = nullptr, *d_err = nullptr;
  int *h_abort = nullptr;
  int64_t persist_steps = 0;

  template <class T> T *alloc(size_t count) {
    void *q = nullptr;
    const size_t b = sizeof(T) * std::max<size_t>(count, 1);
    NLS_HIP(hipMalloc(&q, b));
    bytes += b;
    return static_cast<T *>(q);
  }

  ~nls_batch() {
    if (s) (void)hipStreamSynchronize(s);
    drop_graphs();
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (int k = 1; k < MAXL; ++k)
      if (sl[k]) (void)hipStreamSynchronize(sl[k]);
    void *ptrs[] = {d_traj, d_cmap, d_scal, d_part, d_mf, d_cf, d_V, d_W, d_X,
                    d_state[0], d_state[1], d_state[2], d_snapU, d_snapV, d_cnt, d_err};
    for (void *q : ptrs)
      if (q) (void)hipFree(q);
    for (int k = 0; k < MAXL; ++k) {
      if (d_cmapL[k]) (void)hipFree(d_cmapL[k]);
      if (d_tlistL[k]) (void)hipFree(d_tlistL[k]);
      if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
      if (k > 0 && sl[k]) (void)hipStreamDestroy(sl[k]);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (h_abort) (void)hipHostFree(h_abort);
    if (s) (void)hipStreamDestroy(s);
  }

  // Calls f(std::true_type) for a batch of uniform fields, f(std::false_type) otherwise.
  template <class F> void with_uni(F &&f) const {
    if (uni_all) f(std::true_type{});
    else f(std::false_type{});
  }

  void drop_graphs() {
    for (auto &g : graphs)
      if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  }

  KgeArgs args(int cur) const {
    KgeArgs a{};
    a.traj = d_traj;
    a.cmap = d_cmap;
    a.scal = d_scal;
    a.part = d_part;
    a.total_chunks = lay.total_chunks;
    a.mf = d_mf;
    a.cf = d_cf;
    a.u = d_state[cur % 3];
    a.up = d_state[(cur + 2) % 3];
    a.un = d_state[(cur + 1) % 3];
    a.V = d_V;
    a.W = d_W;
    a.cells = lay.cells;
    a.m = m;
    a.q6 = p.q6;
    a.kind1 = p.method == NLS_KGE_SEWI ? F_SINC2_HALF : F_SINC2_SQRT;
    a.reorth = p.reorth;
    a.reorth_tol = p.reorth_tol > 0.0 ? p.reorth_tol : kReorthTol;
    a.part2 = d_part2;
    a.part3 = d_part3;
    a.any_reorth = d_any;
    return a;
  }

  bool nlse() const { return p.equation == NLS_EQ_NLSE; }
  bool sv() const { return p.equation == NLS_EQ_KGE && p.method == NLS_KGE_SV; }
  int nstates() const { return nlse() ? 2 : 3; }

  NlseArgs nargs(int cur) const {
    NlseArgs a{};
    a.traj = d_traj;
    a.cmap = d_cmap;
    a.scal = d_scal;
    a.part = d_part;
    a.total_chunks = lay.total_chunks;
    a.mf = d_mf;
    a.cf = d_cf;
    a.psi = reinterpret_cast<const c32 *>(d_state[cur & 1]);
    a.psi_next = reinterpret_cast<c32 *>(d_state[(cur + 1) & 1]);
    a.X = reinterpret_cast<c32 *>(d_X);
    a.V = reinterpret_cast<c32 *>(d_V);
    a.W = reinterpret_cast<c32 *>(d_W);
    a.cells = lay.cells;
    a.m = m;
    a.q6 = p.q6;
    a.dt = p.dt;
    a.reorth = p.reorth;
    a.reorth_tol = p.reorth_tol > 0.0 ? p.reorth_tol : kReorthTol;
    a.part2 = d_part2;
    a.part3 = d_part3;
    a.any_reorth = d_any;
    return a;
  }

  template <int DIM, int CPT>
  void launch_nlse_lane(int cur, bool snapshot, int slot, bool prof, hipStream_t s,
                        const int2 *cmap, unsigned ngrid, const int *tlist, int nt) {
    NlseArgs a = nargs(cur);
    a.cmap = cmap;
    a.tlist = tlist;
    const dim3 grid(ngrid), blk(FBS);
    const bool march = pass_kind == 2 && DIM == 2;
    const bool plane = pass_kind == 4 && DIM == 3 && (CPT == 4 || CPT == 8);
    const size_t lds = march2d_lds_bytes(nmax, uni_all);
    const int emax = MSUB + 2 * nmax;
    const size_t lds3 = 2 * march3d_lds_bytes(CPT, nmax, uni_all); // two plane windows
    const int emax3 = FBS * CPT + 2 * nmax;
    // pass j of the 3D plane march (tiles 4 and 8 only)
    auto plane_pass = [&](int stage, int j) {
      if constexpr (CPT == 4 || CPT == 8) {
        with_uni([&](auto u) {
          constexpr bool U = decltype(u)::value;
          if (stage == 0) hipLaunchKernelGGL((k_nlse_march3d<0, CPT, U>), grid, blk, lds3, s, a, j, emax3);
          else if (stage == 1) hipLaunchKernelGGL((k_nlse_march3d<1, CPT, U>), grid, blk, lds3, s, a, j, emax3);
          else if (stage == 2) hipLaunchKernelGGL((k_nlse_march3d<2, CPT, U>), grid, blk, lds3, s, a, j, emax3);
          else hipLaunchKernelGGL((k_nlse_march3d<2, CPT, U, true>), grid, blk, lds3, s, a, j, emax3);
        });
      }
    };
    auto reorth = [&](int j) {
      if (!reorth_live) return;
      hipLaunchKernelGGL(k_nlse_reorth_dots, dim3(ngrid, (j + 3) / 4), blk, 0, s, a, j);
      hipLaunchKernelGGL(k_nlse_reorth_apply, grid, blk, 0, s, a, j);
    };
    auto march_pass = [&](int stage, int j) {
      with_uni([&](auto u) {
        constexpr bool U = decltype(u)::value;
        if (stage == 0) hipLaunchKernelGGL((k_nlse_march2d<0, U>), grid, blk, lds, s, a, j, emax, 1);
        else if (stage == 1) hipLaunchKernelGGL((k_nlse_march2d<1, U>), grid, blk, lds, s, a, j, emax, 1);
        else if (stage == 2) hipLaunchKernelGGL((k_nlse_march2d<2, U>), grid, blk, lds, s, a, j, emax, 1);
        else hipLaunchKernelGGL((k_nlse_march2d<2, U, true>), grid, blk, lds, s, a, j, emax, 1);
      });
    };
    if (march) {
      march_pass(0, 0);
      march_pass(1, 1);
    } else if (plane) {
      plane_pass(0, 0);
      plane_pass(1, 1);
    } else {
      hipLaunchKernelGGL((k_nlse_pass<DIM, 0, CPT>), grid, blk, 0, s, a, 0);
      hipLaunchKernelGGL((k_nlse_pass<DIM, 1, CPT>), grid, blk, 0, s, a, 1);
    }
    reorth(1);
    const bool each = reorth_live;
    prof_open(prof, each, s);
    for (int j = 2; j <= m - 2; ++j) {
      prof_each(prof, each, s, false);
      if (march) march_pass(2, j);
      else if (plane) plane_pass(2, j);
      else hipLaunchKernelGGL((k_nlse_pass<DIM, 2, CPT>), grid, blk, 0, s, a, j);
      prof_each(prof, each, s, true);
      reorth(j);
    }
    prof_close(prof, each, s, m - 3);
    if (p.variant == 1) { // full T: the partials of alpha_{m-1} (stage 3 = last pass)
      if (march) march_pass(3, m - 1);
      else if (plane) plane_pass(3, m - 1);
      else hipLaunchKernelGGL((k_nlse_pass<DIM, 2, CPT, true>), grid, blk, 0, s, a, m - 1);
    }
    hipLaunchKernelGGL(k_nlse_small, dim3(nt), dim3(kWave), 0, s, a, p.variant == 1 ? m - 1 : m - 2);
    SnapC sn{nullptr};
    if (snapshot) sn.psi = d_snapU + (size_t)slot * lay.cells * 2;
    if (march || plane) {
      with_uni([&](auto u) {
        constexpr bool U = decltype(u)::value;
        if (p.reorth) hipLaunchKernelGGL((k_nlse_combine_rt<DIM, true, U>), grid, blk, 0, s, a, sn);
        else hipLaunchKernelGGL((k_nlse_combine_rt<DIM, false, U>), grid, blk, 0, s, a, sn);
      });
    } else {
      with_uni([&](auto u) {
        constexpr bool U = decltype(u)::value;
        if (p.reorth) hipLaunchKernelGGL((k_nlse_combine<DIM, CPT, true, U>), grid, blk, 0, s, a, sn);
        else hipLaunchKernelGGL((k_nlse_combine<DIM, CPT, false, U>), grid, blk, 0, s, a, sn);
      });
    }
    check_launch("fused nlse step");
  }

  template <int DIM, int CPT> void launch_nlse_t(int cur, bool snapshot, int slot, bool prof) {
    if (lanes > 1 && !prof) {
      NLS_HIP(hipEventRecord(ev_fork, s));
      for (int k = 1; k < lanes; ++k) NLS_HIP(hipStreamWaitEvent(sl[k], ev_fork, 0));
      for (int k = 0; k < lanes; ++k)
        launch_nlse_lane<DIM, CPT>(cur, snapshot, slot, false, k == 0 ? s : sl[k], d_cmapL[k],
                                   (unsigned)cmapL_size[k], d_tlistL[k], ntrajL[k]);
      for (int k = 1; k < lanes; ++k) {
        NLS_HIP(hipEventRecord(ev_join[k], sl[k]));
        NLS_HIP(hipStreamWaitEvent(s, ev_join[k], 0));
      }
      return;
    }
    launch_nlse_lane<DIM, CPT>(cur, snapshot, slot, prof, s, d_cmap, (unsigned)lay.cmap.size(),
                               nullptr, ntraj);
  }

  void set_layout(int cpt_) {
    cpt = cpt_;
    const double mu2 = p.mu2;
    lay = make_layout(desc.data(), ntraj, p.dim, cpt, p.dt, mu2, pass_kind >= 2 && pass_kind <= 3,
                      pass_kind == 4);
    apply_uniform();
    nmax = 0;
    for (auto &T : lay.traj) nmax = std::max(nmax, T.n);
    NLS_REQUIRE(lay.cmap.size() <= cmap_cap, "internal: chunk map larger than allocated");
    NLS_HIP(hipMemcpyAsync(d_traj, lay.traj.data(), sizeof(TrajDev) * ntraj,
                           hipMemcpyHostToDevice, s));
    NLS_HIP(hipMemcpyAsync(d_cmap, lay.cmap.data(), sizeof(int2) * lay.cmap.size(),
                           hipMemcpyHostToDevice, s));
    set_lanes(lanes);
  }

  void apply_uniform() {
    bool all = !uni.empty();
    for (int t = 0; t < ntraj && t < (int)uni.size(); ++t) {
      lay.traj[t].uni = uni[t];
      lay.traj[t].cval = uni_c[t];
      lay.traj[t].mval = uni_m[t];
      all = all && uni[t] == (kUniC | kUniM);
    }
    if (all != uni_all && graphs[0]) {
      drop_graphs();
      recapture = true;
    }
    uni_all = all;
  }

  // Records whether trajectory t's FP32 fields are spatially uniform (unless the
  // parameters ask for the fields to be streamed always) and uploads its descriptor.
  void set_uniform(int t, const float *mf, const float *cf, int64_t N) {
    if (uni.empty()) {
      uni.assign(ntraj, 0);
      uni_c.assign(ntraj, 0.0f);
      uni_m.assign(ntraj, 0.0f);
    }
    auto same = [N](const float *f) {
      for (int64_t i = 1; i < N; ++i)
        if (f[i] != f[0]) return false;
      return true;
    };
    const bool allow = p.always_load_fields == 0 && N > 0;
    uni[t] = (allow && same(cf) ? kUniC : 0) | (allow && same(mf) ? kUniM : 0);
    uni_c[t] = N > 0 ? cf[0] : 0.0f;
    uni_m[t] = N > 0 ? mf[0] : 0.0f;
    apply_uniform();
    NLS_HIP(hipMemcpyAsync(d_traj + t, &lay.traj[t], sizeof(TrajDev), hipMemcpyHostToDevice, s));
  }

  // Splits the trajectories into L lanes (largest first onto the lane with fewest
  // cells) and uploads each lane's block map and small-stage trajectory list.
  void set_lanes(int L) {
    lanes = std::max(1, std::min(L, max_lanes));
    if (lanes == 1) {
      NLS_HIP(hipStreamSynchronize(s)); // host layout copies done
      return;
    }
    std::vector<int> order(ntraj);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return lay.traj[x].N > lay.traj[y].N; });
    std::vector<std::vector<int>> sub(lanes);
    std::vector<int64_t> load(lanes, 0);
    for (int t : order) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      sub[k].push_back(t);
      load[k] += lay.traj[t].N;
    }
    for (int k = 0; k < lanes; ++k) {
      std::sort(sub[k].begin(), sub[k].end());
      const std::vector<int2> cm = xcd_cmap(lay.traj, sub[k]);
      NLS_REQUIRE(cm.size() <= cmap_cap, "internal: lane chunk map larger than allocated");
      cmapL_size[k] = cm.size();
      ntrajL[k] = (int)sub[k].size();
      if (!cm.empty())
        NLS_HIP(hipMemcpyAsync(d_cmapL[k], cm.data(), sizeof(int2) * cm.size(),
                               hipMemcpyHostToDevice, s));
      if (!sub[k].empty())
        NLS_HIP(hipMemcpyAsync(d_tlistL[k], sub[k].data(), sizeof(int) * sub[k].size(),
                               hipMemcpyHostToDevice, s));
    }
    NLS_HIP(hipStreamSynchronize(s));
  }


  int nmax_of() const {
    int x = 0;
    for (auto &d : desc) x = std::max(x, d.n);
    return x;
  }

  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      NLS_HIP(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  // Timing of the general passes: one event pair around the back-to-back general passes
  // of a step (events between kernels add their own time to each launch), or one pair
  // per pass when reorthogonalisation kernels run between them.
  void prof_open(bool prof, bool each, hipStream_t st) {
    if (prof && !each) NLS_HIP(hipEventRecord(next_event(), st));
  }
  void prof_close(bool prof, bool each, hipStream_t st, int passes) {
    if (prof && !each && passes > 0) {
      NLS_HIP(hipEventRecord(next_event(), st));
      ev_launches.push_back(passes);
    } else if (prof && !each) {
      --ev_used; // no general pass in this step
    }
  }
  void prof_each(bool prof, bool each, hipStream_t st, bool after) {
    if (prof && each) {
      NLS_HIP(hipEventRecord(next_event(), st));
      if (after) ev_launches.push_back(1);
    }
  }

  // Pass kernels: kind 0 = direct global-memory stencil (any dim), 1 = LDS-tiled
  // chunk + halo rows (2D), 2 = row march with an LDS window (2D), 3 = row march
  // that stores only the basis and recomputes w' (2D, launch_rpass), 4 = plane march
  // (3D, in-plane ranges of 256 * cpt cells).
  // LAST: the extra pass j = m - 1 of the full-T variant (kinds 0, 2 and 4 only).
  template <int DIM, int STAGE, bool LAST = false>
  void launch_pass(const KgeArgs &a, int j, hipStream_t s, unsigned ngrid) {
    const dim3 grid(ngrid), blk(FBS);
    if (pass_kind == 4) {
      const size_t lds = 2 * march3d_lds_bytes(cpt, nmax, uni_all); // two plane windows
      const int emax = FBS * cpt + 2 * nmax;
      with_uni([&](auto u) {
        constexpr bool U = decltype(u)::value;
        if (cpt == 4) hipLaunchKernelGGL((k_kge_march3d<STAGE, 4, U, LAST>), grid, blk, lds, s, a, j, emax);
        else hipLaunchKernelGGL((k_kge_march3d<STAGE, 8, U, LAST>), grid, blk, lds, s, a, j, emax);
      });
    } else if (pass_kind == 2) {
      const size_t lds = march2d_lds_bytes(nmax, uni_all);
      const int emax = MSUB + 2 * nmax;
      with_uni([&](auto u) {
        constexpr bool U = decltype(u)::value;
        hipLaunchKernelGGL((k_kge_march2d<STAGE, U, LAST>), grid, blk, lds, s, a, j, emax, 1);
      });
    } else if (pass_kind == 1) {
      if constexpr (!LAST) {
        const size_t lds = pass2d_lds_bytes(cpt, nmax);
        const int emax = FBS * cpt + 2 * nmax;
        if (cpt == 4) hipLaunchKernelGGL((k_kge_pass2d<STAGE, 4>), grid, blk, lds, s, a, j, emax);
        else hipLaunchKernelGGL((k_kge_pass2d<STAGE, 8>), grid, blk, lds, s, a, j, emax);
      }
    } else {
      if (cpt == 4) hipLaunchKernelGGL((k_kge_pass<DIM, STAGE, 4, LAST>), grid, blk, 0, s, a, j);
      else if (cpt == 8) hipLaunchKernelGGL((k_kge_pass<DIM, STAGE, 8, LAST>), grid, blk, 0, s, a, j);
      else hipLaunchKernelGGL((k_kge_pass<DIM, STAGE, 16, LAST>), grid, blk, 0, s, a, j);
    }
  }

  template <int STAGE> void launch_rpass(const KgeArgs &a, int j, hipStream_t s, unsigned ngrid) {
    const dim3 grid(ngrid), blk(FBS);
    const size_t lds = rmarch2d_lds_bytes(nmax);
    const int eP = MSUB + 4 * nmax, eV = MSUB + 2 * nmax, alt = 1;
    hipLaunchKernelGGL((k_kge_rmarch2d<STAGE>), grid, blk, lds, s, a, j, eP, eV, alt);
  }

  // Full reorthogonalisation of w'_j (modes 1 and 2; the kernels skip trajectories where
  // it is not active).
  void launch_reorth(const KgeArgs &a, int j, hipStream_t st, unsigned ngrid) {
    if (!reorth_live) return;
    hipLaunchKernelGGL(k_kge_reorth_dots, dim3(ngrid, (j + 3) / 4), dim3(FBS), 0, st, a, j);
    hipLaunchKernelGGL(k_kge_reorth_apply, dim3(ngrid), dim3(FBS), 0, st, a, j);
  }

  // One step of the trajectories in a block map (cmap, ngrid workgroups; tlist and
  // nt for the small stage) on stream st.
  template <int DIM>
  void launch_lane(int cur, bool snapshot, int slot, bool prof, hipStream_t st, const int2 *cmap,
                   unsigned ngrid, const int *tlist, int nt) {
    KgeArgs a = args(cur);
    a.cmap = cmap;
    a.tlist = tlist;
    const dim3 grid(ngrid), blk(FBS);
    if (pass_kind == 3) {
      // stored-basis-only passes; events around the general (j >= 3) passes
      launch_rpass<0>(a, 0, st, ngrid);
      launch_rpass<1>(a, 1, st, ngrid);
      if (m - 2 >= 2) launch_rpass<2>(a, 2, st, ngrid);
      prof_open(prof, false, st);
      for (int j = 3; j <= m - 2; ++j) launch_rpass<3>(a, j, st, ngrid);
      prof_close(prof, false, st, m - 4);
    } else {
      launch_pass<DIM, 0>(a, 0, st, ngrid);
      if (m - 2 >= 1) {
        launch_pass<DIM, 1>(a, 1, st, ngrid);
        launch_reorth(a, 1, st, ngrid);
      }
      const bool each = reorth_live;
      prof_open(prof, each, st);
      for (int j = 2; j <= m - 2; ++j) {
        prof_each(prof, each, st, false);
        launch_pass<DIM, 2>(a, j, st, ngrid);
        prof_each(prof, each, st, true);
        launch_reorth(a, j, st, ngrid);
      }
      prof_close(prof, each, st, m - 3);
      if (p.variant == 1) { // full T: the partials of alpha_{m-1}
        if (m - 1 == 1) launch_pass<DIM, 1, true>(a, 1, st, ngrid);
        else launch_pass<DIM, 2, true>(a, m - 1, st, ngrid);
      }
    }
    hipLaunchKernelGGL(k_kge_small, dim3(nt * NSEQ), dim3(kWave), 0, st, a,
                       p.variant == 1 ? m - 1 : m - 2);
    if (snapshot) {
      SnapArgs sn{d_snapU + (size_t)slot * lay.cells, d_snapV + (size_t)slot * lay.cells};
      with_uni([&](auto u) {
        constexpr bool U = decltype(u)::value;
        if (p.reorth) hipLaunchKernelGGL((k_kge_combine_snap<DIM, true, U>), grid, blk, 0, st, a, sn);
        else hipLaunchKernelGGL((k_kge_combine_snap<DIM, false, U>), grid, blk, 0, st, a, sn);
      });
    } else {
      with_uni([&](auto u) {
        constexpr bool U = decltype(u)::value;
        if (p.reorth) hipLaunchKernelGGL((k_kge_combine_rt<DIM, true, U>), grid, blk, 0, st, a);
        else hipLaunchKernelGGL((k_kge_combine_rt<DIM, false, U>), grid, blk, 0, st, a);
      });
    }
    check_launch("fused kge step");
  }

  template <int DIM> void launch_step_t(int cur, bool snapshot, int slot, bool prof) {
    if (lanes > 1 && !prof) {
      // fork s -> lane streams, one part of the batch on each, join back into s
      NLS_HIP(hipEventRecord(ev_fork, s));
      for (int k = 1; k < lanes; ++k) NLS_HIP(hipStreamWaitEvent(sl[k], ev_fork, 0));
      for (int k = 0; k < lanes; ++k)
        launch_lane<DIM>(cur, snapshot, slot, false, k == 0 ? s : sl[k], d_cmapL[k],
                         (unsigned)cmapL_size[k], d_tlistL[k], ntrajL[k]);
      for (int k = 1; k < lanes; ++k) {
        NLS_HIP(hipEventRecord(ev_join[k], sl[k]));
        NLS_HIP(hipStreamWaitEvent(s, ev_join[k], 0));
      }
      return;
    }
    launch_lane<DIM>(cur, snapshot, slot, prof, s, d_cmap, (unsigned)lay.cmap.size(), nullptr,
                     ntraj);
  }

  void launch_step(int cur, bool snapshot, int slot, bool prof) {
    if (nlse()) {
      if (p.dim == 2) {
        if (cpt == 4) launch_nlse_t<2, 4>(cur, snapshot, slot, prof);
        else if (cpt == 8) launch_nlse_t<2, 8>(cur, snapshot, slot, prof);
        else launch_nlse_t<2, 16>(cur, snapshot, slot, prof);
      } else {
        if (cpt == 4) launch_nlse_t<3, 4>(cur, snapshot, slot, prof);
        else if (cpt == 8) launch_nlse_t<3, 8>(cur, snapshot, slot, prof);
        else launch_nlse_t<3, 16>(cur, snapshot, slot, prof);
      }
      return;
    }
    if (sv()) {
      KgeArgs a = args(cur);
      SnapArgs sn{nullptr, nullptr};
      if (snapshot) sn = {d_snapU + (size_t)slot * lay.cells, d_snapV + (size_t)slot * lay.cells};
      const dim3 grid((unsigned)lay.cmap.size()), blk(FBS);
      with_uni([&](auto u) {
        constexpr bool U = decltype(u)::value;
        if (p.dim == 2) hipLaunchKernelGGL((k_kge_sv<2, U>), grid, blk, 0, s, a, sn);
        else hipLaunchKernelGGL((k_kge_sv<3, U>), grid, blk, 0, s, a, sn);
      });
      check_launch("fused kge sv step");
      return;
    }
    if (p.dim == 2) launch_step_t<2>(cur, snapshot, slot, prof);
    else launch_step_t<3>(cur, snapshot, slot, prof);
  }

  // ------------------------------------------------------------ persistent step kernel
  // Workgroups of the persistent kernel resident on the whole chip at once.
  int persist_slots() const {
    int per_cu = 0, dev = 0, ncu = 0;
    const size_t lds = persist2d_lds_bytes(nmax_of(), uni_all);
    with_uni([&](auto u) {
      constexpr bool U = decltype(u)::value;
      if (p.reorth)
        NLS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_kge_persist2d<U, true>, FBS, lds));
      else
        NLS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_kge_persist2d<U, false>, FBS, lds));
    });
    NLS_HIP(hipGetDevice(&dev));
    NLS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    return per_cu * ncu;
  }

  // Tile of the persistent kernel: the given one, or the smallest whose chunk map fits
  // the resident workgroups (one chunk per workgroup for the whole run).
  void setup_persistent() {
    NLS_REQUIRE(p.equation == NLS_EQ_KGE && !sv() && p.dim == 2,
                "persistent: the persistent step kernel runs 2D Klein-Gordon Gautschi / sEWI batches");
    NLS_REQUIRE(p.reorth != 1, "persistent: not with reorthogonalisation in every iteration (reorth 1)");
    NLS_REQUIRE(p.pass_kernel == 0 || p.pass_kernel == 3,
                "persistent: the persistent step kernel runs the row-march pass (pass_kernel 2)");
    int dev = 0, coop = 0;
    NLS_HIP(hipGetDevice(&dev));
    NLS_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    NLS_REQUIRE(coop != 0, "persistent: the device has no cooperative launch");
    NLS_REQUIRE(m >= (p.variant == 1 ? 3 : 4),
                "persistent: Krylov dimension below 4 (reference variant) or 3 (full T)");
    pass_kind = 2;
    NLS_REQUIRE(kind_ok(2, 64), "persistent: a grid side exceeds what the row-march window holds");
    lanes = std::min(p.lanes > 0 ? p.lanes : 2, max_lanes); // snapshot steps
    const size_t slots = (size_t)persist_slots();
    int c = p.cells_per_thread;
    if (c > 0) {
      NLS_REQUIRE(kind_ok(2, c), "persistent: cells_per_thread outside [4, 64]");
      set_layout(c);
    } else {
      for (c = 4; c <= 64; ++c) {
        set_layout(c);
        if (lay.cmap.size() <= slots) break;
      }
    }
    NLS_REQUIRE(lay.cmap.size() <= slots,
                "persistent: " + std::to_string(lay.cmap.size()) + " chunks exceed the " +
                    std::to_string(slots) + " workgroups resident at once (larger tile needed)");
    if (!d_cnt) {
      d_cnt = alloc<int>((size_t)kCntStride * ntraj);
      d_err = alloc<int>(1);
      NLS_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
    }
    if (!h_abort) {
      NLS_HIP(hipHostMalloc((void **)&h_abort, sizeof(int),
                            hipHostMallocCoherent | hipHostMallocMapped));
      __atomic_store_n(h_abort, 0, __ATOMIC_RELEASE);
    }
  }

  bool persist_live() const { return p.persistent && !reorth_live; }

  // nsteps steps without snapshots in one cooperative launch.
  void launch_persist(int nsteps) {
    const size_t slots = (size_t)persist_slots();
    NLS_REQUIRE(lay.cmap.size() <= slots, "persistent: the chunk map exceeds the resident workgroups");
    KgeArgs a = args(0); // the kernel sets the ring pointers per step
    PersistArgs q{};
    for (int k = 0; k < 3; ++k) q.state[k] = d_state[k];
    q.cnt = d_cnt;
    q.err = d_err;
    int *abort_dev = nullptr; // the device's address of the pinned flag
    NLS_HIP(hipHostGetDevicePointer((void **)&abort_dev, h_abort, 0));
    q.abort = abort_dev;
    q.step0 = step;
    q.nsteps = nsteps;
    q.emax = MSUB + 2 * nmax;
    q.last_pass = p.variant == 1 ? m - 1 : m - 2;
    NLS_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int) * kCntStride * ntraj, s));
    void *kargs[] = {&a, &q};
    const size_t lds = persist2d_lds_bytes(nmax, uni_all);
    const dim3 grid((unsigned)lay.cmap.size()), blk(FBS);
    with_uni([&](auto u) {
      constexpr bool U = decltype(u)::value;
      const void *fn = p.reorth ? (const void *)k_kge_persist2d<U, true>
                                : (const void *)k_kge_persist2d<U, false>;
      NLS_HIP(hipLaunchCooperativeKernel(fn, grid, blk, kargs, (unsigned)lds, s));
    });
    check_launch("persistent kge steps");
    persist_steps += nsteps;
  }

  // Tile autotuner: times one step per cells-per-thread candidate on the live
  // buffers (a step only writes the next-state buffer and Krylov scratch, so
  // timing runs leave the trajectory untouched) and keeps the fastest.
  bool kind_ok(int kind, int c) const {
    const int nm = nmax_of();
    if (nlse()) // k_nlse_pass tiles, the NLSE row march (2D) or plane march (3D)
      return (kind == 0 && (c == 4 || c == 8 || c == 16)) ||
             (kind == 2 && p.dim == 2 && 2 * nm <= MSUB && march2d_lds_bytes(nm) <= 65536 &&
              c >= 4 && c <= 64) ||
             (kind == 4 && p.dim == 3 && (c == 4 || c == 8) && 2 * nm <= FBS * M3_HPT);
    // row-march kernels take any range length (tile = target cells per workgroup / 256);
    // kinds 1 and 3 have no last pass of the full-T variant
    if ((kind == 1 || kind == 3) && p.variant != 0) return false;
    if (kind == 3) return p.reorth == 0 && p.dim == 2 && 3 * nm <= RKS * FBS && 2 * nm <= MSUB &&
                          rmarch2d_lds_bytes(nm) <= 65536 && c >= 4 && c <= 64;
    if (kind == 2) return p.dim == 2 && 2 * nm <= MSUB && march2d_lds_bytes(nm) <= 65536 &&
                          c >= 4 && c <= 64;
    if (kind == 1) return p.dim == 2 && (c == 4 || c == 8) && pass2d_lds_bytes(c, nm) <= 65536;
    if (kind == 4) return p.dim == 3 && (c == 4 || c == 8) && 2 * nm <= FBS * M3_HPT;
    return c == 4 || c == 8 || c == 16;
  }

  // Tile autotuner: times steps per (pass kernel, tile) candidate on the live
  // buffers (a step only writes the next-state buffer and Krylov scratch, so
  // timing runs leave the trajectory untouched) and keeps the fastest.
  // Workgroups of a row-march pass (kinds 2, 3) resident on the whole chip at once.
  int march_slots(int kind) const {
    const int nm = nmax_of();
    int per_cu = 0, dev = 0, ncu = 0;
    const size_t lds2 = march2d_lds_bytes(nm, uni_all);
    if (kind == 2 && nlse())
      with_uni([&](auto u) {
        NLS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_nlse_march2d<2, decltype(u)::value>, FBS, lds2));
      });
    else if (kind == 2)
      with_uni([&](auto u) {
        NLS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_kge_march2d<2, decltype(u)::value>, FBS, lds2));
      });
    else if (kind == 3)
      NLS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_kge_rmarch2d<3>, FBS,
                                                           rmarch2d_lds_bytes(nm)));
    NLS_HIP(hipGetDevice(&dev));
    NLS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    return per_cu * ncu;
  }

  int64_t chunks_at(int c) const {
    int64_t k = 0;
    for (auto &d : desc) {
      const int64_t N = (int64_t)d.n * d.n * (p.dim == 3 ? d.n : 1);
      k += (N + (int64_t)FBS * c - 1) / ((int64_t)FBS * c);
    }
    return k;
  }

  void autotune() {
    std::vector<std::array<int, 2>> cands = {
        {3, 12}, {3, 16}, {3, 24}, {3, 32}, {3, 36}, {3, 48}, {3, 64}, {2, 12}, {2, 16},
        {2, 24}, {2, 32}, {2, 36}, {2, 48}, {2, 64}, {1, 4},  {1, 8},  {4, 4},  {4, 8},
        {0, 4},  {0, 8},  {0, 16}};
    // Row-march tiles whose workgroup count just fits one or two rounds of resident
    // workgroups: a pass whose last round is nearly empty takes a whole round longer.
    for (int kind : {2, 3}) {
      if (!kind_ok(kind, 64)) continue;
      const int64_t slots = march_slots(kind);
      for (int rounds = 1; rounds <= 2 && slots > 0; ++rounds)
        for (int c = 4; c <= 64; ++c)
          if (chunks_at(c) <= rounds * slots) {
            if (std::find(cands.begin(), cands.end(), std::array<int, 2>{kind, c}) == cands.end())
              cands.push_back({kind, c});
            break;
          }
    }
    float best = 1e30f;
    int best_cpt = 4, best_kind = 0, best_lanes = 1;
    hipEvent_t e0, e1;
    NLS_HIP(hipEventCreate(&e0));
    NLS_HIP(hipEventCreate(&e1));
    for (auto &cd : cands) {
      if (!kind_ok(cd[0], cd[1])) continue;
      if (p.pass_kernel > 0 && cd[0] != p.pass_kernel - 1) continue;
      pass_kind = cd[0];
      set_layout(cd[1]);
      for (int ln = 1; ln <= max_lanes; ++ln) {
        if (p.lanes > 0 && ln != std::min(p.lanes, max_lanes)) continue;
        set_lanes(ln);
        launch_step((int)(step % nstates()), false, 0, false); // warm-up
        NLS_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) launch_step((int)(step % nstates()), false, 0, false);
        NLS_HIP(hipEventRecord(e1, s));
        NLS_HIP(hipEventSynchronize(e1));
        float ms = 0;
        NLS_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) { best = ms; best_cpt = cd[1]; best_kind = cd[0]; best_lanes = ln; }
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pass_kind = best_kind;
    lanes = best_lanes;
    set_layout(best_cpt);
    // the timing steps were discarded: restart the convergence monitors
    NLS_HIP(hipMemsetAsync(d_scal, 0, sizeof(TrajScal) * ntraj, s));
    if (d_part3) NLS_HIP(hipMemsetAsync(d_part3, 0, sizeof(double) * lay.total_chunks * NSEQ, s));
    if (d_any) NLS_HIP(hipMemsetAsync(d_any, 0, sizeof(int), s));
    tuned = true;
  }

  void capture_graphs() {
    drop_graphs();
    for (int cur = 0; cur < nstates(); ++cur) {
      hipGraph_t g;
      NLS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      launch_step(cur, false, 0, false);
      NLS_HIP(hipStreamEndCapture(s, &g));
      NLS_HIP(hipGraphInstantiate(&graphs[cur], g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
    }
  }

  void check_status() {
    if (d_err) {
      int err = 0;
      NLS_HIP(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
      NLS_HIP(hipStreamSynchronize(s));
      if (err != 0) { // the batch's state is that of a partly advanced step
        NLS_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
        NLS_HIP(hipStreamSynchronize(s));
        if (h_abort) __atomic_store_n(h_abort, 0, __ATOMIC_RELEASE);
      }
      NLS_REQUIRE(err != 2, "persistent step kernel: aborted by the host (nls_batch_abort)");
      NLS_REQUIRE(err == 0, "persistent step kernel: a per-trajectory wait gave up");
    }
    std::vector<TrajScal> h(ntraj);
    NLS_HIP(hipMemcpyAsync(h.data(), d_scal, sizeof(TrajScal) * ntraj, hipMemcpyDeviceToHost, s));
    NLS_HIP(hipStreamSynchronize(s));
    for (int t = 0; t < ntraj; ++t)
      NLS_REQUIRE(h[t].status == 0, "tridiagonal eigensolver did not converge (trajectory " +
                                        std::to_string(t) + ")");
  }

  void run(int nsteps, float *ms) {
    if (!tuned && sv()) {
      // one flat-range kernel per step (chunks of 256 x tile cells, one stream), launched
      // directly: for a single kernel a graph replay costs more than the launch
      // (71.8 against 65.1 us per step on 100 x 300^2)
      pass_kind = 0;
      lanes = 1;
      set_layout(p.cells_per_thread > 0 ? p.cells_per_thread : 16);
      tuned = true;
      recapture = false;
    }
    if (!tuned && p.persistent) {
      setup_persistent();
      tuned = true;
      if (p.use_graph) capture_graphs(); // snapshot steps and the steps after reorth switches on
      recapture = false;
    }
    if (!tuned) {
      if (p.cells_per_thread > 0) {
        const int c = p.cells_per_thread;
        if (p.pass_kernel > 0) pass_kind = p.pass_kernel - 1;
        else pass_kind = kind_ok(3, c) ? 3 : kind_ok(2, c) ? 2 : kind_ok(4, c) ? 4 : kind_ok(1, c) ? 1 : 0;
        NLS_REQUIRE(kind_ok(pass_kind, c), "cells_per_thread / pass_kernel not supported for this batch");
        lanes = std::min(p.lanes > 0 ? p.lanes : 2, max_lanes);
        set_layout(c);
        tuned = true;
      } else {
        autotune();
      }
      if (p.use_graph) capture_graphs();
      recapture = false;
    }
    if (recapture) {
      if (p.use_graph && !sv()) capture_graphs();
      recapture = false;
    }
    hipEvent_t e0, e1;
    NLS_HIP(hipEventCreate(&e0));
    NLS_HIP(hipEventCreate(&e1));
    ev_used = 0;
    ev_launches.clear();
    NLS_HIP(hipEventRecord(e0, s));
    auto snap_at = [&](int64_t idx) {
      return p.snapshot_every > 0 && idx % p.snapshot_every == 0 &&
             idx / p.snapshot_every < p.max_snapshots;
    };
    for (int k = 0; k < nsteps; ++k) {
      const int64_t idx = step + 1; // reference step number i (kg_driver_dev_2d.cpp:144)
      const int cur = (int)(step % nstates());
      const bool snap = snap_at(idx);
      if (!snap && !profiling && persist_live()) {
        // the steps up to the next snapshot step, the end of the run or the next check of
        // the orthogonality flag, in one persistent launch
        int len = 1;
        while (k + len < nsteps && !snap_at(idx + len) &&
               (p.reorth != 2 || (step + len) % kReorthCheck != 0))
          ++len;
        launch_persist(len);
        step += len;
        k += len - 1;
        if (p.reorth == 2 && !reorth_live && step % kReorthCheck == 0) {
          int any = 0;
          NLS_HIP(hipMemcpyAsync(&any, d_any, sizeof(int), hipMemcpyDeviceToHost, s));
          NLS_HIP(hipStreamSynchronize(s));
          if (any) {
            reorth_live = true;
            if (p.use_graph) capture_graphs();
          }
        }
        continue;
      }
      if (snap || !p.use_graph || profiling || !graphs[cur])
        launch_step(cur, snap, snap ? (int)(idx / p.snapshot_every) : 0, profiling);
      else
        NLS_HIP(hipGraphLaunch(graphs[cur], s));
      if (snap) nsnap = std::max(nsnap, (int)(idx / p.snapshot_every) + 1);
      ++step;
      if (p.reorth == 2 && !reorth_live && step % kReorthCheck == 0) {
        // has a trajectory lost semi-orthogonality?  Then launch the reorthogonalisation
        // kernels from the next step on (they act on the trajectories that need it).
        int any = 0;
        NLS_HIP(hipMemcpyAsync(&any, d_any, sizeof(int), hipMemcpyDeviceToHost, s));
        NLS_HIP(hipStreamSynchronize(s));
        if (any) {
          reorth_live = true;
          if (p.use_graph) capture_graphs();
        }
      }
    }
    NLS_HIP(hipEventRecord(e1, s));
    NLS_HIP(hipEventSynchronize(e1));
    float t = 0;
    NLS_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (profiling) {
      pass_ms = 0.0;
      pass_launches = 0;
      for (size_t k = 0; k + 1 < ev_used && k / 2 < ev_launches.size(); k += 2) {
        float d = 0;
        NLS_HIP(hipEventElapsedTime(&d, ev_pool[k], ev_pool[k + 1]));
        pass_ms += d;
        pass_launches += ev_launches[k / 2];
      }
    }
    check_status();
  }

  double pass_bytes() const {
    // Algorithmic bytes of one general Lanczos pass per cell.  Stored w': read
    // w'_{j-1}, v_{j-1} for both sequences and c (4 * 5 B), write v_j, w'_j for both
    // (4 * 4 B).  Recomputed w' (kind 3): read v_{j-1}, v_{j-2} for both sequences
    // and c (4 * 5 B), write v_j for both (4 * 2 B).
    // With uniform fields (row and plane marches, UNI variants) the 4 B of c are not read.
    double b = 0.0;
    for (int t = 0; t < ntraj; ++t) {
      const bool cu = uni_all && (pass_kind == 2 || pass_kind == 4);
      b += ((pass_kind == 3 ? 28.0 : 36.0) - (cu ? 4.0 : 0.0)) * (double)lay.traj[t].N;
    }
    return b;
  }
  int64_t total_cells() const {
    int64_t c = 0;
    for (auto &T : lay.traj) c += T.N;
    return c;
  }
};

static size_t estimate(const nls_batch_params &p, int ntraj, const nls_traj_desc *td, int m) {
  int64_t cells = 0, chunks = 0;
  for (int t = 0; t < ntraj; ++t) {
    const int64_t N = p.dim == 3 ? (int64_t)td[t].n * td[t].n * td[t].n : (int64_t)td[t].n * td[t].n;
    cells += (N + 63) / 64 * 64;
    chunks += (N + FBS * 4 - 1) / (FBS * 4);
  }
  size_t b = 0;
  const bool nl = p.equation == NLS_EQ_NLSE;
  b += sizeof(float) * cells * 2;                         // m, c
  b += sizeof(double) * cells * 3;                        // state ring (NLSE: 2 complex64 + seed)
  b += sizeof(float) * cells * NSEQ * (size_t)(m - 2);    // Krylov basis (NLSE: complex64)
  b += sizeof(float) * cells * NSEQ * 2;                  // w' ping-pong
  b += sizeof(double) * 2 * chunks * NSEQ * PN;           // partials
  if (p.reorth)
    b += sizeof(double) * chunks * NSEQ * (2 * kMaxKrylov + 1); // reorthogonalisation partials
  b += sizeof(TrajScal) * ntraj;
  b += sizeof(double) * cells * 2 * (size_t)std::max(p.max_snapshots, 0) * (nl ? 1 : 1);
  return b;
}

// Krylov dimension: the caller's, or the a-priori bound below.  The reference's variant
// (T[m-1][m-1] = 0) is ill-conditioned at small m on smooth, symmetric data: rounding the
// input to FP32 moves the FP64 result by up to ~1e-4 relative over a few hundred steps at
// m = 8 and ~3e-6 at m = 10, while the full-T variant moves ~3e-8 at every m.  So the
// reference variant is never tuned below the reference drivers' m = 10
// (kg_driver_dev_2d.cpp:138).
static int pick_m(const nls_batch_params &p, int ntraj, const nls_traj_desc *td) {
  if (p.krylov_dim > 0) return p.krylov_dim;
  std::vector<double> r(ntraj);
  for (int t = 0; t < ntraj; ++t) {
    const double dx = 2.0 * td[t].L / (td[t].n - 1);
    // Gershgorin bound of the operator with max(c) <= 4 assumed (the fields are not
    // known when the batch, and with it the Krylov basis, is sized)
    r[t] = (4.0 * p.dim * 4.0 / (dx * dx) + p.mu2) * p.dt * p.dt;
  }
  const int m = autotune_krylov(r);
  return p.variant == 0 ? std::max(m, 10) : m;
}

extern "C" {

int nls_batch_estimate_bytes(const nls_batch_params *p, int ntraj, const nls_traj_desc *trajs,
                             size_t *bytes) {
  return guard([&] {
    NLS_REQUIRE(p && trajs && bytes && ntraj > 0, "nls_batch_estimate_bytes: bad arguments");
    *bytes = estimate(*p, ntraj, trajs, pick_m(*p, ntraj, trajs));
  });
}

int nls_batch_create(const nls_batch_params *p, int ntraj, const nls_traj_desc *trajs,
                     nls_batch **out) {
  return guard([&] {
    NLS_REQUIRE(p && trajs && out && ntraj > 0, "nls_batch_create: bad arguments");
    NLS_REQUIRE(p->equation == NLS_EQ_KGE || p->equation == NLS_EQ_NLSE,
                "nls_batch_create: the fused batch runs Klein-Gordon and NLSE; stochastic phi^4 "
                "is served by nls_server");
    if (p->equation == NLS_EQ_KGE) {
      NLS_REQUIRE(p->method == NLS_KGE_GAUTSCHI || p->method == NLS_KGE_SEWI ||
                      p->method == NLS_KGE_SV,
                  "fused Klein-Gordon batch supports the Gautschi, sEWI and Stormer-Verlet methods");
      NLS_REQUIRE(p->method != NLS_KGE_SV || (p->reorth == 0 && p->pass_kernel <= 1),
                  "the Stormer-Verlet batch has no Krylov passes (reorth 0, pass_kernel 0 or 1)");
      for (int t = 0; t < ntraj && p->method == NLS_KGE_SV; ++t)
        NLS_REQUIRE((p->dim == 3 ? (int64_t)trajs[t].n * trajs[t].n * trajs[t].n
                                 : (int64_t)trajs[t].n * trajs[t].n) < (1 << 24),
                    "Stormer-Verlet batch: a trajectory needs fewer than 2^24 cells");
    } else
      NLS_REQUIRE(p->method == NLS_NLSE_STRANG, "fused NLSE batch supports the Strang method");
    NLS_REQUIRE(p->dim == 2 || p->dim == 3, "dim must be 2 or 3");
    NLS_REQUIRE(p->dt > 0, "dt must be positive");
    NLS_REQUIRE(p->pass_kernel >= 0 && p->pass_kernel <= 5, "pass_kernel must be in [0, 5]");
    NLS_REQUIRE(p->reorth >= 0 && p->reorth <= 2, "reorth must be 0, 1 or 2");
    NLS_REQUIRE(p->always_load_fields == 0 || p->always_load_fields == 1,
                "always_load_fields must be 0 or 1");
    NLS_REQUIRE(p->reorth == 0 || p->pass_kernel != 4,
                "reorthogonalisation needs a pass kernel that stores w' (not kernel 3)");
    NLS_REQUIRE(p->variant == 0 || p->variant == 1, "variant must be 0 or 1");
    NLS_REQUIRE(p->variant == 0 || (p->pass_kernel != 2 && p->pass_kernel != 4),
                "the full-T Lanczos variant runs on pass kernels 0, 2 and 4");
    NLS_REQUIRE(p->lanes >= 0 && p->lanes <= nls_batch::MAXL, "lanes must be in [0, 4]");
    NLS_REQUIRE(p->cells_per_thread == 0 ||
                    (p->cells_per_thread >= 4 && p->cells_per_thread <= 64),
                "cells_per_thread must be 0 (autotune) or in [4, 64]");
    for (int t = 0; t < ntraj; ++t) NLS_REQUIRE(trajs[t].n >= 3 && trajs[t].L > 0, "bad trajectory");
    auto b = std::make_unique<nls_batch>();
    b->p = *p;
    b->ntraj = ntraj;
    b->desc.assign(trajs, trajs + ntraj);
    b->m = b->sv() ? 3 : pick_m(*p, ntraj, trajs);
    NLS_REQUIRE(b->m >= 3 && b->m <= kMaxKrylov, "fused path needs a Krylov dimension in [3, 32]");
    NLS_HIP(hipStreamCreate(&b->s));
    b->sl[0] = b->s;
    if (ntraj >= 2) {
      b->max_lanes = std::min(nls_batch::MAXL, ntraj);
      NLS_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
      for (int k = 1; k < b->max_lanes; ++k) {
        NLS_HIP(hipStreamCreate(&b->sl[k])); // default priority, as the main stream
        NLS_HIP(hipEventCreateWithFlags(&b->ev_join[k], hipEventDisableTiming));
      }
    }
    // layout with the smallest tile (most chunks) sizes the chunk map and partials
    Layout l4 = make_layout(trajs, ntraj, p->dim, 4, p->dt, p->mu2);
    // any tile has at most the chunks of the smallest one; a bin holds at most all of them
    b->cmap_cap = (size_t)NXCD * (l4.total_chunks + 1);
    b->d_traj = b->alloc<TrajDev>(ntraj);
    b->d_cmap = b->alloc<int2>(b->cmap_cap);
    if (b->max_lanes > 1)
      for (int k = 0; k < b->max_lanes; ++k) {
        b->d_cmapL[k] = b->alloc<int2>(b->cmap_cap);
        b->d_tlistL[k] = b->alloc<int>(ntraj);
      }
    b->d_scal = b->alloc<TrajScal>(ntraj);
    b->d_part = b->alloc<double>((size_t)2 * l4.total_chunks * NSEQ * PN);
    if (p->reorth) {
      b->d_part2 = b->alloc<double>((size_t)l4.total_chunks * NSEQ * 2 * kMaxKrylov);
      b->d_part3 = b->alloc<double>((size_t)l4.total_chunks * NSEQ);
      NLS_HIP(hipMemsetAsync(b->d_part3, 0, sizeof(double) * l4.total_chunks * NSEQ, b->s));
      b->d_any = b->alloc<int>(1);
      NLS_HIP(hipMemsetAsync(b->d_any, 0, sizeof(int), b->s));
    }
    b->reorth_live = p->reorth == 1;
    const int64_t cells = l4.cells;
    b->d_mf = b->alloc<float>(cells);
    b->d_cf = b->alloc<float>(cells);
    const int nst = b->nstates();
    for (int k = 0; k < nst; ++k) b->d_state[k] = b->alloc<double>(cells);
    if (!b->sv()) { // Krylov basis and w' (none for Stormer-Verlet)
      b->d_V = b->alloc<float>((size_t)cells * NSEQ * (b->m - 2));
      b->d_W = b->alloc<float>((size_t)cells * NSEQ * 2);
    }
    if (b->nlse()) b->d_X = b->alloc<float>((size_t)cells * 2);
    if (p->max_snapshots > 0) {
      // KGE: u and v (FP64); NLSE: psi (complex128) in d_snapU
      b->d_snapU = b->alloc<double>((size_t)cells * p->max_snapshots * (b->nlse() ? 2 : 1));
      if (!b->nlse()) b->d_snapV = b->alloc<double>((size_t)cells * p->max_snapshots);
    }
    NLS_HIP(hipMemsetAsync(b->d_scal, 0, sizeof(TrajScal) * ntraj, b->s));
    for (int k = 0; k < nst; ++k) NLS_HIP(hipMemsetAsync(b->d_state[k], 0, sizeof(double) * cells, b->s));
    NLS_HIP(hipMemsetAsync(b->d_mf, 0, sizeof(float) * cells, b->s));
    NLS_HIP(hipMemsetAsync(b->d_cf, 0, sizeof(float) * cells, b->s));
    b->set_layout(8);
    *out = b.release();
  });
}

int nls_batch_destroy(nls_batch *b) { return guard([&] { delete b; }); }

int nls_batch_set_initial(nls_batch *b, int t, const double *m, const double *c, const double *u0,
                          const double *v0) {
  return guard([&] {
    NLS_REQUIRE(b && t >= 0 && t < b->ntraj, "nls_batch_set_initial: bad trajectory index");
    NLS_REQUIRE(m && c && u0 && v0, "nls_batch_set_initial: null input");
    NLS_REQUIRE(b->step == 0, "nls_batch_set_initial: the batch has already been advanced");
    const TrajDev &T = b->lay.traj[t];
    const int64_t N = T.N;
    if (b->nlse()) {
      std::vector<float> mf(N), cf(N), psi(2 * N);
      for (int64_t i = 0; i < N; ++i) {
        mf[i] = (float)m[i];
        cf[i] = (float)c[i];
        psi[2 * i] = (float)u0[2 * i];
        psi[2 * i + 1] = (float)u0[2 * i + 1];
      }
      NLS_HIP(hipMemcpyAsync(b->d_mf + T.off, mf.data(), sizeof(float) * N, hipMemcpyHostToDevice, b->s));
      NLS_HIP(hipMemcpyAsync(b->d_cf + T.off, cf.data(), sizeof(float) * N, hipMemcpyHostToDevice, b->s));
      b->set_uniform(t, mf.data(), cf.data(), N);
      NLS_HIP(hipMemcpyAsync(b->d_state[0] + T.off, psi.data(), sizeof(float) * 2 * N,
                             hipMemcpyHostToDevice, b->s));
      if (b->p.max_snapshots > 0) {
        NLS_HIP(hipMemcpyAsync(b->d_snapU + 2 * T.off, u0, sizeof(double) * 2 * N,
                               hipMemcpyHostToDevice, b->s));
        b->nsnap = std::max(b->nsnap, 1);
      }
      NLS_HIP(hipStreamSynchronize(b->s));
      return;
    }
    std::vector<float> mf(N), cf(N);
    std::vector<double> upast(N);
    for (int64_t i = 0; i < N; ++i) {
      mf[i] = (float)m[i];
      cf[i] = (float)c[i];
      upast[i] = u0[i] - b->p.dt * v0[i]; // kg_dev.hpp:55-58
    }
    NLS_HIP(hipMemcpyAsync(b->d_mf + T.off, mf.data(), sizeof(float) * N, hipMemcpyHostToDevice, b->s));
    NLS_HIP(hipMemcpyAsync(b->d_cf + T.off, cf.data(), sizeof(float) * N, hipMemcpyHostToDevice, b->s));
    b->set_uniform(t, mf.data(), cf.data(), N);
    NLS_HIP(hipMemcpyAsync(b->d_state[0] + T.off, u0, sizeof(double) * N, hipMemcpyHostToDevice, b->s));
    NLS_HIP(hipMemcpyAsync(b->d_state[2] + T.off, upast.data(), sizeof(double) * N,
                           hipMemcpyHostToDevice, b->s));
    if (b->p.max_snapshots > 0) {
      NLS_HIP(hipMemcpyAsync(b->d_snapU + T.off, u0, sizeof(double) * N, hipMemcpyHostToDevice, b->s));
      NLS_HIP(hipMemcpyAsync(b->d_snapV + T.off, v0, sizeof(double) * N, hipMemcpyHostToDevice, b->s));
      b->nsnap = std::max(b->nsnap, 1);
    }
    NLS_HIP(hipStreamSynchronize(b->s));
  });
}

int nls_ic_draw(int kind, uint64_t seed, int traj, double L, double *params, int cap, int *count) {
  return guard([&] {
    NLS_REQUIRE(params && count && traj >= 0 && L > 0, "nls_ic_draw: bad arguments");
    double p[ic::kMaxParams];
    const int c = ic::draw(kind, seed, (uint32_t)traj, L, p);
    NLS_REQUIRE(c >= 0, "nls_ic_draw: unknown sampler kind");
    NLS_REQUIRE(c <= cap, "nls_ic_draw: parameter buffer too small");
    std::copy(p, p + c, params);
    *count = c;
  });
}

int nls_batch_set_initial_sampled(nls_batch *b, int t, int kind, uint64_t seed, int traj_index,
                                  double m_const, double c_const) {
  return guard([&] {
    NLS_REQUIRE(b && t >= 0 && t < b->ntraj, "nls_batch_set_initial_sampled: bad trajectory index");
    NLS_REQUIRE(traj_index >= 0, "nls_batch_set_initial_sampled: bad trajectory counter");
    NLS_REQUIRE(b->step == 0, "nls_batch_set_initial_sampled: the batch has already been advanced");
    NLS_REQUIRE(kind >= ic::GAUSSIANS_2D && kind <= ic::BUBBLES_3D, "unknown sampler kind");
    NLS_REQUIRE(ic::dim_of(kind) == b->p.dim, "sampler dimension differs from the batch's");
    NLS_REQUIRE((kind == ic::PACKETS_2D) == b->nlse(),
                "complex samplers (packets) go with the NLSE batch, real ones with Klein-Gordon");
    const TrajDev &T = b->lay.traj[t];
    ic::Params P{};
    P.kind = kind;
    P.n = T.n;
    P.L = b->desc[t].L;
    NLS_REQUIRE(ic::draw(kind, seed, (uint32_t)traj_index, P.L, P.p) >= 0, "sampler draw failed");
    const bool keep = b->p.max_snapshots > 0;
    const int64_t off = T.off;
    float *psi = b->nlse() ? (float *)b->d_state[0] + 2 * off : nullptr;
    double *u = b->nlse() ? nullptr : b->d_state[0] + off;
    double *up = b->nlse() ? nullptr : b->d_state[2] + off;
    double *su = keep ? b->d_snapU + (b->nlse() ? 2 * off : off) : nullptr;
    double *sv = keep && !b->nlse() ? b->d_snapV + off : nullptr;
    const unsigned grid = (unsigned)((T.N + FBS - 1) / FBS);
    hipLaunchKernelGGL(k_ic_sample, dim3(grid), dim3(FBS), 0, b->s, P, T.N, (int)b->nlse(),
                       (float)m_const, (float)c_const, u, up, psi, b->d_mf + off, b->d_cf + off,
                       su, sv);
    NLS_HIP(hipGetLastError());
    const float mc = (float)m_const, cc = (float)c_const;
    b->set_uniform(t, &mc, &cc, 1);
    if (keep) b->nsnap = std::max(b->nsnap, 1);
    NLS_HIP(hipStreamSynchronize(b->s));
  });
}

int nls_batch_run(nls_batch *b, int nsteps, float *ms) {
  return guard([&] {
    NLS_REQUIRE(b && nsteps >= 0, "nls_batch_run: bad arguments");
    b->run(nsteps, ms);
  });
}

int nls_batch_get_snapshots(nls_batch *b, int t, double *u, double *v, int *count) {
  return guard([&] {
    NLS_REQUIRE(b && t >= 0 && t < b->ntraj, "nls_batch_get_snapshots: bad trajectory index");
    const TrajDev &T = b->lay.traj[t];
    if (count) *count = b->nsnap;
    if (b->nlse()) {
      for (int k = 0; k < b->nsnap && u; ++k)
        NLS_HIP(hipMemcpyAsync(u + (size_t)k * 2 * T.N, b->d_snapU + ((size_t)k * b->lay.cells + T.off) * 2,
                               sizeof(double) * 2 * T.N, hipMemcpyDeviceToHost, b->s));
      NLS_HIP(hipStreamSynchronize(b->s));
      return;
    }
    for (int k = 0; k < b->nsnap; ++k) {
      if (u) NLS_HIP(hipMemcpyAsync(u + (size_t)k * T.N, b->d_snapU + (size_t)k * b->lay.cells + T.off,
                                    sizeof(double) * T.N, hipMemcpyDeviceToHost, b->s));
      if (v) NLS_HIP(hipMemcpyAsync(v + (size_t)k * T.N, b->d_snapV + (size_t)k * b->lay.cells + T.off,
                                    sizeof(double) * T.N, hipMemcpyDeviceToHost, b->s));
    }
    NLS_HIP(hipStreamSynchronize(b->s));
  });
}

int nls_batch_get_state(nls_batch *b, int t, double *u, double *v) {
  return guard([&] {
    NLS_REQUIRE(b && t >= 0 && t < b->ntraj, "nls_batch_get_state: bad trajectory index");
    const TrajDev &T = b->lay.traj[t];
    if (b->nlse()) {
      std::vector<float> psi(2 * T.N);
      NLS_HIP(hipMemcpyAsync(psi.data(), b->d_state[b->step & 1] + T.off, sizeof(float) * 2 * T.N,
                             hipMemcpyDeviceToHost, b->s));
      NLS_HIP(hipStreamSynchronize(b->s));
      for (int64_t i = 0; i < 2 * T.N; ++i) u[i] = psi[i];
      return;
    }
    const int cur = (int)(b->step % 3), prev = (cur + 2) % 3;
    NLS_HIP(hipMemcpyAsync(u, b->d_state[cur] + T.off, sizeof(double) * T.N, hipMemcpyDeviceToHost, b->s));
    if (v) {
      std::vector<double> up(T.N);
      NLS_HIP(hipMemcpyAsync(up.data(), b->d_state[prev] + T.off, sizeof(double) * T.N,
                             hipMemcpyDeviceToHost, b->s));
      NLS_HIP(hipStreamSynchronize(b->s));
      for (int64_t i = 0; i < T.N; ++i) v[i] = (u[i] - up[i]) / b->p.dt;
    }
    NLS_HIP(hipStreamSynchronize(b->s));
  });
}

int nls_batch_info(nls_batch *b, int *krylov_dim, int64_t *cells, size_t *device_bytes) {
  return guard([&] {
    NLS_REQUIRE(b, "null batch");
    if (krylov_dim) *krylov_dim = b->m;
    if (cells) *cells = b->total_cells();
    if (device_bytes) *device_bytes = b->bytes;
  });
}

int nls_batch_monitor(nls_batch *b, int t, int *stalls, double *max_tail, double *max_orth,
                      int *reorth_on) {
  return guard([&] {
    NLS_REQUIRE(b && t >= 0 && t < b->ntraj, "nls_batch_monitor: bad trajectory");
    TrajScal h;
    NLS_HIP(hipStreamSynchronize(b->s));
    NLS_HIP(hipMemcpy(&h, b->d_scal + t, sizeof(TrajScal), hipMemcpyDeviceToHost));
    int st = 0;
    double mt = 0.0, mo = 0.0;
    for (int s = 0; s < NSEQ; ++s) {
      st += h.stalls[s];
      mt = std::max(mt, h.max_tail[s]);
      mo = std::max(mo, h.max_orth[s]);
    }
    if (stalls) *stalls = st;
    if (max_tail) *max_tail = mt;
    if (max_orth) *max_orth = mo;
    if (reorth_on) *reorth_on = b->p.reorth == 1 || (b->reorth_live && h.reorth_on);
  });
}

int nls_batch_tile(nls_batch *b, int *cells_per_thread, int *pass_kind, int *lanes) {
  return guard([&] {
    NLS_REQUIRE(b, "nls_batch_tile: null batch");
    if (cells_per_thread) *cells_per_thread = b->cpt;
    if (pass_kind) *pass_kind = b->pass_kind;
    if (lanes) *lanes = b->lanes;
  });
}

int nls_batch_abort(nls_batch *b) {
  return guard([&] {
    NLS_REQUIRE(b, "nls_batch_abort: null batch");
    if (!b->h_abort) // before the first run: the flag the next persistent launch reads
      NLS_HIP(hipHostMalloc((void **)&b->h_abort, sizeof(int),
                            hipHostMallocCoherent | hipHostMallocMapped));
    __atomic_store_n(b->h_abort, 1, __ATOMIC_RELEASE);
  });
}

int nls_batch_persistent_steps(nls_batch *b, int64_t *steps) {
  return guard([&] {
    NLS_REQUIRE(b && steps, "nls_batch_persistent_steps: null argument");
    *steps = b->persist_steps;
  });
}

int nls_batch_set_profiling(nls_batch *b, int on) {
  return guard([&] {
    NLS_REQUIRE(b, "null batch");
    b->profiling = on != 0;
  });
}

int nls_batch_pass_timing(nls_batch *b, double *avg_ms, int *launches, double *bytes_per_launch) {
  return guard([&] {
    NLS_REQUIRE(b, "null batch");
    if (avg_ms) *avg_ms = b->pass_launches ? b->pass_ms / b->pass_launches : 0.0;
    if (launches) *launches = b->pass_launches;
    if (bytes_per_launch) *bytes_per_launch = b->pass_bytes();
  });
}

} // extern "C"
