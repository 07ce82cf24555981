// Host-side CSR assembly of the anisotropic no-flux operator div(c grad .).
//
// Reference: build_anisotropic_laplacian_noflux (common/include/laplacians.hpp:54-103)
// and build_anisotropic_laplacian_noflux_3d (laplacians.hpp:158-218).  The reference
// goes through Eigen triplets; the values it produces are, for every row r:
//   off-diagonal  w(a,b) = (c[a] + c[b]) / 2.0 times s = 1.0/(dx*dy)  (3D: 1.0/(dx*dx))
//   diagonal      -(sum of the row's w, accumulated x-, x+, y-, y+[, z-, z+]) times s
// with the operator spanning the full array, ghost ring included.  Rows are
// emitted directly in ascending column order (Eigen's compressed layout; the
// matrix is symmetric, so the drivers pass CSC as CSR, kg_driver_dev_2d.cpp:132-137).
// negate flips the sign (Gautschi drivers build -L, kg_driver_dev_2d.cpp:105-108);
// shift is then added to the diagonal (the spectral shift mu2 of the Klein-Gordon
// equations u_tt = Lu - mu2 u - m u^3 of the benchmark workloads, no reference
// counterpart).
// q6 keeps the reference's 3D "y" couplings across plane boundaries
// (laplacians.hpp:192-198); q6 = 0 gives the plane-respecting 7-point operator.
#include "nls_common.hpp"
#include "nls.h"

#include <vector>

namespace nls {

int64_t lap2d_nnz(int n) { return (int64_t)n * n + 4LL * n * (n - 1); }

int64_t lap3d_nnz(int n, int q6) {
  const int64_t n2 = (int64_t)n * n, N = n2 * n;
  const int64_t px = n2 * (n - 1), py = q6 ? (N - n) : n2 * (n - 1), pz = N - n2;
  return N + 2 * (px + py + pz);
}

static void build2d(int n, double dx, double dy, const double *c, int negate, double shift,
                    int32_t *rp, int32_t *ci, double *val) {
  const int64_t N = (int64_t)n * n;
  const double s = 1.0 / (dx * dy);
  const double sg = negate ? -1.0 : 1.0;
  int64_t k = 0;
  for (int64_t r = 0; r < N; ++r) {
    const int64_t i = r / n, j = r - i * n;
    const bool hl = j > 0, hr = j < n - 1, hu = i > 0, hd = i < n - 1;
    const double wl = hl ? (c[r - 1] + c[r]) / 2.0 : 0.0;
    const double wr = hr ? (c[r] + c[r + 1]) / 2.0 : 0.0;
    const double wu = hu ? (c[r - n] + c[r]) / 2.0 : 0.0;
    const double wd = hd ? (c[r] + c[r + n]) / 2.0 : 0.0;
    double acc = 0.0;
    if (hl) acc += wl;
    if (hr) acc += wr;
    if (hu) acc += wu;
    if (hd) acc += wd;
    rp[r] = (int32_t)k;
    if (hu) { ci[k] = (int32_t)(r - n); val[k++] = sg * (wu * s); }
    if (hl) { ci[k] = (int32_t)(r - 1); val[k++] = sg * (wl * s); }
    ci[k] = (int32_t)r;
    val[k++] = sg * (-acc * s) + shift;
    if (hr) { ci[k] = (int32_t)(r + 1); val[k++] = sg * (wr * s); }
    if (hd) { ci[k] = (int32_t)(r + n); val[k++] = sg * (wd * s); }
  }
  rp[N] = (int32_t)k;
}

static void build3d(int n, double dx, const double *c, int negate, double shift, int q6,
                    int32_t *rp, int32_t *ci, double *val) {
  const int64_t n2 = (int64_t)n * n, N = n2 * n;
  const double s = 1.0 / (dx * dx);
  const double sg = negate ? -1.0 : 1.0;
  int64_t k = 0;
  for (int64_t r = 0; r < N; ++r) {
    const int64_t i = r % n, inplane = r % n2;
    const bool hxm = i > 0, hxp = i < n - 1;
    const bool hym = q6 ? (r >= n) : (inplane >= n);
    const bool hyp = q6 ? (r + n < N) : (inplane < n2 - n);
    const bool hzm = r >= n2, hzp = r + n2 < N;
    const double wxm = hxm ? (c[r - 1] + c[r]) / 2.0 : 0.0;
    const double wxp = hxp ? (c[r] + c[r + 1]) / 2.0 : 0.0;
    const double wym = hym ? (c[r - n] + c[r]) / 2.0 : 0.0;
    const double wyp = hyp ? (c[r] + c[r + n]) / 2.0 : 0.0;
    const double wzm = hzm ? (c[r - n2] + c[r]) / 2.0 : 0.0;
    const double wzp = hzp ? (c[r] + c[r + n2]) / 2.0 : 0.0;
    double acc = 0.0;
    if (hxm) acc += wxm;
    if (hxp) acc += wxp;
    if (hym) acc += wym;
    if (hyp) acc += wyp;
    if (hzm) acc += wzm;
    if (hzp) acc += wzp;
    rp[r] = (int32_t)k;
    if (hzm) { ci[k] = (int32_t)(r - n2); val[k++] = sg * (wzm * s); }
    if (hym) { ci[k] = (int32_t)(r - n); val[k++] = sg * (wym * s); }
    if (hxm) { ci[k] = (int32_t)(r - 1); val[k++] = sg * (wxm * s); }
    ci[k] = (int32_t)r;
    val[k++] = sg * (-acc * s) + shift;
    if (hxp) { ci[k] = (int32_t)(r + 1); val[k++] = sg * (wxp * s); }
    if (hyp) { ci[k] = (int32_t)(r + n); val[k++] = sg * (wyp * s); }
    if (hzp) { ci[k] = (int32_t)(r + n2); val[k++] = sg * (wzp * s); }
  }
  rp[N] = (int32_t)k;
}

} // namespace nls

extern "C" {

int64_t nls_laplacian_nnz(int dim, int n, int q6) {
  if (n < 2) return 0;
  return dim == 3 ? nls::lap3d_nnz(n, q6) : nls::lap2d_nnz(n);
}

int nls_build_laplacian_2d(int n, double dx, double dy, const double *c, int negate,
                           double shift, int32_t *row_ptr, int32_t *col, double *val,
                           int64_t *nnz) {
  return nls::guard([&] {
    NLS_REQUIRE(n >= 2, "nls_build_laplacian_2d: n must be >= 2");
    NLS_REQUIRE((int64_t)n * n + 4LL * n * (n - 1) < (1LL << 31), "operator too large for int32 CSR");
    if (nnz) *nnz = nls::lap2d_nnz(n);
    if (!row_ptr) return; // size query
    NLS_REQUIRE(c && col && val, "nls_build_laplacian_2d: null buffer");
    nls::build2d(n, dx, dy, c, negate, shift, row_ptr, col, val);
  });
}

int nls_build_laplacian_3d(int n, double dx, const double *c, int negate, double shift, int q6,
                           int32_t *row_ptr, int32_t *col, double *val, int64_t *nnz) {
  return nls::guard([&] {
    NLS_REQUIRE(n >= 2, "nls_build_laplacian_3d: n must be >= 2");
    NLS_REQUIRE(nls::lap3d_nnz(n, q6) < (1LL << 31), "operator too large for int32 CSR");
    if (nnz) *nnz = nls::lap3d_nnz(n, q6);
    if (!row_ptr) return;
    NLS_REQUIRE(c && col && val, "nls_build_laplacian_3d: null buffer");
    nls::build3d(n, dx, c, negate, shift, q6, row_ptr, col, val);
  });
}

} // extern "C"
