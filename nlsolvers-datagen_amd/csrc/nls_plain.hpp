// Internal declarations shared by the plain (FP64 CSR) path and the fused path.
#pragma once

#include "nls_common.hpp"

#include <memory>

namespace nls {

struct c128 {
  double re, im;
};
struct c64 {
  float re, im;
};

// An int32 CSR operator with FP64 values in device memory, owned elsewhere.
struct CsrView {
  int64_t n;
  const int32_t *rp, *ci;
  const double *val;
};

// Device copy of an int32 CSR operator with FP64 values.
struct CsrDev {
  int64_t n = 0, nnz = 0;
  int32_t *rp = nullptr, *ci = nullptr;
  double *val = nullptr;
  CsrDev(int64_t n, int64_t nnz, const int32_t *rp_h, const int32_t *ci_h, const double *val_h);
  ~CsrDev();
  CsrDev(const CsrDev &) = delete;
  CsrDev &operator=(const CsrDev &) = delete;
  CsrView view() const { return CsrView{n, rp, ci, val}; }
};

// Assembles the full-array operator on the host and uploads it.
std::unique_ptr<CsrDev> build_operator(int dim, int n, double L, const double *c, int negate,
                                       double shift, int q6);

// Lanczos workspace of the plain path (KrylovInfo, device/include/lanczos.hpp:23-31,
// plus the matrix-function buffers of matfunc_real.hpp:135-159) for S = double or c128.
// The buffers are carved from one block of workspace_bytes(n, m): allocated by the
// object, or (mem != nullptr) device memory the caller provides and keeps.
template <class S> struct PlainKrylov {
  int64_t n;
  int m, G;
  hipStream_t stream;
  char *block = nullptr;
  bool owns = false;
  S *V = nullptr, *w = nullptr, *T = nullptr, *h = nullptr, *y = nullptr, *scal = nullptr;
  double *part = nullptr, *dscal = nullptr;
  int *status = nullptr;
  // 0: the reference's Lanczos (lanczos.hpp:208-267: recomputed lower subdiagonal, one
  // classical Gram-Schmidt pass, T[m-1][m-1] = 0); 1: textbook (symmetric T from b, no
  // Gram-Schmidt, alpha_{m-1} computed; real alpha for complex vectors)
  int variant = 0;
  PlainKrylov(int64_t n, int m, hipStream_t s, char *mem = nullptr);
  ~PlainKrylov();
  PlainKrylov(const PlainKrylov &) = delete;
  PlainKrylov &operator=(const PlainKrylov &) = delete;
  static size_t workspace_bytes(int64_t n, int m);
  void lanczos(const CsrView &A, const S *in);
  // out = beta0 V f(T) e1 for f(kind; t_re + i t_im)
  void apply(const CsrView &A, S *out, const S *in, double t_re, double t_im, int kind);
  void check_status();
};

void launch_bc_f64(double *u, int n, int dim, hipStream_t s);
void launch_bc_c128(c128 *u, int n, int dim, hipStream_t s);

} // namespace nls
