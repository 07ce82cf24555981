/*
 * nls.h -- C ABI of libnls, the MI355X (gfx950) implementation of the
 * nlsolvers-datagen time steppers: nonlinear Klein-Gordon (Gautschi, sEWI,
 * Stormer-Verlet), cubic NLSE (Strang, SEWI) and stochastic phi^4, with their
 * Krylov (Lanczos) matrix functions on 5-point / 7-point no-flux operators.
 *
 * The reference has no C ABI: its Python layer runs one CUDA executable per
 * trajectory (scripts/real_launcher_2d.py:147-170) and the executables drive C++
 * classes.  Each entry point below names the reference interface it replaces
 * (paths relative to the reference repository root).
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero status on failure;
 *    nls_last_error() then returns a thread-local message.  Nothing throws.
 *  - "host" buffers are caller-owned, C-contiguous; "device" buffers are device
 *    pointers (from nls_dev_malloc or any HIP allocation on the current device).
 *  - Complex arrays are interleaved (re, im) pairs: complex128 = 2 doubles,
 *    complex64 = 2 floats.
 *  - Grids are the FULL array including the ghost ring (n = nx = ny [= nz]),
 *    flat index i*n + j (2D) or (a*n + b)*n + k (3D); the domain is [-L, L]^dim,
 *    dx = 2L/(n-1), dt = T/nt (device/drivers/kg_driver_dev_2d.cpp:51-55).
 *  - Solver objects own their device memory and one HIP stream (default
 *    priority).  Calls are stream-ordered; functions that return host data
 *    synchronise.
 */
#ifndef NLS_H
#define NLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLS_API __attribute__((visibility("default")))

/* ---------------------------------------------------------------- status */
/* Message of the last failing call on this thread ("" if none).
 * Replaces CHECK_CUDA / CHECK_CUSPARSE print-and-throw (device/include/pragmas.hpp:8-26). */
NLS_API const char *nls_last_error(void);
NLS_API int nls_version(void);
/* Device name, compute-unit count and HBM bytes of the current device. */
NLS_API int nls_device_info(char *name, int name_len, int *cu_count, size_t *hbm_bytes);

/* ---------------------------------------------------------------- device memory helpers */
/* Thin wrappers so ctypes callers need no other runtime
 * (the reference allocates with cudaMalloc in its drivers, kg_driver_dev_2d.cpp:125-137). */
NLS_API int nls_dev_malloc(void **ptr, size_t bytes);
NLS_API int nls_dev_free(void *ptr);
NLS_API int nls_memcpy_h2d(void *dst_device, const void *src_host, size_t bytes);
NLS_API int nls_memcpy_d2h(void *dst_host, const void *src_device, size_t bytes);
NLS_API int nls_dev_synchronize(void);

/* ---------------------------------------------------------------- operator */
/* nnz of the full-array operator; q6 only matters for dim 3. */
NLS_API int64_t nls_laplacian_nnz(int dim, int n, int q6);
/* CSR (ascending columns) of div(c grad .) with no-flux ends, scaled by 1/(dx*dy),
 * optionally negated, then `shift` added to the diagonal.  Passing row_ptr = NULL
 * only returns *nnz.
 * Replaces build_anisotropic_laplacian_noflux (common/include/laplacians.hpp:54-103)
 * as called by the drivers (kg_driver_dev_2d.cpp:105-108). */
NLS_API int nls_build_laplacian_2d(int n, double dx, double dy, const double *c, int negate,
                                   double shift, int32_t *row_ptr, int32_t *col, double *val,
                                   int64_t *nnz);
/* 3D 7-point version, scaled by 1/(dx*dx); q6 = 1 keeps the reference's
 * cross-plane "y" couplings.  Replaces build_anisotropic_laplacian_noflux_3d
 * (laplacians.hpp:158-218). */
NLS_API int nls_build_laplacian_3d(int n, double dx, const double *c, int negate, double shift,
                                   int q6, int32_t *row_ptr, int32_t *col, double *val,
                                   int64_t *nnz);

/* ---------------------------------------------------------------- plain CSR SpMV (device pointers) */
/* y = A x with int32 CSR, one call per product, on `stream` (NULL = default).
 * Replaces DeviceSpMV<T>::multiply (device/include/spmv.hpp:65-72, cuSPARSE CSR ALG2). */
NLS_API int nls_csr_spmv_f64(int64_t n, const int32_t *row_ptr, const int32_t *col,
                             const double *val, const double *x, double *y, void *stream);
NLS_API int nls_csr_spmv_f32(int64_t n, const int32_t *row_ptr, const int32_t *col,
                             const float *val, const float *x, float *y, void *stream);
NLS_API int nls_csr_spmv_c128(int64_t n, const int32_t *row_ptr, const int32_t *col,
                              const double *val, const double *x, double *y, void *stream);
NLS_API int nls_csr_spmv_c64(int64_t n, const int32_t *row_ptr, const int32_t *col,
                             const float *val, const float *x, float *y, void *stream);

/* ---------------------------------------------------------------- matrix functions (FP64, plain) */
typedef struct nls_matfunc nls_matfunc;
/* Function kinds.  Real: result = beta0 V f(T) e1 with f of t*sqrt|lambda|.
 * Complex: f of dt*lambda with complex dt. */
enum {
  NLS_F_COS_SQRT = 0,   /* cos(t sqrt|l|)                    matfunc_real.hpp:221-228 */
  NLS_F_SINC2_SQRT = 1, /* (sin x/x)^2, x = t sqrt|l|        matfunc_real.hpp:212-219 */
  NLS_F_ID_SQRT = 2,    /* t sqrt|l|                         matfunc_real.hpp:230-237 */
  NLS_F_SINC2_HALF = 3, /* (sin(x/2)/(x/2))^2 (sEWI filter, no reference counterpart) */
  NLS_F_COSM1_SQRT = 4, /* cos(x) - 1 (increment form, no reference counterpart) */
  NLS_F_EXP = 8,        /* exp(dt l)                         matfunc_complex.hpp:281-287 */
  NLS_F_SINC = 9        /* sin(dt l)/(dt l)                  matfunc_complex.hpp:293-300 */
};
/* Uploads a host CSR and sizes a Krylov workspace of dimension m (2..32).
 * complex_vectors selects complex128 vectors.  Replaces the constructors of
 * MatrixFunctionApplicatorReal (matfunc_real.hpp:127-167) and
 * MatrixFunctionApplicatorComplex (matfunc_complex.hpp:124-176). */
NLS_API int nls_matfunc_create(int64_t n, int64_t nnz, const int32_t *row_ptr, const int32_t *col,
                               const double *val, int m, int complex_vectors, nls_matfunc **out);
NLS_API int nls_matfunc_destroy(nls_matfunc *h);
/* out = f(A) in on device pointers (in may alias out), reference Lanczos
 * (lanczos.hpp:208-267: T[m-1,m-1] = 0, recomputed lower subdiagonal, one
 * classical Gram-Schmidt pass per iteration).
 * Replaces MatrixFunctionApplicatorReal::apply (matfunc_real.hpp:190-313). */
NLS_API int nls_matfunc_apply_real(nls_matfunc *h, double *out, const double *in, double t,
                                   int kind);
/* Same as nls_matfunc_apply_real with no handle and no allocation: the caller passes the
 * device CSR (row_ptr, col, val of an n x n operator), in and out (device, may alias),
 * the Krylov dimension m (2..32) and a stream (NULL = default); the Krylov workspace is a
 * fixed device array of 32 MiB held by each GPU the library runs on (the sizes it takes:
 * 8 n (m + 1) bytes plus ~280 KB, e.g. n = 300^2 up to m = 32), and larger requests fail
 * with a message.  It calls no hipMalloc, hipFree or allocating library routine.  It
 * synchronises the stream to report an eigensolver failure, except while the stream is
 * being captured into a HIP graph.  All calls on one GPU share its array: calls that
 * overlap on one GPU (other streams or threads) race.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). */
NLS_API int nls_matfunc_apply_real_noalloc(int64_t n, const int32_t *row_ptr, const int32_t *col,
                                           const double *val, double *out, const double *in,
                                           double t, int kind, int m, void *stream);
/* Replaces MatrixFunctionApplicatorComplex::apply (matfunc_complex.hpp:254-375). */
NLS_API int nls_matfunc_apply_complex(nls_matfunc *h, double *out, const double *in, double dt_re,
                                      double dt_im, int kind);
/* Copies the m x m column-major T of the last apply to host (complex: 2*m*m doubles)
 * and returns beta0 (test access to KrylovInfo, lanczos.hpp:23-31). */
NLS_API int nls_matfunc_last_krylov(nls_matfunc *h, double *T_host, double *beta0);

/* FP64 GPU small stage on host data: eigen-decomposition of the symmetric
 * tridiagonal (d[m], e[m-1]) and y = beta0 Q f(L) Q^T e1 (yr, yi host arrays of m).
 * kind + 0x100 (real kinds) runs the fused batch's single-wavefront QL instead;
 * kind + 0x200 (kinds 1, 3, 4 on a T with no negative eigenvalue) its power-series route.
 * Replaces cusolverDnDsyevd + f(lambda) + GEMMs (matfunc_real.hpp:252-292,385-392). */
NLS_API int nls_small_matfunc(int m, const double *d, const double *e, int kind, double t_re,
                              double t_im, double beta0, double *yr, double *yi);

/* ---------------------------------------------------------------- Klein-Gordon solver (plain FP64) */
enum { NLS_KGE_GAUTSCHI = 0, NLS_KGE_SV = 1, NLS_KGE_SEWI = 2 };
typedef struct {
  int dim;           /* 2 or 3 */
  int n;             /* full side */
  double L;          /* domain [-L, L]^dim */
  double T;          /* final time */
  int nt;            /* time steps; nt-1 are taken (kg_driver_dev_2d.cpp:144) */
  int num_snapshots; /* snapshot k at step k*(nt/num_snapshots) */
  int krylov_dim;    /* reference: 10 (kg_driver_dev_2d.cpp:138) */
  int method;        /* NLS_KGE_* */
  double mu2;        /* mass term: operator -L + mu2 I (0 = reference equation) */
  int q6;            /* 3D: keep the reference's cross-plane couplings */
  int variant;       /* Lanczos: 0 = the reference's (lanczos.hpp:208-267: recomputed lower
                        subdiagonal, one Gram-Schmidt pass, T[m-1][m-1] = 0); 1 = textbook
                        (symmetric T, no Gram-Schmidt, alpha_{m-1} computed) */
} nls_kge_params;
typedef struct nls_kge nls_kge;
/* Assembles the operator from c (negated for Gautschi/sEWI, as the drivers do),
 * uploads m, u0, v0, sets u_{-1} = u0 - dt v0 and stores snapshot 0.
 * Replaces KGESolverDevice::KGESolverDevice (device/include/kg_dev.hpp:32-80). */
NLS_API int nls_kge_create(const nls_kge_params *p, const double *m, const double *c,
                           const double *u0, const double *v0, nls_kge **out);
/* Same, with a caller-assembled host CSR (the constructor's CSR arguments). */
NLS_API int nls_kge_create_csr(const nls_kge_params *p, int64_t nnz, const int32_t *row_ptr,
                               const int32_t *col, const double *val, const double *m,
                               const double *u0, const double *v0, nls_kge **out);
NLS_API int nls_kge_destroy(nls_kge *h);
/* KGESolverDevice::step (kg_dev.hpp:97-100 -> kg_single.cuh:57-127). */
NLS_API int nls_kge_step(nls_kge *h);
/* KGESolverDevice::step_sv (kg_dev.hpp:102-105 -> kg_single.cuh:129-179). */
NLS_API int nls_kge_step_sv(nls_kge *h);
/* KGESolverDevice::apply_bc (kg_dev.hpp:116-121 -> boundaries.cuh:9-81). */
NLS_API int nls_kge_apply_bc(nls_kge *h);
/* KGESolverDevice::store_snapshot (kg_dev.hpp:123-131). */
NLS_API int nls_kge_store_snapshot(nls_kge *h, int idx);
/* KGESolverDevice::transfer_snapshots (kg_dev.hpp:107-114): host u, v of num_snapshots*N. */
NLS_API int nls_kge_transfer_snapshots(nls_kge *h, double *u, double *v);
/* The driver time loop (kg_driver_dev_2d.cpp:144-153): nt-1 steps of the
 * configured method, Neumann copy after each, snapshot cadence nt/num_snapshots. */
NLS_API int nls_kge_run(nls_kge *h);
/* Current u (and v if non-NULL) to host. */
NLS_API int nls_kge_get_state(nls_kge *h, double *u, double *v);

/* ---------------------------------------------------------------- NLSE solver (plain complex128) */
enum { NLS_NLSE_STRANG = 0, NLS_NLSE_SEWI = 1 };
typedef struct {
  int dim, n;
  double L, T;
  int nt, num_snapshots;
  int krylov_dim; /* reference: 20 (2D Strang), 10, 25, 15 (nlse_cubic_*_driver_*.cpp) */
  int method;     /* NLS_NLSE_* */
  int q6;
  int variant;    /* Lanczos variant as nls_kge_params; 1 also forms the intended SEWI
                     update, where the reference swaps its two terms (nlse_dev.hpp:219-225) */
} nls_nlse_params;
typedef struct nls_nlse nls_nlse;
/* Replaces NLSESolverDevice::NLSESolverDevice (device/include/nlse_dev.hpp:78-163) and
 * the snapshot-0 store of the driver (nlse_cubic_driver_2d.cpp:106). u0 complex128. */
NLS_API int nls_nlse_create(const nls_nlse_params *p, const double *m, const double *c,
                            const double *u0, nls_nlse **out);
NLS_API int nls_nlse_destroy(nls_nlse *h);
/* NLSESolverDevice::step (nlse_dev.hpp:187-203); snapshot when step % freq == 0. */
NLS_API int nls_nlse_step(nls_nlse *h, int step_number);
/* NLSESolverDevice::step_sewi (nlse_dev.hpp:205-239). */
NLS_API int nls_nlse_step_sewi(nls_nlse *h, int step_number);
/* NLSESolverDevice::apply_bc (nlse_dev.hpp:178-185). */
NLS_API int nls_nlse_apply_bc(nls_nlse *h);
/* Driver loop (nlse_cubic_driver_2d.cpp:107-110) and the snapshot array (ns*N complex128). */
NLS_API int nls_nlse_run(nls_nlse *h);
NLS_API int nls_nlse_transfer_snapshots(nls_nlse *h, double *u);
/* NLSESolverDevice::store_snapshot_online (nlse_dev.hpp:324-335): the current state goes to
 * snapshot slot k (k = snapshots stored so far; nothing once num_snapshots are stored), into
 * the library's snapshot array and, if host_dst is non-NULL, to host_dst + k * N
 * (complex128, synchronous, as the reference's cudaMemcpy). */
NLS_API int nls_nlse_store_snapshot_online(nls_nlse *h, double *host_dst);
NLS_API int nls_nlse_get_state(nls_nlse *h, double *u);

/* ---------------------------------------------------------------- stochastic phi^4 (plain FP64) */
enum { NLS_SP4_REFERENCE = 0, NLS_SP4_DAMPED = 1 };
typedef struct {
  int dim, n;
  double L, T;
  int nt, num_snapshots;
  int mode;       /* NLS_SP4_REFERENCE: stochastic_phi4.cuh:19-77 form;
                     NLS_SP4_DAMPED: u_tt = Lu + m(u - u^3) - gamma u_t + sigma xi */
  int q6;
  double gamma, sigma;
  uint64_t seed;  /* Philox key */
  uint32_t stream_id;
} nls_sp4_params;
typedef struct nls_sp4 nls_sp4;
/* Replaces SP4SolverDevice::SP4SolverDevice (device/include/stochastic_phi4_dev.hpp:29-88). */
NLS_API int nls_sp4_create(const nls_sp4_params *p, const double *m, const double *c,
                           const double *u0, const double *v0, nls_sp4 **out);
NLS_API int nls_sp4_destroy(nls_sp4 *h);
/* SP4SolverDevice::step (stochastic_phi4_dev.hpp:105-109 -> stochastic_phi4.cuh:19-77). */
NLS_API int nls_sp4_step(nls_sp4 *h);
NLS_API int nls_sp4_apply_bc(nls_sp4 *h);
NLS_API int nls_sp4_store_snapshot(nls_sp4 *h, int idx);
NLS_API int nls_sp4_transfer_snapshots(nls_sp4 *h, double *u, double *v);
/* Driver loop (device/drivers/sp4_sv_driver_dev_2d.cpp:141-150). */
NLS_API int nls_sp4_run(nls_sp4 *h);

/* Philox4x32-10 normals for `count` cells of (seed, stream, step) to host
 * (replaces the per-cell thrust::default_random_engine draw, stochastic_phi4.cuh:37-53). */
NLS_API int nls_philox_normals(uint64_t seed, uint32_t stream_id, uint32_t step, int64_t count,
                               double *out_host);

/* STREAM copy of `bytes` (read + write counted) repeated `reps` times: achievable
 * HBM bandwidth on this device, reported beside the roofline figures. */
NLS_API int nls_stream_copy_bench(size_t bytes, int reps, double *gbps, double *avg_ms);

/* ---------------------------------------------------------------- fused batched path */
/* Many trajectories (uneven grids allowed) advanced together by the fused
 * FP32 matrix-free stencil+Lanczos kernels with FP64 state and FP64 small
 * stage; steps captured into HIP graphs.  Klein-Gordon with NLS_KGE_SV runs the
 * Stormer-Verlet step instead (FP64, the reference's arithmetic, no Krylov
 * passes: reorth 0, pass_kernel 0 or 1).  No reference counterpart as an
 * object: it replaces running one kg_gautschi_2d_dev / nlse_2d_dev / sp4 process
 * per trajectory (scripts/real_launcher_2d.py:260-306). */
enum { NLS_EQ_KGE = 0, NLS_EQ_NLSE = 1, NLS_EQ_SP4 = 2 };
typedef struct {
  int equation;      /* NLS_EQ_* */
  int method;        /* KGE: NLS_KGE_*; NLSE: NLS_NLSE_STRANG; SP4: NLS_SP4_* */
  int dim;
  double dt;
  int krylov_dim;    /* 0 = pick with the autotuner */
  double mu2;        /* KGE mass term */
  int q6;
  double gamma;      /* SP4 damping */
  int use_graph;     /* capture steps into HIP graphs */
  int snapshot_every;/* steps between snapshots (0 = none kept on device) */
  int max_snapshots; /* snapshot slots per trajectory */
  int cells_per_thread; /* tile: cells per workgroup / 256, in [4, 64]; 0 = autotune.
                           The tiled pass kernels take 4, 8 or 16; the row-march kernels
                           any value (each trajectory is cut into equal ranges of about
                           that size).  A fixed tile makes every trajectory's reduction
                           order, hence its result, independent of the other
                           trajectories in the batch */
  int pass_kernel;   /* 0 = chosen with the tile (autotuner, or the preferred kernel for a
                        fixed tile); k + 1 = use pass kernel k only (0 direct stencil,
                        1 LDS-tiled chunk, 2 row march, 3 row march storing only the
                        basis, 4 plane march (3D, tile 4 or 8)), autotuning its tile
                        when cells_per_thread is 0 */
  int lanes;         /* parts of the batch stepped concurrently on their own streams,
                        1..4; 0 = autotune (2 with a fixed tile) */
  int reorth;        /* full reorthogonalisation of the FP32 Krylov basis: classical
                        Gram-Schmidt with FP64 dot products and coefficients after each
                        Lanczos pass (the reference's one CGS pass per iteration,
                        device/include/lanczos.hpp:249-250).  0 = never; 1 = in every
                        iteration; 2 = for a trajectory from the step after its basis lost
                        semi-orthogonality, |<v_{m-2}, v_0>| > reorth_tol (complex
                        Gram-Schmidt for NLSE).  Not with pass kernel 3 */
  double reorth_tol; /* threshold of mode 2; 0 = 2.5e-4 (about sqrt of the FP32 roundoff) */
  int always_load_fields; /* 0: a trajectory whose m or c is spatially uniform (found at
                        upload) has the row- and plane-march passes and the combines use
                        the value instead of streaming the field (the same numbers, 4 B per
                        cell and pass less); 1: always stream the fields */
  int variant;       /* Lanczos variant of the Krylov passes (KGE Gautschi / sEWI, NLSE):
                        0 = the reference's, T[m-1][m-1] = 0 (device/include/lanczos.hpp:
                        234-263); 1 = full tridiagonal T: one more pass per step forms
                        v_{m-1} on the fly and computes alpha_{m-1} = <A v_{m-1}, v_{m-1}>
                        (it reads 16 B per cell for KGE and stores nothing).  Variant 1 runs
                        on pass kernels 0, 2 and 4.  With krylov_dim = 0 the autotuner
                        picks m >= 10 for variant 0 (the reference drivers' m) and
                        m >= 8 for variant 1 */
  int persistent;    /* 1: run the steps between snapshots as one persistent cooperative
                        launch (2D Klein-Gordon Gautschi / sEWI on the row-march pass, reorth
                        0 or 2): one workgroup per chunk for the whole launch, per-trajectory
                        arrival counters between the phases of a step (bounded waits; an
                        expired wait fails the run), the small stage run by the chunk that
                        arrives last.  The tile is cells_per_thread, or the smallest whose
                        chunks fit the workgroups resident at once.  Results are bitwise those
                        of the graph-captured steps at the same tile.  Snapshot steps, and
                        every step once reorthogonalisation has switched on, run the captured
                        steps.  0: graph-captured steps only */
} nls_batch_params;
typedef struct {
  int n;             /* full side */
  double L;          /* half-width */
  double sigma;      /* SP4 noise amplitude */
  uint64_t seed;     /* SP4 Philox key */
} nls_traj_desc;
typedef struct nls_batch nls_batch;
NLS_API int nls_batch_create(const nls_batch_params *p, int ntraj, const nls_traj_desc *trajs,
                             nls_batch **out);
NLS_API int nls_batch_destroy(nls_batch *b);
/* Host inputs of trajectory t: m, c (N doubles), u0 (N doubles, or N complex128
 * for NLSE), v0 (N doubles, ignored for NLSE). Resets its step counter and
 * stores snapshot 0. */
NLS_API int nls_batch_set_initial(nls_batch *b, int t, const double *m, const double *c,
                                  const double *u0, const double *v0);
/* Seeded initial-condition samplers of the benchmark configurations (the Python
 * samplers nls/samplers.py gaussian_sum_2d, kink_ring_2d, nlse_packets_2d, bubbles_3d;
 * field forms from util.hpp:94-125 (Gaussians, 2D and 3D), real_sampler.py:464-467
 * (kink ring), nlse_sampler.py:43-100 (packets)).  Draw s of trajectory traj is Philox4x32-10 at counter
 * (s, traj, tag, 0) under the 64-bit seed. */
enum { NLS_IC_GAUSSIANS_2D = 0, NLS_IC_KINK_RING_2D = 1, NLS_IC_PACKETS_2D = 2,
       NLS_IC_BUBBLES_3D = 3 };
/* The parameters drawn for one trajectory on [-L, L]^dim, in draw order (host only, no
 * device use): gaussians [A, w, cx, cy] x 3; kink ring [R, w]; packets
 * [K, (A, phi, vx, vy, cx, cy) x K]; bubbles [K, (A, w, cx, cy, cz) x K].  Centres are
 * U(-0.6 L, 0.6 L) (2D) or U(-0.5 L, 0.5 L) (3D) and ring radii U(0.2 L, 0.6 L).  *count
 * receives their number (at most 32). */
NLS_API int nls_ic_draw(int kind, uint64_t seed, int traj, double L, double *params, int cap,
                        int *count);
/* Initial data of trajectory t drawn and evaluated on the GPU, in place of
 * nls_batch_set_initial with host arrays: u0 (psi0 for NLS_IC_PACKETS_2D and an NLSE
 * batch) of sampler `kind` with Philox trajectory counter traj_index, v0 = 0, m and c
 * constant.  Stores snapshot 0. */
NLS_API int nls_batch_set_initial_sampled(nls_batch *b, int t, int kind, uint64_t seed,
                                          int traj_index, double m_const, double c_const);
/* Coefficient field of class `type` on the GPU (replaces the host generators
 * scripts/c_fields_2d.py:169-248 generate_c_fields and scripts/m_fields_2d.py:147-238
 * generate_m_fields, one field per call): family "c" (constant, periodic_structure,
 * piecewise_constant, sign_changing_mass, layered, waveguide, quasiperiodic, turbulent;
 * base = base_value) or "m" (constant, piecewise, gradient, phase, topological, defects,
 * quasiperiodic, multiscale; base = m0; gradient / phase / topological derive from the
 * device field c, n*n doubles, and fall back to constant when c is NULL, as the
 * reference's dispatch does).  out: device, n*n doubles on linspace(-L, L, n)^2, "ij"
 * order.  Parameters come from the trajectory's Philox stream (tags 0xC0C0 for c,
 * 0x3A55 for m), the same draws as nls/fields.py.  stream: a hipStream_t or NULL;
 * returns after the field is written. */
NLS_API int nls_field_sample(const char *family, const char *type, int n, double L,
                             uint64_t seed, int traj, double base, const double *c,
                             double *out, void *stream);
/* The reference launchers' phenomena on the GPU, their parameters drawn from the
 * reference's spaces: "kink_field" (real_sampler.py:223-271 with the anisotropic random
 * velocity field :48-65; u, v device n*n doubles: u0, v0) or "multi_soliton"
 * (nlse_sampler.py:102-219; u device 2*n*n doubles, interleaved complex psi0, divided by
 * max |psi0| when normalize != 0; v unused).  choices (host, may be NULL) receives the
 * index of the value chosen for each key of the space, in key order (kink_field: 5;
 * multi_soliton: 9, then the soliton count); *count their number. */
NLS_API int nls_phenomenon_sample(const char *name, int n, double L, uint64_t seed, int traj,
                                  int normalize, double *u, double *v, int *choices, int cap,
                                  int *count, void *stream);
/* Advances every trajectory by nsteps; *ms receives the HIP-event time of the
 * region on the batch stream (NULL to skip). */
NLS_API int nls_batch_run(nls_batch *b, int nsteps, float *ms);
/* Snapshots taken so far of trajectory t (count*N doubles each; complex128 for NLSE). */
NLS_API int nls_batch_get_snapshots(nls_batch *b, int t, double *u, double *v, int *count);
/* Current state of trajectory t. */
NLS_API int nls_batch_get_state(nls_batch *b, int t, double *u, double *v);
/* Krylov dimension in use, total cells, bytes of device memory held. */
NLS_API int nls_batch_info(nls_batch *b, int *krylov_dim, int64_t *cells, size_t *device_bytes);
/* Per-kernel timing of the last nls_batch_run with profiling on: average
 * duration (ms) and launches of the dominant Lanczos pass kernel. */
NLS_API int nls_batch_set_profiling(nls_batch *b, int on);
NLS_API int nls_batch_pass_timing(nls_batch *b, double *avg_ms, int *launches,
                                  double *bytes_per_launch);
/* Tile in use after the first nls_batch_run: cells per thread, the pass kernel
 * (0 direct stencil, 1 LDS-tiled chunk, 2 row march with LDS window, 3 row march
 * storing only the basis vectors, w' recomputed) and the number of lanes (parts of
 * the batch stepped concurrently on their own streams); any output may be NULL. */
NLS_API int nls_batch_tile(nls_batch *b, int *cells_per_thread, int *pass_kind, int *lanes);
/* Steps run so far by the persistent step kernel (nls_batch_params.persistent). */
NLS_API int nls_batch_persistent_steps(nls_batch *b, int64_t *steps);
/* Ends the persistent launch in progress, or the next one, from any host thread: one
 * workgroup reads this flag (pinned host memory) at the start of every step and turns
 * it into the kernel's error flag, which every wait reads while it polls (a waiting
 * workgroup also reads the host flag every 4096 polls), so every workgroup exits and
 * that nls_batch_run fails with "aborted by the host".
 * The flag is cleared when the run returns.  The batch then holds a partly advanced
 * step; destroy it.  Captured (non-persistent) steps are not affected.  No reference
 * counterpart (the reference runs one process per trajectory and stops it). */
NLS_API int nls_batch_abort(nls_batch *b);
/* Krylov monitor of trajectory t since creation: steps in which the weight of the last
 * Lanczos vector in f(T) e1, |y_{m-1}| / |y|, exceeded 1e-6 (the residual did not fall
 * within the Krylov dimension), the largest such weight, the largest loss of
 * orthogonality |<v_{m-2}, v_0>| seen (reorth modes 1 and 2; 0 otherwise) and whether
 * full reorthogonalisation is active for it.  Any output may be NULL. */
NLS_API int nls_batch_monitor(nls_batch *b, int t, int *stalls, double *max_tail,
                              double *max_orth, int *reorth_on);
/* Device bytes a batch with these trajectories would hold (admission control). */
NLS_API int nls_batch_estimate_bytes(const nls_batch_params *p, int ntraj,
                                     const nls_traj_desc *trajs, size_t *bytes);

/* ---------------------------------------------------------------- HDF5 trajectory files */
/* One file per trajectory in the launchers' layout: groups metadata (problem_type,
 * boundary_condition = "noflux", run_id, run_index, timestamp, elapsed_time,
 * phenomenon, phenomenon_<key>), grid (nx, ny[, nz], Lx, Ly[, Lz]), time (T, nt,
 * num_snapshots), initial_condition/u0[, v0], focusing/m (attr type) and c (real
 * runs: focusing/c; complex runs: root c), root u[, v], X, Y[, Z].  Replaces
 * Launcher.save_to_hdf5 (scripts/real_launcher_2d.py:201-245,
 * scripts/complex_launcher_2d.py:193-236).  libhdf5 is loaded at run time
 * ($NLS_HDF5_LIB, libhdf5.so, libhdf5.so.103, /opt/conda/lib/libhdf5.so); without it
 * these calls fail with a message and the rest of the library is unaffected. */
NLS_API int nls_h5_available(char *version, int len);
typedef struct {
  const char *path;
  int complex_field;   /* 1: NLSE run (u0, u complex128; no v0, v) */
  int dim;             /* 2 or 3 */
  int nx, ny, nz;      /* solver grid (nz = 1 in 2D) */
  double Lx, Ly, Lz;
  double T;
  int nt, num_snapshots;
  const char *problem_type, *run_id, *phenomenon, *m_type;
  const char *timestamp; /* NULL: current local time */
  int run_index;
  double elapsed_time;
  int n_params;        /* metadata attributes phenomenon_<key> = value (strings) */
  const char *const *param_keys;
  const char *const *param_values;
  const double *u0, *v0, *m, *c; /* solver-grid arrays; v0 ignored for complex runs */
  int ns_out, ox, oy, oz;         /* stored trajectory shape (after any downsampling) */
  const double *u, *v;            /* ns_out * ox * oy * oz values (u complex128 if complex) */
  const char *c_type;             /* conductivity class, metadata attribute c_type (NULL or
                                     "" = none; the reference keeps it in the directory name) */
} nls_h5_run;
/* Writes the file now. */
NLS_API int nls_h5_write_run(const nls_h5_run *r);
/* Asynchronous writer: submit copies the arrays and returns (waiting only while
 * more than max_queued_bytes are queued); a background thread writes the files in
 * submission order.  drain waits for the queue and reports the first failure. */
typedef struct nls_h5_writer nls_h5_writer;
NLS_API int nls_h5_writer_create(size_t max_queued_bytes, nls_h5_writer **out);
NLS_API int nls_h5_writer_submit(nls_h5_writer *w, const nls_h5_run *r);
NLS_API int nls_h5_writer_drain(nls_h5_writer *w, int64_t *files_written);
NLS_API int nls_h5_writer_destroy(nls_h5_writer *w);
/* Read-back: dataset `name` (e.g. "u", "focusing/m") as doubles (complex: interleaved
 * re, im, *is_complex = 1); dims receives up to 4 extents.  cap = capacity of buf in
 * doubles (buf may be NULL to query dims and type). */
NLS_API int nls_h5_read_dataset(const char *path, const char *name, double *buf, int64_t cap,
                                int64_t *dims, int *ndim, int *is_complex);
/* Elements [offset, offset + count) of the 1-D real dataset `name` as float32 (the flat
 * per-class files' u, v, m, c are read one trajectory at a time this way). */
NLS_API int nls_h5_read_range_f32(const char *path, const char *name, int64_t offset,
                                  int64_t count, float *buf);
/* Attribute `name` of object `obj` ("metadata", "grid", ..., "/" for the file's global
 * attributes): numeric into *num, or string (variable- or fixed-length, as NetCDF-4 text
 * attributes are) into str (len bytes, NUL-terminated); *is_string tells which. */
NLS_API int nls_h5_read_attr(const char *path, const char *obj, const char *name, double *num,
                             char *str, int len, int *is_string);

/* ---------------------------------------------------------------- dataset packaging */
/* Per-class NetCDF-4 file of many trajectories, as the reference's aggregation writes it
 * (aggregate_single_core.py:167-206, netCDF4.Dataset(format="NETCDF4")): dimensions
 * total_temporal = num_trajectories * num_snapshots * cells and total_spatial =
 * num_trajectories * cells; float variables u, v (complex classes: re_u, im_u) over
 * total_temporal with chunks of 65536, u0, v0 (re_u0, im_u0) over total_spatial with
 * chunks of 65536, m and c over total_spatial; global attributes num_snapshots,
 * num_trajectories, spatial_shape ("(nx, ny)"), Lx, Ly[, Lz], is_complex, dims.  Written
 * through libhdf5 (+ libhdf5_hl for the dimension scales) with the NetCDF-4 conventions;
 * trajectory t's arrays go to offset t * their per-trajectory size. */
typedef struct nls_nc_class nls_nc_class;
NLS_API int nls_nc_class_create(const char *path, int dims, const int *shape, int num_snapshots,
                                int num_trajectories, int is_complex, double Lx, double Ly,
                                double Lz, nls_nc_class **out);
/* One trajectory (float32 host arrays; complex classes: u = Re u, v = Im u, u0 = Re u0,
 * v0 = Im u0); a NULL array leaves that part at netCDF's fill value. */
NLS_API int nls_nc_class_put(nls_nc_class *h, int traj, const float *u, const float *v,
                             const float *u0, const float *v0, const float *m, const float *c);
NLS_API int nls_nc_class_close(nls_nc_class *h);
/* The canonical per-class NetCDF-4 file that canonicalize_data.py:174-262 makes from a
 * class file: dimensions member, time, x, y[, z] with coordinate variables (member int
 * 0..num_trajectories-1, time = linspace(0, T, num_snapshots), x/y/z = linspace(-L, L, n)),
 * float variables u, v (complex classes: u_real, u_imag) over (member, time, x, y[, z])
 * chunked one trajectory per chunk, m and c over (member, x, y[, z]), and the global
 * attributes pde_type, boundary_condition, num_trajectories, num_snapshots, nt,
 * spatial_shape, dims, is_complex, Lx, Ly[, Lz], T and, when not NULL, m_classes and
 * c_classes.  Same NetCDF-4 conventions as nls_nc_class. */
typedef struct nls_nc_std nls_nc_std;
NLS_API int nls_nc_std_create(const char *path, int dims, const int *shape, int num_snapshots,
                              int num_trajectories, int is_complex, double Lx, double Ly,
                              double Lz, double T, int nt, const char *pde_type,
                              const char *boundary_condition, const char *m_classes,
                              const char *c_classes, nls_nc_std **out);
/* One trajectory (float32 host arrays): u, v of num_snapshots x cells (complex classes:
 * the real and imaginary parts of u), m, c of cells; a NULL array leaves that part at
 * netCDF's fill value. */
NLS_API int nls_nc_std_put(nls_nc_std *h, int traj, const float *u, const float *v,
                           const float *m, const float *c);
NLS_API int nls_nc_std_close(nls_nc_std *h);

/* ---------------------------------------------------------------- serving loop */
/* Continuous batching of stochastic phi^4 trajectory requests of different
 * sizes on one GPU.  No reference counterpart as a service: it replaces the
 * SLURM job arrays that run one executable per trajectory
 * (scripts/kg_2d_launch.sh, scripts/real_launcher_2d.py:260-306) and the
 * per-step noise of SP4Solver::step (device/include/stochastic_phi4.cuh:19-77).
 * Requests are admitted first-in first-out while they fit the HBM budget (one
 * pre-allocated slab) and the slot limit; every launch advances every active
 * request by one Stormer-Verlet step with Philox noise keyed by the request's
 * seed, so a request's result does not depend on its batch-mates. */
typedef struct {
  size_t hbm_budget_bytes; /* device slab holding all admitted requests */
  int max_active;          /* concurrent requests */
  double L;                /* domain [-L, L]^2 */
  int mode;                /* NLS_SP4_DAMPED or NLS_SP4_REFERENCE */
  double gamma;            /* damping */
  int max_steps_per_group; /* steps launched between admission checks (0 = 64) */
  const char *h5_dir;      /* if set: each request streams its snapshots (u and
                              v = (u_k - u_{k-1}) / dt) to h5_dir/run_<run_id>_<id:04d>.h5
                              in the launchers' layout, written by a background thread */
  const char *run_id;      /* file name / metadata run_id (NULL: "serve") */
} nls_server_params;
typedef struct {
  int n;             /* grid side */
  int nt;            /* nt-1 steps are taken */
  double dt;
  double sigma;      /* noise amplitude */
  uint64_t seed;     /* Philox key */
  int num_snapshots; /* snapshot k at step k*(nt/num_snapshots); 0 = final state only */
} nls_request;
typedef struct {
  int64_t completed, failed, active, queued;
  double p50_ms, p99_ms, mean_ms; /* completion - submission */
  double cell_steps;              /* total cell-steps executed */
  double busy_s;                  /* time spent in launched step groups */
  size_t peak_bytes;
  int peak_active;
} nls_server_stats;
typedef struct nls_server nls_server;
NLS_API int nls_server_create(const nls_server_params *p, nls_server **out);
NLS_API int nls_server_destroy(nls_server *s);
/* m, c, u0, v0: host n*n doubles or NULL (m = c = 1, u0 = v0 = 0). */
NLS_API int nls_server_submit(nls_server *s, const nls_request *q, const double *m, const double *c,
                              const double *u0, const double *v0, int64_t *id);
/* state: 0 queued, 1 running, 2 done, 3 failed */
NLS_API int nls_server_status(nls_server *s, int64_t id, int *state);
NLS_API int nls_server_wait(nls_server *s, int64_t id, double timeout_s);
NLS_API int nls_server_drain(nls_server *s, double timeout_s);
/* snapshots (count*n*n doubles), final state (n*n), latency; any output may be NULL. */
NLS_API int nls_server_fetch(nls_server *s, int64_t id, double *snapshots, int *count,
                             double *u_final, double *latency_ms);
NLS_API int nls_server_release(nls_server *s, int64_t id);
NLS_API int nls_server_get_stats(nls_server *s, nls_server_stats *st);
/* Host-only check of the admission allocator (first-fit over one slab):
 * ops[2k] = 0 allocates ops[2k+1] bytes, 1 frees the block of op ops[2k+1];
 * results[k] = offset (-1 if it does not fit) or bytes in use after a free. */
NLS_API int nls_pool_simulate(size_t capacity, int nops, const int64_t *ops, int64_t *results);

#ifdef __cplusplus
}
#endif
#endif /* NLS_H */
