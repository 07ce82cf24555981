// Serving loop for stochastic phi^4 trajectory requests (continuous batching).
//
// Each request is one trajectory of the damped stochastic phi^4 equation
//   u_tt = div(c grad u) + m (u - u^3) - gamma u_t + sigma xi
// (or the reference's form, stochastic_phi4.cuh:19-77) stepped by
// Stormer-Verlet with Philox noise that is a pure function of (seed, step,
// cell), so a request's result does not depend on what else is on the GPU.
//
// A worker thread owns the device.  It admits queued requests (FIFO) while the
// HBM budget (one pre-allocated slab, first-fit, nls_pool.hpp) and the slot
// limit allow, then launches groups of steps: every launch advances every
// active request by one step (its own step counter, grid size and buffers),
// with the Neumann ghost copy folded in.  A group ends at the next snapshot
// or completion of any active request; snapshots and final states go to host
// memory, finished requests release their memory, and new ones are admitted
// between groups.  Latency = completion - submission, wall clock.
//
// With an HDF5 directory set, each request's file (the launchers' per-run layout)
// is opened at admission and every snapshot is appended as it reaches host memory;
// the HDF5 calls run on a writer thread of their own, so disk never stalls the
// device loop.
#include "nls.h"
#include "nls_common.hpp"
#include "nls_h5.hpp"
#include "nls_philox.hpp"
#include "nls_pool.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

namespace nls {
namespace serve {

constexpr int SBS = 256;
// One cell per thread: a step of a small group is latency-bound (dependent loads of the
// block map, the request, the state; Philox and the FP64 Box-Muller), so short threads
// and many workgroups finish it sooner (4 cells per thread: p50 62 ms, 1 cell: 30 ms on
// the Poisson 50 req/s workload).
constexpr int SCPT = 1;
// Largest grid side of a request: its cells are indexed with 32-bit integers.
constexpr int kMaxSide = 46340;
constexpr int NXCD = 8;

struct ReqDev {
  int n, nchunks, mode;
  int uni; // m and c spatially uniform: mval, cval instead of the fields (the same numbers)
  int64_t N;
  double *u0, *u1, *u2; // state ring
  const float *mf, *cf;
  int64_t base_step;    // steps completed before this launch group
  uint64_t seed;
  uint32_t stream, pad2;
  double tau, s, gp_inv, gm, t2, namp, s_eff;
  float mval, cval;
};

// UNI: the request's m and c are uniform (values from the descriptor, no field loads).
// g: the cell's index (row, col), which also keys its noise.
template <bool UNI>
__device__ inline double sp4_new(const ReqDev &R, const double *u, const double *up, int g,
                                 int row, int col, uint32_t it) {
  const int n = R.n;
  const double ui = u[g];
  const double ci = UNI ? (double)R.cval : (double)R.cf[g];
  auto cn = [&](int64_t o) { return UNI ? ci : (double)R.cf[o]; };
  double lap = 0.0;
  // div(c grad u) with edge weights (c_i + c_j)/2 (laplacians.hpp:75-91), 5-point
  if (row > 0) lap += 0.5 * (ci + cn(g - n)) * (u[g - n] - ui);
  if (col > 0) lap += 0.5 * (ci + cn(g - 1)) * (u[g - 1] - ui);
  if (col < n - 1) lap += 0.5 * (ci + cn(g + 1)) * (u[g + 1] - ui);
  if (row < n - 1) lap += 0.5 * (ci + cn(g + n)) * (u[g + n] - ui);
  lap *= R.s;
  const double z = noise1(R.seed, R.stream, it, (uint64_t)g);
  const double mv = UNI ? (double)R.mval : (double)R.mf[g];
  if (R.mode == NLS_SP4_REFERENCE) {
    const double nl = -mv * (ui - ui * ui * ui + R.s_eff * z);
    double x = 2.0 * ui - up[g];
    x = x + R.t2 * lap;
    return x + R.t2 * nl;
  }
  const double f = lap + mv * (ui - ui * ui * ui);
  return (2.0 * ui - R.gm * up[g] + R.t2 * f + R.namp * z) * R.gp_inv;
}

__global__ __launch_bounds__(SBS) void k_sp4_group_step(const ReqDev *__restrict__ reqs,
                                                         const int2 *__restrict__ cmap, int k) {
  const int2 cm = cmap[blockIdx.x];
  if (cm.x < 0) return;
  const ReqDev R = reqs[cm.x];
  const int64_t st = R.base_step + k;
  double *ring[3] = {R.u0, R.u1, R.u2};
  const double *u = ring[st % 3], *up = ring[(st + 2) % 3];
  double *un = ring[(st + 1) % 3];
  const uint32_t it = (uint32_t)(st + 1);
  const int n = R.n;
  auto cl = [n](int a) { return a < 1 ? 1 : (a > n - 2 ? n - 2 : a); };
  // 32-bit cell indices (a request has at most kMaxSide^2 cells): the division that gives
  // the row is on the chain from the thread's index to its loads
  const int base = cm.y * (SBS * SCPT);
#pragma unroll
  for (int kk = 0; kk < SCPT; ++kk) {
    const int i = base + (int)threadIdx.x + kk * SBS;
    if (i >= (int)R.N) break;
    const int r = (int)((unsigned)i / (unsigned)n), c = i - r * n;
    const int rs = cl(r), cs = cl(c);
    const int src = rs * n + cs; // Neumann ghost copy (boundaries.cuh:9-19)
    un[i] = R.uni ? sp4_new<true>(R, u, up, src, rs, cs, it)
                  : sp4_new<false>(R, u, up, src, rs, cs, it);
  }
}

struct Request {
  int64_t id;
  nls_request q;
  std::vector<float> mf, cf;
  bool uni = false; // mf and cf spatially uniform, with the values mval, cval
  float mval = 1.0f, cval = 1.0f;
  std::vector<double> u0, v0;
  std::vector<double> m64, c64;        // caller's fields, kept for the HDF5 header
  std::shared_ptr<h5::Stream> stream;  // touched by the writer thread only
  int64_t N = 0;
  size_t bytes = 0, off = SIZE_MAX;
  int64_t steps_done = 0;
  int snap_every = 0, nsnap = 0;
  std::vector<double> snaps, ufinal;
  int state = 0; // 0 queued 1 active 2 done 3 failed
  std::string error;
  std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
  double latency_ms = 0.0;
};

} // namespace serve
} // namespace nls

using namespace nls;
using namespace nls::serve;

struct nls_server {
  nls_server_params p{};
  hipStream_t s = nullptr;
  char *slab = nullptr;
  std::unique_ptr<SlabPool> pool;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<std::shared_ptr<Request>> queue;
  std::vector<std::shared_ptr<Request>> active;
  std::map<int64_t, std::shared_ptr<Request>> all;
  std::atomic<bool> stop{false};
  std::thread worker;
  int64_t next_id = 1;
  // stats
  double cell_steps = 0.0, busy_s = 0.0;
  size_t peak_bytes = 0;
  int peak_active = 0;
  int64_t completed = 0, failed = 0;
  std::vector<double> latencies;
  std::string fatal;
  // HDF5 streaming: jobs run in order on their own thread
  std::string h5_dir, run_id;
  std::thread h5_thread;
  std::mutex h5_mu;
  std::condition_variable h5_cv, h5_idle;
  std::deque<std::function<void()>> h5_jobs;
  bool h5_stop = false, h5_busy = false;
  std::string h5_error;
  // device work lists
  ReqDev *d_reqs = nullptr;
  int2 *d_cmap = nullptr;
  size_t reqs_cap = 0, cmap_cap = 0;
  ReqDev *h_reqs = nullptr; // pinned
  int2 *h_cmap = nullptr;   // pinned

  ~nls_server() {
    stop = true;
    cv.notify_all();
    if (worker.joinable()) worker.join();
    {
      std::lock_guard<std::mutex> lk(h5_mu);
      h5_stop = true;
    }
    h5_cv.notify_all();
    if (h5_thread.joinable()) h5_thread.join();
    if (s) (void)hipStreamSynchronize(s);
    if (slab) (void)hipFree(slab);
    if (d_reqs) (void)hipFree(d_reqs);
    if (d_cmap) (void)hipFree(d_cmap);
    if (h_reqs) (void)hipHostFree(h_reqs);
    if (h_cmap) (void)hipHostFree(h_cmap);
    if (s) (void)hipStreamDestroy(s);
  }

  void h5_loop() {
    for (;;) {
      std::function<void()> job;
      {
        std::unique_lock<std::mutex> lk(h5_mu);
        h5_cv.wait(lk, [&] { return h5_stop || !h5_jobs.empty(); });
        if (h5_jobs.empty()) return;
        job = std::move(h5_jobs.front());
        h5_jobs.pop_front();
        h5_busy = true;
      }
      std::string err;
      try {
        job();
      } catch (const std::exception &e) {
        err = e.what();
      }
      {
        std::lock_guard<std::mutex> lk(h5_mu);
        h5_busy = false;
        if (!err.empty() && h5_error.empty()) h5_error = err;
      }
      h5_idle.notify_all();
    }
  }

  void h5_push(std::function<void()> job) {
    {
      std::lock_guard<std::mutex> lk(h5_mu);
      h5_jobs.push_back(std::move(job));
    }
    h5_cv.notify_one();
  }

  // Opens the request's file (header, snapshot 0 = initial state) on the writer thread.
  void h5_open(const std::shared_ptr<Request> &r) {
    auto hdr = std::make_shared<h5::Run>();
    char name[64];
    std::snprintf(name, sizeof name, "_%04lld.h5", (long long)r->id);
    hdr->path = h5_dir + "/run_" + run_id + name;
    hdr->dim = 2;
    hdr->nx = hdr->ny = hdr->ox = hdr->oy = r->q.n;
    hdr->Lx = hdr->Ly = p.L;
    hdr->T = r->q.dt * r->q.nt;
    hdr->nt = r->q.nt;
    hdr->num_snapshots = r->q.num_snapshots;
    hdr->problem_type = "phi4";
    hdr->run_id = run_id;
    hdr->run_index = (int)r->id;
    hdr->phenomenon = "stochastic";
    hdr->m_type = "custom";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", r->q.sigma);
    hdr->params.emplace_back("sigma", buf);
    std::snprintf(buf, sizeof buf, "%llu", (unsigned long long)r->q.seed);
    hdr->params.emplace_back("seed", buf);
    std::snprintf(buf, sizeof buf, "%.17g", p.gamma);
    hdr->params.emplace_back("gamma", buf);
    std::snprintf(buf, sizeof buf, "%.17g", r->q.dt);
    hdr->params.emplace_back("dt", buf);
    hdr->u0 = r->u0;
    hdr->v0 = r->v0;
    hdr->m = std::move(r->m64);
    hdr->c = std::move(r->c64);
    const bool snap0 = r->q.num_snapshots > 0;
    h5_push([r, hdr, snap0] {
      r->stream = std::make_shared<h5::Stream>(*hdr);
      if (snap0) r->stream->append(hdr->u0.data(), hdr->v0.data());
    });
  }

  static size_t req_bytes(int64_t N) {
    // FP64 state ring (3 x 8 B) + FP32 m and c (2 x 4 B) per cell, 256 B aligned each
    auto r = [](size_t b) { return (b + 255) / 256 * 256; };
    return 3 * r(sizeof(double) * N) + 2 * r(sizeof(float) * N);
  }

  void ensure_capacity(size_t nreq, size_t nblk) {
    if (nreq > reqs_cap) {
      if (d_reqs) NLS_HIP(hipFree(d_reqs));
      if (h_reqs) NLS_HIP(hipHostFree(h_reqs));
      reqs_cap = std::max<size_t>(nreq, 2 * reqs_cap);
      NLS_HIP(hipMalloc(&d_reqs, sizeof(ReqDev) * reqs_cap));
      NLS_HIP(hipHostMalloc(&h_reqs, sizeof(ReqDev) * reqs_cap, hipHostMallocDefault));
    }
    if (nblk > cmap_cap) {
      if (d_cmap) NLS_HIP(hipFree(d_cmap));
      if (h_cmap) NLS_HIP(hipHostFree(h_cmap));
      cmap_cap = std::max<size_t>(nblk, 2 * cmap_cap);
      NLS_HIP(hipMalloc(&d_cmap, sizeof(int2) * cmap_cap));
      NLS_HIP(hipHostMalloc(&h_cmap, sizeof(int2) * cmap_cap, hipHostMallocDefault));
    }
  }

  // Called with mu held.  Moves queued requests that fit into `active`.
  void admit_locked() {
    while (!queue.empty() && (int)active.size() < p.max_active) {
      auto r = queue.front();
      const size_t off = pool->alloc(r->bytes);
      if (off == SIZE_MAX) break; // FIFO: wait for memory to free up
      queue.pop_front();
      r->off = off;
      r->state = 1;
      r->t_admit = std::chrono::steady_clock::now();
      active.push_back(r);
      peak_bytes = std::max(peak_bytes, pool->in_use());
      peak_active = std::max(peak_active, (int)active.size());
    }
  }

  void upload_initial(Request &r) {
    const int64_t N = r.N;
    char *base = slab + r.off;
    auto rr = [](size_t b) { return (b + 255) / 256 * 256; };
    double *u0 = (double *)base, *u1 = (double *)(base + rr(8 * N)), *u2 = (double *)(base + 2 * rr(8 * N));
    float *mf = (float *)(base + 3 * rr(8 * N)), *cf = (float *)(base + 3 * rr(8 * N) + rr(4 * N));
    (void)u1;
    const double dt = r.q.dt;
    std::vector<double> up(N);
    for (int64_t i = 0; i < N; ++i) up[i] = r.u0[i] - dt * r.v0[i];
    NLS_HIP(hipMemcpyAsync(u0, r.u0.data(), 8 * N, hipMemcpyHostToDevice, s));
    NLS_HIP(hipMemcpyAsync(u2, up.data(), 8 * N, hipMemcpyHostToDevice, s));
    NLS_HIP(hipMemcpyAsync(mf, r.mf.data(), 4 * N, hipMemcpyHostToDevice, s));
    NLS_HIP(hipMemcpyAsync(cf, r.cf.data(), 4 * N, hipMemcpyHostToDevice, s));
    NLS_HIP(hipStreamSynchronize(s));
    // snapshot 0 = initial state
    if (r.q.num_snapshots > 0) std::copy(r.u0.begin(), r.u0.end(), r.snaps.begin());
    r.nsnap = r.q.num_snapshots > 0 ? 1 : 0;
    r.u0.clear(); r.u0.shrink_to_fit();
    r.v0.clear(); r.v0.shrink_to_fit();
  }

  ReqDev desc(const Request &r) const {
    ReqDev d{};
    const int64_t N = r.N;
    auto rr = [](size_t b) { return (b + 255) / 256 * 256; };
    char *base = slab + r.off;
    d.n = r.q.n;
    d.N = N;
    d.nchunks = (int)((N + SBS * SCPT - 1) / (SBS * SCPT));
    d.mode = p.mode;
    d.u0 = (double *)base;
    d.u1 = (double *)(base + rr(8 * N));
    d.u2 = (double *)(base + 2 * rr(8 * N));
    d.mf = (const float *)(base + 3 * rr(8 * N));
    d.cf = (const float *)(base + 3 * rr(8 * N) + rr(4 * N));
    d.base_step = r.steps_done;
    d.uni = r.uni ? 1 : 0;
    d.mval = r.mval;
    d.cval = r.cval;
    d.seed = r.q.seed;
    d.stream = 0;
    const double dt = r.q.dt, dx = 2.0 * p.L / (r.q.n - 1);
    d.tau = dt;
    d.s = 1.0 / (dx * dx);
    d.gp_inv = 1.0 / (1.0 + 0.5 * p.gamma * dt);
    d.gm = 1.0 - 0.5 * p.gamma * dt;
    d.t2 = dt * dt;
    d.namp = r.q.sigma * dt * std::sqrt(dt) / dx;          // 2D space-time white noise
    d.s_eff = 0.01 * std::sqrt(2.0 * dx) / dt;              // stochastic_phi4_dev.hpp:47-51
    return d;
  }

  // steps until the request's next snapshot or its end
  int64_t steps_to_event(const Request &r) const {
    const int64_t left = (int64_t)(r.q.nt - 1) - r.steps_done;
    int64_t e = left;
    if (r.snap_every > 0 && r.nsnap < r.q.num_snapshots) {
      const int64_t next = (r.steps_done / r.snap_every + 1) * r.snap_every;
      e = std::min(e, next - r.steps_done);
    }
    return e;
  }

  void copy_state(const Request &r, double *dst, int back = 0) {
    const ReqDev d = desc(r);
    double *ring[3] = {d.u0, d.u1, d.u2};
    NLS_HIP(hipMemcpyAsync(dst, ring[(r.steps_done + 3 - back) % 3], 8 * r.N,
                           hipMemcpyDeviceToHost, s));
  }


  void loop() {
    try {
      while (!stop) {
        std::vector<std::shared_ptr<Request>> act;
        std::vector<std::shared_ptr<Request>> fresh;
        {
          std::unique_lock<std::mutex> lk(mu);
          const size_t before = active.size();
          admit_locked();
          for (size_t k = before; k < active.size(); ++k) fresh.push_back(active[k]);
          if (active.empty()) {
            cv.wait_for(lk, std::chrono::milliseconds(2));
            continue;
          }
          act = active;
        }
        for (auto &r : fresh) {
          if (!h5_dir.empty()) h5_open(r); // before the upload releases u0, v0
          upload_initial(*r);
        }
        // group length: until the first snapshot/completion among active requests
        int64_t K = p.max_steps_per_group > 0 ? p.max_steps_per_group : 64;
        for (auto &r : act) K = std::min(K, steps_to_event(*r));
        const auto t0 = std::chrono::steady_clock::now();
        if (K > 0) {
          // XCD-aware block map (workgroup b runs on XCD b % 8): the chunks, request by
          // request, cut into 8 contiguous runs of equal length, one per XCD, so a
          // request's neighbouring chunks share an XCD's L2 and no XCD holds more
          // workgroups than another (a large request is split across XCDs)
          std::vector<int2> all;
          for (size_t k = 0; k < act.size(); ++k) {
            const int nch = (int)((act[k]->N + SBS * SCPT - 1) / (SBS * SCPT));
            for (int c = 0; c < nch; ++c) all.push_back(int2{(int)k, c});
          }
          const size_t per = (all.size() + NXCD - 1) / NXCD;
          const size_t nblk = per * NXCD;
          ensure_capacity(act.size(), nblk);
          for (size_t k = 0; k < nblk; ++k) h_cmap[k] = int2{-1, -1};
          for (size_t k = 0; k < all.size(); ++k) h_cmap[(k % per) * NXCD + k / per] = all[k];
          double cells = 0;
          for (size_t k = 0; k < act.size(); ++k) {
            h_reqs[k] = desc(*act[k]);
            cells += (double)act[k]->N;
          }
          NLS_HIP(hipMemcpyAsync(d_reqs, h_reqs, sizeof(ReqDev) * act.size(), hipMemcpyHostToDevice, s));
          NLS_HIP(hipMemcpyAsync(d_cmap, h_cmap, sizeof(int2) * nblk, hipMemcpyHostToDevice, s));
          for (int64_t k = 0; k < K; ++k)
            hipLaunchKernelGGL(k_sp4_group_step, dim3((unsigned)nblk), dim3(SBS), 0, s, d_reqs,
                               d_cmap, (int)k);
          check_launch("serve step");
          NLS_HIP(hipStreamSynchronize(s));
          cell_steps += cells * (double)K;
          for (auto &r : act) r->steps_done += K;
        }
        busy_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // events: snapshots and completions
        std::vector<std::shared_ptr<Request>> finished;
        std::vector<std::pair<std::shared_ptr<Request>, std::shared_ptr<std::vector<double>>>> prevs;
        for (auto &r : act) {
          if (r->snap_every > 0 && r->nsnap < r->q.num_snapshots && r->steps_done > 0 &&
              r->steps_done % r->snap_every == 0) {
            copy_state(*r, r->snaps.data() + (size_t)r->nsnap * r->N);
            if (!h5_dir.empty()) { // u_{k-1} for the velocity (u_k - u_{k-1}) / dt
              auto prev = std::make_shared<std::vector<double>>(r->N);
              copy_state(*r, prev->data(), 1);
              prevs.emplace_back(r, prev);
            }
            r->nsnap++;
          }
          if (r->steps_done >= (int64_t)r->q.nt - 1) {
            r->ufinal.resize(r->N);
            copy_state(*r, r->ufinal.data());
            finished.push_back(r);
          }
        }
        NLS_HIP(hipStreamSynchronize(s));
        for (auto &pr : prevs) { // append the snapshot just copied (writer thread)
          auto r = pr.first;
          auto u = std::make_shared<std::vector<double>>(
              r->snaps.begin() + (size_t)(r->nsnap - 1) * r->N,
              r->snaps.begin() + (size_t)r->nsnap * r->N);
          auto up = pr.second;
          const double dt = r->q.dt;
          h5_push([r, u, up, dt] {
            std::vector<double> v(u->size());
            for (size_t i = 0; i < v.size(); ++i) v[i] = ((*u)[i] - (*up)[i]) / dt;
            r->stream->append(u->data(), v.data());
          });
        }
        if (!finished.empty()) {
          std::lock_guard<std::mutex> lk(mu);
          const auto now = std::chrono::steady_clock::now();
          for (auto &r : finished) {
            bool ok = true;
            for (double x : r->ufinal) if (!std::isfinite(x)) { ok = false; break; }
            r->t_done = now;
            r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
            r->state = ok ? 2 : 3;
            if (!ok) { r->error = "non-finite state"; ++failed; }
            else { ++completed; latencies.push_back(r->latency_ms); }
            pool->release(r->off);
            active.erase(std::find(active.begin(), active.end(), r));
            if (!h5_dir.empty()) {
              const double el = r->latency_ms / 1e3;
              h5_push([r, el] {
                if (r->stream) r->stream->close(el);
              });
            }
          }
          done_cv.notify_all();
        }
      }
    } catch (const std::exception &e) {
      std::lock_guard<std::mutex> lk(mu);
      fatal = e.what();
      for (auto &r : active) { r->state = 3; r->error = fatal; }
      for (auto &r : queue) { r->state = 3; r->error = fatal; }
      done_cv.notify_all();
    }
  }
};

extern "C" {

int nls_server_create(const nls_server_params *p, nls_server **out) {
  return guard([&] {
    NLS_REQUIRE(p && out, "nls_server_create: null argument");
    NLS_REQUIRE(p->hbm_budget_bytes >= (1u << 20), "nls_server_create: budget below 1 MiB");
    NLS_REQUIRE(p->max_active >= 1, "nls_server_create: max_active must be >= 1");
    NLS_REQUIRE(p->mode == NLS_SP4_REFERENCE || p->mode == NLS_SP4_DAMPED, "unknown sp4 mode");
    NLS_REQUIRE(p->L > 0, "L must be positive");
    auto sv = std::make_unique<nls_server>();
    sv->p = *p;
    if (p->h5_dir && *p->h5_dir) {
      std::string ver;
      NLS_REQUIRE(h5::available(&ver), "h5_dir set but libhdf5 could not be loaded");
      sv->h5_dir = p->h5_dir;
      sv->run_id = p->run_id && *p->run_id ? p->run_id : "serve";
      sv->h5_thread = std::thread([raw = sv.get()] { raw->h5_loop(); });
    }
    sv->p.h5_dir = nullptr; // the strings live in sv->h5_dir / sv->run_id
    sv->p.run_id = nullptr;
    NLS_HIP(hipStreamCreate(&sv->s));
    NLS_HIP(hipMalloc((void **)&sv->slab, p->hbm_budget_bytes));
    sv->pool = std::make_unique<SlabPool>(p->hbm_budget_bytes);
    sv->worker = std::thread([raw = sv.get()] { raw->loop(); });
    *out = sv.release();
  });
}

int nls_server_destroy(nls_server *s) { return guard([&] { delete s; }); }

int nls_server_submit(nls_server *sv, const nls_request *q, const double *m, const double *c,
                      const double *u0, const double *v0, int64_t *id) {
  return guard([&] {
    NLS_REQUIRE(sv && q && id, "nls_server_submit: null argument");
    NLS_REQUIRE(q->n >= 3 && q->nt >= 2 && q->dt > 0, "nls_server_submit: bad request");
    NLS_REQUIRE(q->n <= kMaxSide, "nls_server_submit: grid side above " + std::to_string(kMaxSide));
    NLS_REQUIRE(q->num_snapshots >= 0 && q->num_snapshots <= q->nt, "bad num_snapshots");
    auto r = std::make_shared<Request>();
    r->q = *q;
    r->N = (int64_t)q->n * q->n;
    r->bytes = nls_server::req_bytes(r->N);
    NLS_REQUIRE(r->bytes <= sv->p.hbm_budget_bytes,
                "request needs more device memory than the whole serving budget");
    r->mf.resize(r->N);
    r->cf.resize(r->N);
    r->uni = true;
    for (int64_t i = 0; i < r->N; ++i) {
      r->mf[i] = m ? (float)m[i] : 1.0f;
      r->cf[i] = c ? (float)c[i] : 1.0f;
      r->uni = r->uni && r->mf[i] == r->mf[0] && r->cf[i] == r->cf[0];
    }
    if (r->N > 0) {
      r->mval = r->mf[0];
      r->cval = r->cf[0];
    }
    if (!sv->h5_dir.empty()) {
      r->m64.assign(r->N, 1.0);
      r->c64.assign(r->N, 1.0);
      if (m) std::copy(m, m + r->N, r->m64.begin());
      if (c) std::copy(c, c + r->N, r->c64.begin());
    }
    r->u0.assign(u0 ? u0 : nullptr, u0 ? u0 + r->N : nullptr);
    if (!u0) r->u0.assign(r->N, 0.0);
    r->v0.assign(v0 ? v0 : nullptr, v0 ? v0 + r->N : nullptr);
    if (!v0) r->v0.assign(r->N, 0.0);
    r->snap_every = q->num_snapshots > 0 ? q->nt / q->num_snapshots : 0;
    r->snaps.assign((size_t)r->N * std::max(q->num_snapshots, 0), 0.0);
    r->t_submit = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      NLS_REQUIRE(sv->fatal.empty(), "server stopped after a device error: " + sv->fatal);
      r->id = sv->next_id++;
      sv->queue.push_back(r);
      sv->all[r->id] = r;
    }
    sv->cv.notify_all();
    *id = r->id;
  });
}

int nls_server_status(nls_server *sv, int64_t id, int *state) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    NLS_REQUIRE(it != sv->all.end(), "unknown request id");
    *state = it->second->state;
  });
}

int nls_server_wait(nls_server *sv, int64_t id, double timeout_s) {
  return guard([&] {
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    NLS_REQUIRE(it != sv->all.end(), "unknown request id");
    auto r = it->second;
    const bool ok = sv->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                         [&] { return r->state >= 2; });
    NLS_REQUIRE(ok, "timeout waiting for request");
    NLS_REQUIRE(r->state == 2, "request failed: " + r->error);
  });
}

int nls_server_drain(nls_server *sv, double timeout_s) {
  return guard([&] {
    std::unique_lock<std::mutex> lk(sv->mu);
    const bool ok = sv->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
      return (sv->queue.empty() && sv->active.empty()) || !sv->fatal.empty();
    });
    NLS_REQUIRE(ok, "timeout draining the server");
    NLS_REQUIRE(sv->fatal.empty(), "server stopped after a device error: " + sv->fatal);
    lk.unlock();
    if (!sv->h5_dir.empty()) { // files complete on disk
      std::unique_lock<std::mutex> hl(sv->h5_mu);
      const bool done = sv->h5_idle.wait_for(hl, std::chrono::duration<double>(timeout_s), [&] {
        return sv->h5_jobs.empty() && !sv->h5_busy;
      });
      NLS_REQUIRE(done, "timeout waiting for the HDF5 writer");
      NLS_REQUIRE(sv->h5_error.empty(), "HDF5 writer: " + sv->h5_error);
    }
  });
}

int nls_server_fetch(nls_server *sv, int64_t id, double *snapshots, int *count, double *u_final,
                     double *latency_ms) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    NLS_REQUIRE(it != sv->all.end(), "unknown request id");
    auto r = it->second;
    NLS_REQUIRE(r->state == 2, r->state == 3 ? "request failed: " + r->error : "request not finished");
    if (count) *count = r->nsnap;
    if (snapshots) std::copy(r->snaps.begin(), r->snaps.begin() + (size_t)r->nsnap * r->N, snapshots);
    if (u_final) std::copy(r->ufinal.begin(), r->ufinal.end(), u_final);
    if (latency_ms) *latency_ms = r->latency_ms;
  });
}

int nls_server_release(nls_server *sv, int64_t id) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    NLS_REQUIRE(it != sv->all.end(), "unknown request id");
    NLS_REQUIRE(it->second->state >= 2, "request still in flight");
    sv->all.erase(it);
  });
}

int nls_server_get_stats(nls_server *sv, nls_server_stats *st) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(sv->mu);
    nls_server_stats o{};
    o.completed = sv->completed;
    o.failed = sv->failed;
    o.active = (int64_t)sv->active.size();
    o.queued = (int64_t)sv->queue.size();
    std::vector<double> l = sv->latencies;
    std::sort(l.begin(), l.end());
    auto q = [&](double f) {
      if (l.empty()) return 0.0;
      const double pos = f * (l.size() - 1);
      const size_t lo = (size_t)pos;
      const size_t hi = std::min(lo + 1, l.size() - 1);
      return l[lo] + (pos - lo) * (l[hi] - l[lo]);
    };
    o.p50_ms = q(0.50);
    o.p99_ms = q(0.99);
    o.mean_ms = l.empty() ? 0.0 : std::accumulate(l.begin(), l.end(), 0.0) / l.size();
    o.cell_steps = sv->cell_steps;
    o.busy_s = sv->busy_s;
    o.peak_bytes = sv->peak_bytes;
    o.peak_active = sv->peak_active;
    *st = o;
  });
}

// Admission allocator on the host, no device: ops[2k] = 0 (alloc ops[2k+1] bytes)
// or 1 (free the block returned by op ops[2k+1]); results[k] = offset / -1 for an
// alloc, bytes in use after a free.
int nls_pool_simulate(size_t capacity, int nops, const int64_t *ops, int64_t *results) {
  return guard([&] {
    SlabPool pool(capacity);
    for (int k = 0; k < nops; ++k) {
      if (ops[2 * k] == 0) {
        const size_t off = pool.alloc((size_t)ops[2 * k + 1]);
        results[k] = off == SIZE_MAX ? -1 : (int64_t)off;
      } else {
        const int64_t ref = ops[2 * k + 1];
        NLS_REQUIRE(ref >= 0 && ref < k && results[ref] >= 0, "free of an invalid block");
        pool.release((size_t)results[ref]);
        results[k] = (int64_t)pool.in_use();
      }
    }
  });
}

} // extern "C"
