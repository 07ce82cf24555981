// Shared host/device helpers for libnls: error handling at the C-ABI boundary,
// deterministic wavefront/workgroup reductions and small launch utilities.
// All device code targets gfx950 (CDNA4, 64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace nls {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_error(const std::string &msg);
void clear_error();

#define NLS_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::nls::Error(std::string(#call) + " failed: " + hipGetErrorString(e_) +      \
                         " (" __FILE__ ":" + std::to_string(__LINE__) + ")");            \
  } while (0)

#define NLS_REQUIRE(cond, msg)                                                           \
  do {                                                                                   \
    if (!(cond)) throw ::nls::Error(std::string(msg));                                   \
  } while (0)

// Runs f and converts any exception into a non-zero status plus a thread-local
// message (nls_last_error); nothing throws across the C ABI.
template <class F> int guard(F &&f) {
  try {
    clear_error();
    f();
    return 0;
  } catch (const std::exception &e) {
    set_error(e.what());
    return -1;
  } catch (...) {
    set_error("unknown exception");
    return -1;
  }
}

inline void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw Error(std::string("kernel launch ") + what + ": " + hipGetErrorString(e));
}

constexpr int kWave = 64;

// Butterfly sum across a 64-lane wavefront.  The combination order is fixed, so
// the result is bitwise reproducible for identical inputs.
template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// FP64 sum across a 64-lane wavefront from data-parallel-primitive (DPP) lane moves inside
// each 16-lane row (lane pairs i^1, i^2, then the mirrored lanes of each 8 and of each 16)
// and the four row sums read from lanes 0, 16, 32, 48 and added in that order.  Every lane
// returns the same total, bitwise reproducible for identical inputs; no LDS traffic (the
// butterfly of wave_sum moves each level through LDS with ds_bpermute).
template <int CTRL> __device__ inline double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double wave_sum_dpp(double v) {
  v += dpp_move<0xB1>(v);  // quad_perm [1,0,3,2]: lane i^1
  v += dpp_move<0x4E>(v);  // quad_perm [2,3,0,1]: lane i^2
  v += dpp_move<0x141>(v); // row_half_mirror: lane 7-i of each 8
  v += dpp_move<0x140>(v); // row_mirror: lane 15-i of each 16
  auto lane_d = [&](int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
  };
  return ((lane_d(0) + lane_d(16)) + lane_d(32)) + lane_d(48);
}

// Sum over a workgroup of BS threads (BS a multiple of 64).  Fixed order: each
// wave reduces by butterfly, wave results are added in wave-index order by
// wave 0.  Every thread returns the total.  sh must hold BS/64 entries.
template <int BS, class T> __device__ inline T block_sum(T v, T *sh) {
  constexpr int NW = BS / kWave;
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T tot = sh[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) tot += sh[i];
  return tot;
}

inline unsigned blocks_for(int64_t n, int per_block) {
  return (unsigned)((n + per_block - 1) / per_block);
}

} // namespace nls
