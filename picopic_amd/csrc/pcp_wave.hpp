// Wave-level aggregation of scatter-adds (64-lane CDNA wavefronts).
//
// Particles are stored in cell order, so the lanes of a wavefront mostly
// deposit into the same few grid nodes.  wave_add() groups the lanes of the
// (converged) wavefront by target index, sums each group across the wave and
// issues one device-scope atomic per distinct node, instead of one per lane.
// It must be called by all lanes of the wavefront (converged control flow);
// `active` masks the lanes that carry a contribution.
#pragma once

#include <hip/hip_runtime.h>

namespace pcp {

// DPP lane moves of a 64-bit value (every lane of the wavefront active)
template <int CTRL>
__device__ __forceinline__ double dpp64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double readlane64(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes in a fixed order, returned to every lane: pairs and
// quads by quad permutes, the 16-lane rows by row rotations (4 and 8), then
// the four row sums.  All lanes must be active.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp64<0xB1>(x);    // quad_perm [1, 0, 3, 2]
  x += dpp64<0x4E>(x);    // quad_perm [2, 3, 0, 1]
  x += dpp64<0x124>(x);   // row_ror 4
  x += dpp64<0x128>(x);   // row_ror 8
  return (readlane64(x, 0) + readlane64(x, 16)) + (readlane64(x, 32) + readlane64(x, 48));
}

// Scatter-add with wave aggregation: lanes with the same target index are
// summed (in FP64, also for FP32 targets, so the aggregated FP32 current is
// the correctly rounded sum of up to 64 contributions up to a 2^-53
// relative term) and one lane per distinct index issues the atomic.
template <class T>
__device__ __forceinline__ void wave_add(T *base, int idx, T v, bool active) {
  unsigned long long pending = __ballot(active);
  const int lane = __lane_id();
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lk = __builtin_amdgcn_readlane(idx, leader);
    const bool m = active && idx == lk;
    const double s = wave_sum(m ? (double)v : 0.0);
    if (lane == leader) unsafeAtomicAdd(base + lk, (T)s);
    pending &= ~__ballot(m);
    active = active && !m;
  }
}

// one atomic per distinct key in the wavefront; each active lane receives a
// distinct slot counted from the value the counter held
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *ctr, uint32_t key, bool active) {
  uint32_t slot = 0;
  unsigned long long pending = __ballot(active);
  const int lane = __lane_id();
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t lk = __shfl(key, leader, 64);
    const bool m = active && key == lk;
    const unsigned long long mask = __ballot(m);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(ctr + lk, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (m) slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    pending &= ~mask;
    active = active && !m;
  }
  return slot;
}

}  // namespace pcp
