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

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// FP32 wave sum carrying the rounding error of every pairwise addition
// (Knuth's TwoSum), so the result is the correctly rounded sum of the 64
// lanes up to a second-order term
__device__ __forceinline__ float wave_sum(float x) {
  float c = 0.f;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    float y = __shfl_xor(x, off, 64);
    float yc = __shfl_xor(c, off, 64);
    float s = x + y;
    float bp = s - x;
    float err = (x - (s - bp)) + (y - bp);
    c = c + yc + err;
    x = s;
  }
  return x + c;
}

template <class T>
__device__ __forceinline__ void wave_add(T *base, int idx, T v, bool active) {
  unsigned long long pending = __ballot(active);
  const int lane = __lane_id();
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lk = __shfl(idx, leader, 64);
    const bool m = active && idx == lk;
    const T s = wave_sum(m ? v : (T)0);
    if (lane == leader) unsafeAtomicAdd(base + lk, s);
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
