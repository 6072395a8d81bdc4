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

__device__ __forceinline__ void wave_add(double *base, int idx, double v, bool active) {
  unsigned long long pending = __ballot(active);
  const int lane = __lane_id();
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lk = __shfl(idx, leader, 64);
    const bool m = active && idx == lk;
    const double s = wave_sum(m ? v : 0.0);
    if (lane == leader) unsafeAtomicAdd(base + lk, s);
    pending &= ~__ballot(m);
    active = active && !m;
  }
}

}  // namespace pcp
