// Counter-based random numbers: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC11), keyed by the seed and
// counted by (ctr, draw), so every draw is independent of the launch shape.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcp {

__device__ __forceinline__ void philox4x32(uint64_t ctr, uint32_t draw, uint64_t seed, uint32_t out[4]) {
  uint32_t c[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), draw, 0x5eedu};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c[0];
    uint64_t p1 = (uint64_t)M1 * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k[0];
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k[1];
    c[0] = n0; c[1] = (uint32_t)p1; c[2] = n2; c[3] = (uint32_t)p0;
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
  }
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

// two doubles in [0, 1) (53 random bits each) from one Philox block
__device__ __forceinline__ void uniform2(uint64_t ctr, uint32_t draw, uint64_t seed, double &u0, double &u1) {
  uint32_t o[4];
  philox4x32(ctr, draw, seed, o);
  const double s = 1.0 / 9007199254740992.0;   // 2^-53
  u0 = (double)(((uint64_t)(o[0] >> 5) << 26) | (o[1] >> 6)) * s;
  u1 = (double)(((uint64_t)(o[2] >> 5) << 26) | (o[3] >> 6)) * s;
}

// Box-Muller on (0,1] x [0,1)
__device__ __forceinline__ void normal2(uint64_t ctr, uint32_t draw, uint64_t seed, double &n0, double &n1) {
  double u0, u1;
  uniform2(ctr, draw, seed, u0, u1);
  const double two_pi = 2.0 * 3.1415926535897932;
  double rad = sqrt(-2.0 * log(1.0 - u0));
  n0 = rad * cos(two_pi * u1);
  n1 = rad * sin(two_pi * u1);
}

}  // namespace pcp
