// Binary Coulomb collisions on the device (defined in pcp_collide.hip).
#pragma once

#include "pcp_internal.hpp"
#include "pcp_mlp.hpp"

namespace pcp {

// largest number of macro-particles of one species in one cell that a
// collision pass stages in LDS; larger cells take the crowded path (a
// global stage of 2^22 particles per pass; more raises ERR_COLL_OVERFLOW)
constexpr int COLL_MAXC = 256;

struct CollBuffers;

// the pass number a collision pass draws with: a value, or (p set) a device
// counter that the step advances after its collision phase, so that a
// captured step graph draws new numbers at every replay
struct PassNo {
  const uint64_t *p;
  uint64_t v;
  __host__ __device__ __forceinline__ uint64_t get() const { return p ? *p : v; }
};

// the collision surrogate: network on the cells it covers, P12 on the rest
constexpr int COLL_SURROGATE = 3;
// a cell goes to the network only with at least this many macro-particles
// of each species, normal moments, and every normalised feature within
// SURR_MAX_Z standard deviations of the training set
constexpr double SURR_MIN_PARTICLES = 8;
constexpr float SURR_MAX_Z = 4.0f;
// smallest squared rescaling factor applied to a thermal velocity spread
constexpr double SURR_MIN_FACTOR = 0.25;


// One collision pass (scheme P12 / TA77S / SK98) between electron species se
// and ion species si, enqueued on the context stream.  `vslot` selects the
// FP32 velocity slot (ignored for FP64 contexts).  Returns -1 on allocation
// or launch failure.
int launch_collide(pcp_ctx *c, int scheme, int se, int si, uint64_t seed, PassNo pass, int vslot);
// allocates the pass's buffers for species se, si (at their capacities)
int coll_prepare(pcp_ctx *c, int se, int si);
void coll_free(pcp_ctx *c);
// one P12 pass whose collided cells append training rows (raw features,
// relative temperature changes) to the network's training set
int surrogate_collect(pcp_ctx *c, int se, int si, uint64_t seed, uint64_t step);
// cells of the last surrogate pass that went to the P12 fallback
long surrogate_fallback_cells(pcp_ctx *c);

// Fused surrogate collision phase of FP32 steps: buffers (moments zeroed),
// the moment array the push fills, the per-cell rescalings the move applies,
// and the per-cell work between them (features, network, rescalings, P12
// on the cells the network does not cover).  Return -1 on allocation or
// launch failure, -2 without a trained network.
int fsurr_prepare(pcp_ctx *c, int se, int si);
double *fsurr_mom(pcp_ctx *c);
const float *fsurr_par(pcp_ctx *c);
int fsurr_cells(pcp_ctx *c, int se, int si, uint64_t seed, PassNo pass);

}  // namespace pcp
