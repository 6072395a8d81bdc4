// Binary Coulomb collisions on the device (defined in pcp_collide.hip).
#pragma once

#include "pcp_internal.hpp"

namespace pcp {

// largest number of macro-particles of one species in one cell that a
// collision pass accepts (the cell is staged in LDS); more raises
// ERR_COLL_OVERFLOW
constexpr int COLL_MAXC = 256;

struct CollBuffers;

// One collision pass (scheme P12 / TA77S / SK98) between electron species se
// and ion species si, enqueued on the context stream.  `vslot` selects the
// FP32 velocity slot (ignored for FP64 contexts).  Returns -1 on allocation
// or launch failure.
int launch_collide(pcp_ctx *c, int scheme, int se, int si, uint64_t seed, uint64_t step, int vslot);
void coll_free(pcp_ctx *c);

}  // namespace pcp
