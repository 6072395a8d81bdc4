// Per-species counting sort by cell (defined in pcp_sort.hip).
#pragma once

#include <algorithm>

#include "pcp_internal.hpp"

namespace pcp {

// size of the scan's temporary storage for ncells + 1 counters
size_t sort_temp_bytes(long ncells);
// Sorts species sp by (domain, cell) and drops removed particles (their
// slots at the end are marked removed).  Returns 1 when sorted, 0 if the
// species is empty, -1 on failure.  Positions before the last move are not
// kept across a sort.
int launch_sort(pcp_ctx *c, int sp);

}  // namespace pcp
