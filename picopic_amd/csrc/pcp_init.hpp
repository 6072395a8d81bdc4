// Device-side synthetic input generation (defined in pcp_init.hip).
#pragma once

#include <algorithm>

#include "pcp_internal.hpp"

namespace pcp {

void launch_init_plasma(pcp_ctx *c, int sp, long base, long n_new, const long *dev_cell_off, long n_cells,
                        double weight, double vth_r, double vth_phi, double vth_z, uint64_t seed,
                        const long *dev_dom_off);
void launch_init_bunch(pcp_ctx *c, int sp, long base, long n_new, double sigma_r, double sigma_z, double z0,
                       double vz, double weight, uint64_t seed);
void launch_count_live(pcp_ctx *c, int sp, unsigned long long *dev_out);
// device-to-device copy of `bytes` (read + write counted), iters launches
int bench_copy(pcp_ctx *c, long bytes, int iters, double *gbps);

}  // namespace pcp
