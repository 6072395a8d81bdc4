// Device-side synthetic input generation (defined in pcp_init.hip).
#pragma once

#include <algorithm>

#include "pcp_internal.hpp"
#include "pcp_refinit.hpp"

namespace pcp {

void launch_init_plasma(pcp_ctx *c, int sp, long base, long n_new, const long *dev_cell_off, long n_cells,
                        double weight, double vth_r, double vth_phi, double vth_z, uint64_t seed,
                        const long *dev_dom_off);
void launch_init_bunch(pcp_ctx *c, int sp, long base, long n_new, double sigma_r, double sigma_z, double z0,
                       double vz, double weight, uint64_t seed);
// the reference's flat placement + weights + thermal velocities of one
// plasma species, M particles per domain (dev_doms: one RefInitDom per
// domain); particles placed beyond their domain are appended as migrants
// (the caller runs the migration key pass afterwards)
void launch_init_reference(pcp_ctx *c, int sp, long base, uint32_t M, const RefInitDom *dev_doms,
                           const RadicalPowers &pw, double temperature, uint64_t seed);
// BeamP::inject of n particles of species sp in domain d (slots base ..);
// with `migrants`, the particles placed past the domain's cells are
// appended to a fresh migrant list (the caller runs the key pass)
void launch_inject(pcp_ctx *c, int sp, long base, long n, int d, double domain_radius, double dl, double velocity,
                   double v_macro_avg, double n_per_macro_avg, uint64_t seed, uint64_t ctr0, bool migrants);
void launch_count_live(pcp_ctx *c, int sp, unsigned long long *dev_out);
// device-to-device copy of `bytes` (read + write counted), iters launches
int bench_copy(pcp_ctx *c, long bytes, int iters, double *gbps);

}  // namespace pcp
