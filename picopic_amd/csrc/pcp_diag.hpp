// FP64 diagnostic reductions (defined in pcp_diag.hip).
#pragma once

#include "pcp_internal.hpp"

namespace pcp {

// discrete continuity residual of the last step into dev_out[3] (device):
// max |dQ + dt div J| over the nodes, max |dQ|, largest |q w|; synchronises
int launch_continuity_residual(pcp_ctx *c, double *dev_out);

// the same with the charge array in a fixed device array of
// CONTINUITY_SCRATCH_DOUBLES doubles (80 MiB per GPU): no allocation.
// Returns -2 when the grid's (Nr + 4) (Nz + 4) nodes exceed it, -3 before
// continuity_scratch_init ran for the context's device.
constexpr long CONTINUITY_SCRATCH_DOUBLES = 10L << 20;
int launch_continuity_residual_noalloc(pcp_ctx *c, double *dev_out);
// resolves the fixed array's address on `device` (current); once per device
int continuity_scratch_init(int device);

// probe window: component comp of a [D][ncomp][P] grid (float when f32)
// over global cells [r0, r0 + nr_w) x [z0, z0 + nz_w), row-major doubles
void launch_probe(pcp_ctx *c, const void *grid, bool f32, int ncomp, int comp, long r0, long z0, long nr_w, long nz_w,
                  double *dev_out);

// species map `kind` (0 density, 1 temperature, 2 energy, 3 momentum) of
// species sp into OUT (ncomp x D P; interior nodes): N is scratch (D P),
// PW scratch (3 D P), PABS the species' momentum_abs_w map, refreshed only by
// the energy map (the temperature map reads what it holds)
int launch_species_map(pcp_ctx *c, int sp, int kind, double *N, double *PW, double *PABS, double *OUT);

}  // namespace pcp
