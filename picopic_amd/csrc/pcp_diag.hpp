// FP64 diagnostic reductions (defined in pcp_diag.hip).
#pragma once

#include "pcp_internal.hpp"

namespace pcp {

// discrete continuity residual of the last step into dev_out[3] (device):
// max |dQ + dt div J| over the nodes, max |dQ|, largest |q w|; synchronises
int launch_continuity_residual(pcp_ctx *c, double *dev_out);

}  // namespace pcp
