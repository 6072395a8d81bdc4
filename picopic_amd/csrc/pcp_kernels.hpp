// Launch wrappers of the FP64 kernels (defined in pcp_kernels.hip).
#pragma once

#include "pcp_internal.hpp"

namespace pcp {

struct SeqTable {
  uint64_t *p[MAX_SPECIES];
};

template <class T>
void launch_yee_e_t(pcp_ctx *c, FieldsT<T> f, int clear_j);
template <class T>
void launch_yee_h_t(pcp_ctx *c, FieldsT<T> f);
template <class T>
void launch_stitch_t(pcp_ctx *c, FieldsT<T> f, int which);
void launch_migrant_keys_t(pcp_ctx *c, SeqTable t);

void launch_yee_e(pcp_ctx *c, int clear_j);
void launch_yee_h(pcp_ctx *c);
void launch_stitch(pcp_ctx *c, int which);
void launch_fill(pcp_ctx *c, double *p, long n, double v);
void launch_step_begin(pcp_ctx *c);
// one-thread kernels on a device scalar (set / add one)
void launch_set_u64(pcp_ctx *c, uint64_t *p, uint64_t v);
void launch_inc_u64(pcp_ctx *c, uint64_t *p);
void launch_stamp(pcp_ctx *c);
void launch_push(pcp_ctx *c, int sp);
void launch_move(pcp_ctx *c, int sp);
void launch_migrate_mark(pcp_ctx *c, int sp);   // removal + migrant list (phase API)
void launch_step_fused(pcp_ctx *c, int sp, bool push = true);
void launch_mark(pcp_ctx *c, int sp, int only_if_exact);
void launch_deposit_exact(pcp_ctx *c, int sp, int only_if_exact);
// fix-up pass after a step whose strict motions reset J_r nodes (no-op
// otherwise, decided on the device)
void launch_fix(pcp_ctx *c, int sp);
void launch_migrant_keys(pcp_ctx *c);
void launch_gather_point(pcp_ctx *c, int d, double r, double z, int which, double *dev_out);

}  // namespace pcp
