// FP32 split-precision mode (defined in pcp32.hip).
//
// Particles: a packed global cell index (u32, r_cell * Nz + z_cell) plus FP32
// in-cell offsets (0 <= f < 1, in units of the cell size) -- 12 B of position
// with the absolute resolution of a 2^-24 fraction of a cell everywhere in
// the grid -- and FP32 velocities and weight: 28 B read, 24 B written per
// particle-step.  Fields and J are FP32.  The current is deposited from the
// unrounded FP64 end point of the move (old split position + FP32
// displacement), so that a displacement below the offset's resolution still
// deposits, as in FP64.  FP64 is used for the near-axis
// radial weights, for wall reflections, for the Villasenor-Buneman terms of
// the axis cell and of steps longer than 1/32 of the radius (elsewhere they
// are evaluated in FP32 in a cell-local form whose size-i parts cancel
// analytically) and for the energy reductions.
#pragma once

#include "pcp_internal.hpp"

namespace pcp {

constexpr uint32_t DEAD_CELL = 0xFFFFFFFFu;

// fused surrogate collision phase of an FP32 step: per global cell and
// collision slot (0 electrons, 1 ions), 8 moments of the pushed velocities
// (count, sum w, sum w v_r, sum w v_phi, sum w v_z, sum w v_z^2,
// sum w (v_r^2 + v_phi^2), -) and 8 rescale parameters (V_r, V_phi, V_z,
// a_perp, a_par, -, -, -) of v -> V + a (v - V)
constexpr int SURR_PAR_STRIDE = 16;

struct Species32 {
  uint32_t *cell[2];
  float *fr[2], *fz[2];
  // velocities are double-buffered like the positions: a step reads slot
  // cur and writes slot cur ^ 1, so the pre-push state stays available to
  // the ordered re-deposition of the same step
  float *vr[2], *vphi[2], *vz[2];
  float *w;
  uint64_t *seq, *id;
  long n, cap;
  double q, m;
  int beam;
};

struct SortBuffers32 {
  float *w;
  uint64_t *seq, *id;
};

// position (r, z) in split form, written to both position slots of q
__device__ __forceinline__ void store_split(Species32 &s, long q, const Geom &g, double r, double z) {
  double ur = r / g.dr, uz = z / g.dz;
  double fcr = floor(ur), fcz = floor(uz);
  float fr = (float)(ur - fcr), fz = (float)(uz - fcz);
  int cr = (int)fcr, cz = (int)fcz;
  if (fr >= 1.f) { fr = 0.f; cr += 1; }
  if (fz >= 1.f) { fz = 0.f; cz += 1; }
  uint32_t cell = (uint32_t)cr * (uint32_t)g.Nz + (uint32_t)cz;
  s.cell[0][q] = cell; s.cell[1][q] = cell;
  s.fr[0][q] = fr; s.fr[1][q] = fr;
  s.fz[0][q] = fz; s.fz[1][q] = fz;
}

// state of an FP32 context (owned by pcp_ctx)
struct State32 {
  FieldsT<float> f;
  Species32 sp[MAX_SPECIES];
  SortBuffers32 sb[MAX_SPECIES];
  // compensated deposit: the rounding errors of J's atomics, same layout as
  // J, folded into J after the step's deposit (null: not compensated)
  float *Jc = nullptr;
  // the fix-up's list of flagged particles (batch species << 40 | index)
  uint64_t *fix_list = nullptr;
  unsigned long long *fix_count = nullptr;
};

int alloc32(pcp_ctx *c);
void free32(pcp_ctx *c);
// allocates (zeroed) or frees the compensation array; -1 on failure
int set_compensated32(pcp_ctx *c, bool on);
// folds the compensation array into J (no-op when not compensated)
void launch32_fold_j(pcp_ctx *c);
void launch32_yee_e(pcp_ctx *c, int clear_j);
void launch32_yee_h(pcp_ctx *c);
void launch32_stitch(pcp_ctx *c, int which);
// push = false: the step's velocities were pushed (launch32_push) and
// collided beforehand; the launcher exchanges the velocity slots after the
// move so that slot cur ^ 1 keeps the pre-move velocities for the exact path
void launch32_push(pcp_ctx *c, int sp);
// push of a step with the fused surrogate collision phase; slot >= 0 adds
// the pushed velocities' per-cell moments of the species to mom (layout of
// SURR_PAR_STRIDE doubles per global cell, 8 per collision slot)
void launch32_push_mom(pcp_ctx *c, int sp, double *mom, int slot);
// the fused step of all species in one persistent work-queue launch; with
// cpar (and !push) the electron (se) and ion (si) velocities take their
// cell's surrogate rescaling before the move
void launch32_step_batch(pcp_ctx *c, bool push = true, const float *cpar = nullptr, int se = -1, int si = -1);
// fix-up pass after a step whose strict motions reset J_r nodes (no-op
// otherwise, decided on the device); push: the step pushed (no collision
// phase)
void launch32_fix(pcp_ctx *c, bool push, const float *cpar = nullptr, int se = -1, int si = -1);
void launch32_migrant_keys(pcp_ctx *c);
int launch32_sort(pcp_ctx *c, int sp);
void launch32_count_live(pcp_ctx *c, int sp, unsigned long long *dev_out);
void launch32_init_plasma(pcp_ctx *c, int sp, long base, long n_new, const long *dev_cell_off, long n_cells,
                          double weight, double vth, uint64_t seed, const long *dev_dom_off);
void launch32_init_bunch(pcp_ctx *c, int sp, long base, long n_new, double sigma_r, double sigma_z, double z0,
                         double vz, double weight, uint64_t seed);
void launch32_gather_point(pcp_ctx *c, int d, double r, double z, int which, double *dev_out);
// FP64 reductions: field energy sum over interior nodes of
// eps0 E^2 / 2 + mu0 H^2 / 2 weighted by the cell volume; kinetic energy
// sum of w m c^2 (gamma - 1) of a species
void launch_field_energy(pcp_ctx *c, double *dev_out, bool to_log = false);
void launch_kinetic_energy(pcp_ctx *c, int sp, double *dev_out);
// FP64 contexts with the energy log on: the kinetic energy of species sp
// into the current step's row (FP32 contexts sum it inside the step kernel)
void launch_kinetic_energy_log(pcp_ctx *c, int sp);
// out[7] += sum w, sum w v_r, sum w v_phi, sum w v_z, sum w v_z^2,
// sum w (v_r^2 + v_phi^2), number of live particles
void launch_species_moments(pcp_ctx *c, int sp, double *dev_out);

}  // namespace pcp
