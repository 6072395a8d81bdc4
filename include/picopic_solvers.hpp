/*
 * C++ drop-in layer over the C ABI of picopic.h: device-backed objects with
 * the member functions of the reference's solver interfaces, one object per
 * domain as Domain::Domain builds them (src/domain.cpp:26-84), all domains
 * of one shared-memory block (SMB) sharing one device context.
 *
 *   DevicePusher        Pusher::operator()             include/pusher.hpp:33-59
 *   DeviceCurrent       Current::current_distribution  include/current.hpp:35-54
 *   DeviceMaxwellSolver MaxwellSolver::set_pml, calc_field_h, calc_field_e,
 *                       get_field_h, get_field_e       include/maxwellSolver.hpp:35-317
 *   DeviceCollisions    Collisions::operator()         include/collisions.hpp:32-78
 *   DeviceBlock         the SMB phases that are not solver virtuals:
 *                       overlays (src/SMB.cpp:566-596, 784-823, 1010-1044),
 *                       the per-particle kinematics of advance_particles
 *                       (src/SMB.cpp:1290-1303) and particles_runaway_collector
 *                       (src/SMB.cpp:194-564)
 *
 * The reference calls each solver once per domain from an OpenMP loop over
 * the domain grid and then reaches an SMB phase boundary.  Here every such
 * per-domain call "arrives" at the block's phase barrier; the call that
 * completes the set (one arrival from every domain) launches the phase for all
 * domains at once, the others return immediately.  Launches are queued on the
 * context's stream; nothing waits for the device until the host reads data.
 *
 * Errors: the reference aborts through LOG_S(FATAL); these objects throw
 * picopic_hip::Error with the library's message instead.  A block is safe to
 * call from the threads of an OpenMP loop (one mutex per block).
 *
 * Two modes:
 *   PHASES      every reference phase maps to its entry point (FP64 contexts)
 *   WHOLE_STEP  the first phase of a step (calc_field_e) runs the whole step
 *               -- Yee, stitches, push+move+deposit fused, migration, J
 *               stitch -- and the remaining phases of that step are no-ops;
 *               the only mode of FP32 split-precision contexts
 */
#ifndef PICOPIC_SOLVERS_HPP
#define PICOPIC_SOLVERS_HPP

#include <array>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "picopic.h"

namespace picopic_hip {

struct Error : std::runtime_error {
  int code;
  Error(int rc, const std::string &what) : std::runtime_error(what), code(rc) {}
};

inline void check(int rc, const char *where) {
  if (rc != PCP_OK) throw Error(rc, std::string(where) + ": " + pcp_last_error());
}

enum class Mode { PHASES, WHOLE_STEP };

// the phases of one reference time step, in order (src/PiCoPiC.cpp:299-345)
// (the collision pass sits between the push and the move, as in
// SMB::advance_particles)
enum Phase { FIELD_E = 0, STITCH_E, FIELD_H, STITCH_H, PUSH, COLLIDE, KINEMATICS, MIGRATE, DEPOSIT, STITCH_J, N_PHASES };

class DeviceBlock {
 public:
  DeviceBlock(const pcp_config &cfg, Mode mode) : mode_(mode) {
    if (cfg.precision == PCP_PREC_FP32_SPLIT && mode == Mode::PHASES)
      throw Error(PCP_E_ARG, "FP32 split-precision blocks run in WHOLE_STEP mode");
    check(pcp_create(&cfg, &ctx_), "pcp_create");
    domains_ = (int)(cfg.domains_r * cfg.domains_z);
    for (auto &s : seen_) s.assign((size_t)domains_, 0);
    count_.fill(0);
  }
  ~DeviceBlock() { pcp_destroy(ctx_); }
  DeviceBlock(const DeviceBlock &) = delete;
  DeviceBlock &operator=(const DeviceBlock &) = delete;

  pcp_ctx *ctx() const { return ctx_; }
  int domains() const { return domains_; }
  Mode mode() const { return mode_; }
  long steps_launched() const { return steps_; }

  // per-domain arrival at a phase; the completing arrival launches it
  void arrive(Phase ph, int domain) {
    std::lock_guard<std::mutex> lock(mu_);
    if (domain < 0 || domain >= domains_) throw Error(PCP_E_ARG, "domain index out of range");
    if (seen_[ph][(size_t)domain]) throw Error(PCP_E_STATE, "domain arrived twice at one phase");
    seen_[ph][(size_t)domain] = 1;
    if (++count_[ph] < domains_) return;
    count_[ph] = 0;
    std::fill(seen_[ph].begin(), seen_[ph].end(), 0);
    launch(ph);
  }
  // SMB-level phases, called once per block (the reference's overlay and
  // runaway-collector calls after the OpenMP loops)
  void field_e_overlay() { once(STITCH_E); }
  void field_h_overlay() { once(STITCH_H); }
  void current_overlay() { once(STITCH_J); }
  void particles_runaway_collector() { once(MIGRATE); }
  // the per-domain kinematics wrappers of Domain (dump_particle_positions_to_old,
  // update_particles_coords, particles_back_position_to_rz, reflect,
  // particles_back_velocity_to_rz, bind_cell_numbers) as one arrival
  void kinematics(int domain) { arrive(KINEMATICS, domain); }

  void sync() { check(pcp_sync(ctx_), "pcp_sync"); }

  // the block's collision operator (every domain's Collisions object names
  // the same one); in WHOLE_STEP mode it becomes the collision phase of
  // pcp_step, in PHASES mode the COLLIDE phase calls pcp_collide
  void set_collisions(int scheme, int electrons, int ions, uint64_t seed) {
    std::lock_guard<std::mutex> lock(mu_);
    if (coll_on_ && scheme == coll_[0] && electrons == coll_[1] && ions == coll_[2] && seed == coll_seed_) return;
    if (mode_ == Mode::WHOLE_STEP)
      check(pcp_set_collisions(ctx_, scheme, electrons, ions, seed), "pcp_set_collisions");
    coll_on_ = true;
    coll_[0] = scheme;
    coll_[1] = electrons;
    coll_[2] = ions;
    coll_seed_ = seed;
    coll_pass_ = 0;
  }

 private:
  void once(Phase ph) {
    std::lock_guard<std::mutex> lock(mu_);
    launch(ph);
  }
  void launch(Phase ph) {
    if (mode_ == Mode::WHOLE_STEP) {
      if (ph == FIELD_E) {
        check(pcp_step(ctx_, 1), "pcp_step");
        ++steps_;
      }
      return;
    }
    switch (ph) {
      case FIELD_E: check(pcp_yee_e(ctx_), "pcp_yee_e"); break;
      case STITCH_E: check(pcp_stitch(ctx_, PCP_GRID_E), "pcp_stitch E"); break;
      case FIELD_H: check(pcp_yee_h(ctx_), "pcp_yee_h"); break;
      case STITCH_H: check(pcp_stitch(ctx_, PCP_GRID_H), "pcp_stitch H"); break;
      case PUSH: check(pcp_push(ctx_), "pcp_push"); break;
      case COLLIDE:
        if (!coll_on_) throw Error(PCP_E_STATE, "collision phase without a collision operator");
        check(pcp_collide(ctx_, coll_[0], coll_[1], coll_[2], coll_seed_, coll_pass_++), "pcp_collide");
        break;
      case KINEMATICS: check(pcp_move(ctx_), "pcp_move"); break;
      case MIGRATE: check(pcp_migrate(ctx_), "pcp_migrate"); break;
      case DEPOSIT: check(pcp_deposit(ctx_), "pcp_deposit"); break;
      case STITCH_J:
        check(pcp_stitch(ctx_, PCP_GRID_J), "pcp_stitch J");
        ++steps_;
        break;
      default: break;
    }
  }

  pcp_ctx *ctx_ = nullptr;
  bool coll_on_ = false;
  int coll_[3] = {0, 0, 0};
  uint64_t coll_seed_ = 0, coll_pass_ = 0;
  int domains_ = 0;
  Mode mode_;
  long steps_ = 0;
  std::mutex mu_;
  std::array<std::vector<char>, N_PHASES> seen_;
  std::array<int, N_PHASES> count_;
};

// Pusher (include/pusher.hpp:33-59): the pusher kind is the context's
// (configure.ac --with-pusher as pcp_config::pusher)
class DevicePusher {
 public:
  DevicePusher(DeviceBlock &b, int domain) : b_(b), d_(domain) {}
  void operator()() { b_.arrive(PUSH, d_); }

 private:
  DeviceBlock &b_;
  int d_;
};

// Current (include/current.hpp:35-54); Domain::reset_current is folded in
// (the device deposit clears J first)
class DeviceCurrent {
 public:
  DeviceCurrent(DeviceBlock &b, int domain) : b_(b), d_(domain) {}
  void current_distribution() { b_.arrive(DEPOSIT, d_); }

 private:
  DeviceBlock &b_;
  int d_;
};

// MaxwellSolver / MaxwellSolverYee (include/maxwellSolver.hpp:35-317,
// src/maxwellSolver/maxwellSolverYee.cpp:22-257)
class DeviceMaxwellSolver {
 public:
  DeviceMaxwellSolver(DeviceBlock &b, int domain) : b_(b), d_(domain) {}
  // sigma is set at context creation from pcp_config's PML fields
  void set_pml() {}
  void calc_field_e() { b_.arrive(FIELD_E, d_); }
  void calc_field_h() { b_.arrive(FIELD_H, d_); }
  std::array<double, 3> get_field_e(double radius, double longitude) {
    std::array<double, 3> out{};
    check(pcp_get_field_e(b_.ctx(), d_, radius, longitude, out.data()), "pcp_get_field_e");
    return out;
  }
  std::array<double, 3> get_field_h(double radius, double longitude) {
    std::array<double, 3> out{};
    check(pcp_get_field_h(b_.ctx(), d_, radius, longitude, out.data()), "pcp_get_field_h");
    return out;
  }

 private:
  DeviceBlock &b_;
  int d_;
};

// Collisions (include/collisions.hpp:32-78; CollisionsP12 / TA77S / SK98
// operator(), src/collisions/*.cpp): one object per domain as the reference
// builds them; the pass runs over all domains of the block when the last
// domain arrives.  scheme: PCP_COLL_P12, PCP_COLL_TA77S, PCP_COLL_SK98 or
// PCP_COLL_SURROGATE; electrons / ions: species indices.
class DeviceCollisions {
 public:
  DeviceCollisions(DeviceBlock &b, int domain, int scheme, int electrons, int ions, uint64_t seed)
      : b_(b), d_(domain) {
    b_.set_collisions(scheme, electrons, ions, seed);
  }
  void operator()() {
    if (b_.mode() == Mode::PHASES) b_.arrive(COLLIDE, d_);
  }

 private:
  DeviceBlock &b_;
  int d_;
};

}  // namespace picopic_hip

#endif
