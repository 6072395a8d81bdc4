// A shared-memory-block time loop written the way the reference's SMB drives
// its domains (src/SMB.cpp:1231-1306, src/PiCoPiC.cpp:299-345): OpenMP loops
// over the domain grid calling each domain's solver objects, then the SMB
// phase boundary (overlay, runaway collector).  The solvers are the
// device-backed drop-ins of picopic_solvers.hpp.
//
// usage: smb_driver [steps] [p12|ta77s|sk98]
// Runs the same seeded two-species input twice on the device -- once with
// every reference phase mapped to its entry point (PHASES) and once with one
// fused step per time step (WHOLE_STEP) -- and prints one JSON line with the
// largest relative differences between the two end states.  With a
// collision scheme, every domain holds a Collisions object and the pass
// runs between the push and the move.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "picopic_solvers.hpp"

namespace {

using picopic_hip::DeviceBlock;

// one domain of the block, holding its solver objects like Domain
// (include/domain.hpp:77-143)
struct Domain {
  picopic_hip::DevicePusher pusher;
  picopic_hip::DeviceCurrent current;
  picopic_hip::DeviceMaxwellSolver maxwell_solver;
  std::unique_ptr<picopic_hip::DeviceCollisions> collisions;
  DeviceBlock &block;
  int index;
  Domain(DeviceBlock &b, int d, int scheme)
      : pusher(b, d), current(b, d), maxwell_solver(b, d), block(b), index(d) {
    if (scheme >= 0) collisions.reset(new picopic_hip::DeviceCollisions(b, d, scheme, 0, 1, 77));
  }
  void weight_field_e() { maxwell_solver.calc_field_e(); }
  void weight_field_h() { maxwell_solver.calc_field_h(); }
  void push_particles() { pusher(); }
  void collide() {
    if (collisions) (*collisions)();
  }
  void advance_kinematics() { block.kinematics(index); }
  void weight_current() { current.current_distribution(); }
};

struct SMB {
  DeviceBlock block;
  int r_domains, z_domains;
  std::vector<std::unique_ptr<Domain>> domains;
  SMB(const pcp_config &cfg, picopic_hip::Mode mode, int scheme)
      : block(cfg, mode), r_domains((int)cfg.domains_r), z_domains((int)cfg.domains_z) {
    for (int d = 0; d < r_domains * z_domains; ++d) domains.emplace_back(new Domain(block, d, scheme));
  }
  Domain &at(int i, int j) { return *domains[(size_t)(i * z_domains + j)]; }
  void solve_maxvell() {
#pragma omp parallel for collapse(2)
    for (int i = 0; i < r_domains; ++i)
      for (int j = 0; j < z_domains; ++j) at(i, j).weight_field_e();
    block.field_e_overlay();
#pragma omp parallel for collapse(2)
    for (int i = 0; i < r_domains; ++i)
      for (int j = 0; j < z_domains; ++j) at(i, j).weight_field_h();
    block.field_h_overlay();
  }
  void advance_particles() {
#pragma omp parallel for collapse(2)
    for (int i = 0; i < r_domains; ++i)
      for (int j = 0; j < z_domains; ++j) {
        at(i, j).push_particles();
        at(i, j).collide();
        at(i, j).advance_kinematics();
      }
    block.particles_runaway_collector();
  }
  void solve_current() {
#pragma omp parallel for collapse(2)
    for (int i = 0; i < r_domains; ++i)
      for (int j = 0; j < z_domains; ++j) at(i, j).weight_current();
    block.current_overlay();
  }
};

// deterministic host input (64-bit LCG)
struct Lcg {
  uint64_t s;
  double next() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
};

pcp_config make_config() {
  pcp_config c{};
  c.size_r = 32e-3;
  c.size_z = 64e-3;
  c.grid_r = 32;
  c.grid_z = 64;
  c.domains_r = 2;
  c.domains_z = 2;
  c.dt = 0.5 * 1e-3 / 3e8;
  c.pusher = PCP_PUSHER_BORIS_ADAPTIVE;
  c.ccs = PCP_CCS_VB;
  c.precision = PCP_PREC_FP64;
  c.n_species = 2;
  c.species[0] = pcp_species_desc{-1.6e-19, 9.1e-31, 0, 40000};
  c.species[1] = pcp_species_desc{1.6e-19, 1836 * 9.1e-31, 0, 40000};
  return c;
}

void upload(DeviceBlock &b, const pcp_config &c, long n) {
  for (int sp = 0; sp < c.n_species; ++sp) {
    Lcg g{(uint64_t)(1234 + sp)};
    std::vector<double> r(n), z(n), vr(n), vp(n), vz(n), w(n, 1e6);
    std::vector<uint64_t> id(n);
    const double vmax = sp == 0 ? 2e7 : 2e5;
    for (long p = 0; p < n; ++p) {
      r[p] = std::sqrt(g.next()) * (c.size_r - 1e-3) + 0.5e-3;
      z[p] = 0.5e-3 + g.next() * (c.size_z - 1e-3);
      vr[p] = (2 * g.next() - 1) * vmax;
      vp[p] = (2 * g.next() - 1) * vmax;
      vz[p] = (2 * g.next() - 1) * vmax;
      id[p] = (uint64_t)(sp * 1000000 + p);
    }
    picopic_hip::check(pcp_upload_species(b.ctx(), sp, n, r.data(), z.data(), vr.data(), vp.data(), vz.data(),
                                          w.data(), id.data()),
                       "pcp_upload_species");
  }
}

struct State {
  std::vector<std::vector<double>> cols[2];   // per species: r, z, vr, vphi, vz sorted by id
  std::vector<double> grids[3];               // E, H, J
};

State read_state(DeviceBlock &b, const pcp_config &c) {
  State s;
  for (int sp = 0; sp < c.n_species; ++sp) {
    long n = 0;
    picopic_hip::check(pcp_count(b.ctx(), sp, &n), "pcp_count");
    std::vector<double> a[5];
    for (auto &x : a) x.resize((size_t)n);
    std::vector<uint64_t> id((size_t)n);
    picopic_hip::check(pcp_download_species(b.ctx(), sp, a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                                            a[4].data(), nullptr, id.data(), nullptr, nullptr, nullptr),
                       "pcp_download_species");
    std::vector<size_t> o((size_t)n);
    for (size_t k = 0; k < o.size(); ++k) o[k] = k;
    std::sort(o.begin(), o.end(), [&](size_t x, size_t y) { return id[x] < id[y]; });
    for (auto &x : a) {
      std::vector<double> t(x.size());
      for (size_t k = 0; k < o.size(); ++k) t[k] = x[o[k]];
      s.cols[sp].push_back(t);
    }
  }
  const int which[3] = {PCP_GRID_E, PCP_GRID_H, PCP_GRID_J};
  for (int k = 0; k < 3; ++k) {
    // [domain][component][Nr_d + 4][Nz_d + 4]
    s.grids[k].resize((size_t)pcp_grid_points(b.ctx()) * 3 * (size_t)b.domains());
    picopic_hip::check(pcp_download_grid(b.ctx(), which[k], s.grids[k].data()), "pcp_download_grid");
  }
  return s;
}

// max |a - b| / max |b| over interior nodes of a grid (halo J is undefined
// after the J stitch)
double grid_err(const std::vector<double> &a, const std::vector<double> &b, const pcp_config &c) {
  const long nr = c.grid_r / c.domains_r, nz = c.grid_z / c.domains_z, yr = nz + 4;
  const long per = (nr + 4) * yr;
  double num = 0, den = 0;
  for (size_t blk = 0; blk < a.size() / (size_t)per; ++blk)
    for (long i = 2; i < nr + 2; ++i)
      for (long k = 2; k < nz + 2; ++k) {
        size_t o = blk * (size_t)per + (size_t)(i * yr + k);
        num = std::max(num, std::fabs(a[o] - b[o]));
        den = std::max(den, std::fabs(b[o]));
      }
  return den > 0 ? num / den : num;
}

double col_err(const std::vector<double> &a, const std::vector<double> &b) {
  if (a.size() != b.size()) return INFINITY;
  double num = 0, den = 0;
  for (size_t k = 0; k < a.size(); ++k) {
    num = std::max(num, std::fabs(a[k] - b[k]));
    den = std::max(den, std::fabs(b[k]));
  }
  return den > 0 ? num / den : num;
}

}  // namespace

int main(int argc, char **argv) {
  const int steps = argc > 1 ? std::atoi(argv[1]) : 5;
  int scheme = -1;
  if (argc > 2) {
    const std::string s = argv[2];
    scheme = s == "p12" ? PCP_COLL_P12 : s == "ta77s" ? PCP_COLL_TA77S : s == "sk98" ? PCP_COLL_SK98 : -2;
    if (scheme == -2) {
      std::fprintf(stderr, "smb_driver: unknown collision scheme %s\n", argv[2]);
      return 2;
    }
  }
  const pcp_config cfg = make_config();
  const long n = 8000;
  try {
    SMB phases(cfg, picopic_hip::Mode::PHASES, scheme);
    SMB fused(cfg, picopic_hip::Mode::WHOLE_STEP, scheme);
    upload(phases.block, cfg, n);
    upload(fused.block, cfg, n);
    for (int t = 0; t < steps; ++t) {
      for (SMB *s : {&phases, &fused}) {
        s->solve_maxvell();
        s->advance_particles();
        s->solve_current();
      }
    }
    State a = read_state(phases.block, cfg), b = read_state(fused.block, cfg);
    double part = 0;
    for (int sp = 0; sp < cfg.n_species; ++sp)
      for (int k = 0; k < 5; ++k) part = std::max(part, col_err(a.cols[sp][(size_t)k], b.cols[sp][(size_t)k]));
    std::printf("{\"steps\": %d, \"collisions\": %d, \"threads\": %d, \"launched_phases\": %ld, \"launched_fused\": %ld, "
                "\"particles\": %.3e, \"E\": %.3e, \"H\": %.3e, \"J\": %.3e}\n",
                steps, scheme, omp_get_max_threads(), phases.block.steps_launched(), fused.block.steps_launched(), part,
                grid_err(a.grids[0], b.grids[0], cfg), grid_err(a.grids[1], b.grids[1], cfg),
                grid_err(a.grids[2], b.grids[2], cfg));
  } catch (const picopic_hip::Error &e) {
    std::fprintf(stderr, "smb_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
