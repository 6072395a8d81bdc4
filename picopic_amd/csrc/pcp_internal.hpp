// Internal state and device helpers of the picopic_amd HIP library.
//
// Device layout (HBM):
//   field grids  : one array per kind, [domain][component][Nr_d+4][Nz_d+4],
//                  z contiguous -- the reference Grid<double> layout with the
//                  row pointers flattened (include/algo/grid.hpp:24-58)
//   particles    : per species SoA. Positions are double-buffered (pos[cur]
//                  is the current position, pos[cur^1] the position before
//                  the last move), velocities/weights/keys single-buffered.
//   order keys   : per particle, the position in the reference's per-domain
//                  vector (seq); removed particles carry r < 0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/picopic.h"

namespace pcp {

// reference include/constant.hpp:28-58, configure.ac:224
constexpr double PI = 3.1415926535897932;
constexpr double EPSILON0 = 8.85E-12;
constexpr double EL_CHARGE = 1.6E-19;
constexpr double EL_MASS = 9.1E-31;
constexpr double LIGHT_VEL = 3.0E8;
constexpr double LIGHT_VEL_POW_2 = LIGHT_VEL * LIGHT_VEL;
constexpr double MAGN_CONST = 1.26E-6;
constexpr double MNZL = 1e-15;
constexpr double REL_LIMIT = 5e7;

constexpr int MAX_SPECIES = 8;

// device-side error bits (reported as PCP_E_STATE)
enum : int {
  ERR_NAN_POSITION = 1,   // pusherBoris.cpp:51-57
  ERR_BETA_GT_1 = 2,      // rel.cpp:32-34
  ERR_NAN_FIELD = 4,      // maxwellSolverYee.cpp:253-255
  ERR_MIGRANT_OVERFLOW = 8,
  ERR_MOVED_TOO_FAR = 16, // currentVB.cpp:199-204 (LOG_S(ERROR), not fatal)
  ERR_COLL_OVERFLOW = 32, // crowded collision cells hold more macro-particles than a pass stages
};

// Geometry shared by every domain of the SMB (domains are equal).
// Division by an invariant divisor d >= 1 for dividends n < 2^31:
// q = floor(n m / 2^p) with p = 31 + ceil(log2 d), m = ceil(2^p / d)
// (m < 2^32).  Exact: n m / 2^p = n / d + n e / 2^p with 0 <= e < 1, and
// n e / 2^p < 2^-ceil(log2 d) <= 1 / d never carries n / d past an integer.
struct FastDiv {
  uint32_t d, m;
  int p;
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const { return (uint32_t)(((uint64_t)n * m) >> p); }
};

inline FastDiv make_fastdiv(uint32_t d) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  FastDiv f;
  f.d = d;
  f.p = 31 + l;
  f.m = (uint32_t)(((1ull << f.p) + d - 1) / d);
  return f;
}

struct Geom {
  int Dr, Dz;        // domains
  int nr, nz;        // cells per domain
  int xr, yr;        // nr + 4, nz + 4
  long P;            // xr * yr
  long Nr, Nz;       // global cells
  double dr, dz;     // domain cell size (size / domains / cells)
  double dt;
  double size_r, size_z;             // domain size
  double gsize_r, gsize_z;           // global size
  long gca_r, gca_z;                 // domain Geometry::global_cell_amount
  int pml_on;
  int ccs, pusher;
  int bin_r, bin_z;  // sort bin (cells; divide nr and nz)
  int nbz;           // sort bins along z per domain (nz / bin_z)
  FastDiv div_nr, div_nz, div_Dz, div_Nz, div_bin_r, div_bin_z;
  double inv_vol0;   // 1 / (pi dz dr^2): CELL_VOLUME(i) = (2 i + 1) / inv_vol0
  float dtr_f, dtz_f, inv_dt_f;   // dt / dr, dt / dz, 1 / dt rounded to FP32
};

// Per-domain values derived from the domain index (src/SMB.cpp:49-121).
struct Dom {
  int I, J;          // domain indices (r, z)
  long bot_r, left_z;
  int w0, w1, w2, w3; // walls r=0, z=0, r=R, z=L
};

__host__ __device__ inline Dom make_dom(const Geom &g, int d) {
  Dom o;
  o.I = (int)g.div_Dz.div((uint32_t)d);
  o.J = d - o.I * g.Dz;
  o.bot_r = (long)g.nr * o.I;
  o.left_z = (long)g.nz * o.J;
  o.w0 = o.I == 0;
  o.w1 = o.J == 0;
  o.w2 = o.I == g.Dr - 1;
  o.w3 = o.J == g.Dz - 1;
  return o;
}

// Sort key of local cell (lr, lz) of domain d: the cells of a domain in
// bins of bin_r x bin_z cells, bins row-major, cells row-major inside a bin
// (bin_r = 1, bin_z = nz is plain row-major order).  A permutation of
// [0, D nr nz), so that every cell keeps one contiguous run of particles.
__host__ __device__ inline uint32_t sort_cell_key(const Geom &g, int d, long lr, long lz) {
  const long bsz = (long)g.bin_r * g.bin_z;
  const long bi = g.div_bin_r.div((uint32_t)lr), bk = g.div_bin_z.div((uint32_t)lz);
  return (uint32_t)((long)d * g.nr * g.nz + (bi * g.nbz + bk) * bsz + (lr - bi * g.bin_r) * g.bin_z +
                    (lz - bk * g.bin_z));
}

// halo-grid index of local (i, k) (interior-relative, may be -2..n+1)
__host__ __device__ inline long gidx(const Geom &g, int d, int comp, int ncomp, int i, int k) {
  return ((long)d * ncomp + comp) * g.P + (long)(i + 2) * g.yr + (k + 2);
}

template <class T>
struct FieldsT {
  T *E, *H, *Hat, *J;   // [D][3][P]
  T *eps, *sigma;       // [D][P]
};
using Fields = FieldsT<double>;

struct SpeciesDev {
  double *r[2], *z[2];
  double *vr, *vphi, *vz, *w;
  uint64_t *seq;     // order key within (domain, species)
  uint64_t *id;
  long n;            // slots in use (live + removed)
  long cap;
  double q, m;
  int beam;
};

// one step's bookkeeping kept in device memory
struct StepState {
  unsigned long long step;     // step counter (stamps the strict-motion markers)
  int err;
  int need_exact;              // a strict-axis J_r reset happened in this step
  unsigned int n_mig;          // migrants appended in this step
  unsigned int queue;          // next trip of the batched particle launch
};

// second set of per-particle arrays used as the sort's scatter target
struct SortBuffers {
  double *vr, *vphi, *vz, *w;
  uint64_t *seq, *id;
};

struct Migrants {
  uint64_t *key;      // colour << 47 | srcdom << 35 | old seq
  uint64_t *ref;      // species << 40 | particle index
  uint32_t *group;    // dst domain * MAX_SPECIES + species
  uint32_t *bucket;   // migrant indices grouped by group, each group's tiles sorted by key
  uint64_t *skey;     // the keys in bucket order
  uint32_t *gcount, *gstart, *gfill;
  uint32_t *gtile;    // exclusive scan of the groups' MIG_TILE-entry tile counts, total last
  uint64_t *gbase;    // next free seq per (domain, species)
  long cap;
  int ngroups;
};
// migrants of one group are ranked by sorting tiles of MIG_TILE keys in LDS
// and binary-searching each key in the group's other tiles
constexpr int MIG_TILE_THREADS = 256, MIG_TILE_ITEMS = 8, MIG_TILE = MIG_TILE_THREADS * MIG_TILE_ITEMS;

// In-step energy log (pcp_set_energy_log): a ring of EHIST rows of
// EHIST_COLS doubles -- step number, field energy, kinetic energy of each
// species -- filled on the device during every logged step.
constexpr int EHIST = 4096;
constexpr int DIAG_DOUBLES = 16;   // per-context scratch of the diagnostic reductions
constexpr int EHIST_COLS = 2 + MAX_SPECIES;

__host__ __device__ inline double *ehist_row(double *h, unsigned long long step) {
  return h + (long)(step % EHIST) * EHIST_COLS;
}

// domain owning global cell (cr, cz)
__device__ __forceinline__ int domain_of(const Geom &g, int cr, int cz) { return (cr / g.nr) * g.Dz + (cz / g.nz); }

// Appends particle p of species sp, moving from domain d_old to d_new, to
// the step's migrant list.  Its key orders the arrivals of a destination as
// particles_runaway_collector's push_backs do: colour pass of the source
// (SMB.cpp:211-340 visits domains in 2 x 2 colour order), source domain,
// position in the source vector.
__device__ __forceinline__ void append_migrant(Migrants mg, StepState *st, int sp, long p, int d_old, int d_new,
                                               uint64_t seq, const Geom &g, int *err) {
  unsigned int slot = atomicAdd(&st->n_mig, 1u);
  if (slot >= (unsigned long)mg.cap) { *err |= ERR_MIGRANT_OVERFLOW; return; }
  const int I = (int)g.div_Dz.div((uint32_t)d_old), J = d_old - I * g.Dz;
  const uint64_t colour = (uint64_t)((I % 2) * 2 + (J % 2));
  mg.key[slot] = (colour << 47) | ((uint64_t)d_old << 35) | (seq & ((1ull << 35) - 1));
  mg.ref[slot] = ((uint64_t)sp << 40) | (uint64_t)p;
  mg.group[slot] = (uint32_t)(d_new * MAX_SPECIES + sp);
}

// BeamP (src/beamP.cpp:24-66) per species: the configuration and, per
// domain, the constructor's injection radius and rate and the injection
// state (bunch counter, particles injected so far)
struct BeamDomHost {
  double domain_radius = 0, macro_per_step = 0;
  unsigned int bunch = 0;
  uint64_t injected = 0;
};
struct BeamHost {
  int on = 0;
  unsigned int macro_amount = 0, bunches_amount = 0;
  double start_time = 0, radius = 0, density = 0, bunch_length = 0, bunches_distance = 0, velocity = 0;
  uint64_t seed = 0;
  std::vector<BeamDomHost> dom;
};

struct State32;      // FP32 split-precision state (pcp32.hpp)
struct CollBuffers;  // collision pass buffers (pcp_collide.hip)
struct MlpState;     // collision surrogate network (pcp_mlp.hpp)

}  // namespace pcp

struct pcp_ctx {
  pcp_config cfg;
  pcp::Geom g;
  int D;
  hipStream_t stream;
  pcp::Fields f;
  pcp::SpeciesDev sp[pcp::MAX_SPECIES];
  int cur;                         // current position buffer
  pcp::StepState *st;              // device
  pcp::Migrants mig;
  uint64_t *marker;                // [D][P] strict-motion J_r reset markers
  uint32_t *cell_flag;             // [Nr][Nz] cells near reset nodes (step stamp)
  pcp::SortBuffers sortb[pcp::MAX_SPECIES];
  uint32_t *sort_key, *sort_count, *sort_offset, *sort_fill;
  void *sort_temp;
  size_t sort_temp_bytes;
  long sort_key_cap;
  int sort_every[pcp::MAX_SPECIES]; // 0: never; k: before every k-th step
  long steps_done;
  pcp::State32 *s32 = nullptr;    // non-null in FP32 split mode
  pcp::CollBuffers *coll = nullptr;
  pcp::MlpState *mlp = nullptr;
  // launch shape of the batched FP32 particle kernel (autotuned)
  struct {
    int block = 256;                // threads per block: 128, 256 or 512
    int blocks_per_cu = 0;          // resident blocks per CU (0: occupancy)
  } tune;
  int graph_on = 0;                 // replay captured step graphs
  // captured step graphs, one per buffer state (the particle buffer
  // pointers, counts and `cur` that the step's kernels take); the sorts and
  // the velocity-slot exchange cycle through a few states
  struct StepGraph {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec = nullptr;
    unsigned long long used = 0;
  };
  std::vector<StepGraph> graphs;
  unsigned long long graph_tick = 0;
  long graph_launches = 0, graph_captures = 0;
  int coll_scheme = -1;            // collision phase of pcp_step (-1: off)
  int coll_se = 0, coll_si = 1;
  uint64_t coll_seed = 0;
  uint64_t *coll_pass = nullptr;    // device: the step's collision pass number
  int n_cu;
  int device = -1;                  // the HIP device the context runs on
  int timing;
  double t_now = 0;                 // TimeSim::current: advances by dt after every step
  pcp::BeamHost beam[pcp::MAX_SPECIES];
  double *diag = nullptr;           // scratch of the diagnostic reductions (DIAG_DOUBLES)
  double *probe_buf = nullptr;      // staging of probe windows (grown on demand)
  long probe_cap = 0;
  // species maps: scratch density and momentum_w, the last map (3 x D P),
  // and each species' momentum_abs_w map (kept between calls, as the
  // reference's member is)
  double *map_n = nullptr, *map_pw = nullptr, *map_out = nullptr;
  double *map_pabs[pcp::MAX_SPECIES] = {};
  int map_ncomp = 0;
  double *ehist = nullptr;          // energy log [EHIST][EHIST_COLS]
  int elog = 0;                     // log energies in every step
  std::vector<std::pair<std::string, double>> ktimes;
};
