// Binary Coulomb collisions on the device: the Collisions base pass
// (src/collisions.cpp:87-236: sort_to_cells, random_sort,
// collect_weighted_params_tot_grid) and the pairing loop shared by
// CollisionsP12 / TA77S / SK98 (collisionsP12.cpp:330-530), plus TA77S
// correct_velocities (collisionsTA77S.cpp:419-566).
//
// Pipeline per pass (two species: electrons `se`, ions `si`):
//   k_coll_count    particles per (domain, cell) bucket
//   scan            bucket offsets (hipCUB)
//   k_coll_fill     bucket lists of particle indices
//   k_coll_classify cells with both species into lists by size (at most 16,
//                   64, 256 or more particles of either species; a cell
//                   missing a species ends its r-row as the reference's
//                   `break` does)
//   k_coll_prep     per listed cell (one lane per cell of at most 16, one
//                   wavefront with an LDS stage per cell of at most 64 or
//                   256, one workgroup with a global stage per larger cell):
//                   shuffle each bucket by sorting on
//                   a Philox key of (particle id, pass), per-cell sums in the
//                   shuffled order, density / temperature / Debye length, and
//                   the first cell of each r-row whose values are not normal
//                   (the reference's `break` ends the rest of that row)
//   k_coll_pairs    per listed cell, grouped the same way: the reference's
//                   pair schedule with lanes on independent pair chains,
//                   velocities staged in FP64; pair q of global cell c draws
//                   the uniforms Philox(seed; c << 24 | q, 2 pass + 1)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

#include "pcp32.hpp"
#include "pcp_collide.hpp"
#include "pcp_collide_physics.hpp"
#include "pcp_mlp.hpp"
#include "pcp_rng.hpp"
#include "pcp_wave.hpp"

namespace pcp {

namespace {

constexpr int COLL_BLOCK = 64;   // one wavefront per cell

// particle access for FP64 and FP32 split storage
struct View64 {
  SpeciesDev s;
  int cur;
  __device__ __forceinline__ bool cell_of(const Geom &g, long p, int &cr, int &cz) const {
    double r = s.r[cur][p];
    if (!(r >= 0)) return false;
    cr = cell_number(r, g.dr);
    cz = cell_number(s.z[cur][p], g.dz);
    return true;
  }
  __device__ __forceinline__ void load(long p, double v[3]) const {
    v[0] = s.vr[p]; v[1] = s.vphi[p]; v[2] = s.vz[p];
  }
  __device__ __forceinline__ void store(long p, const double v[3]) const {
    s.vr[p] = v[0]; s.vphi[p] = v[1]; s.vz[p] = v[2];
  }
  __device__ __forceinline__ double w(long p) const { return s.w[p]; }
  __device__ __forceinline__ uint64_t id(long p) const { return s.id[p]; }
  __device__ __forceinline__ long n() const { return s.n; }
};

struct View32 {
  Species32 s;
  int cur;    // position slot
  int vslot;  // velocity slot
  __device__ __forceinline__ bool cell_of(const Geom &g, long p, int &cr, int &cz) const {
    uint32_t c = s.cell[cur][p];
    if (c == DEAD_CELL) return false;
    cr = (int)g.div_Nz.div(c);
    cz = (int)(c - (uint32_t)cr * (uint32_t)g.Nz);
    return true;
  }
  __device__ __forceinline__ void load(long p, double v[3]) const {
    v[0] = s.vr[vslot][p]; v[1] = s.vphi[vslot][p]; v[2] = s.vz[vslot][p];
  }
  __device__ __forceinline__ void store(long p, const double v[3]) const {
    s.vr[vslot][p] = (float)v[0]; s.vphi[vslot][p] = (float)v[1]; s.vz[vslot][p] = (float)v[2];
  }
  __device__ __forceinline__ double w(long p) const { return (double)s.w[p]; }
  __device__ __forceinline__ uint64_t id(long p) const { return s.id[p]; }
  __device__ __forceinline__ long n() const { return s.n; }
};

// domain-major key of global cell (cr, cz): domain d = I Dz + J, then the
// cell's row-major index inside the domain (multiply-high divisions)
__device__ __forceinline__ uint32_t local_cell_key(const Geom &g, int cr, int cz) {
  const int I = (int)g.div_nr.div((uint32_t)cr), J = (int)g.div_nz.div((uint32_t)cz);
  const int lr = cr - I * g.nr, lz = cz - J * g.nz;
  return (uint32_t)(((long)(I * g.Dz + J) * g.nr + lr) * g.nz + lz);
}

// gate (optional): a device counter; the pass does nothing while it is 0.
// mask (optional): only particles of cells with mask[key] != 0 are bucketed.
__device__ __forceinline__ bool gate_closed(const unsigned long long *gate) { return gate && *gate == 0ull; }

// bucket of particle p (sentinel: removed, or its cell not in mask)
template <class V>
__device__ __forceinline__ uint32_t bucket_of(const Geom &g, const V &v, long p, long n, uint32_t sentinel,
                                              const int *mask) {
  uint32_t k = sentinel;
  int cr, cz;
  if (p < n && v.cell_of(g, p, cr, cz)) k = local_cell_key(g, cr, cz);
  if (mask && k != sentinel && !mask[k]) k = sentinel;
  return k;
}

// pass 1: particles per bucket (one atomic per distinct key of a wavefront)
template <class V>
__global__ void __launch_bounds__(256) k_coll_count(Geom g, V v, uint32_t *cnt, uint32_t sentinel, const int *mask,
                                                    const unsigned long long *gate) {
  if (gate_closed(gate)) return;
  const long n = v.n();
  const long stride = (long)gridDim.x * blockDim.x;
  for (long b0 = (long)blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
    const long p = b0 + threadIdx.x;
    const uint32_t k = bucket_of(g, v, p, n, sentinel, mask);
    const bool a = k != sentinel;
    (void)wave_reserve(cnt, a ? k : 0, a);
  }
}

// pass 2: each particle's index at its bucket offset + a slot (fill counts
// from zero); the order inside a bucket is the atomics' -- the shuffle
// orders every bucket by its particles' keys
template <class V>
__global__ void __launch_bounds__(256) k_coll_fill(Geom g, V v, const uint32_t *off, uint32_t *fill, uint32_t *perm,
                                                   uint32_t sentinel, const int *mask,
                                                   const unsigned long long *gate) {
  if (gate_closed(gate)) return;
  const long n = v.n();
  const long stride = (long)gridDim.x * blockDim.x;
  for (long b0 = (long)blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
    const long p = b0 + threadIdx.x;
    const uint32_t k = bucket_of(g, v, p, n, sentinel, mask);
    const bool a = k != sentinel;
    const uint32_t slot = wave_reserve(fill, a ? k : 0, a);
    if (a) perm[off[k] + slot] = (uint32_t)p;
  }
}

__global__ void k_fill_i32(int *p, long n, int v) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = v;
}

__global__ void k_fill_u32(uint32_t *p, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = 0;
}

// cell c (domain-major key) -> domain, local (i, j), global (cr, cz)
struct CellPos {
  int d, i, j;
  long cr, cz;
};
__device__ __forceinline__ CellPos cell_pos(const Geom &g, long c) {
  CellPos q;
  long per = (long)g.nr * g.nz;
  q.d = (int)(c / per);
  long lc = c - (long)q.d * per;
  q.i = (int)(lc / g.nz);
  q.j = (int)(lc - (long)q.i * g.nz);
  q.cr = (long)(q.d / g.Dz) * g.nr + q.i;
  q.cz = (long)(q.d % g.Dz) * g.nz + q.j;
  return q;
}

// ascending bitonic sort of n (<= COLL_MAXC) keys with their indices, by
// one wavefront
__device__ void wave_sort(uint64_t *key, uint32_t *idx, int n) {
  int P = 1;
  while (P < n) P <<= 1;
  for (int t = n + (int)threadIdx.x; t < P; t += COLL_BLOCK) {
    key[t] = ~0ull;
    idx[t] = 0xFFFFFFFFu;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < P; t += COLL_BLOCK) {
        int l = t ^ j;
        if (l > t) {
          bool asc = (t & k) == 0;
          uint64_t a = key[t], b = key[l];
          // (key, index) order: padding (key ~0, index ~0) stays behind
          // every real entry
          bool gt = a > b || (a == b && idx[t] > idx[l]);
          if (asc ? gt : !gt) {
            key[t] = b; key[l] = a;
            uint32_t x = idx[t]; idx[t] = idx[l]; idx[l] = x;
          }
        }
      }
      __syncthreads();
    }
}

struct CellSums {
  double amount, energy, m[3];
};

// sequential sums in bucket order (collect_weighted_params_tot_grid)
__device__ CellSums cell_sums(const double *vx, const double *vy, const double *vz, const double *w, int n,
                              double mass) {
  CellSums t{0, 0, {0, 0, 0}};
  for (int k = 0; k < n; ++k) {
    double vr = vx[k], vphi = vy[k], vzz = vz[k];
    double v_sq = vr * vr + vphi * vphi + vzz * vzz;
    double weighted_m = w[k] * mass;
    t.m[0] += weighted_m * vr;
    t.m[1] += weighted_m * vphi;
    t.m[2] += weighted_m * vzz;
    t.amount += w[k];
    t.energy += weighted_m * v_sq / 2;
  }
  return t;
}

// per-cell values kept between the two kernels
constexpr int TOT = 16;   // a_e E_e m_e[3] a_i E_i m_i[3] dens_e dens_i debye dens_e^(2/3) dens_i^(2/3)

// a cell's particles staged in LDS (FP64 velocities), up to MAXC per species
template <int MAXC>
struct CellStageT {
  double vx[2][MAXC], vy[2][MAXC], vz[2][MAXC], w[2][MAXC];
  uint32_t idx[2][MAXC];
};
using CellStage = CellStageT<COLL_MAXC>;

// cells are processed by size class: at most 16, 64 or COLL_MAXC particles
// of either species (smaller stages leave room for more wavefronts per CU)
constexpr int COLL_CLASSES = 4;
// class 0: one lane per cell; 1, 2: a wavefront per cell, LDS stage; 3
// (crowded): a workgroup per cell, stage in global memory
constexpr int COLL_CLASS_MAXC[COLL_CLASSES] = {16, 64, COLL_MAXC, INT_MAX};
// particles of all crowded cells of one pass that the global stage holds;
// beyond it a pass raises ERR_COLL_OVERFLOW
constexpr long COLL_CROWD_CAP = 1L << 22;
constexpr long COLL_CROWD_CELLS = COLL_CROWD_CAP / (COLL_MAXC + 1) + 1;
constexpr int COLL_CROWD_BLOCK = 256;

template <class V, int MAXC>
__device__ __forceinline__ void stage_species(const V &v, const uint32_t *perm, uint32_t o, int n, int s,
                                              CellStageT<MAXC> &S) {
  for (int k = threadIdx.x; k < n; k += COLL_BLOCK) {
    uint32_t p = perm[o + k];
    double vv[3];
    v.load(p, vv);
    S.idx[s][k] = p;
    S.vx[s][k] = vv[0]; S.vy[s][k] = vv[1]; S.vz[s][k] = vv[2];
    S.w[s][k] = v.w(p);
  }
}

// shuffle of one bucket: order by (Philox key, low id bits)
template <class V>
__device__ __forceinline__ void shuffle_bucket(const V &v, uint32_t *perm, uint32_t o, int n, uint64_t seed,
                                               uint64_t step, uint64_t *skey, uint32_t *sidx) {
  for (int k = threadIdx.x; k < n; k += COLL_BLOCK) {
    uint32_t p = perm[o + k];
    uint64_t id = v.id(p);
    uint32_t ph[4];
    philox4x32(id, (uint32_t)(2 * step), seed, ph);
    skey[k] = ((uint64_t)ph[0] << 32) | (id & 0xFFFFFFFFull);
    sidx[k] = p;
  }
  __syncthreads();
  wave_sort(skey, sidx, n);
  for (int k = threadIdx.x; k < n; k += COLL_BLOCK) perm[o + k] = sidx[k];
  __syncthreads();
}

// Cells with particles of both species go to the list of their size class;
// a cell missing a species has zero density and temperature, which ends
// the rest of its r-row (collisionsP12.cpp:351-354).
__global__ void k_coll_classify(Geom g, const uint32_t *off_e, const uint32_t *off_i, const int *mask,
                                uint32_t *lists, uint32_t *counts, uint32_t *crowd_off, int *row_bad, StepState *st,
                                const unsigned long long *gate) {
  if (gate_closed(gate)) return;
  const long ncells = (long)g.Dr * g.Dz * g.nr * g.nz;
  const long stride = (long)gridDim.x * blockDim.x;
  int err = 0;
  for (long c0 = (long)blockIdx.x * blockDim.x; c0 < ncells; c0 += stride) {
    const long c = c0 + threadIdx.x;
    int cls = -1;
    uint32_t crowd_at = 0;
    if (c < ncells && (!mask || mask[c])) {
      const int ne = (int)(off_e[c + 1] - off_e[c]), ni = (int)(off_i[c + 1] - off_i[c]);
      if (ne == 0 || ni == 0) {
        const CellPos q = cell_pos(g, c);
        atomicMin(&row_bad[(long)q.d * g.nr + q.i], q.j);
      } else {
        const int m = max(ne, ni);
        cls = 0;
        while (m > COLL_CLASS_MAXC[cls]) ++cls;
        if (cls == COLL_CLASSES - 1) {
          // a range of the crowded stage (counts[COLL_CLASSES] particles used)
          const uint32_t need = (uint32_t)(ne + ni);
          crowd_at = atomicAdd(&counts[COLL_CLASSES], need);
          if ((long)crowd_at + need > COLL_CROWD_CAP) {
            err |= ERR_COLL_OVERFLOW;
            cls = -1;
          }
        }
      }
    }
    // one counter atomic per class present in the wavefront
    const bool a = cls >= 0;
    const uint32_t at = wave_reserve(counts, a ? (uint32_t)cls : 0u, a);
    if (a) lists[(long)cls * ncells + at] = (uint32_t)c;
    if (cls == COLL_CLASSES - 1) crowd_off[at] = crowd_at;
  }
  if (err) atomicOr(&st->err, err);
}

// Collisions::collect_weighted_params_tot_grid's per-cell values from the
// cell's sums in shuffled order: density, temperature, Debye length, and
// the row break of a cell whose values are not normal
__device__ __forceinline__ void cell_stats(const Geom &g, long c, const CellSums &te, const CellSums &ti,
                                           double sum_el, double sum_ion, double *tot, int *row_bad) {
  const CellPos q = cell_pos(g, c);
  const double inv_e = 1. / EL_CHARGE;
  double temperature_el = te.energy / te.amount * inv_e;
  double temperature_ion = ti.energy / ti.amount * inv_e;
  // Collisions::geometry_cell_volume (collisions.cpp:162-170) passes
  // `i+shift` to CELL_VOLUME, which expands to 2.0 * i + shift + 1
  // (quirk kept: the macro argument is not parenthesised)
  const double shift = (double)((long)(q.d / g.Dz) * g.nr);
  const int i = q.i;
  double cell_volume = PCP_CELL_VOLUME(i+shift, g.dr, g.dz);
  double density_el = sum_el / cell_volume;
  double density_ion = sum_ion / cell_volume;
  double *t = tot + c * TOT;
  t[0] = te.amount; t[1] = te.energy; t[2] = te.m[0]; t[3] = te.m[1]; t[4] = te.m[2];
  t[5] = ti.amount; t[6] = ti.energy; t[7] = ti.m[0]; t[8] = ti.m[1]; t[9] = ti.m[2];
  t[10] = density_el; t[11] = density_ion;
  bool bad = !isnormal(temperature_ion) || !isnormal(temperature_el) || !isnormal(density_ion) ||
             !isnormal(density_el);
  t[12] = bad ? 0.0 : debye_length(density_el, density_ion, temperature_el, temperature_ion);
  // P12's s12 cap (collisionsP12.cpp) takes n^(2/3) of both species: once per cell here
  t[13] = pow(density_el, 2. / 3.);
  t[14] = pow(density_ion, 2. / 3.);
  if (bad) atomicMin(&row_bad[(long)q.d * g.nr + q.i], q.j);
}

template <class V, int MAXC>
__global__ void __launch_bounds__(COLL_BLOCK) k_coll_prep(Geom g, V ve, V vi, const uint32_t *off_e,
                                                          const uint32_t *off_i, uint32_t *perm_e, uint32_t *perm_i,
                                                          double m_e, double m_i, uint64_t seed, PassNo pass,
                                                          double *tot, int *row_bad, const uint32_t *list,
                                                          const uint32_t *count, StepState *st,
                                                          const unsigned long long *gate) {
  const uint64_t step = pass.get();
  __shared__ uint64_t skey[MAXC];
  __shared__ CellStageT<MAXC> S;
  if (gate_closed(gate)) return;
  const uint32_t nlist = *count;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const long c = list[li];
    int nn[2] = {(int)(off_e[c + 1] - off_e[c]), (int)(off_i[c + 1] - off_i[c])};
    shuffle_bucket(ve, perm_e, off_e[c], nn[0], seed, step, skey, S.idx[0]);
    shuffle_bucket(vi, perm_i, off_i[c], nn[1], seed, step, skey, S.idx[1]);
    stage_species(ve, perm_e, off_e[c], nn[0], 0, S);
    stage_species(vi, perm_i, off_i[c], nn[1], 1, S);
    __syncthreads();
    if (threadIdx.x == 0) {
      CellSums te = cell_sums(S.vx[0], S.vy[0], S.vz[0], S.w[0], nn[0], m_e);
      CellSums ti = cell_sums(S.vx[1], S.vy[1], S.vz[1], S.w[1], nn[1], m_i);
      double sum_el = 0, sum_ion = 0;
      for (int k = 0; k < nn[0]; ++k) sum_el += S.w[0][k];
      for (int k = 0; k < nn[1]; ++k) sum_ion += S.w[1][k];
      cell_stats(g, c, te, ti, sum_el, sum_ion, tot, row_bad);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Tiny cells (at most COLL_TINY particles of either species, e.g. the sparse
// rows near the axis of a plasma uniform in r^2): one lane per cell, so a
// wavefront works on 64 cells.  The shuffle sorts the keys in registers
// (odd-even transposition network, fully unrolled); the pairs run the same
// schedule as the wavefront kernels, one pair after another, on velocities
// staged in FP64 in a global per-lane array laid out [slot][lane] (adjacent
// lanes, adjacent addresses).  Pairs of one schedule phase touch disjoint
// particles and draw by ordinal, so the results equal the wavefront path's.
// ---------------------------------------------------------------------------
constexpr int COLL_TINY = 16;
constexpr int COLL_LANE_BLOCK = 256;
constexpr int COLL_LANE_SLOTS = 4 * 2 * COLL_TINY;   // vx, vy, vz, w of both species
static_assert(COLL_CLASS_MAXC[0] == COLL_TINY, "class 0 is the lane-per-cell class");

template <class V>
__device__ __forceinline__ void lane_shuffle(const V &v, uint32_t *perm, uint32_t o, int n, uint64_t seed,
                                             uint64_t step, uint32_t idx[COLL_TINY]) {
  uint64_t key[COLL_TINY];
#pragma unroll
  for (int k = 0; k < COLL_TINY; ++k) {
    key[k] = ~0ull;
    idx[k] = 0xFFFFFFFFu;
    if (k < n) {
      const uint32_t p = perm[o + k];
      const uint64_t id = v.id(p);
      uint32_t ph[4];
      philox4x32(id, (uint32_t)(2 * step), seed, ph);
      key[k] = ((uint64_t)ph[0] << 32) | (id & 0xFFFFFFFFull);
      idx[k] = p;
    }
  }
  // (key, index) ascending, padding last: the order wave_sort gives
#pragma unroll
  for (int r = 0; r < COLL_TINY; ++r)
#pragma unroll
    for (int k = r & 1; k + 1 < COLL_TINY; k += 2) {
      const bool gt = key[k] > key[k + 1] || (key[k] == key[k + 1] && idx[k] > idx[k + 1]);
      if (gt) {
        const uint64_t tk = key[k]; key[k] = key[k + 1]; key[k + 1] = tk;
        const uint32_t ti = idx[k]; idx[k] = idx[k + 1]; idx[k + 1] = ti;
      }
    }
#pragma unroll
  for (int k = 0; k < COLL_TINY; ++k)
    if (k < n) perm[o + k] = idx[k];
}

// cell_sums of the shuffled particles, and their weight sum
template <class V>
__device__ __forceinline__ CellSums lane_sums(const V &v, const uint32_t idx[COLL_TINY], int n, double mass,
                                              double &sum_w) {
  CellSums t{0, 0, {0, 0, 0}};
  sum_w = 0;
#pragma unroll
  for (int k = 0; k < COLL_TINY; ++k) {
    if (k >= n) break;
    double vv[3];
    v.load(idx[k], vv);
    const double w = v.w(idx[k]);
    double vr = vv[0], vphi = vv[1], vzz = vv[2];
    double v_sq = vr * vr + vphi * vphi + vzz * vzz;
    double weighted_m = w * mass;
    t.m[0] += weighted_m * vr;
    t.m[1] += weighted_m * vphi;
    t.m[2] += weighted_m * vzz;
    t.amount += w;
    t.energy += weighted_m * v_sq / 2;
    sum_w += w;
  }
  return t;
}

template <class V>
__global__ void __launch_bounds__(COLL_LANE_BLOCK) k_coll_prep_lane(Geom g, V ve, V vi, const uint32_t *off_e,
                                                                    const uint32_t *off_i, uint32_t *perm_e,
                                                                    uint32_t *perm_i, double m_e, double m_i,
                                                                    uint64_t seed, PassNo pass, double *tot,
                                                                    int *row_bad, const uint32_t *list,
                                                                    const uint32_t *count,
                                                                    const unsigned long long *gate) {
  const uint64_t step = pass.get();
  if (gate_closed(gate)) return;
  const uint32_t nlist = *count;
  for (uint32_t li = blockIdx.x * blockDim.x + threadIdx.x; li < nlist; li += gridDim.x * blockDim.x) {
    const long c = list[li];
    const int ne = (int)(off_e[c + 1] - off_e[c]), ni = (int)(off_i[c + 1] - off_i[c]);
    uint32_t ie[COLL_TINY], ii[COLL_TINY];
    lane_shuffle(ve, perm_e, off_e[c], ne, seed, step, ie);
    lane_shuffle(vi, perm_i, off_i[c], ni, seed, step, ii);
    double sum_el, sum_ion;
    const CellSums te = lane_sums(ve, ie, ne, m_e, sum_el);
    const CellSums ti = lane_sums(vi, ii, ni, m_i, sum_ion);
    cell_stats(g, c, te, ti, sum_el, sum_ion, tot, row_bad);
  }
}

// who runs a cell's loops: a wavefront (lanes split the independent pairs
// of a phase) or a single lane (everything in order)
template <int N>
struct GroupTeam {
  static constexpr int step = N;
  __device__ __forceinline__ int first() const { return (int)threadIdx.x; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};
using WaveTeam = GroupTeam<COLL_BLOCK>;
using BlockTeam = GroupTeam<COLL_CROWD_BLOCK>;
struct LaneTeam {
  static constexpr int step = 1;
  __device__ __forceinline__ int first() const { return 0; }
  __device__ __forceinline__ void sync() const {}
};

// a cell's staged FP64 particles: the wavefront's LDS stage, or a lane's
// slots of the global array
template <int MAXC>
struct WaveStage {
  CellStageT<MAXC> *S;
  __device__ __forceinline__ double &vx(int s, int k) const { return S->vx[s][k]; }
  __device__ __forceinline__ double &vy(int s, int k) const { return S->vy[s][k]; }
  __device__ __forceinline__ double &vz(int s, int k) const { return S->vz[s][k]; }
  __device__ __forceinline__ double &w(int s, int k) const { return S->w[s][k]; }
};
struct LaneStage {
  double *b;     // this lane's slot 0
  long stride;   // lanes of the launch
  __device__ __forceinline__ double &at(int c, int s, int k) const {
    return b[(long)((c * 2 + s) * COLL_TINY + k) * stride];
  }
  __device__ __forceinline__ double &vx(int s, int k) const { return at(0, s, k); }
  __device__ __forceinline__ double &vy(int s, int k) const { return at(1, s, k); }
  __device__ __forceinline__ double &vz(int s, int k) const { return at(2, s, k); }
  __device__ __forceinline__ double &w(int s, int k) const { return at(3, s, k); }
};
struct CrowdStage {
  double *b;     // the cell's range of the crowded stage
  int n0;        // electrons first, then ions
  __device__ __forceinline__ double &at(int c, int s, int k) const { return b[c * COLL_CROWD_CAP + s * n0 + k]; }
  __device__ __forceinline__ double &vx(int s, int k) const { return at(0, s, k); }
  __device__ __forceinline__ double &vy(int s, int k) const { return at(1, s, k); }
  __device__ __forceinline__ double &vz(int s, int k) const { return at(2, s, k); }
  __device__ __forceinline__ double &w(int s, int k) const { return at(3, s, k); }
};

// the pair schedule of one cell (collisionsP12.cpp:356-528)
template <int SCHEME, class Stage, class Team>
struct PairRunner {
  Stage S;
  Team T;
  double m[2], q[2];   // electrons, ions
  double dens[2], debye, dt;
  uint64_t seed, step, gcell;
  int *err;
  double dens23[2];   // n^(2/3) of electrons and ions
  __device__ __forceinline__ void pair(int sa, int ka, int sb, int kb, uint32_t ord) {
    PairDraws R;
    uint64_t ctr = (gcell << 24) | (uint64_t)(ord & 0xFFFFFFu);
    uint32_t draw = (uint32_t)(2 * step + 1);
    uniform2(ctr, draw, seed, R.u[0], R.u[1]);
    uniform2(ctr ^ (1ull << 63), draw, seed, R.u[2], R.u[3]);
    R.pos = 0;
    double va[3] = {S.vx(sa, ka), S.vy(sa, ka), S.vz(sa, ka)};
    double vb[3] = {S.vx(sb, kb), S.vy(sb, kb), S.vz(sb, kb)};
    collide_pair<SCHEME>(m[sa], m[sb], q[sa], q[sb], va, S.w(sa, ka), vb, S.w(sb, kb), dens[sa], dens[sb],
                         dens23[sa], dens23[sb], debye, dt, R, err);
    S.vx(sa, ka) = va[0]; S.vy(sa, ka) = va[1]; S.vz(sa, ka) = va[2];
    S.vx(sb, kb) = vb[0]; S.vy(sb, kb) = vb[1]; S.vz(sb, kb) = vb[2];
  }
  // like-particle pairs of species s (cases 1a / 1b), ordinals from qb
  __device__ void like(int s, int n, uint32_t qb) {
    if (n % 2 == 0) {
      for (int k = T.first(); k < n / 2; k += Team::step) pair(s, 2 * k, s, 2 * k + 1, qb + (uint32_t)k);
    } else {
      if (n >= 3 && T.first() == 0) {
        pair(s, 0, s, 1, qb);
        pair(s, 1, s, 2, qb + 1);
        pair(s, 2, s, 0, qb + 2);
      }
      if (n >= 5)
        for (int k = T.first(); k < (n - 3) / 2; k += Team::step)
          pair(s, 3 + 2 * k, s, 4 + 2 * k, qb + 3 + (uint32_t)k);
    }
  }
};

__device__ __forceinline__ uint32_t like_pairs(int n) {
  if (n % 2 == 0) return (uint32_t)(n / 2);
  return (uint32_t)((n >= 3 ? 3 : 0) + (n >= 5 ? (n - 3) / 2 : 0));
}

__device__ __forceinline__ int grp(unsigned f, float div) { return (int)floorf((float)f / div); }

// cases 1a/1b then 2a/2b of one cell
template <class Runner>
__device__ void cell_pairs(Runner &P, int nel, int nion) {
  const auto &T = P.T;
  constexpr int STEP = decltype(P.T)::step;
  // cases 1a/1b in reference order: ions even, electrons even, ions odd, electrons odd
  const uint32_t n_ion = like_pairs(nion), n_el = like_pairs(nel);
  // blocks in reference order: ion pairs if nion is even, electron pairs
  // if nel is even, ion pairs if nion is odd, electron pairs if nel is odd
  uint32_t q_ion = 0, q_el = n_ion;
  if (nion % 2 != 0 && nel % 2 == 0) { q_el = 0; q_ion = n_el; }
  P.like(1, nion, q_ion);
  T.sync();
  P.like(0, nel, q_el);
  T.sync();
  const uint32_t qc = n_ion + n_el;
  // case 2a / 2b
  if (nion == nel) {
    for (int k = T.first(); k < nel; k += STEP) P.pair(0, k, 1, k, qc + (uint32_t)k);
  } else if (nion > nel && nel > 0) {
    unsigned int c_i = (unsigned int)floorf((float)nion / (float)nel);
    double c_r = (double)((float)nion / (float)nel - (float)c_i);
    unsigned int ions_1st = (unsigned int)((c_i + 1) * c_r * nel);
    unsigned int els_1st = (unsigned int)(c_r * nel);
    unsigned int ions_2nd = (unsigned int)(c_i * (1 - c_r) * nel);
    const float d1 = (float)(c_i + 1), d2 = (float)c_i;
    // 1st group: electron e meets ions fgi with floor(fgi / (c_i + 1)) == e
    for (int e = T.first(); e < nel; e += STEP) {
      unsigned lo = (unsigned)e * (c_i + 1);
      lo = lo > 0 ? lo - 1 : 0;
      for (unsigned fgi = lo; fgi < ions_1st && fgi <= (unsigned)(e + 1) * (c_i + 1); ++fgi)
        if (grp(fgi, d1) == e && (int)fgi < nion) P.pair(0, e, 1, (int)fgi, qc + fgi);
    }
    T.sync();
    for (int e = T.first(); e < nel; e += STEP) {
      if (e < (int)els_1st) continue;
      int le = e - (int)els_1st;
      unsigned lo = (unsigned)le * c_i;
      lo = lo > 0 ? lo - 1 : 0;
      for (unsigned fgi = lo; fgi < ions_2nd && fgi <= (unsigned)(le + 1) * c_i; ++fgi)
        if (grp(fgi, d2) == le && (int)(fgi + ions_1st) < nion)
          P.pair(0, e, 1, (int)(fgi + ions_1st), qc + ions_1st + fgi);
    }
  } else if (nion < nel && nion > 0) {
    double c_i = floorf((float)nel / (float)nion);
    double c_r = (double)((float)nel / (float)nion) - c_i;
    unsigned int els_1st = (unsigned int)((c_i + 1) * c_r * nion);
    unsigned int ions_1st = (unsigned int)(c_r * nion);
    unsigned int els_2nd = (unsigned int)(c_i * (1 - c_r) * nion);
    const float d1 = (float)(c_i + 1), d2 = (float)c_i;
    const unsigned ci1 = (unsigned)c_i + 1, ci0 = (unsigned)c_i;
    for (int i = T.first(); i < nion; i += STEP) {
      unsigned lo = (unsigned)i * ci1;
      lo = lo > 0 ? lo - 1 : 0;
      for (unsigned fge = lo; fge < els_1st && fge <= (unsigned)(i + 1) * ci1; ++fge)
        if (grp(fge, d1) == i && (int)fge < nel) P.pair(0, (int)fge, 1, i, qc + fge);
    }
    T.sync();
    for (int i = T.first(); i < nion; i += STEP) {
      if (i < (int)ions_1st) continue;
      int li = i - (int)ions_1st;
      unsigned lo = (unsigned)li * ci0;
      lo = lo > 0 ? lo - 1 : 0;
      for (unsigned fge = lo; fge < els_2nd && fge <= (unsigned)(li + 1) * ci0; ++fge)
        if (grp(fge, d2) == li && (int)(fge + els_1st) < nel)
          P.pair(0, (int)(fge + els_1st), 1, i, qc + els_1st + fge);
    }
  }
  T.sync();
}

// TA77S correct_velocities for species s of the staged cell; t = the cell's
// {amount, energy, moment[3]} from before the collisions
struct CorrParams {
  double ok, alpha, V0[3], dV[3];
};
template <class Stage>
__device__ CorrParams correct_params(const Stage &S, int s, int n, double mass, const double *t) {
  CorrParams P;
  double m_new[3] = {0, 0, 0}, E_new = 0;
  for (int k = 0; k < n; ++k) {
    double vr = S.vx(s, k), vphi = S.vy(s, k), vz = S.vz(s, k);
    double v_sq = vr * vr + vphi * vphi + vz * vz;
    double mass_m = mass * S.w(s, k);
    m_new[0] += mass_m * vr;
    m_new[1] += mass_m * vphi;
    m_new[2] += mass_m * vz;
    E_new += mass_m * v_sq / 2;
  }
  const double amount = t[0], energy = t[1];
  double mass_tot = amount * mass;
  double dV[3], V0[3];
  for (int k = 0; k < 3; ++k) dV[k] = (m_new[k] - t[2 + k]) / mass_tot;
  for (int k = 0; k < 3; ++k) V0[k] = t[2 + k] / mass_tot;
  double delta_E = E_new - energy;
  double V0_sq = V0[0] * V0[0] + V0[1] * V0[1] + V0[2] * V0[2];
  double alpha = (energy - mass_tot * V0_sq / 2) /
                 (energy + delta_E -
                  mass_tot * (pow(V0[0] + dV[0], 2) + pow(V0[1] + dV[1], 2) + pow(V0[2] + dV[2], 2)) / 2);
  bool ok = isnormal(alpha);
  if (ok) {
    if (alpha < 0) alpha = -alpha;
    if (alpha == 0) alpha = 1;
    alpha = sqrt(alpha);
  }
  P.ok = ok ? 1.0 : 0.0;
  P.alpha = alpha;
  for (int k = 0; k < 3; ++k) { P.V0[k] = V0[k]; P.dV[k] = dV[k]; }
  return P;
}

template <class Stage, class Team>
__device__ void correct_species(const Stage &S, const Team &T, int s, int n, double mass, const double *t) {
  CorrParams P;
  if constexpr (Team::step == 1) {
    P = correct_params(S, s, n, mass, t);
  } else {
    __shared__ CorrParams shared_p;
    if (threadIdx.x == 0) shared_p = correct_params(S, s, n, mass, t);
    __syncthreads();
    P = shared_p;
  }
  if (P.ok != 0.0) {
    const double alpha = P.alpha;
    for (int k = T.first(); k < n; k += Team::step) {
      double vr = S.vx(s, k), vphi = S.vy(s, k), vz = S.vz(s, k);
      S.vx(s, k) = P.V0[0] + alpha * (vr - P.V0[0] - P.dV[0]);
      S.vy(s, k) = P.V0[1] + alpha * (vphi - P.V0[1] - P.dV[1]);
      S.vz(s, k) = P.V0[2] + alpha * (vz - P.V0[2] - P.dV[2]);
    }
  }
  T.sync();
}

template <int SCHEME, class V, int MAXC>
__global__ void __launch_bounds__(COLL_BLOCK) k_coll_pairs(Geom g, V ve, V vi, const uint32_t *off_e,
                                                           const uint32_t *off_i, const uint32_t *perm_e,
                                                           const uint32_t *perm_i, double m_e, double m_i,
                                                           double q_e, double q_i, uint64_t seed, PassNo pass,
                                                           const double *tot, const int *row_bad,
                                                           const uint32_t *list, const uint32_t *count,
                                                           StepState *st, const unsigned long long *gate) {
  const uint64_t step = pass.get();
  __shared__ CellStageT<MAXC> S;
  if (gate_closed(gate)) return;
  const uint32_t nlist = *count;
  int err = 0;
  const WaveStage<MAXC> WS{&S};
  const WaveTeam T;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const long c = list[li];
    int nn[2] = {(int)(off_e[c + 1] - off_e[c]), (int)(off_i[c + 1] - off_i[c])};
    stage_species(ve, perm_e, off_e[c], nn[0], 0, S);
    stage_species(vi, perm_i, off_i[c], nn[1], 1, S);
    __syncthreads();
    const CellPos q = cell_pos(g, c);
    const double *t = tot + c * TOT;
    const bool active = q.j < row_bad[(long)q.d * g.nr + q.i];
    if (active) {
      PairRunner<SCHEME, WaveStage<MAXC>, WaveTeam> P{WS, T, {m_e, m_i}, {q_e, q_i}, {t[10], t[11]}, t[12], g.dt,
                                                      seed, step, (uint64_t)q.cr * (uint64_t)g.Nz + (uint64_t)q.cz,
                                                      &err, {t[13], t[14]}};
      cell_pairs(P, nn[0], nn[1]);
    }
    if (SCHEME == COLL_TA77S) {
      // correct_velocities runs over every cell, ions first
      correct_species(WS, T, 1, nn[1], m_i, t + 5);
      correct_species(WS, T, 0, nn[0], m_e, t + 0);
    }
    for (int k = threadIdx.x; k < nn[0]; k += COLL_BLOCK) {
      double v3[3] = {S.vx[0][k], S.vy[0][k], S.vz[0][k]};
      ve.store(S.idx[0][k], v3);
    }
    for (int k = threadIdx.x; k < nn[1]; k += COLL_BLOCK) {
      double v3[3] = {S.vx[1][k], S.vy[1][k], S.vz[1][k]};
      vi.store(S.idx[1][k], v3);
    }
    __syncthreads();
  }
  if (err && threadIdx.x == 0) atomicOr(&st->err, err);
}

// the pairs of tiny cells, one lane per cell; `stage` holds
// COLL_LANE_SLOTS doubles per lane of the launch
template <int SCHEME, class V>
__global__ void __launch_bounds__(COLL_LANE_BLOCK) k_coll_pairs_lane(
    Geom g, V ve, V vi, const uint32_t *off_e, const uint32_t *off_i, const uint32_t *perm_e, const uint32_t *perm_i,
    double m_e, double m_i, double q_e, double q_i, uint64_t seed, PassNo pass, const double *tot,
    const int *row_bad, const uint32_t *list, const uint32_t *count, double *stage, StepState *st,
    const unsigned long long *gate) {
  const uint64_t step = pass.get();
  if (gate_closed(gate)) return;
  const uint32_t nlist = *count;
  const long lanes = (long)gridDim.x * blockDim.x;
  const long lane = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const LaneStage S{stage + lane, lanes};
  const LaneTeam T;
  int err = 0;
  for (long li = lane; li < (long)nlist; li += lanes) {
    const long c = list[li];
    const int nn[2] = {(int)(off_e[c + 1] - off_e[c]), (int)(off_i[c + 1] - off_i[c])};
    const uint32_t o[2] = {off_e[c], off_i[c]};
    for (int k = 0; k < nn[0]; ++k) {
      const uint32_t p = perm_e[o[0] + k];
      double vv[3];
      ve.load(p, vv);
      S.vx(0, k) = vv[0]; S.vy(0, k) = vv[1]; S.vz(0, k) = vv[2];
      S.w(0, k) = ve.w(p);
    }
    for (int k = 0; k < nn[1]; ++k) {
      const uint32_t p = perm_i[o[1] + k];
      double vv[3];
      vi.load(p, vv);
      S.vx(1, k) = vv[0]; S.vy(1, k) = vv[1]; S.vz(1, k) = vv[2];
      S.w(1, k) = vi.w(p);
    }
    const CellPos q = cell_pos(g, c);
    const double *t = tot + c * TOT;
    if (q.j < row_bad[(long)q.d * g.nr + q.i]) {
      PairRunner<SCHEME, LaneStage, LaneTeam> P{S, T, {m_e, m_i}, {q_e, q_i}, {t[10], t[11]}, t[12], g.dt,
                                                seed, step, (uint64_t)q.cr * (uint64_t)g.Nz + (uint64_t)q.cz, &err,
                                                {t[13], t[14]}};
      cell_pairs(P, nn[0], nn[1]);
    }
    if (SCHEME == COLL_TA77S) {
      correct_species(S, T, 1, nn[1], m_i, t + 5);
      correct_species(S, T, 0, nn[0], m_e, t + 0);
    }
    for (int k = 0; k < nn[0]; ++k) {
      double v3[3] = {S.vx(0, k), S.vy(0, k), S.vz(0, k)};
      ve.store(perm_e[o[0] + k], v3);
    }
    for (int k = 0; k < nn[1]; ++k) {
      double v3[3] = {S.vx(1, k), S.vy(1, k), S.vz(1, k)};
      vi.store(perm_i[o[1] + k], v3);
    }
  }
  if (err) atomicOr(&st->err, err);
}

// ---------------------------------------------------------------------------
// Crowded cells (more than COLL_MAXC particles of a species): a workgroup per
// cell.  The shuffle sorts the cell's keys in the global key array with a
// bitonic network whose comparators all run ascending (the first merge step
// of each stage mirrors), so entries past n act as +inf and are skipped; the
// pairs run the same schedule on an FP64 stage in global memory.
// ---------------------------------------------------------------------------
template <class V>
__device__ void crowd_shuffle(const V &v, uint32_t *perm, uint32_t o, int n, uint64_t seed, uint64_t step,
                              uint64_t *key, uint32_t *idx) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const uint32_t p = perm[o + k];
    const uint64_t id = v.id(p);
    uint32_t ph[4];
    philox4x32(id, (uint32_t)(2 * step), seed, ph);
    key[k] = ((uint64_t)ph[0] << 32) | (id & 0xFFFFFFFFull);
    idx[k] = p;
  }
  __syncthreads();
  int P = 1;
  while (P < n) P <<= 1;
  for (int kk = 2; kk <= P; kk <<= 1)
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int l = j == (kk >> 1) ? (i ^ (kk - 1)) : (i ^ j);
        if (l > i && l < n) {
          const uint64_t a = key[i], b = key[l];
          if (a > b || (a == b && idx[i] > idx[l])) {
            key[i] = b;
            key[l] = a;
            const uint32_t x = idx[i];
            idx[i] = idx[l];
            idx[l] = x;
          }
        }
      }
      __syncthreads();
    }
  for (int k = threadIdx.x; k < n; k += blockDim.x) perm[o + k] = idx[k];
  __syncthreads();
}

template <class V>
__device__ __forceinline__ void crowd_stage_species(const V &v, const uint32_t *perm, uint32_t o, int n, int s,
                                                    const CrowdStage &S) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const uint32_t p = perm[o + k];
    double vv[3];
    v.load(p, vv);
    S.vx(s, k) = vv[0]; S.vy(s, k) = vv[1]; S.vz(s, k) = vv[2];
    S.w(s, k) = v.w(p);
  }
}

template <class V>
__global__ void __launch_bounds__(COLL_CROWD_BLOCK) k_coll_prep_crowd(
    Geom g, V ve, V vi, const uint32_t *off_e, const uint32_t *off_i, uint32_t *perm_e, uint32_t *perm_i, double m_e,
    double m_i, uint64_t seed, PassNo pass, double *tot, int *row_bad, const uint32_t *list, const uint32_t *count,
    const uint32_t *crowd_off, uint64_t *ckey, uint32_t *cidx, double *cstage, const unsigned long long *gate) {
  const uint64_t step = pass.get();
  if (gate_closed(gate)) return;
  const uint32_t nlist = *count;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const long c = list[li];
    const int nn[2] = {(int)(off_e[c + 1] - off_e[c]), (int)(off_i[c + 1] - off_i[c])};
    const uint32_t base = crowd_off[li];
    crowd_shuffle(ve, perm_e, off_e[c], nn[0], seed, step, ckey + base, cidx + base);
    crowd_shuffle(vi, perm_i, off_i[c], nn[1], seed, step, ckey + base + nn[0], cidx + base + nn[0]);
    const CrowdStage S{cstage + base, nn[0]};
    crowd_stage_species(ve, perm_e, off_e[c], nn[0], 0, S);
    crowd_stage_species(vi, perm_i, off_i[c], nn[1], 1, S);
    __syncthreads();
    if (threadIdx.x == 0) {
      // cell_sums in the shuffled order, from the stage
      CellSums t[2];
      double sum_w[2];
      const double mass[2] = {m_e, m_i};
      for (int sp = 0; sp < 2; ++sp) {
        t[sp] = CellSums{0, 0, {0, 0, 0}};
        sum_w[sp] = 0;
        for (int k = 0; k < nn[sp]; ++k) {
          double vr = S.vx(sp, k), vphi = S.vy(sp, k), vzz = S.vz(sp, k), w = S.w(sp, k);
          double v_sq = vr * vr + vphi * vphi + vzz * vzz;
          double weighted_m = w * mass[sp];
          t[sp].m[0] += weighted_m * vr;
          t[sp].m[1] += weighted_m * vphi;
          t[sp].m[2] += weighted_m * vzz;
          t[sp].amount += w;
          t[sp].energy += weighted_m * v_sq / 2;
        }
        for (int k = 0; k < nn[sp]; ++k) sum_w[sp] += S.w(sp, k);
      }
      cell_stats(g, c, t[0], t[1], sum_w[0], sum_w[1], tot, row_bad);
    }
    __syncthreads();
  }
}

template <int SCHEME, class V>
__global__ void __launch_bounds__(COLL_CROWD_BLOCK) k_coll_pairs_crowd(
    Geom g, V ve, V vi, const uint32_t *off_e, const uint32_t *off_i, const uint32_t *perm_e, const uint32_t *perm_i,
    double m_e, double m_i, double q_e, double q_i, uint64_t seed, PassNo pass, const double *tot,
    const int *row_bad, const uint32_t *list, const uint32_t *count, const uint32_t *crowd_off, double *cstage,
    StepState *st, const unsigned long long *gate) {
  const uint64_t step = pass.get();
  if (gate_closed(gate)) return;
  const uint32_t nlist = *count;
  const BlockTeam T;
  int err = 0;
  for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const long c = list[li];
    const int nn[2] = {(int)(off_e[c + 1] - off_e[c]), (int)(off_i[c + 1] - off_i[c])};
    const CrowdStage S{cstage + crowd_off[li], nn[0]};
    crowd_stage_species(ve, perm_e, off_e[c], nn[0], 0, S);
    crowd_stage_species(vi, perm_i, off_i[c], nn[1], 1, S);
    __syncthreads();
    const CellPos q = cell_pos(g, c);
    const double *t = tot + c * TOT;
    if (q.j < row_bad[(long)q.d * g.nr + q.i]) {
      PairRunner<SCHEME, CrowdStage, BlockTeam> P{S, T, {m_e, m_i}, {q_e, q_i}, {t[10], t[11]}, t[12], g.dt, seed,
                                                  step, (uint64_t)q.cr * (uint64_t)g.Nz + (uint64_t)q.cz, &err,
                                                  {t[13], t[14]}};
      cell_pairs(P, nn[0], nn[1]);
    }
    if (SCHEME == COLL_TA77S) {
      correct_species(S, T, 1, nn[1], m_i, t + 5);
      correct_species(S, T, 0, nn[0], m_e, t + 0);
    }
    for (int k = threadIdx.x; k < nn[0]; k += COLL_CROWD_BLOCK) {
      double v3[3] = {S.vx(0, k), S.vy(0, k), S.vz(0, k)};
      ve.store(perm_e[off_e[c] + k], v3);
    }
    for (int k = threadIdx.x; k < nn[1]; k += COLL_CROWD_BLOCK) {
      double v3[3] = {S.vx(1, k), S.vy(1, k), S.vz(1, k)};
      vi.store(perm_i[off_i[c] + k], v3);
    }
    __syncthreads();
  }
  if (err) atomicOr(&st->err, err);
}

// ---------------------------------------------------------------------------
// collision surrogate (no reference counterpart)
// ---------------------------------------------------------------------------
// per-cell moments, per species: count, sum w, sum w v (3), sum w v_z^2,
// sum w (v_r^2 + v_phi^2)
constexpr int SST = 16;

template <class V>
__device__ __forceinline__ void species_moments(const V &v, const uint32_t *perm, uint32_t o, int n, double *out) {
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int k = threadIdx.x; k < n; k += COLL_BLOCK) {
    const uint32_t p = perm[o + k];
    double vv[3];
    v.load(p, vv);
    const double w = v.w(p);
    a[0] += 1.0;
    a[1] += w;
    a[2] += w * vv[0];
    a[3] += w * vv[1];
    a[4] += w * vv[2];
    a[5] += w * vv[2] * vv[2];
    a[6] += w * (vv[0] * vv[0] + vv[1] * vv[1]);
  }
#pragma unroll
  for (int q = 0; q < 7; ++q) a[q] = wave_sum(a[q]);
  if (threadIdx.x == 0)
    for (int q = 0; q < 7; ++q) out[q] = a[q];
}

template <class V>
__global__ void __launch_bounds__(COLL_BLOCK) k_surr_moments(Geom g, V ve, V vi, const uint32_t *off_e,
                                                             const uint32_t *off_i, const uint32_t *perm_e,
                                                             const uint32_t *perm_i, double *mom) {
  const long ncells = (long)g.Dr * g.Dz * g.nr * g.nz;
  for (long c = blockIdx.x; c < ncells; c += gridDim.x) {
    species_moments(ve, perm_e, off_e[c], (int)(off_e[c + 1] - off_e[c]), mom + c * SST);
    species_moments(vi, perm_i, off_i[c], (int)(off_i[c + 1] - off_i[c]), mom + c * SST + 8);
  }
}

// temperatures [eV] of one species from its moments: T_par (z), T_perp
// (per degree of freedom of r, phi), and the mean velocity
__device__ __forceinline__ bool species_temps(const double *m, double mass, double &tpar, double &tperp, double V[3]) {
  const double sw = m[1];
  if (!(m[0] >= SURR_MIN_PARTICLES) || !(sw > 0)) return false;
  V[0] = m[2] / sw; V[1] = m[3] / sw; V[2] = m[4] / sw;
  tpar = mass * (m[5] / sw - V[2] * V[2]) / EL_CHARGE;
  tperp = mass * (m[6] / sw - V[0] * V[0] - V[1] * V[1]) / (2.0 * EL_CHARGE);
  return isnormal(tpar) && tpar > 0 && isnormal(tperp) && tperp > 0;
}

// Collision-rate scale of each target (relative changes of Te_par,
// Te_perp, Ti_par, Ti_perp over one pass): the network learns the relative
// change divided by dt times the species' collision frequency up to a
// constant -- electrons n_e / (Te_perp sqrt(Te_par)) (the isotropisation
// rate of a bi-Maxwellian), ions n_i / Ti^1.5 -- so that its outputs stay
// of order one across the density and temperature range.
__device__ __forceinline__ bool surr_rate_scale(const double *me, const double *mi, double m_e, double m_i,
                                                double vol, double dt, double sc[MLP_OUT]) {
  double tep, tet, tip, tit, V[3];
  if (!species_temps(me, m_e, tep, tet, V) || !species_temps(mi, m_i, tip, tit, V)) return false;
  const double t = dt * 1e10;
  const double se = (me[1] / vol) * 1e-20 * t / (tet * sqrt(tep));
  const double ti = (tip + 2.0 * tit) / 3.0;
  const double si = (mi[1] / vol) * 1e-20 * t / (ti * sqrt(ti));
  sc[0] = sc[1] = se;
  sc[2] = sc[3] = si;
  return isnormal(se) && isnormal(si);
}

// features: log10 n_e, log10 n_i, log10 Te_par, log10 Te_perp, log10 Ti,
// log10 dt; normalised by the fitted mean / std when `norm` is given.
// ok[c] = 1 for a cell the network may handle.
__global__ void k_surr_features(Geom g, const double *mom, double m_e, double m_i, SurrNorm norm, int normalise,
                                float *X, int *ok) {
  const long ncells = (long)g.Dr * g.Dz * g.nr * g.nz;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (long)gridDim.x * blockDim.x) {
    const double *me = mom + c * SST, *mi = me + 8;
    double tep, tet, tip, tit, Ve[3], Vi[3];
    bool good = species_temps(me, m_e, tep, tet, Ve) && species_temps(mi, m_i, tip, tit, Vi);
    const CellPos q = cell_pos(g, c);
    const double vol = PCP_CELL_VOLUME((double)q.cr, g.dr, g.dz);
    double f[MLP_IN] = {log10(me[1] / vol), log10(mi[1] / vol), log10(tep), log10(tet),
                        log10((tip + 2.0 * tit) / 3.0), log10(g.dt)};
    for (int k = 0; k < MLP_IN; ++k) {
      float x = (float)f[k];
      if (normalise) {
        x = (x - norm.mean[k]) / norm.std[k];
        if (!(fabsf(x) <= SURR_MAX_Z)) good = false;
      }
      X[c * MLP_XSTRIDE + k] = good ? x : 0.f;
    }
    X[c * MLP_XSTRIDE + 6] = 0.f;
    X[c * MLP_XSTRIDE + 7] = 0.f;
    ok[c] = good ? 1 : 0;
  }
}

// training rows of the cells that collided: relative changes of Te_par,
// Te_perp, Ti_par, Ti_perp over the pass, each over its rate scale
// a training row of cell c: valid when the cell collided (not after a row
// break) and its moments are normal before and after the pass
__device__ __forceinline__ bool surr_row(const Geom &g, long c, const double *before, const double *after, double m_e,
                                         double m_i, const int *ok, const int *row_bad, double rel[MLP_OUT]) {
  if (!ok[c]) return false;
  const CellPos q = cell_pos(g, c);
  if (q.j >= row_bad[(long)q.d * g.nr + q.i]) return false;
  double b[4], a[4], V[3], sc[MLP_OUT];
  if (!species_temps(before + c * SST, m_e, b[0], b[1], V) || !species_temps(before + c * SST + 8, m_i, b[2], b[3], V) ||
      !species_temps(after + c * SST, m_e, a[0], a[1], V) || !species_temps(after + c * SST + 8, m_i, a[2], a[3], V))
    return false;
  const double vol = PCP_CELL_VOLUME((double)q.cr, g.dr, g.dz);
  if (!surr_rate_scale(before + c * SST, before + c * SST + 8, m_e, m_i, vol, g.dt, sc)) return false;
  for (int k = 0; k < MLP_OUT; ++k) rel[k] = (a[k] - b[k]) / b[k] / sc[k];
  return true;
}

__global__ void k_surr_rowflag(Geom g, const double *before, const double *after, double m_e, double m_i,
                               const int *ok, const int *row_bad, uint32_t *flag) {
  const long ncells = (long)g.Dr * g.Dz * g.nr * g.nz;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c <= ncells; c += (long)gridDim.x * blockDim.x) {
    double rel[MLP_OUT];
    flag[c] = c < ncells && surr_row(g, c, before, after, m_e, m_i, ok, row_bad, rel) ? 1u : 0u;
  }
}

// training rows of the cells that collided, in cell order from row n0 on:
// relative changes of Te_par, Te_perp, Ti_par, Ti_perp over the pass, each
// over its rate scale
__global__ void k_surr_targets(Geom g, const double *before, const double *after, double m_e, double m_i,
                               const int *ok, const int *row_bad, const float *X, const uint32_t *offset, long n0,
                               float *DX, float *DT, long cap) {
  const long ncells = (long)g.Dr * g.Dz * g.nr * g.nz;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (long)gridDim.x * blockDim.x) {
    double rel[MLP_OUT];
    if (!surr_row(g, c, before, after, m_e, m_i, ok, row_bad, rel)) continue;
    const long slot = n0 + (long)offset[c];
    if (slot >= cap) continue;
    for (int k = 0; k < MLP_XSTRIDE; ++k) DX[slot * MLP_XSTRIDE + k] = X[c * MLP_XSTRIDE + k];
    for (int k = 0; k < MLP_OUT; ++k) DT[slot * MLP_YSTRIDE + k] = (float)rel[k];
  }
}

// The network's relative temperature changes applied per species as
// rescalings of the thermal velocity about the species' mean velocity (z
// by alpha_par, r and phi by alpha_perp), then one common factor on both
// species' thermal velocities restores the cell's total kinetic energy:
// each species' momentum and the cell's energy are conserved exactly (up to
// rounding).  Cells the network may not handle are flagged for the binary
// collision fallback.
template <class V>
__global__ void __launch_bounds__(COLL_BLOCK) k_surr_apply(Geom g, V ve, V vi, const uint32_t *off_e,
                                                           const uint32_t *off_i, const uint32_t *perm_e,
                                                           const uint32_t *perm_i, const double *mom, double m_e,
                                                           double m_i, const float *Y, SurrNorm norm, const int *ok,
                                                           int *fallback, unsigned long long *n_fallback) {
  __shared__ CellStage S;
  __shared__ double red[4];
  const long ncells = (long)g.Dr * g.Dz * g.nr * g.nz;
  for (long c = blockIdx.x; c < ncells; c += gridDim.x) {
    const int nn[2] = {(int)(off_e[c + 1] - off_e[c]), (int)(off_i[c + 1] - off_i[c])};
    const bool use = ok[c] && nn[0] <= COLL_MAXC && nn[1] <= COLL_MAXC;
    if (threadIdx.x == 0) {
      fallback[c] = use ? 0 : 1;
      if (!use && (nn[0] > 0 || nn[1] > 0)) atomicAdd(n_fallback, 1ull);
    }
    if (!use) continue;
    stage_species(ve, perm_e, off_e[c], nn[0], 0, S);
    stage_species(vi, perm_i, off_i[c], nn[1], 1, S);
    __syncthreads();
    double Vs[2][3], apar[2], aperp[2], sc[MLP_OUT];
    const double mass[2] = {m_e, m_i};
    const CellPos q = cell_pos(g, c);
    // ok[c] guarantees normal moments, hence a normal rate scale
    (void)surr_rate_scale(mom + c * SST, mom + c * SST + 8, m_e, m_i, PCP_CELL_VOLUME((double)q.cr, g.dr, g.dz), g.dt,
                          sc);
    for (int s = 0; s < 2; ++s) {
      const double *m = mom + c * SST + 8 * s;
      for (int k = 0; k < 3; ++k) Vs[s][k] = m[2 + k] / m[1];
      const double ypar = (double)Y[c * MLP_YSTRIDE + 2 * s] * norm.tscale[2 * s] * sc[2 * s];
      const double yperp = (double)Y[c * MLP_YSTRIDE + 2 * s + 1] * norm.tscale[2 * s + 1] * sc[2 * s + 1];
      apar[s] = sqrt(fmax(1.0 + ypar, SURR_MIN_FACTOR));
      aperp[s] = sqrt(fmax(1.0 + yperp, SURR_MIN_FACTOR));
    }
    // thermal energy before and after the per-species rescaling
    double e0 = 0, e1 = 0;
    for (int s = 0; s < 2; ++s)
      for (int k = threadIdx.x; k < nn[s]; k += COLL_BLOCK) {
        const double dx = S.vx[s][k] - Vs[s][0], dy = S.vy[s][k] - Vs[s][1], dz = S.vz[s][k] - Vs[s][2];
        const double wm = S.w[s][k] * mass[s];
        e0 += wm * (dx * dx + dy * dy + dz * dz);
        const double nx = aperp[s] * dx, ny = aperp[s] * dy, nz = apar[s] * dz;
        e1 += wm * (nx * nx + ny * ny + nz * nz);
      }
    e0 = wave_sum(e0);
    e1 = wave_sum(e1);
    if (threadIdx.x == 0) red[0] = e1 > 0 ? sqrt(e0 / e1) : 1.0;
    __syncthreads();
    const double beta = red[0];
    for (int s = 0; s < 2; ++s)
      for (int k = threadIdx.x; k < nn[s]; k += COLL_BLOCK) {
        double v3[3];
        v3[0] = Vs[s][0] + beta * (aperp[s] * (S.vx[s][k] - Vs[s][0]));
        v3[1] = Vs[s][1] + beta * (aperp[s] * (S.vy[s][k] - Vs[s][1]));
        v3[2] = Vs[s][2] + beta * (apar[s] * (S.vz[s][k] - Vs[s][2]));
        if (s == 0) ve.store(S.idx[0][k], v3);
        else vi.store(S.idx[1][k], v3);
      }
    __syncthreads();
  }
}

__global__ void k_fill_zero_ull(unsigned long long *p) { *p = 0ull; }

// ---------------------------------------------------------------------------
// Fused surrogate phase of an FP32 step.  The push kernel (k32_push_mom)
// leaves each global cell's moments of the pushed velocities in mom; the
// network runs on the features of every cell; k_fsurr_params turns its
// outputs into per-cell rescalings that the move kernel applies to each
// particle on the fly (no buckets, no staging).  The rescaling is the one of
// k_surr_apply, with the energy-restoring factor beta computed from the
// moments: for species s with thermal sums P_perp = sum w (v_r^2 + v_phi^2)
// - |sum w v_perp|^2 / sum w and P_par likewise,
//   beta^2 = sum_s m_s (P_perp + P_par) / sum_s m_s (a_perp^2 P_perp + a_par^2 P_par).
// Cells the network does not cover keep their velocities here and are
// flagged (by local cell key) for the P12 fallback.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void global_cell(const Geom &g, long c, long &cr, long &cz) {
  cr = (long)g.div_Nz.div((uint32_t)c);
  cz = c - cr * g.Nz;
}

__global__ void k_fsurr_features(Geom g, const double *mom, double m_e, double m_i, SurrNorm norm, float *X,
                                 int *ok) {
  const long ncells = g.Nr * g.Nz;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (long)gridDim.x * blockDim.x) {
    const double *me = mom + c * SURR_PAR_STRIDE, *mi = me + 8;
    double tep, tet, tip, tit, Ve[3], Vi[3];
    bool good = species_temps(me, m_e, tep, tet, Ve) && species_temps(mi, m_i, tip, tit, Vi);
    long cr, cz;
    global_cell(g, c, cr, cz);
    const double vol = PCP_CELL_VOLUME((double)cr, g.dr, g.dz);
    const double f[MLP_IN] = {log10(me[1] / vol), log10(mi[1] / vol), log10(tep), log10(tet),
                              log10((tip + 2.0 * tit) / 3.0), log10(g.dt)};
    for (int k = 0; k < MLP_IN; ++k) {
      float x = ((float)f[k] - norm.mean[k]) / norm.std[k];
      if (!(fabsf(x) <= SURR_MAX_Z)) good = false;
      X[c * MLP_XSTRIDE + k] = good ? x : 0.f;
    }
    X[c * MLP_XSTRIDE + 6] = 0.f;
    X[c * MLP_XSTRIDE + 7] = 0.f;
    ok[c] = good ? 1 : 0;
  }
}

// also writes the P12 fallback's bucket counts (local cell keys) from the
// moments' particle counts: the particles are still in the cells the push
// counted them in
__global__ void k_fsurr_params(Geom g, double *mom, double m_e, double m_i, const float *Y, SurrNorm norm,
                               const int *ok, float *par, int *fallback, unsigned long long *n_fallback,
                               uint32_t *cnt_e, uint32_t *cnt_i) {
  const long ncells = g.Nr * g.Nz;
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt_e[ncells] = cnt_i[ncells] = 0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (long)gridDim.x * blockDim.x) {
    double *m = mom + c * SURR_PAR_STRIDE;
    float *P = par + c * SURR_PAR_STRIDE;
    long cr, cz;
    global_cell(g, c, cr, cz);
    const uint32_t lk = local_cell_key(g, (int)cr, (int)cz);
    if (ok[c]) {
      double sc[MLP_OUT];
      (void)surr_rate_scale(m, m + 8, m_e, m_i, PCP_CELL_VOLUME((double)cr, g.dr, g.dz), g.dt, sc);
      const double mass[2] = {m_e, m_i};
      double V[2][3], ap[2], az[2], pp[2], pz[2];
      double e0 = 0, e1 = 0;
      for (int s = 0; s < 2; ++s) {
        const double *q = m + 8 * s;
        for (int k = 0; k < 3; ++k) V[s][k] = q[2 + k] / q[1];
        const double ypar = (double)Y[c * MLP_YSTRIDE + 2 * s] * norm.tscale[2 * s] * sc[2 * s];
        const double yperp = (double)Y[c * MLP_YSTRIDE + 2 * s + 1] * norm.tscale[2 * s + 1] * sc[2 * s + 1];
        az[s] = sqrt(fmax(1.0 + ypar, SURR_MIN_FACTOR));
        ap[s] = sqrt(fmax(1.0 + yperp, SURR_MIN_FACTOR));
        pp[s] = fmax(q[6] - (q[2] * q[2] + q[3] * q[3]) / q[1], 0.0);
        pz[s] = fmax(q[5] - q[4] * q[4] / q[1], 0.0);
        e0 += mass[s] * (pp[s] + pz[s]);
        e1 += mass[s] * (ap[s] * ap[s] * pp[s] + az[s] * az[s] * pz[s]);
      }
      const double beta = e1 > 0 ? sqrt(e0 / e1) : 1.0;
      for (int s = 0; s < 2; ++s) {
        P[8 * s + 0] = (float)V[s][0];
        P[8 * s + 1] = (float)V[s][1];
        P[8 * s + 2] = (float)V[s][2];
        P[8 * s + 3] = (float)(beta * ap[s]);
        P[8 * s + 4] = (float)(beta * az[s]);
      }
      fallback[lk] = 0;
      cnt_e[lk] = cnt_i[lk] = 0;
    } else {
      for (int s = 0; s < 2; ++s) {
        P[8 * s + 0] = P[8 * s + 1] = P[8 * s + 2] = 0.f;
        P[8 * s + 3] = P[8 * s + 4] = 1.f;
      }
      const bool any = m[0] > 0 || m[8] > 0;
      fallback[lk] = any ? 1 : 0;
      cnt_e[lk] = any ? (uint32_t)m[0] : 0u;
      cnt_i[lk] = any ? (uint32_t)m[8] : 0u;
    }
    // fallback cells counted with one atomic per wavefront
    const bool fb = !ok[c] && (m[0] > 0 || m[8] > 0);
    const unsigned long long bal = __ballot(fb);
    if (bal && __lane_id() == __ffsll((long long)__ballot(1)) - 1) atomicAdd(n_fallback, (unsigned long long)__popcll(bal));
    // the next step's push accumulates into zeroed moments
    for (int k = 0; k < SURR_PAR_STRIDE; ++k) m[k] = 0.0;
  }
}

__global__ void k_fill_zero_f64(double *p, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = 0.0;
}

}  // namespace

struct CollBuffers {
  uint32_t *cnt[2] = {nullptr, nullptr}, *off[2] = {nullptr, nullptr};
  uint32_t *fill[2] = {nullptr, nullptr}, *perm[2] = {nullptr, nullptr};
  uint32_t *lists = nullptr, *lcount = nullptr;   // cells per size class
  double *lane_stage = nullptr;                   // per-lane FP64 stage of the tiny-cell pairs
  uint32_t *crowd_off = nullptr;                  // crowded cells: range of each listed cell
  uint64_t *crowd_key = nullptr;                  // shuffle keys, indices and FP64 stage of
  uint32_t *crowd_idx = nullptr;                  // the crowded cells (COLL_CROWD_CAP particles)
  double *crowd_stage = nullptr;
  long lane_blocks = 0;

  long cap[2] = {0, 0};
  double *tot = nullptr;
  int *row_bad = nullptr;
  void *scan_tmp = nullptr;
  size_t scan_bytes = 0;
  long ncells = 0;
  // surrogate
  double *mom = nullptr, *mom2 = nullptr;
  float *feat = nullptr, *pred = nullptr;
  int *ok = nullptr, *fallback = nullptr;
  unsigned long long *counter = nullptr;   // [0] fallback cells
  // fused surrogate phase (per global cell)
  double *fmom = nullptr;
  float *fX = nullptr, *fY = nullptr, *fpar = nullptr;
  int *fok = nullptr;
  long fcells = 0;
};

void coll_free(pcp_ctx *c) {
  CollBuffers *b = c->coll;
  if (!b) return;
  for (int s = 0; s < 2; ++s) {
    (void)hipFree(b->cnt[s]); (void)hipFree(b->off[s]); (void)hipFree(b->fill[s]);
    (void)hipFree(b->perm[s]);
  }
  (void)hipFree(b->lists); (void)hipFree(b->lcount); (void)hipFree(b->lane_stage);
  (void)hipFree(b->crowd_off); (void)hipFree(b->crowd_key); (void)hipFree(b->crowd_idx); (void)hipFree(b->crowd_stage);
  (void)hipFree(b->tot); (void)hipFree(b->row_bad); (void)hipFree(b->scan_tmp);
  (void)hipFree(b->mom); (void)hipFree(b->mom2); (void)hipFree(b->feat); (void)hipFree(b->pred);
  (void)hipFree(b->ok); (void)hipFree(b->fallback); (void)hipFree(b->counter);
  (void)hipFree(b->fmom); (void)hipFree(b->fX); (void)hipFree(b->fY); (void)hipFree(b->fpar); (void)hipFree(b->fok);
  delete b;
  c->coll = nullptr;
}

static int coll_alloc(pcp_ctx *c, long cap_e, long cap_i) {
  CollBuffers *b = c->coll;
  const long ncells = (long)c->D * c->g.nr * c->g.nz;
  if (b && b->cap[0] >= cap_e && b->cap[1] >= cap_i) return 0;
  coll_free(c);
  b = new CollBuffers();
  c->coll = b;
  b->ncells = ncells;
  const long caps[2] = {std::max(cap_e, 1L), std::max(cap_i, 1L)};
  for (int s = 0; s < 2; ++s) {
    b->cap[s] = caps[s];
    if (hipMalloc(&b->cnt[s], (ncells + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&b->off[s], (ncells + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&b->fill[s], (ncells + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&b->perm[s], caps[s] * sizeof(uint32_t)) != hipSuccess)
      return -1;
  }
  // the lane-per-cell kernels: four workgroups per CU, COLL_LANE_SLOTS doubles per lane
  b->lane_blocks = (long)c->n_cu * 4;
  if (hipMalloc(&b->lists, (size_t)COLL_CLASSES * ncells * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&b->lcount, (COLL_CLASSES + 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&b->crowd_off, COLL_CROWD_CELLS * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&b->crowd_key, COLL_CROWD_CAP * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(&b->crowd_idx, COLL_CROWD_CAP * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&b->crowd_stage, 4 * COLL_CROWD_CAP * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->lane_stage, (size_t)COLL_LANE_SLOTS * b->lane_blocks * COLL_LANE_BLOCK * sizeof(double)) !=
          hipSuccess)
    return -1;
  if (hipMalloc(&b->tot, ncells * TOT * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->row_bad, (long)c->D * c->g.nr * sizeof(int)) != hipSuccess ||
      hipMalloc(&b->mom, ncells * SST * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->mom2, ncells * SST * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->feat, ncells * MLP_XSTRIDE * sizeof(float)) != hipSuccess ||
      hipMalloc(&b->pred, ncells * MLP_YSTRIDE * sizeof(float)) != hipSuccess ||
      hipMalloc(&b->ok, ncells * sizeof(int)) != hipSuccess ||
      hipMalloc(&b->fallback, ncells * sizeof(int)) != hipSuccess ||
      hipMalloc(&b->counter, 2 * sizeof(unsigned long long)) != hipSuccess)
    return -1;
  if (hipMemset(b->counter, 0, 2 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, b->scan_bytes, b->cnt[0], b->off[0], (int)(ncells + 1)) !=
      hipSuccess)
    return -1;
  if (hipMalloc(&b->scan_tmp, std::max<size_t>(b->scan_bytes, 16)) != hipSuccess) return -1;
  return 0;
}

// bucket lists of both species (sort_to_cells)
// `counted`: the bucket counts are already in cnt (the fused surrogate phase
// writes them from the push's per-cell particle counts)
template <class V>
static int coll_buckets(pcp_ctx *c, V ve, V vi, long n_e, long n_i, const int *mask = nullptr,
                        const unsigned long long *gate = nullptr, bool counted = false) {
  CollBuffers *b = c->coll;
  const Geom &g = c->g;
  const long ncells = b->ncells;
  const uint32_t sentinel = (uint32_t)ncells;
  const long nn[2] = {n_e, n_i};
  unsigned gz = (unsigned)std::min<long>((ncells + 256) / 256, (long)c->n_cu * 16);
  for (int s = 0; s < 2; ++s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(gz), dim3(256), 0, c->stream, b->fill[s], ncells + 1);
    unsigned gp = (unsigned)std::max(1L, std::min<long>((nn[s] + 255) / 256, (long)c->n_cu * 16));
    if (!counted) {
      hipLaunchKernelGGL(k_fill_u32, dim3(gz), dim3(256), 0, c->stream, b->cnt[s], ncells + 1);
      if (nn[s] > 0)
        hipLaunchKernelGGL(k_coll_count<V>, dim3(gp), dim3(256), 0, c->stream, g, s == 0 ? ve : vi, b->cnt[s],
                           sentinel, mask, gate);
    }
    size_t bytes = b->scan_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(b->scan_tmp, bytes, b->cnt[s], b->off[s], (int)(ncells + 1), c->stream) !=
        hipSuccess)
      return -1;
    if (nn[s] > 0)
      hipLaunchKernelGGL(k_coll_fill<V>, dim3(gp), dim3(256), 0, c->stream, g, s == 0 ? ve : vi, b->off[s],
                         b->fill[s], b->perm[s], sentinel, mask, gate);
  }
  return 0;
}

// the binary pass (shuffle + statistics, then pairs) on the cells of `mask`
// (all cells when null)
template <class V, int CLS>
static void coll_class(pcp_ctx *c, int scheme, V ve, V vi, double m_e, double m_i, double q_e, double q_i,
                       uint64_t seed, PassNo step, bool pairs, const unsigned long long *gate) {
  CollBuffers *b = c->coll;
  const Geom &g = c->g;
  constexpr int MAXC = COLL_CLASS_MAXC[CLS];
  const uint32_t *list = b->lists + (size_t)CLS * b->ncells, *count = b->lcount + CLS;
  const unsigned gc = (unsigned)std::min<long>(b->ncells, (long)c->n_cu * 32);
  if constexpr (CLS == 0) {
    const unsigned gl = (unsigned)b->lane_blocks;
    if (!pairs)
      hipLaunchKernelGGL((k_coll_prep_lane<V>), dim3(gl), dim3(COLL_LANE_BLOCK), 0, c->stream, g, ve, vi, b->off[0],
                         b->off[1], b->perm[0], b->perm[1], m_e, m_i, seed, step, b->tot, b->row_bad, list, count,
                         gate);
    else if (scheme == COLL_TA77S)
      hipLaunchKernelGGL((k_coll_pairs_lane<COLL_TA77S, V>), dim3(gl), dim3(COLL_LANE_BLOCK), 0, c->stream, g, ve,
                         vi, b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                         b->row_bad, list, count, b->lane_stage, c->st, gate);
    else if (scheme == COLL_SK98)
      hipLaunchKernelGGL((k_coll_pairs_lane<COLL_SK98, V>), dim3(gl), dim3(COLL_LANE_BLOCK), 0, c->stream, g, ve,
                         vi, b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                         b->row_bad, list, count, b->lane_stage, c->st, gate);
    else
      hipLaunchKernelGGL((k_coll_pairs_lane<COLL_P12, V>), dim3(gl), dim3(COLL_LANE_BLOCK), 0, c->stream, g, ve,
                         vi, b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                         b->row_bad, list, count, b->lane_stage, c->st, gate);
  } else if constexpr (CLS == COLL_CLASSES - 1) {
    const unsigned gw = (unsigned)(c->n_cu * 2);
    if (!pairs)
      hipLaunchKernelGGL((k_coll_prep_crowd<V>), dim3(gw), dim3(COLL_CROWD_BLOCK), 0, c->stream, g, ve, vi,
                         b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, seed, step, b->tot, b->row_bad,
                         list, count, b->crowd_off, b->crowd_key, b->crowd_idx, b->crowd_stage, gate);
    else if (scheme == COLL_TA77S)
      hipLaunchKernelGGL((k_coll_pairs_crowd<COLL_TA77S, V>), dim3(gw), dim3(COLL_CROWD_BLOCK), 0, c->stream, g, ve,
                         vi, b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                         b->row_bad, list, count, b->crowd_off, b->crowd_stage, c->st, gate);
    else if (scheme == COLL_SK98)
      hipLaunchKernelGGL((k_coll_pairs_crowd<COLL_SK98, V>), dim3(gw), dim3(COLL_CROWD_BLOCK), 0, c->stream, g, ve,
                         vi, b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                         b->row_bad, list, count, b->crowd_off, b->crowd_stage, c->st, gate);
    else
      hipLaunchKernelGGL((k_coll_pairs_crowd<COLL_P12, V>), dim3(gw), dim3(COLL_CROWD_BLOCK), 0, c->stream, g, ve,
                         vi, b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                         b->row_bad, list, count, b->crowd_off, b->crowd_stage, c->st, gate);
  } else {
  if (!pairs)
    hipLaunchKernelGGL((k_coll_prep<V, MAXC>), dim3(gc), dim3(COLL_BLOCK), 0, c->stream, g, ve, vi, b->off[0],
                       b->off[1], b->perm[0], b->perm[1], m_e, m_i, seed, step, b->tot, b->row_bad, list, count, c->st,
                       gate);
  else if (scheme == COLL_TA77S)
    hipLaunchKernelGGL((k_coll_pairs<COLL_TA77S, V, MAXC>), dim3(gc), dim3(COLL_BLOCK), 0, c->stream, g, ve, vi,
                       b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                       b->row_bad, list, count, c->st, gate);
  else if (scheme == COLL_SK98)
    hipLaunchKernelGGL((k_coll_pairs<COLL_SK98, V, MAXC>), dim3(gc), dim3(COLL_BLOCK), 0, c->stream, g, ve, vi,
                       b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                       b->row_bad, list, count, c->st, gate);
  else
    hipLaunchKernelGGL((k_coll_pairs<COLL_P12, V, MAXC>), dim3(gc), dim3(COLL_BLOCK), 0, c->stream, g, ve, vi,
                       b->off[0], b->off[1], b->perm[0], b->perm[1], m_e, m_i, q_e, q_i, seed, step, b->tot,
                       b->row_bad, list, count, c->st, gate);
  }
}

// the binary pass on the cells of `mask` (all cells when null): cells
// classified by size, then shuffle + statistics of every cell (the row
// breaks are final only after all of them), then the pairs
template <class V>
static void coll_binary(pcp_ctx *c, int scheme, V ve, V vi, double m_e, double m_i, double q_e, double q_i,
                        uint64_t seed, PassNo step, const int *mask, const unsigned long long *gate = nullptr) {
  CollBuffers *b = c->coll;
  const Geom &g = c->g;
  const long ncells = b->ncells;
  const long rows = (long)c->D * g.nr;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::max<long>(1, std::min<long>((rows + 255) / 256, 1024))),
                     dim3(256), 0, c->stream, b->row_bad, rows, INT_MAX);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, b->lcount, (long)COLL_CLASSES + 1);
  const unsigned gz = (unsigned)std::max(1L, std::min<long>((ncells + 255) / 256, (long)c->n_cu * 16));
  hipLaunchKernelGGL(k_coll_classify, dim3(gz), dim3(256), 0, c->stream, g, b->off[0], b->off[1], mask, b->lists,
                     b->lcount, b->crowd_off, b->row_bad, c->st, gate);
  coll_class<V, 0>(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, false, gate);
  coll_class<V, 1>(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, false, gate);
  coll_class<V, 2>(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, false, gate);
  coll_class<V, 3>(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, false, gate);
  coll_class<V, 0>(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, true, gate);
  coll_class<V, 1>(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, true, gate);
  coll_class<V, 2>(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, true, gate);
  coll_class<V, 3>(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, true, gate);
}

template <class V>
static void surr_moments(pcp_ctx *c, V ve, V vi, double *mom) {
  CollBuffers *b = c->coll;
  unsigned gc = (unsigned)std::min<long>(b->ncells, (long)c->n_cu * 32);
  hipLaunchKernelGGL(k_surr_moments<V>, dim3(gc), dim3(COLL_BLOCK), 0, c->stream, c->g, ve, vi, b->off[0], b->off[1],
                     b->perm[0], b->perm[1], mom);
}

static SurrNorm surr_norm(const pcp_ctx *c) {
  SurrNorm nm;
  const MlpState *m = c->mlp;
  for (int k = 0; k < MLP_IN; ++k) {
    nm.mean[k] = m ? m->feat_mean[k] : 0.f;
    nm.std[k] = m ? m->feat_std[k] : 1.f;
  }
  for (int k = 0; k < MLP_OUT; ++k) nm.tscale[k] = m ? m->targ_scale[k] : 1.f;
  return nm;
}

template <class V>
static int coll_run(pcp_ctx *c, int scheme, V ve, V vi, long n_e, long n_i, double m_e, double m_i, double q_e,
                    double q_i, uint64_t seed, PassNo step) {
  CollBuffers *b = c->coll;
  if (coll_buckets(c, ve, vi, n_e, n_i) < 0) return -1;
  if (scheme != COLL_SURROGATE) {
    coll_binary(c, scheme, ve, vi, m_e, m_i, q_e, q_i, seed, step, nullptr);
    return 0;
  }
  // network on the cells it covers, P12 on the others
  if (!c->mlp || !c->mlp->trained) return -2;
  const long ncells = b->ncells;
  unsigned gc = (unsigned)std::min<long>(ncells, (long)c->n_cu * 32);
  unsigned gt = (unsigned)std::min<long>((ncells + 255) / 256, (long)c->n_cu * 16);
  surr_moments(c, ve, vi, b->mom);
  hipLaunchKernelGGL(k_surr_features, dim3(gt), dim3(256), 0, c->stream, c->g, b->mom, m_e, m_i, surr_norm(c), 1,
                     b->feat, b->ok);
  mlp_forward(c, b->feat, ncells, b->pred);
  hipLaunchKernelGGL(k_fill_zero_ull, dim3(1), dim3(1), 0, c->stream, b->counter);
  hipLaunchKernelGGL(k_surr_apply<V>, dim3(gc), dim3(COLL_BLOCK), 0, c->stream, c->g, ve, vi, b->off[0], b->off[1],
                     b->perm[0], b->perm[1], b->mom, m_e, m_i, b->pred, surr_norm(c), b->ok, b->fallback,
                     b->counter);
  coll_binary(c, COLL_P12, ve, vi, m_e, m_i, q_e, q_i, seed, step, b->fallback);
  return 0;
}

// one P12 pass that also appends (features, relative temperature changes)
// rows of the cells it collided to the network's training set
template <class V>
static int coll_collect(pcp_ctx *c, V ve, V vi, long n_e, long n_i, double m_e, double m_i, double q_e, double q_i,
                        uint64_t seed, PassNo step) {
  CollBuffers *b = c->coll;
  MlpState *m = c->mlp;
  if (coll_buckets(c, ve, vi, n_e, n_i) < 0) return -1;
  const long ncells = b->ncells;
  unsigned gt = (unsigned)std::min<long>((ncells + 255) / 256, (long)c->n_cu * 16);
  surr_moments(c, ve, vi, b->mom);
  hipLaunchKernelGGL(k_surr_features, dim3(gt), dim3(256), 0, c->stream, c->g, b->mom, m_e, m_i, surr_norm(c), 0,
                     b->feat, b->ok);
  coll_binary(c, COLL_P12, ve, vi, m_e, m_i, q_e, q_i, seed, step, nullptr);
  surr_moments(c, ve, vi, b->mom2);
  // rows in cell order (the bucket arrays of species 0 are free again):
  // flags, their exclusive sum, then the rows at their offsets
  const long n0 = m->n;
  unsigned gf = (unsigned)std::min<long>((ncells + 256) / 256, (long)c->n_cu * 16);
  hipLaunchKernelGGL(k_surr_rowflag, dim3(gf), dim3(256), 0, c->stream, c->g, b->mom, b->mom2, m_e, m_i, b->ok,
                     b->row_bad, b->cnt[0]);
  size_t bytes = b->scan_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(b->scan_tmp, bytes, b->cnt[0], b->off[0], (int)(ncells + 1), c->stream) !=
      hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_surr_targets, dim3(gt), dim3(256), 0, c->stream, c->g, b->mom, b->mom2, m_e, m_i, b->ok,
                     b->row_bad, b->feat, b->off[0], n0, m->X, m->T, m->cap);
  uint32_t rows = 0;
  if (hipMemcpyAsync(&rows, b->off[0] + ncells, sizeof rows, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    return -1;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
  m->n = std::min<long>(n0 + (long)rows, m->cap);
  return 0;
}

int coll_prepare(pcp_ctx *c, int se, int si) {
  if (c->s32) return coll_alloc(c, c->s32->sp[se].cap, c->s32->sp[si].cap);
  return coll_alloc(c, c->sp[se].cap, c->sp[si].cap);
}

int launch_collide(pcp_ctx *c, int scheme, int se, int si, uint64_t seed, PassNo step, int vslot) {
  if (c->s32) {
    Species32 &e = c->s32->sp[se], &i = c->s32->sp[si];
    if (coll_alloc(c, e.cap, i.cap) < 0) return -1;
    View32 ve{e, c->cur, vslot}, vi{i, c->cur, vslot};
    return coll_run(c, scheme, ve, vi, e.n, i.n, e.m, i.m, e.q, i.q, seed, step);
  }
  SpeciesDev &e = c->sp[se], &i = c->sp[si];
  if (coll_alloc(c, e.cap, i.cap) < 0) return -1;
  View64 ve{e, c->cur}, vi{i, c->cur};
  return coll_run(c, scheme, ve, vi, e.n, i.n, e.m, i.m, e.q, i.q, seed, step);
}

int fsurr_prepare(pcp_ctx *c, int se, int si) {
  if (!c->s32) return -1;
  if (coll_alloc(c, c->s32->sp[se].cap, c->s32->sp[si].cap) < 0) return -1;
  CollBuffers *b = c->coll;
  const long nc = c->g.Nr * c->g.Nz;
  if (b->fcells == nc) return 0;
  if (hipMalloc(&b->fmom, nc * SURR_PAR_STRIDE * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->fX, nc * MLP_XSTRIDE * sizeof(float)) != hipSuccess ||
      hipMalloc(&b->fY, nc * MLP_YSTRIDE * sizeof(float)) != hipSuccess ||
      hipMalloc(&b->fpar, nc * SURR_PAR_STRIDE * sizeof(float)) != hipSuccess ||
      hipMalloc(&b->fok, nc * sizeof(int)) != hipSuccess)
    return -1;
  unsigned gz = (unsigned)std::min<long>((nc * SURR_PAR_STRIDE + 255) / 256, (long)c->n_cu * 16);
  hipLaunchKernelGGL(k_fill_zero_f64, dim3(gz), dim3(256), 0, c->stream, b->fmom, nc * SURR_PAR_STRIDE);
  b->fcells = nc;
  return 0;
}

double *fsurr_mom(pcp_ctx *c) { return c->coll ? c->coll->fmom : nullptr; }
const float *fsurr_par(pcp_ctx *c) { return c->coll ? c->coll->fpar : nullptr; }

int fsurr_cells(pcp_ctx *c, int se, int si, uint64_t seed, PassNo step) {
  CollBuffers *b = c->coll;
  if (!b || !b->fmom || !c->mlp || !c->mlp->trained) return -2;
  Species32 &e = c->s32->sp[se], &i = c->s32->sp[si];
  const long nc = b->fcells;
  const unsigned gt = (unsigned)std::min<long>((nc + 255) / 256, (long)c->n_cu * 16);
  hipLaunchKernelGGL(k_fsurr_features, dim3(gt), dim3(256), 0, c->stream, c->g, b->fmom, e.m, i.m, surr_norm(c),
                     b->fX, b->fok);
  mlp_forward(c, b->fX, nc, b->fY);
  hipLaunchKernelGGL(k_fill_zero_ull, dim3(1), dim3(1), 0, c->stream, b->counter);
  hipLaunchKernelGGL(k_fsurr_params, dim3(gt), dim3(256), 0, c->stream, c->g, b->fmom, e.m, i.m, b->fY,
                     surr_norm(c), b->fok, b->fpar, b->fallback, b->counter, b->cnt[0], b->cnt[1]);
  // P12 on the cells the network does not cover, on the pushed velocities
  // (every kernel returns at once while no cell is flagged)
  View32 ve{e, c->cur, c->cur ^ 1}, vi{i, c->cur, c->cur ^ 1};
  if (coll_buckets(c, ve, vi, e.n, i.n, b->fallback, b->counter, true) < 0) return -1;
  coll_binary(c, COLL_P12, ve, vi, e.m, i.m, e.q, i.q, seed, step, b->fallback, b->counter);
  return 0;
}

int surrogate_collect(pcp_ctx *c, int se, int si, uint64_t seed, uint64_t pass) {
  const PassNo step{nullptr, pass};
  if (c->s32) {
    Species32 &e = c->s32->sp[se], &i = c->s32->sp[si];
    if (coll_alloc(c, e.cap, i.cap) < 0) return -1;
    View32 ve{e, c->cur, c->cur}, vi{i, c->cur, c->cur};
    return coll_collect(c, ve, vi, e.n, i.n, e.m, i.m, e.q, i.q, seed, step);
  }
  SpeciesDev &e = c->sp[se], &i = c->sp[si];
  if (coll_alloc(c, e.cap, i.cap) < 0) return -1;
  View64 ve{e, c->cur}, vi{i, c->cur};
  return coll_collect(c, ve, vi, e.n, i.n, e.m, i.m, e.q, i.q, seed, step);
}

long surrogate_fallback_cells(pcp_ctx *c) {
  if (!c->coll) return 0;
  unsigned long long n = 0;
  if (hipMemcpy(&n, c->coll->counter, sizeof n, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return (long)n;
}

}  // namespace pcp
