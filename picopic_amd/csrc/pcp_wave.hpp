// Wavefront-level reductions and run sums (64-lane CDNA wavefronts).
//
// Particles are stored in cell order, so the lanes of a wavefront mostly
// deposit into the same few grid nodes.  The run sums below add each
// contribution along the runs of lanes that share a node and let the last
// lane of a run issue the one atomic; wave_sum is a fixed-order reduction
// over the whole wavefront.  Both must be called by all 64 lanes.
#pragma once

#include <hip/hip_runtime.h>

namespace pcp {

// DPP lane moves of a 64-bit value (every lane of the wavefront active)
template <int CTRL>
__device__ __forceinline__ double dpp64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double readlane64(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes in a fixed order, returned to every lane: pairs and
// quads by quad permutes, the 16-lane rows by row rotations (4 and 8), then
// the four row sums.  All lanes must be active.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp64<0xB1>(x);    // quad_perm [1, 0, 3, 2]
  x += dpp64<0x4E>(x);    // quad_perm [2, 3, 0, 1]
  x += dpp64<0x124>(x);   // row_ror 4
  x += dpp64<0x128>(x);   // row_ror 8
  return (readlane64(x, 0) + readlane64(x, 16)) + (readlane64(x, 32) + readlane64(x, 48));
}

// Runs of equal keys in lane order.  Cell-sorted particles put the lanes of
// a wavefront into a few long runs (one per cell they share).  A segmented
// inclusive scan -- row shifts by 1, 2, 4 and 8 lanes, then the broadcasts of
// lane 15 of each row into the next row and of lane 31 into rows 2 and 3 --
// leaves the total of each run in its last lane, which issues one atomic for
// the run.  m[s] = ~0 where the lane adds its step-s source (no run head
// between the lane and its source, and the step reaches the lane's row),
// else 0; computed once per key and shared by every value deposited at nodes
// fixed relative to that key.  All 64 lanes must be active.
struct Runs {
  int m[6];
  bool tail;   // last lane of its run
};

// DPP move with lanes that have no source reading 0
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int x) {
  return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true);
}

__device__ __forceinline__ Runs runs_of(uint32_t key) {
  const int lane = __lane_id();
  const int row = lane >> 4;
  const uint32_t prev = (uint32_t)dpp_mov<0x138>((int)key);   // wave_shr:1 -- key of lane - 1
  const int h = (lane == 0 || prev != key) ? 1 : 0;            // first lane of a run
  Runs R;
  // masks as F - 1 (F is 0 or 1): ~0 where no head lies between
  int F = h;   // a run head in (lane - d, lane] before the step of distance d
  R.m[0] = F - 1;
  F |= dpp_mov<0x111>(F);   // row_shr:1
  R.m[1] = F - 1;
  F |= dpp_mov<0x112>(F);   // row_shr:2
  R.m[2] = F - 1;
  F |= dpp_mov<0x114>(F);   // row_shr:4
  R.m[3] = F - 1;
  F |= dpp_mov<0x118>(F);   // row_shr:8 -- now: a head in this row up to the lane
  R.m[4] = (F | (~row & 1)) - 1;                     // lane 15 of the previous row: rows 1, 3
  const int Fb = dpp_mov<0x142>(F);
  F |= Fb & row;                                     // heads of row 2 seen from row 3 (row & 1)
  R.m[5] = (F | (row < 2 ? 1 : 0)) - 1;              // lane 31: rows 2, 3
  // wave_shl:1 -- head flag of lane + 1; evaluated by every lane (a DPP
  // move inside a short-circuit would read lane 63 as inactive)
  const int hn = dpp_mov<0x130>(h);
  R.tail = (lane == 63) | (hn != 0);
  return R;
}

template <int CTRL>
__device__ __forceinline__ double run_step(double x, int m) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_mov<CTRL>((int)b) & m;
  const int hi = dpp_mov<CTRL>((int)(b >> 32)) & m;
  return x + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL>
__device__ __forceinline__ float run_step_f(float x, int m) {
  return x + __int_as_float(dpp_mov<CTRL>(__float_as_int(x)) & m);
}

// run_sum in FP32 (the FP32 mode's currents): one DPP-and plus one add per
// step; the scan's rounding error over a run of up to 64 terms stays below
// 6 x 2^-24 of the run's magnitude
__device__ __forceinline__ float run_sum_f(float x, const Runs &R) {
  x = run_step_f<0x111>(x, R.m[0]);   // row_shr:1
  x = run_step_f<0x112>(x, R.m[1]);   // row_shr:2
  x = run_step_f<0x114>(x, R.m[2]);   // row_shr:4
  x = run_step_f<0x118>(x, R.m[3]);   // row_shr:8
  x = run_step_f<0x142>(x, R.m[4]);   // row_bcast:15
  x = run_step_f<0x143>(x, R.m[5]);   // row_bcast:31
  return x;
}

// Compensated summation in FP32.  two_sum returns fl(a + b) and sets e to
// its rounding error, a + b - fl(a + b), exactly (Knuth's branch-free
// TwoSum; needs round-to-nearest and no reassociation, which the build's
// flags keep).
__device__ __forceinline__ float two_sum(float a, float b, float &e) {
  const float s = a + b;
  const float bp = s - a;
  e = (a - (s - bp)) + (b - bp);
  return s;
}

template <int CTRL>
__device__ __forceinline__ void run_step_fc(float &x, float &c, int m) {
  const float t = __int_as_float(dpp_mov<CTRL>(__float_as_int(x)) & m);
  const float u = __int_as_float(dpp_mov<CTRL>(__float_as_int(c)) & m);
  float e;
  x = two_sum(x, t, e);
  c = c + (u + e);
}

// run_sum_f with Kahan-style compensation: each scan step adds the partner's
// (sum, compensation) pair and keeps the step's rounding error in c, so the
// last lane of a run holds the run's total as x + c with an error of a few
// units of c's last place instead of up to 6 x 2^-24 of the run's magnitude
__device__ __forceinline__ float run_sum_fc(float x, const Runs &R, float &c) {
  c = 0.f;
  run_step_fc<0x111>(x, c, R.m[0]);   // row_shr:1
  run_step_fc<0x112>(x, c, R.m[1]);   // row_shr:2
  run_step_fc<0x114>(x, c, R.m[2]);   // row_shr:4
  run_step_fc<0x118>(x, c, R.m[3]);   // row_shr:8
  run_step_fc<0x142>(x, c, R.m[4]);   // row_bcast:15
  run_step_fc<0x143>(x, c, R.m[5]);   // row_bcast:31
  return x;
}

// Compensated accumulation of (s, c) into the FP32 node *a: the atomic adds
// s and returns the node's previous value, from which the lane recovers the
// atomic's own rounding error exactly; that error and c go to the node's
// companion a[jc] (an array of the same layout), whose contents the step
// later folds into the node.  Every atomic on *a is exact up to what its
// companion records, so the final J = fl(J + Jc) carries one rounding plus
// the companion's own (second-order) error, whatever the number of atomics.
__device__ __forceinline__ void comp_atomic_add(float *a, long jc, float s, float c) {
  const float old = atomicAdd(a, s);
  float e;
  (void)two_sum(old, s, e);
  const float r = e + c;
  if (r != 0.f) unsafeAtomicAdd(a + jc, r);
}

// partial sums of x along the runs; the last lane of a run holds its total
__device__ __forceinline__ double run_sum(double x, const Runs &R) {
  x = run_step<0x111>(x, R.m[0]);   // row_shr:1
  x = run_step<0x112>(x, R.m[1]);   // row_shr:2
  x = run_step<0x114>(x, R.m[2]);   // row_shr:4
  x = run_step<0x118>(x, R.m[3]);   // row_shr:8
  x = run_step<0x142>(x, R.m[4]);   // row_bcast:15
  x = run_step<0x143>(x, R.m[5]);   // row_bcast:31
  return x;
}

// Neighbouring runs.  Cell-ordered lanes put consecutive runs in cells
// next to each other along z, whose anchors' keys differ by one; the right
// nodes of one run's cell are then the left nodes of the next run's.  For
// the last lane of each run: `prev` is the last lane of the run before it
// (its own lane when there is none), `take` that run's key is this key - 1,
// `give` the next run's key is this key + 1.  All 64 lanes must call it.
struct RunLinks {
  int prev;
  bool take, give;
};

__device__ __forceinline__ RunLinks run_links(uint32_t key, const Runs &R) {
  const int lane = __lane_id();
  const unsigned long long tails = __ballot(R.tail);
  const unsigned long long below = tails & ((1ull << lane) - 1ull);
  RunLinks L;
  L.prev = below ? 63 - __clzll((long long)below) : lane;
  const uint32_t kprev = (uint32_t)__shfl((int)key, L.prev, 64);
  const uint32_t knext = (uint32_t)dpp_mov<0x130>((int)key);   // wave_shl:1 -- key of lane + 1
  L.take = R.tail && below != 0ull && kprev + 1u == key;
  L.give = R.tail && lane < 63 && knext == key + 1u;
  return L;
}

// one atomic per distinct key in the wavefront; each active lane receives a
// distinct slot counted from the value the counter held
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *ctr, uint32_t key, bool active) {
  uint32_t slot = 0;
  unsigned long long pending = __ballot(active);
  const int lane = __lane_id();
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t lk = __shfl(key, leader, 64);
    const bool m = active && key == lk;
    const unsigned long long mask = __ballot(m);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(ctr + lk, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (m) slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    pending &= ~mask;
    active = active && !m;
  }
  return slot;
}

// ---------------------------------------------------------------------------
// Workgroup-level key table of the counting sort.  A workgroup of
// SORT_BLOCK threads takes chunks of SORT_CHUNK consecutive particles; the
// particles of one chunk are counted per key in an LDS table first, and the
// table issues one global atomic per distinct key of the chunk.  Sorted or
// nearly sorted particles span a few keys per chunk, so the global counters
// of crowded cells (thousands of beam particles) see about SORT_CHUNK times
// fewer atomics than with one per particle.  Keys that find no free entry
// within SORT_PROBES probes are counted directly in global memory.
// ---------------------------------------------------------------------------
constexpr int SORT_BLOCK = 256;
constexpr int SORT_ROUNDS = 4;
constexpr long SORT_CHUNK = (long)SORT_BLOCK * SORT_ROUNDS;
constexpr int SORT_TABLE = 256;          // a power of two
constexpr int SORT_PROBES = 4;
constexpr uint32_t SORT_EMPTY = 0xffffffffu;

struct SortTable {
  uint32_t key[SORT_TABLE];
  uint32_t cnt[SORT_TABLE];
};

// every thread empties the entries it owns (no barrier)
__device__ __forceinline__ void table_clear(SortTable &t) {
  for (int h = (int)threadIdx.x; h < SORT_TABLE; h += SORT_BLOCK) {
    t.key[h] = SORT_EMPTY;
    t.cnt[h] = 0;
  }
}

// Adds the active lanes' keys to the table, one LDS atomic per distinct key
// of the wavefront.  Each active lane receives its entry (-1 when the table
// had no room for its key) and its rank among the entry's particles.  Must be
// called by all 64 lanes.
__device__ __forceinline__ void table_add(SortTable &t, uint32_t key, bool active, int &entry, uint32_t &rank) {
  entry = -1;
  rank = 0;
  unsigned long long pending = __ballot(active);
  const int lane = __lane_id();
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t lk = __shfl(key, leader, 64);
    const bool m = active && key == lk;
    const unsigned long long mask = __ballot(m);
    int e = -1;
    uint32_t base = 0;
    if (lane == leader) {
      uint32_t h = lk & (uint32_t)(SORT_TABLE - 1);
      for (int i = 0; i < SORT_PROBES; ++i) {
        const uint32_t old = atomicCAS(&t.key[h], SORT_EMPTY, lk);
        if (old == SORT_EMPTY || old == lk) {
          e = (int)h;
          base = atomicAdd(&t.cnt[h], (uint32_t)__popcll(mask));
          break;
        }
        h = (h + 1) & (uint32_t)(SORT_TABLE - 1);
      }
    }
    e = __shfl(e, leader, 64);
    base = __shfl(base, leader, 64);
    if (m) {
      entry = e;
      rank = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    }
    pending &= ~mask;
    active = active && !m;
  }
}

// Counting pass over one chunk: key_of(p) gives particle p's key (the
// sentinel when it is dropped) and is called once per particle below n.
template <class KeyOf>
__device__ __forceinline__ void table_count_chunk(SortTable &t, long c0, long n, uint32_t sentinel, uint32_t *count,
                                                  KeyOf key_of) {
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const long p = c0 + (long)r * SORT_BLOCK + threadIdx.x;
    const uint32_t k = p < n ? key_of(p) : sentinel;
    const bool a = k != sentinel;
    int e;
    uint32_t rank;
    table_add(t, k, a, e, rank);
    (void)wave_reserve(count, a && e < 0 ? k : 0, a && e < 0);
  }
  __syncthreads();
  for (int h = (int)threadIdx.x; h < SORT_TABLE; h += SORT_BLOCK) {
    if (t.key[h] != SORT_EMPTY) atomicAdd(count + t.key[h], t.cnt[h]);
    t.key[h] = SORT_EMPTY;
    t.cnt[h] = 0;
  }
  __syncthreads();
}

// Scatter pass over one chunk: every particle below n whose key is not the
// sentinel receives its destination offset[key] + slot, the slots of a key
// reserved from fill[key] one range per table entry; move(p, q) copies
// particle p to q.
template <class Move>
__device__ __forceinline__ void table_scatter_chunk(SortTable &t, long c0, long n, uint32_t sentinel,
                                                    const uint32_t *key, const uint32_t *offset, uint32_t *fill,
                                                    Move move) {
  uint32_t kk[SORT_ROUNDS], rr[SORT_ROUNDS];
  int ee[SORT_ROUNDS];
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const long p = c0 + (long)r * SORT_BLOCK + threadIdx.x;
    const uint32_t k = p < n ? key[p] : sentinel;
    const bool a = k != sentinel;
    int e;
    uint32_t rank;
    table_add(t, k, a, e, rank);
    const uint32_t direct = wave_reserve(fill, a && e < 0 ? k : 0, a && e < 0);
    kk[r] = k;
    ee[r] = a ? e : -2;
    rr[r] = e >= 0 ? rank : direct;
  }
  __syncthreads();
  // each entry's count becomes the first slot of its range
  for (int h = (int)threadIdx.x; h < SORT_TABLE; h += SORT_BLOCK)
    if (t.key[h] != SORT_EMPTY) t.cnt[h] = atomicAdd(fill + t.key[h], t.cnt[h]);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    if (ee[r] == -2) continue;
    const long p = c0 + (long)r * SORT_BLOCK + threadIdx.x;
    const uint32_t slot = ee[r] >= 0 ? t.cnt[ee[r]] + rr[r] : rr[r];
    move(p, (long)offset[kk[r]] + slot);
  }
  __syncthreads();
  table_clear(t);
  __syncthreads();
}

}  // namespace pcp
