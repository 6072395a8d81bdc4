// Kernels of the FP64 per-step particle loop on gfx950.
//
//  k_yee_e / k_yee_h    Yee stencils, one thread per node of the halo-extended
//                       grid of every domain (coalesced along z)
//  k_stitch_copy        E/H overlay: halo <- neighbour interior
//  k_stitch_sum_j       J overlay: interior += neighbour halo contributions
//  k_step_fused         push + move + VB/ZigZag deposit, one thread per particle
//  k_push, k_move, k_migrate, k_mark, k_deposit_exact
//                       the same work split into the reference's phases; the
//                       exact deposit honours the order-dependent J_r reset of
//                       strict axial motion through per-node markers
//  k_mig_*              order keys of particles that changed domain
#include <hip/hip_runtime.h>
#include <hipcub/block/block_radix_sort.hpp>

#include <climits>

#include "pcp_kernels.hpp"
#include "pcp_physics.hpp"
#include "pcp_wave.hpp"

namespace pcp {

// ---------------------------------------------------------------------------
// sinks for the deposition templates
// ---------------------------------------------------------------------------
struct AtomicSink {
  double *J;   // component 0 of the destination domain
  long P;
  int yr, nr, nz;
  int *err;
  int *need_exact;   // set when a strict-axis J_r reset happens (fused mode)
  __device__ __forceinline__ bool inb(int i, int k) const { return i >= -2 && i < nr + 2 && k >= -2 && k < nz + 2; }
  __device__ __forceinline__ void add(int c, int i, int k, double v) {
    if (!inb(i, k)) { *err |= ERR_MOVED_TOO_FAR; return; }
    unsafeAtomicAdd(&J[c * P + (long)(i + 2) * yr + (k + 2)], v);
  }
  __device__ __forceinline__ void strict_reset(int, int) {
    if (need_exact) *need_exact = 1;
  }
};

struct ExactSink {
  double *J;
  const uint64_t *marker;   // J_r reset markers of the destination domain
  long P;
  int yr, nr, nz;
  uint64_t stamp, key;      // this step's stamp, this particle's order key
  int *err;
  __device__ __forceinline__ bool inb(int i, int k) const { return i >= -2 && i < nr + 2 && k >= -2 && k < nz + 2; }
  __device__ __forceinline__ void add(int c, int i, int k, double v) {
    if (!inb(i, k)) { *err |= ERR_MOVED_TOO_FAR; return; }
    long o = (long)(i + 2) * yr + (k + 2);
    if (c == 0) {
      uint64_t m = marker[o];
      // a particle later in the reference order reset this node: drop
      if ((m >> 40) == stamp && (m & ((1ull << 40) - 1)) > key) return;
    }
    unsafeAtomicAdd(&J[c * P + o], v);
  }
  __device__ __forceinline__ void strict_reset(int, int) {}
};

struct MarkSink {
  uint64_t *marker;
  int yr, nr, nz;
  uint64_t val;   // stamp << 40 | key
  __device__ __forceinline__ void add(int, int, int, double) {}
  __device__ __forceinline__ void strict_reset(int i, int k) {
    if (i < -2 || i >= nr + 2 || k < -2 || k >= nz + 2) return;
    atomicMax((unsigned long long *)&marker[(long)(i + 2) * yr + (k + 2)], (unsigned long long)val);
  }
};

__device__ __forceinline__ uint64_t order_key(int sp, uint64_t seq) { return ((uint64_t)sp << 36) | seq; }

// ---------------------------------------------------------------------------
// Yee E: maxwellSolverYee.cpp:161-257. Halo nodes are zeroed
// (field_e.overlay_set(0)); with clear_j the J node is zeroed after it is
// read (Domain::reset_current folded into the only reader of J).  T is the
// storage/arithmetic type: double reproduces the reference operation for
// operation, float is the FP32 split-precision mode.
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_yee_e(Geom g, FieldsT<T> f, int clear_j, StepState *st) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)g.Dr * g.Dz * g.P;
  if (t >= total) return;
  int d = (int)(t / g.P);
  long o = t - (long)d * g.P;
  int i = (int)(o / g.yr) - 2, k = (int)(o % g.yr) - 2;
  T *E = f.E + (long)d * 3 * g.P;
  T *J = f.J + (long)d * 3 * g.P;
  const T *Hat = f.Hat + (long)d * 3 * g.P;
  const long P = g.P;
  const int yr = g.yr;
  if (i < 0 || k < 0 || i >= g.nr || k >= g.nz) {
    E[o] = 0; E[P + o] = 0; E[2 * P + o] = 0;
    if (clear_j) { J[o] = 0; J[P + o] = 0; J[2 * P + o] = 0; }
    return;
  }
  Dom dm = make_dom(g, d);
  const int r_begin = dm.w0 ? 1 : 0, z_begin = dm.w1 ? 1 : 0;
  const int r_end = dm.w2 ? g.nr - 1 : g.nr, z_end = dm.w3 ? g.nz - 1 : g.nz;
  const T dr = (T)g.dr, dz = (T)g.dz, dt = (T)g.dt;
  T epsilonx2 = 2 * f.eps[(long)d * P + o];
  T sigma_t = g.pml_on ? f.sigma[(long)d * P + o] * dt : (T)0;
  T koef_e = (epsilonx2 - sigma_t) / (epsilonx2 + sigma_t);
  T koef_h = 2 * dt / (epsilonx2 + sigma_t);
  const long om = o - 1;    // k - 1
  const long oi = o - yr;   // i - 1
  T jr = J[o], jp = J[P + o], jz = J[2 * P + o];
  if (i == 0 && dm.w0) {
    if (k >= z_begin && k < z_end) {
      T er = E[o] * koef_e;
      er -= (jr + (Hat[P + o] - Hat[P + om]) / dz) * koef_h;
      E[o] = er;
      T ez = E[2 * P + o] * koef_e;
      ez -= (jz - Hat[P + o] * (T)4. / dr) * koef_h;
      E[2 * P + o] = ez;
    }
  } else if (k == 0 && dm.w1) {
    if (i >= r_begin && i < r_end) {
      T ez = E[2 * P + o] * koef_e;
      ez -= (jz - (Hat[P + o] - Hat[P + oi]) / dr
             - (Hat[P + o] + Hat[P + oi]) / ((T)2. * dr * (T)(double)(i + dm.bot_r))) * koef_h;
      E[2 * P + o] = ez;
    }
  } else if (i >= r_begin && i < r_end && k >= z_begin && k < z_end) {
    T er = E[o] * koef_e;
    er -= (jr + (Hat[P + o] - Hat[P + om]) / dz) * koef_h;
    T ep = E[P + o] * koef_e;
    ep -= (jp - (Hat[o] - Hat[om]) / dz + (Hat[2 * P + o] - Hat[2 * P + oi]) / dr) * koef_h;
    T ez = E[2 * P + o] * koef_e;
    ez -= (jz - (Hat[P + o] - Hat[P + oi]) / dr
           - (Hat[P + o] + Hat[P + oi]) / ((T)2. * dr * (T)(double)(i + dm.bot_r))) * koef_h;
    E[o] = er; E[P + o] = ep; E[2 * P + o] = ez;
    if (isnan(er) || isnan(ep) || isnan(ez)) atomicOr(&st->err, ERR_NAN_FIELD);
  }
  if (clear_j) { J[o] = 0; J[P + o] = 0; J[2 * P + o] = 0; }
}

// Yee H: maxwellSolverYee.cpp:111-159 (the outer-wall H_r block runs before
// the regular loop, so H_r at i = Nr-1 is advanced twice, quirk Q12)
template <class T>
__global__ void k_yee_h(Geom g, FieldsT<T> f) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)g.Dr * g.Dz * g.P;
  if (t >= total) return;
  int d = (int)(t / g.P);
  long o = t - (long)d * g.P;
  int i = (int)(o / g.yr) - 2, k = (int)(o % g.yr) - 2;
  T *H = f.H + (long)d * 3 * g.P;
  T *Hat = f.Hat + (long)d * 3 * g.P;
  const T *E = f.E + (long)d * 3 * g.P;
  const long P = g.P;
  const int yr = g.yr;
  if (i < 0 || k < 0 || i >= g.nr || k >= g.nz) {
    H[o] = 0; H[P + o] = 0; H[2 * P + o] = 0;
    Hat[o] = 0; Hat[P + o] = 0; Hat[2 * P + o] = 0;
    return;
  }
  Dom dm = make_dom(g, d);
  const T dr = (T)g.dr, dz = (T)g.dz, dt = (T)g.dt, mu0 = (T)MAGN_CONST;
  const long op = o + 1;    // k + 1
  const long oi = o + yr;   // i + 1
  T hat_r = Hat[o];
  if (dm.w2 && i == g.nr - 1) {
    T alpha_t = dt * (E[P + op] - E[P + o]) / (dz * mu0);
    hat_r += alpha_t;
  }
  T alpha_t_r = dt * (E[P + op] - E[P + o]) / (dz * mu0);
  H[o] = hat_r + alpha_t_r / 2;
  Hat[o] = hat_r + alpha_t_r;
  T alpha_t_phi = dt * ((E[2 * P + oi] - E[2 * P + o]) / dr - (E[op] - E[o]) / dz) / mu0;
  T hp = Hat[P + o];
  H[P + o] = hp + alpha_t_phi / 2;
  Hat[P + o] = hp + alpha_t_phi;
  T alpha_t_z = dt * ((E[P + oi] + E[P + o]) / ((T)2. * dr * (T)(i + 0.5 + (double)dm.bot_r))
                      + (E[P + oi] - E[P + o]) / dr) / mu0;
  T hz = Hat[2 * P + o];
  H[2 * P + o] = hz - alpha_t_z / 2;
  Hat[2 * P + o] = hz - alpha_t_z;
}

// E/H overlay (SMB.cpp:784-823, 1010-1044): halos hold zeros before the
// stitch, so every stitched halo node is a copy of the neighbour's interior
// node; the anti-diagonal corners are never stitched (quirk Q14).
template <class T>
__global__ void k_stitch_copy(Geom g, T *G, int ncomp) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)g.Dr * g.Dz * g.P;
  if (t >= total) return;
  int d = (int)(t / g.P);
  long o = t - (long)d * g.P;
  int i = (int)(o / g.yr) - 2, k = (int)(o % g.yr) - 2;
  int di = i < 0 ? -1 : (i >= g.nr ? 1 : 0);
  int dk = k < 0 ? -1 : (k >= g.nz ? 1 : 0);
  if ((di == 0 && dk == 0) || di * dk == -1) return;
  int I = d / g.Dz + di, Jd = d % g.Dz + dk;
  if (I < 0 || I >= g.Dr || Jd < 0 || Jd >= g.Dz) return;
  int dn = I * g.Dz + Jd;
  int in = i - di * g.nr, kn = k - dk * g.nz;
  long on = (long)(in + 2) * g.yr + (kn + 2);
  for (int c = 0; c < ncomp; ++c) G[((long)d * ncomp + c) * g.P + o] = G[((long)dn * ncomp + c) * g.P + on];
}

// J overlay (SMB.cpp:566-596): every interior node within two cells of a
// domain edge gains the halo deposits of the r, z and main-diagonal
// neighbours; anti-diagonal deposits are lost (quirk Q14).
template <class T>
__global__ void k_stitch_sum_j(Geom g, T *J) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)g.Dr * g.Dz * g.nr * g.nz;
  if (t >= total) return;
  long cells = (long)g.nr * g.nz;
  int d = (int)(t / cells);
  long c = t - (long)d * cells;
  int i = (int)(c / g.nz), k = (int)(c % g.nz);
  int lo_i = i < 2, hi_i = i >= g.nr - 2, lo_k = k < 2, hi_k = k >= g.nz - 2;
  if (!(lo_i || hi_i || lo_k || hi_k)) return;
  int I0 = d / g.Dz, J0 = d % g.Dz;
  long o = (long)(i + 2) * g.yr + (k + 2);
  T acc[3] = {0, 0, 0};
  bool any = false;
  for (int di = -1; di <= 1; ++di) {
    if (di == -1 && !lo_i) continue;
    if (di == 1 && !hi_i) continue;
    for (int dk = -1; dk <= 1; ++dk) {
      if (di == 0 && dk == 0) continue;
      if (di * dk == -1) continue;
      if (dk == -1 && !lo_k) continue;
      if (dk == 1 && !hi_k) continue;
      int I = I0 + di, Jd = J0 + dk;
      if (I < 0 || I >= g.Dr || Jd < 0 || Jd >= g.Dz) continue;
      int dn = I * g.Dz + Jd;
      int in = i - di * g.nr, kn = k - dk * g.nz;
      long on = (long)(in + 2) * g.yr + (kn + 2);
      for (int cc = 0; cc < 3; ++cc) acc[cc] += J[((long)dn * 3 + cc) * g.P + on];
      any = true;
    }
  }
  if (any)
    for (int cc = 0; cc < 3; ++cc) J[((long)d * 3 + cc) * g.P + o] += acc[cc];
}

__global__ void k_fill(double *p, long n, double v) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = v;
}


// ehist (optional): the energy log, whose row of the new step is cleared
// and stamped with the step number
__global__ void k_step_begin(StepState *st, double *ehist) {
  if (threadIdx.x == 0) {
    st->step += 1;
    st->need_exact = 0;
    st->n_mig = 0;
    st->queue = 0;
  }
  __syncthreads();
  if (ehist && threadIdx.x < EHIST_COLS) {
    double *row = ehist_row(ehist, st->step);
    row[threadIdx.x] = threadIdx.x == 0 ? (double)st->step : 0.0;
  }
}

// ---------------------------------------------------------------------------
// particle kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ GridView view(const double *G, const Geom &g, int d) {
  GridView v;
  v.G = G + (long)d * 3 * g.P;
  v.P = g.P;
  v.yr = g.yr;
  return v;
}

// gather + push for one particle (Pusher::operator() body)
__device__ __forceinline__ void push_one(const Geom &g, const Fields &f, double q, double m, double r, double z,
                                         double v[3], int *err) {
  int d = domain_of(g, cell_number(r, g.dr), cell_number(z, g.dz));
  Dom dm = make_dom(g, d);
  double e[3], h[3];
  gather_e(view(f.E, g, d), g.dr, g.dz, dm.bot_r, dm.left_z, r, z, e);
  gather_h(view(f.H, g, d), g.dr, g.dz, dm.bot_r, dm.left_z, r, z, h);
  push_velocity(g.pusher, q, m, g.dt, e, h, v, err);
}

__device__ __forceinline__ bool bad_position(double r, double z) {
  return isnan(r) || isinf(r) || isnan(z) || isinf(z);
}

__global__ void __launch_bounds__(256) k_push(Geom g, Fields f, SpeciesDev s, int cur, StepState *st) {
  int err = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x) {
    double r = s.r[cur][p], z = s.z[cur][p];
    if (r < 0) continue;
    if (bad_position(r, z)) { err |= ERR_NAN_POSITION; continue; }
    double v[3] = {s.vr[p], s.vphi[p], s.vz[p]};
    push_one(g, f, s.q, s.m, r, z, v, &err);
    s.vr[p] = v[0]; s.vphi[p] = v[1]; s.vz[p] = v[2];
  }
  if (err) atomicOr(&st->err, err);
}

// move one particle; returns false if it left the simulation area (removed
// by the runaway collector, SMB.cpp:243-320)
__device__ __forceinline__ bool move_one(const Geom &g, int beam, double &r, double &z, double &vr, double &vphi,
                                         double &vz, int &d_old, int &d_new) {
  d_old = domain_of(g, cell_number(r, g.dr), cell_number(z, g.dz));
  Dom dm = make_dom(g, d_old);
  double vzo;
  move_particle(g, dm, beam, r, z, vr, vphi, vz, vzo);
  vz = vzo;
  if (bad_position(r, z)) return false;
  long cr = (long)floor(r / g.dr), cz = (long)floor(z / g.dz);
  if (cr < 0 || cz < 0 || cr >= g.Nr || cz >= g.Nz) return false;
  d_new = domain_of(g, (int)cr, (int)cz);
  return true;
}

// phase "move": pos[cur] -> pos[cur^1]; velocities in place.  Removal and
// migration are left to k_migrate (the reference collects after binding).
__global__ void __launch_bounds__(256) k_move(Geom g, SpeciesDev s, int cur) {
  const int nxt = cur ^ 1;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x) {
    double r = s.r[cur][p], z = s.z[cur][p];
    if (r < 0 || bad_position(r, z)) { s.r[nxt][p] = -1; s.z[nxt][p] = z; continue; }
    double vr = s.vr[p], vphi = s.vphi[p], vz = s.vz[p];
    Dom dm = make_dom(g, domain_of(g, cell_number(r, g.dr), cell_number(z, g.dz)));
    double vzo;
    move_particle(g, dm, s.beam, r, z, vr, vphi, vz, vzo);
    s.r[nxt][p] = r; s.z[nxt][p] = z;
    s.vr[p] = vr; s.vphi[p] = vphi; s.vz[p] = vzo;
  }
}

// phase "migrate": pos[cur] is the new position, pos[cur^1] the old one
__global__ void k_migrate(Geom g, SpeciesDev s, int sp, int cur, StepState *st, Migrants mg) {
  int err = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x) {
    double r = s.r[cur][p], z = s.z[cur][p];
    double ro = s.r[cur ^ 1][p], zo = s.z[cur ^ 1][p];
    if (r < 0 || ro < 0) { s.r[cur][p] = -1; continue; }
    long cr = bad_position(r, z) ? -1 : (long)floor(r / g.dr);
    long cz = bad_position(r, z) ? -1 : (long)floor(z / g.dz);
    if (cr < 0 || cz < 0 || cr >= g.Nr || cz >= g.Nz) { s.r[cur][p] = -1; continue; }
    int d_new = domain_of(g, (int)cr, (int)cz);
    int d_old = domain_of(g, cell_number(ro, g.dr), cell_number(zo, g.dz));
    if (d_new != d_old) append_migrant(mg, st, sp, p, d_old, d_new, s.seq[p], g, &err);
  }
  if (err) atomicOr(&st->err, err);
}

// the deposit of one lane's particle; every lane of the wavefront calls it
// (inactive lanes with live = false)
template <int CCS>
__device__ __forceinline__ void deposit64(bool live, const Geom &g, const Dom &dn, double r, double z,
                                          double ro, double zo, double vphi, double q, double w, const Fields &f,
                                          int d_new, StepState *st, int *err, uint64_t *marker, uint32_t *cflag,
                                          const uint64_t *seq, long p, int sp) {
  if (CCS == PCP_CCS_ZIGZAG) {
    zigzag_runs(f.J, (long)d_new * 3 * g.P, g.P, g.yr, g.nr, g.nz,
                zigzag_terms(g.dr, g.dz, g.dt, dn.bot_r, dn.left_z, r, z, ro, zo, vphi, q, w), live, err);
  } else {
    const double p_charge = q * w;
    // azimuthal part (4 interior nodes)
    AzTerms A = vb_azimuthal_terms(g.nr, g.nz, g.dr, g.dz, dn.bot_r, dn.left_z, r, z, ro, zo, vphi, p_charge);
    // r-z part: branch selection per lane, arithmetic and atomics converged
    SegRecorder R;
#pragma unroll
    for (int q2 = 0; q2 < 3; ++q2) {
      R.rn[q2] = 1.5 * g.dr; R.zn[q2] = 1.25 * g.dz; R.ro[q2] = 1.25 * g.dr; R.zo[q2] = 1.5 * g.dz;
      R.i[q2] = 1; R.k[q2] = 1;
    }
    if (live) vb_rz_dispatch(R, g.dr, g.dz, r, z, ro, zo, err);
    if (live && R.is_strict) {
      // strict motion is rare: plain atomics; a J_r reset marks the node for
      // the fix-up pass of this step
      const ResetMark M{marker + (long)d_new * g.P, cflag, ((st->step & 0xFFFFFF) << 40) | order_key(sp, seq[p]),
                        (uint32_t)st->step, dn.bot_r, dn.left_z, g.Nr, g.Nz, g.yr, g.nr, g.nz, &st->need_exact};
      ResetSink<double> A1{f.J + (long)d_new * 3 * g.P, g.P, g.yr, g.nr, g.nz, err, M};
      vb_strict(A1, g.dr, g.dz, g.dt, dn.bot_r, dn.left_z, r, z, ro, zo, p_charge);
    }
    vb_runs(f.J, (long)d_new * 3 * g.P, g.P, g.yr, g, dn, live, A, R, p_charge, err);
  }
}

// One particle of the fused push + move + deposit step (all lanes of the
// wavefront take part, live or not, since the deposit's run sums move data
// between lanes; the grid-stride loops run the same trips in every lane).
// PUSH = false skips the push (steps with a collision phase push and
// collide first).
template <int CCS, bool PUSH>
__device__ __forceinline__ void fused_trip(const Geom &g, const Fields &f, const SpeciesDev &s, int sp, int cur,
                                           StepState *st, const Migrants &mg, uint64_t *marker, uint32_t *cflag,
                                           long p, int *err) {
  const int nxt = cur ^ 1;
  bool live = false;
  double r = g.dr, z = g.dz, ro = g.dr, zo = g.dz, vphi = 0, w = 0;
  int d_new = 0;
  if (p < s.n) {
    r = s.r[cur][p];
    z = s.z[cur][p];
    if (r < 0) {
      s.r[nxt][p] = -1; s.z[nxt][p] = z;
    } else if (bad_position(r, z)) {
      *err |= ERR_NAN_POSITION;
      s.r[nxt][p] = -1; s.z[nxt][p] = z;
    } else {
      double v[3] = {s.vr[p], s.vphi[p], s.vz[p]};
      if (PUSH) push_one(g, f, s.q, s.m, r, z, v, err);
      ro = r; zo = z;
      double vr = v[0], vz = v[2];
      vphi = v[1];
      int d_old;
      bool alive = move_one(g, s.beam, r, z, vr, vphi, vz, d_old, d_new);
      s.vr[p] = vr; s.vphi[p] = vphi; s.vz[p] = vz;
      if (!alive) {
        s.r[nxt][p] = -1; s.z[nxt][p] = z;
      } else {
        s.r[nxt][p] = r; s.z[nxt][p] = z;
        if (d_new != d_old) append_migrant(mg, st, sp, p, d_old, d_new, s.seq[p], g, err);
        w = s.w[p];
        live = true;
      }
    }
  }
  if (!live) { r = ro = g.dr; z = zo = g.dz; d_new = 0; vphi = 0; w = 0; }
  const Dom dn = make_dom(g, d_new);
  deposit64<CCS>(live, g, dn, r, z, ro, zo, vphi, s.q, w, f, d_new, st, err, marker, cflag, s.seq, p, sp);
}

template <int CCS, bool PUSH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CCS == PCP_CCS_ZIGZAG ? 3 : 2)))
k_step_fused(Geom g, Fields f, SpeciesDev s, int sp, int cur, StepState *st, Migrants mg, uint64_t *marker,
             uint32_t *cflag) {
  int err = 0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long b0 = (long)blockIdx.x * blockDim.x; b0 < s.n; b0 += stride)
    fused_trip<CCS, PUSH>(g, f, s, sp, cur, st, mg, marker, cflag, b0 + threadIdx.x, &err);
  if (err) atomicOr(&st->err, err);
}

// The VB form takes its arguments as one struct and re-reads them every
// iteration through the kernel-argument segment pointer made opaque to the
// optimiser: held across the loop they spilled 116 scalar registers into
// vector lanes (15 % faster on the batched case without).  The ZigZag
// kernel measured 5 % slower in this form and keeps the plain arguments.
struct StepArgs64 {
  Geom g;
  Fields f;
  SpeciesDev s;
  StepState *st;
  Migrants mg;
  uint64_t *marker;
  uint32_t *cflag;
  int sp, cur;
};
typedef __attribute__((address_space(4))) const StepArgs64 KernargStep64;

template <bool PUSH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_step_fused_vb(StepArgs64 args) {
  (void)args;
  int err = 0;
  KernargStep64 *A = (KernargStep64 *)__builtin_amdgcn_kernarg_segment_ptr();
  const long n = A->s.n;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long b0 = (long)blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
    KernargStep64 *q = A;
    asm volatile("" : "+s"(q));
    const StepArgs64 &a = *(const StepArgs64 *)q;
    fused_trip<PCP_CCS_VB, PUSH>(a.g, a.f, a.s, a.sp, a.cur, a.st, a.mg, a.marker, a.cflag, b0 + threadIdx.x, &err);
  }
  if (err) atomicOr(&A->st->err, err);
}

// strict-motion J_r reset markers (exact path)
__global__ void k_mark(Geom g, SpeciesDev s, int sp, int cur, StepState *st, uint64_t *marker, int only_if_exact) {
  if (only_if_exact && !st->need_exact) return;
  const uint64_t stamp = st->step & 0xFFFFFF;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x) {
    double r = s.r[cur][p], z = s.z[cur][p];
    double ro = s.r[cur ^ 1][p], zo = s.z[cur ^ 1][p];
    if (r < 0 || ro < 0) continue;
    if (!vb_is_strict_reset(g.dr, g.dz, r, z, ro, zo)) continue;
    int d = domain_of(g, cell_number(r, g.dr), cell_number(z, g.dz));
    Dom dm = make_dom(g, d);
    MarkSink S{marker + (long)d * g.P, g.yr, g.nr, g.nz, (stamp << 40) | order_key(sp, s.seq[p])};
    int i_n = cell_number(r, g.dr) - (int)dm.bot_r, k_n = cell_number(z, g.dz) - (int)dm.left_z;
    S.strict_reset(i_n, k_n);
    S.strict_reset(i_n, k_n + 1);
  }
}

// exact deposition of one species: Current::current_distribution semantics
__global__ void __launch_bounds__(256) k_deposit_exact(Geom g, Fields f, SpeciesDev s, int sp, int cur, StepState *st,
                                const uint64_t *marker, int only_if_exact) {
  if (only_if_exact && !st->need_exact) return;
  int err = 0;
  const uint64_t stamp = st->step & 0xFFFFFF;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x) {
    double r = s.r[cur][p], z = s.z[cur][p];
    double ro = s.r[cur ^ 1][p], zo = s.z[cur ^ 1][p];
    if (r < 0 || ro < 0) continue;
    int d = domain_of(g, cell_number(r, g.dr), cell_number(z, g.dz));
    Dom dm = make_dom(g, d);
    double w = s.w[p];
    ExactSink S{f.J + (long)d * 3 * g.P, marker + (long)d * g.P, g.P, g.yr, g.nr, g.nz, stamp,
                order_key(sp, s.seq[p]), &err};
    if (g.ccs == PCP_CCS_VB) {
      double p_charge = s.q * w;
      vb_azimuthal(S, g.nr, g.nz, g.dr, g.dz, dm.bot_r, dm.left_z, r, z, ro, zo, s.vphi[p], p_charge);
      vb_rz(S, g.dr, g.dz, g.dt, dm.bot_r, dm.left_z, r, z, ro, zo, p_charge, &err);
    } else {
      zigzag(S, g.dr, g.dz, g.dt, dm.bot_r, dm.left_z, r, z, ro, zo, s.vphi[p], s.q, w);
    }
  }
  if (err) atomicOr(&st->err, err);
}

// fix-up pass of a step with strict-motion resets: particles whose old cell
// is flagged recompute their r-z current and take back the J_r
// contributions at nodes that a later particle reset
__global__ void __launch_bounds__(256) k_fix(Geom g, Fields f, SpeciesDev s, int sp, int cur, StepState *st,
                                             const uint64_t *marker, const uint32_t *cflag) {
  if (!st->need_exact) return;
  int err = 0;
  const uint64_t stamp = st->step & 0xFFFFFF;
  const uint32_t fval = (uint32_t)st->step;
  // a thread screens four consecutive particles by their pre-step cell
  // (independent loads) and redoes the few in flagged cells
  const long groups = (s.n + 3) / 4;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < groups; t += (long)gridDim.x * blockDim.x) {
    bool hit[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long p = t * 4 + j;
      const double ro = p < s.n ? s.r[cur ^ 1][p] : -1.0, zo = p < s.n ? s.z[cur ^ 1][p] : 0.0;
      const long ci = min(max((long)cell_number(ro, g.dr), 0L), g.Nr - 1);
      const long ck = min(max((long)cell_number(zo, g.dz), 0L), g.Nz - 1);
      hit[j] = ro >= 0 && cflag[ci * g.Nz + ck] == fval;
    }
    for (int j = 0; j < 4; ++j) {
      if (!hit[j]) continue;
      const long p = t * 4 + j;
      const double r = s.r[cur][p], ro = s.r[cur ^ 1][p];
      if (r < 0) continue;
      const double z = s.z[cur][p], zo = s.z[cur ^ 1][p];
      const int d = domain_of(g, cell_number(r, g.dr), cell_number(z, g.dz));
      const Dom dm = make_dom(g, d);
      FixSink<double> S{f.J + (long)d * 3 * g.P, marker + (long)d * g.P, g.yr, g.nr, g.nz, stamp,
                        order_key(sp, s.seq[p])};
      vb_rz(S, g.dr, g.dz, g.dt, dm.bot_r, dm.left_z, r, z, ro, zo, s.q * s.w[p], &err);
    }
  }
}

__global__ void k_stamp(StepState *st) { st->step += 1; }
// device scalars a captured step reads and advances (no memset nodes)
__global__ void k_set_u64(uint64_t *p, uint64_t v) { *p = v; }
__global__ void k_inc_u64(uint64_t *p) { *p += 1; }

// ---------------------------------------------------------------------------
// migrant order keys: arrivals follow every particle already in the
// destination, ordered by (colour pass of the source, source domain, source
// vector position) -- the sequential push_back order of
// particles_runaway_collector (SMB.cpp:211-340)
// ---------------------------------------------------------------------------
__global__ void k_mig_count(StepState *st, Migrants mg) {
  unsigned int n = min(st->n_mig, (unsigned int)mg.cap);
  for (unsigned int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    atomicAdd(&mg.gcount[mg.group[t]], 1u);
}

__global__ void k_mig_scan(StepState *st, Migrants mg) {
  // one block: exclusive scans over the groups (few) of the migrant counts
  // and of their tile counts, then reset the fill counters
  if (st->n_mig == 0) return;
  __shared__ uint32_t part[1024], tpart[1024];
  int tid = threadIdx.x, nt = blockDim.x;
  int per = (mg.ngroups + nt - 1) / nt;
  int b = tid * per, e = min(mg.ngroups, b + per);
  uint32_t s = 0, ts = 0;
  for (int i = b; i < e; ++i) {
    s += mg.gcount[i];
    ts += (mg.gcount[i] + MIG_TILE - 1) / MIG_TILE;
  }
  part[tid] = s;
  tpart[tid] = ts;
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0, tacc = 0;
    for (int i = 0; i < nt; ++i) {
      uint32_t v = part[i], w = tpart[i];
      part[i] = acc;
      tpart[i] = tacc;
      acc += v;
      tacc += w;
    }
    mg.gtile[mg.ngroups] = tacc;
  }
  __syncthreads();
  uint32_t acc = part[tid], tacc = tpart[tid];
  for (int i = b; i < e; ++i) {
    mg.gstart[i] = acc;
    mg.gtile[i] = tacc;
    acc += mg.gcount[i];
    tacc += (mg.gcount[i] + MIG_TILE - 1) / MIG_TILE;
    mg.gfill[i] = 0;
  }
}

__global__ void k_mig_scatter(StepState *st, Migrants mg) {
  unsigned int n = min(st->n_mig, (unsigned int)mg.cap);
  for (unsigned int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    uint32_t gi = mg.group[t];
    uint32_t pos = mg.gstart[gi] + atomicAdd(&mg.gfill[gi], 1u);
    mg.bucket[pos] = t;
  }
}

// sorts every MIG_TILE-entry tile of every group by key in LDS (a block per
// tile); the sorted keys go to skey, the migrant indices back to bucket
__global__ void __launch_bounds__(MIG_TILE_THREADS) k_mig_sort_tiles(StepState *st, Migrants mg) {
  using Sort = hipcub::BlockRadixSort<uint64_t, MIG_TILE_THREADS, MIG_TILE_ITEMS, uint32_t>;
  __shared__ typename Sort::TempStorage tmp;
  if (st->n_mig == 0) return;
  const uint32_t ntiles = mg.gtile[mg.ngroups];
  for (uint32_t vt = blockIdx.x; vt < ntiles; vt += gridDim.x) {
    // the group of tile vt: gtile[lo] <= vt < gtile[lo + 1]
    int lo = 0, hi = mg.ngroups;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (mg.gtile[mid] <= vt) lo = mid;
      else hi = mid;
    }
    const uint32_t first = (vt - mg.gtile[lo]) * MIG_TILE;
    const uint32_t base = mg.gstart[lo] + first;
    const uint32_t len = min((uint32_t)MIG_TILE, mg.gcount[lo] - first);
    uint64_t k[MIG_TILE_ITEMS];
    uint32_t v[MIG_TILE_ITEMS];
#pragma unroll
    for (int i = 0; i < MIG_TILE_ITEMS; ++i) {
      const uint32_t pos = threadIdx.x * MIG_TILE_ITEMS + i;
      v[i] = pos < len ? mg.bucket[base + pos] : 0xFFFFFFFFu;
      k[i] = pos < len ? mg.key[v[i]] : ~0ull;   // padding sorts last
    }
    Sort(tmp).Sort(k, v);
#pragma unroll
    for (int i = 0; i < MIG_TILE_ITEMS; ++i) {
      const uint32_t pos = threadIdx.x * MIG_TILE_ITEMS + i;
      if (pos < len) {
        mg.skey[base + pos] = k[i];
        mg.bucket[base + pos] = v[i];
      }
    }
    __syncthreads();   // tmp is reused by the next tile
  }
}

// rank of a migrant in its group = its place in its own sorted tile plus the
// keys below it in each other tile of the group (keys are unique: the
// source domain and its vector position); O(n log MIG_TILE) per tile pair
__global__ void k_mig_rank(StepState *st, Migrants mg, SeqTable seqs) {
  unsigned int n = min(st->n_mig, (unsigned int)mg.cap);
  for (unsigned int u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const uint32_t m = mg.bucket[u];
    const uint32_t gi = mg.group[m];
    const uint64_t k = mg.skey[u];
    const uint32_t b = mg.gstart[gi], cnt = mg.gcount[gi];
    const uint32_t off = u - b, own = off / MIG_TILE;
    uint64_t rank = off - own * MIG_TILE;
    for (uint32_t t0 = 0, tile = 0; t0 < cnt; t0 += MIG_TILE, ++tile) {
      if (tile == own) continue;
      const uint64_t *tk = mg.skey + b + t0;
      uint32_t lo = 0, len = min((uint32_t)MIG_TILE, cnt - t0);
      while (len > 0) {   // lower bound of k in the tile
        const uint32_t half = len >> 1;
        if (tk[lo + half] < k) {
          lo += half + 1;
          len -= half + 1;
        } else {
          len = half;
        }
      }
      rank += lo;
    }
    uint64_t ref = mg.ref[m];
    int sp = (int)(ref >> 40);
    long p = (long)(ref & ((1ull << 40) - 1));
    seqs.p[sp][p] = mg.gbase[gi] + rank;
  }
}

__global__ void k_mig_finish(StepState *st, Migrants mg) {
  if (st->n_mig == 0) return;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < mg.ngroups; t += gridDim.x * blockDim.x) {
    mg.gbase[t] += mg.gcount[t];
    mg.gcount[t] = 0;
  }
}

// point gathers for the host-side get_field_e/h
__global__ void k_gather_point(Geom g, Fields f, int d, double r, double z, int which, double *out) {
  Dom dm = make_dom(g, d);
  double c[3];
  if (which == 0) gather_e(view(f.E, g, d), g.dr, g.dz, dm.bot_r, dm.left_z, r, z, c);
  else gather_h(view(f.H, g, d), g.dr, g.dz, dm.bot_r, dm.left_z, r, z, c);
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
}

}  // namespace pcp

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
namespace pcp {

static inline unsigned grid_for(long n, int block) { return (unsigned)((n + block - 1) / block); }
// grid-stride particle kernels: enough waves to fill every CU several times
static inline unsigned pgrid(pcp_ctx *c, long n, int block) {
  long want = (n + block - 1) / block;
  long cap = (long)c->n_cu * 16;
  if (want > cap) want = cap;
  return (unsigned)(want < 1 ? 1 : want);
}

template <class T>
void launch_yee_e_t(pcp_ctx *c, FieldsT<T> f, int clear_j) {
  long total = (long)c->D * c->g.P;
  hipLaunchKernelGGL(k_yee_e<T>, dim3(grid_for(total, 256)), dim3(256), 0, c->stream, c->g, f, clear_j, c->st);
}
template <class T>
void launch_yee_h_t(pcp_ctx *c, FieldsT<T> f) {
  long total = (long)c->D * c->g.P;
  hipLaunchKernelGGL(k_yee_h<T>, dim3(grid_for(total, 256)), dim3(256), 0, c->stream, c->g, f);
}
template <class T>
void launch_stitch_t(pcp_ctx *c, FieldsT<T> f, int which) {
  if (c->D == 1) return;
  long total = (long)c->D * c->g.P;
  if (which == PCP_GRID_E) {
    hipLaunchKernelGGL(k_stitch_copy<T>, dim3(grid_for(total, 256)), dim3(256), 0, c->stream, c->g, f.E, 3);
  } else if (which == PCP_GRID_H || which == PCP_GRID_H_AT_ET) {
    hipLaunchKernelGGL(k_stitch_copy<T>, dim3(grid_for(total, 256)), dim3(256), 0, c->stream, c->g, f.H, 3);
    hipLaunchKernelGGL(k_stitch_copy<T>, dim3(grid_for(total, 256)), dim3(256), 0, c->stream, c->g, f.Hat, 3);
  } else {
    long cells = (long)c->D * c->g.nr * c->g.nz;
    hipLaunchKernelGGL(k_stitch_sum_j<T>, dim3(grid_for(cells, 256)), dim3(256), 0, c->stream, c->g, f.J);
  }
}
template void launch_yee_e_t<double>(pcp_ctx *, FieldsT<double>, int);
template void launch_yee_e_t<float>(pcp_ctx *, FieldsT<float>, int);
template void launch_yee_h_t<double>(pcp_ctx *, FieldsT<double>);
template void launch_yee_h_t<float>(pcp_ctx *, FieldsT<float>);
template void launch_stitch_t<double>(pcp_ctx *, FieldsT<double>, int);
template void launch_stitch_t<float>(pcp_ctx *, FieldsT<float>, int);

void launch_yee_e(pcp_ctx *c, int clear_j) { launch_yee_e_t<double>(c, c->f, clear_j); }
void launch_yee_h(pcp_ctx *c) { launch_yee_h_t<double>(c, c->f); }
void launch_stitch(pcp_ctx *c, int which) { launch_stitch_t<double>(c, c->f, which); }
void launch_fill(pcp_ctx *c, double *p, long n, double v) {
  hipLaunchKernelGGL(k_fill, dim3(pgrid(c, n, 256)), dim3(256), 0, c->stream, p, n, v);
}
void launch_step_begin(pcp_ctx *c) {
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, c->stream, c->st, c->elog ? c->ehist : nullptr);
}
void launch_stamp(pcp_ctx *c) { hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, c->stream, c->st); }
void launch_set_u64(pcp_ctx *c, uint64_t *p, uint64_t v) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, c->stream, p, v);
}
void launch_inc_u64(pcp_ctx *c, uint64_t *p) { hipLaunchKernelGGL(k_inc_u64, dim3(1), dim3(1), 0, c->stream, p); }
void launch_push(pcp_ctx *c, int sp) {
  SpeciesDev &s = c->sp[sp];
  if (!s.n) return;
  hipLaunchKernelGGL(k_push, dim3(pgrid(c, s.n, 256)), dim3(256), 0, c->stream, c->g, c->f, s, c->cur, c->st);
}
void launch_move(pcp_ctx *c, int sp) {
  SpeciesDev &s = c->sp[sp];
  if (!s.n) return;
  hipLaunchKernelGGL(k_move, dim3(pgrid(c, s.n, 256)), dim3(256), 0, c->stream, c->g, s, c->cur);
}
void launch_migrate_mark(pcp_ctx *c, int sp) {
  SpeciesDev &s = c->sp[sp];
  if (!s.n) return;
  hipLaunchKernelGGL(k_migrate, dim3(pgrid(c, s.n, 256)), dim3(256), 0, c->stream, c->g, s, sp, c->cur, c->st,
                     c->mig);
}
void launch_step_fused(pcp_ctx *c, int sp, bool push) {
  SpeciesDev &s = c->sp[sp];
  if (!s.n) return;
  const dim3 grid(pgrid(c, s.n, 256)), block(256);
  if (c->g.ccs == PCP_CCS_ZIGZAG) {
    if (push)
      hipLaunchKernelGGL((k_step_fused<PCP_CCS_ZIGZAG, true>), grid, block, 0, c->stream, c->g, c->f, s, sp, c->cur,
                         c->st, c->mig, c->marker, c->cell_flag);
    else
      hipLaunchKernelGGL((k_step_fused<PCP_CCS_ZIGZAG, false>), grid, block, 0, c->stream, c->g, c->f, s, sp, c->cur,
                         c->st, c->mig, c->marker, c->cell_flag);
  } else {
    const StepArgs64 args{c->g, c->f, s, c->st, c->mig, c->marker, c->cell_flag, sp, c->cur};
    if (push) hipLaunchKernelGGL((k_step_fused_vb<true>), grid, block, 0, c->stream, args);
    else hipLaunchKernelGGL((k_step_fused_vb<false>), grid, block, 0, c->stream, args);
  }
}
void launch_mark(pcp_ctx *c, int sp, int only_if_exact) {
  SpeciesDev &s = c->sp[sp];
  if (!s.n || c->g.ccs != PCP_CCS_VB) return;
  hipLaunchKernelGGL(k_mark, dim3(pgrid(c, s.n, 256)), dim3(256), 0, c->stream, c->g, s, sp, c->cur, c->st,
                     c->marker, only_if_exact);
}
void launch_deposit_exact(pcp_ctx *c, int sp, int only_if_exact) {
  SpeciesDev &s = c->sp[sp];
  if (!s.n) return;
  hipLaunchKernelGGL(k_deposit_exact, dim3(pgrid(c, s.n, 256)), dim3(256), 0, c->stream, c->g, c->f, s, sp, c->cur,
                     c->st, c->marker, only_if_exact);
}
void launch_fix(pcp_ctx *c, int sp) {
  SpeciesDev &s = c->sp[sp];
  if (!s.n || c->g.ccs != PCP_CCS_VB) return;
  hipLaunchKernelGGL(k_fix, dim3(pgrid(c, (s.n + 3) / 4, 256)), dim3(256), 0, c->stream, c->g, c->f, s, sp, c->cur, c->st,
                     c->marker, c->cell_flag);
}
void launch_migrant_keys_t(pcp_ctx *c, SeqTable t) {
  if (c->D == 1) return;
  unsigned gm = (unsigned)(c->n_cu * 4);
  hipLaunchKernelGGL(k_mig_count, dim3(gm), dim3(256), 0, c->stream, c->st, c->mig);
  hipLaunchKernelGGL(k_mig_scan, dim3(1), dim3(1024), 0, c->stream, c->st, c->mig);
  hipLaunchKernelGGL(k_mig_scatter, dim3(gm), dim3(256), 0, c->stream, c->st, c->mig);
  hipLaunchKernelGGL(k_mig_sort_tiles, dim3(gm), dim3(MIG_TILE_THREADS), 0, c->stream, c->st, c->mig);
  hipLaunchKernelGGL(k_mig_rank, dim3(gm), dim3(256), 0, c->stream, c->st, c->mig, t);
  hipLaunchKernelGGL(k_mig_finish, dim3(grid_for(c->mig.ngroups, 256)), dim3(256), 0, c->stream, c->st, c->mig);
}
void launch_migrant_keys(pcp_ctx *c) {
  SeqTable t;
  for (int i = 0; i < MAX_SPECIES; ++i) t.p[i] = c->sp[i].seq;
  launch_migrant_keys_t(c, t);
}
void launch_gather_point(pcp_ctx *c, int d, double r, double z, int which, double *dev_out) {
  hipLaunchKernelGGL(k_gather_point, dim3(1), dim3(1), 0, c->stream, c->g, c->f, d, r, z, which, dev_out);
}

}  // namespace pcp
