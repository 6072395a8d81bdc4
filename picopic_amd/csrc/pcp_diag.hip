// FP64 diagnostic reductions over the device state.
//
// Discrete continuity residual of the last step: for every node (i dr, k dz)
// of the global grid, the change of the charge that the macro-particles'
// ring clouds ([r -+ dr/2] x [z -+ dz/2], uniform volume density) put in the
// node's control volume [(i -+ 1/2) dr] x [(k -+ 1/2) dz], plus dt times the
// current flowing out of it: J_r(i, k) through the cylinder r = (i + 1/2) dr
// over the band around z = k dz, J_z(i, k) through z = (k + 1/2) dz over the
// annulus around r = i dr (pi dr^2 / 4 at the axis).  The Villasenor-Buneman
// deposit (currentVB.cpp) makes it vanish up to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pcp32.hpp"
#include "pcp_diag.hpp"
#include "pcp_physics.hpp"

namespace pcp {

namespace {

// old and new position of live particle p (false if removed)
struct Pos64 {
  SpeciesDev s;
  int cur;
  __device__ bool at(long p, double &r, double &z, double &ro, double &zo, double &qw) const {
    r = s.r[cur][p];
    ro = s.r[cur ^ 1][p];
    if (!(r >= 0) || !(ro >= 0)) return false;
    z = s.z[cur][p];
    zo = s.z[cur ^ 1][p];
    qw = s.q * s.w[p];
    return true;
  }
};

struct Pos32 {
  Species32 s;
  int cur;
  long Nz;
  double dr, dz;
  __device__ bool at(long p, double &r, double &z, double &ro, double &zo, double &qw) const {
    const uint32_t c = s.cell[cur][p], co = s.cell[cur ^ 1][p];
    if (c == DEAD_CELL || co == DEAD_CELL) return false;
    const long cr = c / (uint32_t)Nz, cz = c - (uint32_t)cr * (uint32_t)Nz;
    const long ocr = co / (uint32_t)Nz, ocz = co - (uint32_t)ocr * (uint32_t)Nz;
    r = ((double)cr + (double)s.fr[cur][p]) * dr;
    z = ((double)cz + (double)s.fz[cur][p]) * dz;
    ro = ((double)ocr + (double)s.fr[cur ^ 1][p]) * dr;
    zo = ((double)ocz + (double)s.fz[cur ^ 1][p]) * dz;
    qw = s.q * (double)s.w[p];
    return true;
  }
};

// add sign * q of the ring cloud at (r, z) to the control volumes
__device__ void add_cloud(double *Q, long ldq, double r, double z, double dr, double dz, double q) {
  const double vc = PI * ((r + dr / 2) * (r + dr / 2) - (r - dr / 2) * (r - dr / 2)) * dz;
  const int i0 = (int)floor(r / dr), k0 = (int)floor(z / dz);
  for (int i = i0; i <= i0 + 1; ++i) {
    const double a0 = fmax(0.0, (i - 0.5) * dr), a1 = (i + 0.5) * dr;
    const double lo = fmax(a0, r - dr / 2), hi = fmin(a1, r + dr / 2);
    if (a1 <= 0 || hi <= lo) continue;
    const double vr = PI * (hi * hi - lo * lo);
    for (int k = k0; k <= k0 + 1; ++k) {
      const double hz = fmax(0.0, fmin((k + 0.5) * dz, z + dz / 2) - fmax((k - 0.5) * dz, z - dz / 2));
      if (hz > 0) atomicAdd(&Q[(long)(i + 2) * ldq + (k + 2)], q * vr * hz / vc);
    }
  }
}

template <class V>
__global__ void k_cont_charge(V v, long n, double dr, double dz, double *Q, long ldq, double *qmax) {
  double m = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    double r, z, ro, zo, qw;
    if (!v.at(p, r, z, ro, zo, qw)) continue;
    add_cloud(Q, ldq, r, z, dr, dz, qw);
    add_cloud(Q, ldq, ro, zo, dr, dz, -qw);
    m = fmax(m, fabs(qw));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax((unsigned long long *)qmax, (unsigned long long)__double_as_longlong(m));
}

// J of global node (i, k), read from the domain that owns it
template <class T>
__device__ __forceinline__ double jnode(const Geom &g, const T *J, int c, long i, long k) {
  const int I = (int)(i / g.nr), Jd = (int)(k / g.nz);
  const int d = I * g.Dz + Jd;
  const long li = i - (long)I * g.nr, lk = k - (long)Jd * g.nz;
  return (double)J[((long)d * 3 + c) * g.P + (li + 2) * g.yr + (lk + 2)];
}

template <class T>
__global__ void k_cont_residual(Geom g, const T *J, const double *Q, long ldq, double *out) {
  double mr = 0, mq = 0;
  const long ni = g.Nr - 1, nk = g.Nz - 2;   // nodes 0..Nr-2, 1..Nz-2
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ni * nk; t += (long)gridDim.x * blockDim.x) {
    const long i = t / nk, k = 1 + t % nk;
    const double dr = g.dr, dz = g.dz, dt = g.dt;
    auto ar = [&](long ii) { return 2 * PI * (ii + 0.5) * dr * dz; };
    const double az = i == 0 ? PI * dr * dr / 4 : 2 * PI * (double)i * dr * dr;
    double div = jnode(g, J, 0, i, k) * ar(i) * dt;
    if (i > 0) div -= jnode(g, J, 0, i - 1, k) * ar(i - 1) * dt;
    div += jnode(g, J, 2, i, k) * az * dt;
    div -= jnode(g, J, 2, i, k - 1) * az * dt;
    const double q = Q[(i + 2) * ldq + (k + 2)];
    mr = fmax(mr, fabs(q + div));
    mq = fmax(mq, fabs(q));
  }
  for (int off = 32; off > 0; off >>= 1) {
    mr = fmax(mr, __shfl_xor(mr, off, 64));
    mq = fmax(mq, __shfl_xor(mq, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax((unsigned long long *)&out[0], (unsigned long long)__double_as_longlong(mr));
    atomicMax((unsigned long long *)&out[1], (unsigned long long)__double_as_longlong(mq));
  }
}

__global__ void k_zero_d(double *p, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = 0.0;
}

}  // namespace

// out[0] = max |residual|, out[1] = max |charge change| over the nodes,
// out[2] = largest |q w| of a macro-particle (non-negative doubles compare
// as their bit patterns, hence the integer maxima)
static int continuity_residual_into(pcp_ctx *c, double *Q, double *dev_out) {
  const Geom &g = c->g;
  const long ldq = g.Nz + 4, nq = (g.Nr + 4) * ldq;
  const unsigned gz = (unsigned)std::min<long>((nq + 255) / 256, (long)c->n_cu * 16);
  hipLaunchKernelGGL(k_zero_d, dim3(gz), dim3(256), 0, c->stream, Q, nq);
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, c->stream, dev_out, 3L);
  for (int sp = 0; sp < c->cfg.n_species; ++sp) {
    if (c->s32) {
      Species32 &s = c->s32->sp[sp];
      if (!s.n) continue;
      unsigned gp = (unsigned)std::min<long>((s.n + 255) / 256, (long)c->n_cu * 16);
      hipLaunchKernelGGL(k_cont_charge<Pos32>, dim3(gp), dim3(256), 0, c->stream, Pos32{s, c->cur, g.Nz, g.dr, g.dz},
                         s.n, g.dr, g.dz, Q, ldq, dev_out + 2);
    } else {
      SpeciesDev &s = c->sp[sp];
      if (!s.n) continue;
      unsigned gp = (unsigned)std::min<long>((s.n + 255) / 256, (long)c->n_cu * 16);
      hipLaunchKernelGGL(k_cont_charge<Pos64>, dim3(gp), dim3(256), 0, c->stream, Pos64{s, c->cur}, s.n, g.dr, g.dz,
                         Q, ldq, dev_out + 2);
    }
  }
  const long nodes = (g.Nr - 1) * (g.Nz - 2);
  const unsigned gr = (unsigned)std::max<long>(1, std::min<long>((nodes + 255) / 256, (long)c->n_cu * 16));
  if (c->s32)
    hipLaunchKernelGGL(k_cont_residual<float>, dim3(gr), dim3(256), 0, c->stream, g, c->s32->f.J, Q, ldq, dev_out);
  else
    hipLaunchKernelGGL(k_cont_residual<double>, dim3(gr), dim3(256), 0, c->stream, g, c->f.J, Q, ldq, dev_out);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
  return 0;
}

int launch_continuity_residual(pcp_ctx *c, double *dev_out) {
  const long nq = (c->g.Nr + 4) * (c->g.Nz + 4);
  struct Scratch {
    double *p = nullptr;
    ~Scratch() {
      if (p) (void)hipFree(p);
    }
  } q;
  if (hipMalloc(&q.p, nq * sizeof(double)) != hipSuccess) return -1;
  return continuity_residual_into(c, q.p, dev_out);
}

// The allocation-free variant's charge array: a fixed device array, one per
// GPU, held until the process ends.  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device (the array's
// name is a host-side stand-in).  Calls on one GPU share it: calls that
// overlap on one GPU race.
static __device__ __align__(256) double scratch[CONTINUITY_SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};
}  // namespace

int continuity_scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return -1;
  if (g_scratch[device]) return 0;
  double *ptr = nullptr;
  if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
  g_scratch[device] = ptr;
  return 0;
}

int launch_continuity_residual_noalloc(pcp_ctx *c, double *dev_out) {
  const long nq = (c->g.Nr + 4) * (c->g.Nz + 4);
  if (nq > CONTINUITY_SCRATCH_DOUBLES) return -2;
  if (c->device < 0 || c->device >= MAX_DEVICES || !g_scratch[c->device]) return -3;
  return continuity_residual_into(c, g_scratch[c->device], dev_out);
}

// ---------------------------------------------------------------------------
// Probe windows (OutWriter::operator(), src/outWriter.cpp:49-100): node
// values of one component over global cells [r0, r0 + nr_w) x [z0, z0 + nz_w),
// each read from the domain that owns the cell (its interior node), written
// row-major as double.
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_probe(Geom g, const T *grid, int ncomp, int comp, long r0, long z0, long nr_w, long nz_w,
                        double *out) {
  const long n = nr_w * nz_w;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long i = r0 + t / nz_w, j = z0 + t % nz_w;
    const int d = (int)(i / g.nr) * g.Dz + (int)(j / g.nz);
    const int li = (int)(i - (i / g.nr) * g.nr), lj = (int)(j - (j / g.nz) * g.nz);
    out[t] = (double)grid[gidx(g, d, comp, ncomp, li, lj)];
  }
}

void launch_probe(pcp_ctx *c, const void *grid, bool f32, int ncomp, int comp, long r0, long z0, long nr_w, long nz_w,
                  double *dev_out) {
  const long n = nr_w * nz_w;
  const unsigned grid_n = (unsigned)std::max(1L, std::min((n + 255) / 256, (long)c->n_cu * 8));
  if (f32)
    hipLaunchKernelGGL(k_probe<float>, dim3(grid_n), dim3(256), 0, c->stream, c->g, (const float *)grid, ncomp, comp,
                       r0, z0, nr_w, nz_w, dev_out);
  else
    hipLaunchKernelGGL(k_probe<double>, dim3(grid_n), dim3(256), 0, c->stream, c->g, (const double *)grid, ncomp,
                       comp, r0, z0, nr_w, nz_w, dev_out);
}

// ---------------------------------------------------------------------------
// Species maps of the probes (src/specieP.cpp:665-839): weight_cylindrical
// (include/algo/weighter.hpp:27-132) of each particle into its domain's
// halo grids, with FP64 atomics, then the derived maps per node, and for the
// density the pre-dump overlay's halo sums (src/outController.cpp:409-491:
// overlay_x and overlay_y over interior rows / columns, no diagonal; the
// derived maps' halos are zero, so their overlay leaves the interior as it
// is).  The momentum factors take gamma when |v| is below REL_LIMIT
// (src/specieP.cpp:731,763, as written).
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ void add_node(double *G, const Geom &g, int d, int li, int lk, double v) {
  atomicAdd(&G[(long)d * g.P + (long)(li + 2) * g.yr + (lk + 2)], v);
}

// weighter.hpp:27-132 for a particle at (pos_r, pos_z) of domain d (whose
// first cells are bot_r, left_z)
__device__ void weight_cylindrical_dev(double *G, const Geom &g, int d, long bot_r, long left_z, double pos_r,
                                       double pos_z, double weight_value) {
  const double dr = g.dr, dz = g.dz;
  double r1, r2, r3, dz1, dz2, rho_w_value, v_0;
  unsigned int r_i = (unsigned int)cell_number(pos_r, dr);
  const unsigned int z_k = (unsigned int)cell_number(pos_z, dz);
  const int r_i_shift = (int)(r_i - (unsigned int)bot_r);
  const int z_k_shift = (int)(z_k - (unsigned int)left_z);
  double v_1 = PCP_CELL_VOLUME(r_i, dr, dz);
  double v_2 = PCP_CELL_VOLUME(r_i + 1, dr, dz);
  double a[4];
  if (pos_r > dr) {
    r1 = pos_r - 0.5 * dr;
    r2 = (r_i + 0.5) * dr;
    r3 = pos_r + 0.5 * dr;
    v_0 = 2. * PI * dz * dr * pos_r;
    dz1 = (z_k + 0.5) * dz - (pos_z - 0.5 * dz);
    dz2 = (pos_z + 0.5 * dz) - (z_k + 0.5) * dz;
    rho_w_value = weight_value / v_0;
    a[0] = PCP_CYL_RNG_VOL(dz1, r1, r2) / v_1;
    a[1] = PCP_CYL_RNG_VOL(dz1, r2, r3) / v_2;
    a[2] = PCP_CYL_RNG_VOL(dz2, r1, r2) / v_1;
    a[3] = PCP_CYL_RNG_VOL(dz2, r2, r3) / v_2;
  } else if (pos_r <= dr / 2.) {
    r_i = 0;
    r2 = (r_i + 0.5) * dr;
    r3 = pos_r + 0.5 * dr;
    dz1 = (z_k + 0.5) * dz - (pos_z - 0.5 * dz);
    dz2 = (pos_z + 0.5 * dz) - (z_k + 0.5) * dz;
    v_0 = PI * dz * (2. * pos_r * pos_r + dr * dr / 2.);
    rho_w_value = weight_value / v_0;
    a[0] = PI * dz1 * (dr * dr / 2. - pos_r * dr + pos_r * pos_r) / v_1;
    a[1] = PCP_CYL_RNG_VOL(dz1, r2, r3) / v_2;
    a[2] = PI * dz2 * (dr * dr / 2. - pos_r * dr + pos_r * pos_r) / v_1;
    a[3] = PCP_CYL_RNG_VOL(dz2, r2, r3) / v_2;
  } else {
    r1 = pos_r - 0.5 * dr;
    r2 = (r_i + 0.5) * dr;
    r3 = pos_r + 0.5 * dr;
    dz1 = (z_k + 0.5) * dz - (pos_z - 0.5 * dz);
    dz2 = (pos_z + 0.5 * dz) - (z_k + 0.5) * dz;
    v_0 = 2. * PI * dz * dr * pos_r;
    v_1 = PI * (dz) * (dr) * (dr);   // CYL_VOL(dz, dr), geometry.hpp:192
    v_2 = PCP_CELL_VOLUME(r_i + 1, dr, dz);
    rho_w_value = weight_value / v_0;
    a[0] = PCP_CYL_RNG_VOL(dz1, r1, r2) / v_1;
    a[1] = PCP_CYL_RNG_VOL(dz1, r2, r3) / v_2;
    a[2] = PCP_CYL_RNG_VOL(dz2, r1, r2) / v_1;
    a[3] = PCP_CYL_RNG_VOL(dz2, r2, r3) / v_2;
  }
  add_node(G, g, d, r_i_shift, z_k_shift, rho_w_value * a[0]);
  add_node(G, g, d, r_i_shift + 1, z_k_shift, rho_w_value * a[1]);
  add_node(G, g, d, r_i_shift, z_k_shift + 1, rho_w_value * a[2]);
  add_node(G, g, d, r_i_shift + 1, z_k_shift + 1, rho_w_value * a[3]);
}

// live particle p: position and velocity (false when removed)
struct MapParts64 {
  SpeciesDev s;
  int cur;
  __device__ bool at(long p, double &r, double &z, double v[3], double &w) const {
    r = s.r[cur][p];
    if (!(r >= 0)) return false;
    z = s.z[cur][p];
    v[0] = s.vr[p]; v[1] = s.vphi[p]; v[2] = s.vz[p];
    w = s.w[p];
    return true;
  }
};
struct MapParts32 {
  Species32 s;
  int cur;
  long Nz;
  double dr, dz;
  __device__ bool at(long p, double &r, double &z, double v[3], double &w) const {
    const uint32_t c = s.cell[cur][p];
    if (c == DEAD_CELL) return false;
    const long cr = c / (uint32_t)Nz, cz = c - (uint32_t)cr * (uint32_t)Nz;
    r = ((double)cr + (double)s.fr[cur][p]) * dr;
    z = ((double)cz + (double)s.fz[cur][p]) * dz;
    v[0] = s.vr[cur][p]; v[1] = s.vphi[cur][p]; v[2] = s.vz[cur][p];
    w = s.w[p];
    return true;
  }
};

// kind: 0 density only, 1 density + momentum_w (temperature, momentum),
// 2 density + momentum_abs_w (energy)
template <class V>
__global__ void k_map_deposit(Geom g, V parts, long n, double mass, int kind, double *N, double *PW, double *PABS) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    double r, z, v[3], w;
    if (!parts.at(p, r, z, v, w)) continue;
    const int d = domain_of(g, cell_number(r, g.dr), cell_number(z, g.dz));
    const Dom dm = make_dom(g, d);
    weight_cylindrical_dev(N, g, d, dm.bot_r, dm.left_z, r, z, w);
    if (kind == 1) {
      const double m_weighted = mass * w;
      double pc[3] = {m_weighted * v[0], m_weighted * v[1], m_weighted * v[2]};
      const double v_abs_2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      if (v_abs_2 < REL_LIMIT * REL_LIMIT) {
        const double gamma = 1 / sqrt(1.0 - v_abs_2 / LIGHT_VEL_POW_2);
        for (int c = 0; c < 3; ++c) pc[c] *= gamma;
      }
      for (int c = 0; c < 3; ++c)
        weight_cylindrical_dev(PW + (long)c * g.Dr * g.Dz * g.P, g, d, dm.bot_r, dm.left_z, r, z, pc[c]);
    } else if (kind == 2) {
      const double v_abs = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      double p_abs = mass * w * v_abs;
      if (v_abs < REL_LIMIT) p_abs *= 1 / sqrt(1.0 - v_abs * v_abs / LIGHT_VEL_POW_2);
      weight_cylindrical_dev(PABS, g, d, dm.bot_r, dm.left_z, r, z, p_abs);
    }
  }
}

__device__ __forceinline__ double energy_m_dev(double mass, double momentum_2) {
  if (momentum_2 > REL_LIMIT * REL_LIMIT * mass * mass) {
    const double e_rest = mass * LIGHT_VEL_POW_2;
    return sqrt(momentum_2 * LIGHT_VEL_POW_2 + e_rest * e_rest) - e_rest;
  }
  return 0.5 * momentum_2 / mass;
}

// map kinds: 0 density, 1 temperature, 2 energy, 3 momentum.  OUT is
// [D][ncomp][P] (ncomp 3 for the momentum); interior nodes only are written
// (halos zero).  PW is [3][D][P].
__global__ void k_map_derive(Geom g, int kind, double mass, const double *N, const double *PW, const double *PABS,
                             double *OUT) {
  const long DP = (long)g.Dr * g.Dz * g.P;
  const long total = (long)g.Dr * g.Dz * g.nr * g.nz;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int d = (int)(t / ((long)g.nr * g.nz));
    const long lc = t - (long)d * g.nr * g.nz;
    const int li = (int)(lc / g.nz), lk = (int)(lc - (long)li * g.nz);
    const long loc = (long)(li + 2) * g.yr + (lk + 2);
    const long ix = (long)d * g.P + loc;
    if (kind == 0) {
      // own deposits plus the neighbours' halo deposits (overlay_x / _y)
      const Dom dm = make_dom(g, d);
      double v = N[ix];
      if (li < 2 && dm.I > 0) v += N[(long)(d - g.Dz) * g.P + (long)(li + g.nr + 2) * g.yr + (lk + 2)];
      if (li >= g.nr - 2 && dm.I < g.Dr - 1) v += N[(long)(d + g.Dz) * g.P + (long)(li - g.nr + 2) * g.yr + (lk + 2)];
      if (lk < 2 && dm.J > 0) v += N[(long)(d - 1) * g.P + (long)(li + 2) * g.yr + (lk + g.nz + 2)];
      if (lk >= g.nz - 2 && dm.J < g.Dz - 1) v += N[(long)(d + 1) * g.P + (long)(li + 2) * g.yr + (lk - g.nz + 2)];
      OUT[ix] = v;
      continue;
    }
    const double density_local = N[ix];
    if (kind == 3) {
      for (int c = 0; c < 3; ++c)
        OUT[((long)d * 3 + c) * g.P + loc] = density_local != 0 ? PW[c * DP + ix] / density_local : 0.0;
      continue;
    }
    if (density_local == 0) {
      OUT[ix] = 0;
      continue;
    }
    double p2;
    const double pa = PABS[ix];
    if (kind == 1) {
      const double p_vec_sum_2 = PW[ix] * PW[ix] + PW[DP + ix] * PW[DP + ix] + PW[2 * DP + ix] * PW[2 * DP + ix];
      p2 = pa * pa - p_vec_sum_2;
    } else {
      p2 = pa * pa;
    }
    p2 /= density_local * density_local;
    OUT[ix] = energy_m_dev(mass, p2) * (1. / EL_CHARGE);
  }
}

__global__ void k_zero_f64(double *p, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = 0.0;
}

}  // namespace

int launch_species_map(pcp_ctx *c, int sp, int kind, double *N, double *PW, double *PABS, double *OUT) {
  const Geom &g = c->g;
  const long DP = (long)c->D * g.P;
  const unsigned gz = (unsigned)std::min<long>((4 * DP + 255) / 256, (long)c->n_cu * 16);
  hipLaunchKernelGGL(k_zero_f64, dim3(gz), dim3(256), 0, c->stream, N, DP);
  hipLaunchKernelGGL(k_zero_f64, dim3(gz), dim3(256), 0, c->stream, OUT, 3 * DP);
  const int dep = kind == 0 ? 0 : kind == 2 ? 2 : 1;
  if (dep == 1) hipLaunchKernelGGL(k_zero_f64, dim3(gz), dim3(256), 0, c->stream, PW, 3 * DP);
  if (dep == 2) hipLaunchKernelGGL(k_zero_f64, dim3(gz), dim3(256), 0, c->stream, PABS, DP);
  long n;
  double mass;
  if (c->s32) {
    Species32 &s = c->s32->sp[sp];
    n = s.n;
    mass = s.m;
    if (n > 0) {
      MapParts32 v{s, c->cur, g.Nz, g.dr, g.dz};
      const unsigned gp = (unsigned)std::max(1L, std::min((n + 255) / 256, (long)c->n_cu * 16));
      hipLaunchKernelGGL(k_map_deposit<MapParts32>, dim3(gp), dim3(256), 0, c->stream, g, v, n, mass, dep, N, PW,
                         PABS);
    }
  } else {
    SpeciesDev &s = c->sp[sp];
    n = s.n;
    mass = s.m;
    if (n > 0) {
      MapParts64 v{s, c->cur};
      const unsigned gp = (unsigned)std::max(1L, std::min((n + 255) / 256, (long)c->n_cu * 16));
      hipLaunchKernelGGL(k_map_deposit<MapParts64>, dim3(gp), dim3(256), 0, c->stream, g, v, n, mass, dep, N, PW,
                         PABS);
    }
  }
  const long cells = (long)c->D * g.nr * g.nz;
  const unsigned gc = (unsigned)std::max(1L, std::min((cells + 255) / 256, (long)c->n_cu * 16));
  hipLaunchKernelGGL(k_map_derive, dim3(gc), dim3(256), 0, c->stream, g, kind, mass, N, PW, PABS, OUT);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace pcp
