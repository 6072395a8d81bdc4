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
int launch_continuity_residual(pcp_ctx *c, double *dev_out) {
  const Geom &g = c->g;
  const long ldq = g.Nz + 4, nq = (g.Nr + 4) * ldq;
  struct Scratch {
    double *p = nullptr;
    ~Scratch() {
      if (p) (void)hipFree(p);
    }
  } q;
  if (hipMalloc(&q.p, nq * sizeof(double)) != hipSuccess) return -1;
  double *Q = q.p;
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

}  // namespace pcp
