// Seeded synthetic inputs generated on the device.
//
// Random numbers come from Philox-4x32-10 (pcp_rng.hpp) counted by
// (particle index, draw), so every particle's draws are independent of the
// launch shape.
#include <hip/hip_runtime.h>

#include "pcp32.hpp"
#include "pcp_init.hpp"
#include "pcp_physics.hpp"
#include "pcp_refinit.hpp"
#include "pcp_rng.hpp"

namespace pcp {

// One thread per particle; cell_of[p] found by binary search over the
// per-cell exclusive offsets (cells in domain-major order).
__global__ void k_init_plasma(Geom g, SpeciesDev s, long base, long n_new, const long *cell_off, long n_cells,
                              double weight, double vth_r, double vth_phi, double vth_z, uint64_t seed,
                              const long *dom_off) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n_new; p += (long)gridDim.x * blockDim.x) {
    long lo = 0, hi = n_cells - 1;
    while (lo < hi) {   // last cell with cell_off[c] <= p
      long mid = (lo + hi + 1) >> 1;
      if (cell_off[mid] <= p) lo = mid; else hi = mid - 1;
    }
    long c = lo;
    long cells_per_dom = (long)g.nr * g.nz;
    int d = (int)(c / cells_per_dom);
    long lc = c - (long)d * cells_per_dom;
    int li = (int)(lc / g.nz), lk = (int)(lc % g.nz);
    Dom dm = make_dom(g, d);
    long gi = dm.bot_r + li, gk = dm.left_z + lk;
    double u0, u1, n0, n1, n2, n3;
    uniform2((uint64_t)(base + p), 0, seed, u0, u1);
    normal2((uint64_t)(base + p), 1, seed, n0, n1);
    normal2((uint64_t)(base + p), 2, seed, n2, n3);
    double r0 = gi * g.dr, r1 = (gi + 1) * g.dr;
    double r = sqrt(r0 * r0 + u0 * (r1 * r1 - r0 * r0));
    double z = (gk + u1) * g.dz;
    // keep the particle inside its cell
    if (r < r0) r = r0;
    if (r >= r1) r = nextafter(r1, 0.0);
    if (z >= (gk + 1) * g.dz) z = nextafter((gk + 1) * g.dz, 0.0);
    long q = base + p;
    s.r[0][q] = r; s.r[1][q] = r;
    s.z[0][q] = z; s.z[1][q] = z;
    s.vr[q] = n0 * vth_r;
    s.vphi[q] = n1 * vth_phi;
    s.vz[q] = n2 * vth_z;
    s.w[q] = weight;
    s.seq[q] = (uint64_t)(p - dom_off[d]);
    s.id[q] = (uint64_t)q;
  }
}

__global__ void k_init_bunch(Geom g, SpeciesDev s, long base, long n_new, double sigma_r, double sigma_z, double z0,
                             double vz, double weight, uint64_t seed, const uint64_t *gbase, int sp) {
  double R = g.gsize_r, L = g.gsize_z;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n_new; p += (long)gridDim.x * blockDim.x) {
    double n0, n1, n2, n3;
    normal2((uint64_t)(base + p), 7, seed, n0, n1);
    normal2((uint64_t)(base + p), 8, seed, n2, n3);
    double x = n0 * sigma_r, y = n1 * sigma_r;
    double r = sqrt(x * x + y * y);
    double z = z0 + n2 * sigma_z;
    double rmin = 0.5 * g.dr, rmax = R - 0.5 * g.dr, zmin = 0.5 * g.dz, zmax = L - 0.5 * g.dz;
    if (r < rmin) r = rmin;
    if (r > rmax) r = rmax;
    if (z < zmin) z = zmin;
    if (z > zmax) z = zmax;
    long q = base + p;
    s.r[0][q] = r; s.r[1][q] = r;
    s.z[0][q] = z; s.z[1][q] = z;
    s.vr[q] = 0;
    s.vphi[q] = 0;
    s.vz[q] = vz;
    s.w[q] = weight;
    int d = (int)((long)floor(r / g.dr) / g.nr) * g.Dz + (int)((long)floor(z / g.dz) / g.nz);
    s.seq[q] = gbase[(long)d * MAX_SPECIES + sp] + (uint64_t)p;
    s.id[q] = (uint64_t)q;
  }
}

// The reference's initial distribution of one plasma species: particle
// k of domain d (M per domain) goes to slot base + d M + k.  A particle that
// the domain places beyond its own cells (inner domain edges are not walls,
// src/specieP.cpp:148-149) is a migrant of the first runaway collector: it
// is appended to the migrant list with its generating domain's key and
// takes its order key in the owning domain from the migration pass.
template <bool F32>
__global__ void k_init_reference(Geom g, SpeciesDev s64, Species32 s32, int sp, long base, uint32_t M,
                                 const RefInitDom *doms, RadicalPowers pw, double temperature, double mass,
                                 uint64_t seed, Migrants mg, StepState *st) {
  const long total = (long)M * g.Dr * g.Dz;
  int err = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (long)gridDim.x * blockDim.x) {
    const int d = (int)(p / M);
    const uint32_t k = (uint32_t)(p - (long)d * M);
    const RefInitDom P = doms[d];
    double r, z, v[3], w;
    if (!ref_init_particle(g, P, pw, M, k, d, temperature, mass, seed, r, z, v, w)) err |= ERR_NAN_POSITION;
    const long q = base + p;
    const uint64_t seq = P.seq_base + k;
    if (F32) {
      store_split(s32, q, g, r, z);
      s32.vr[0][q] = s32.vr[1][q] = (float)v[0];
      s32.vphi[0][q] = s32.vphi[1][q] = (float)v[1];
      s32.vz[0][q] = s32.vz[1][q] = (float)v[2];
      s32.w[q] = (float)w;
      s32.seq[q] = seq;
      s32.id[q] = (uint64_t)q;
    } else {
      s64.r[0][q] = r; s64.r[1][q] = r;
      s64.z[0][q] = z; s64.z[1][q] = z;
      s64.vr[q] = v[0];
      s64.vphi[q] = v[1];
      s64.vz[q] = v[2];
      s64.w[q] = w;
      s64.seq[q] = seq;
      s64.id[q] = (uint64_t)q;
    }
    int d_own;
    if (F32) {
      const uint32_t cell = s32.cell[0][q];
      const int cr = (int)g.div_Nz.div(cell);
      d_own = domain_of(g, cr, (int)(cell - (uint32_t)cr * (uint32_t)g.Nz));
    } else {
      d_own = domain_of(g, cell_number(r, g.dr), cell_number(z, g.dz));
    }
    if (d_own != d) append_migrant(mg, st, sp, q, d, d_own, seq, g, &err);
  }
  if (err) atomicOr(&st->err, err);
}

__global__ void k_reset_migrants(StepState *st) { st->n_mig = 0; }

// BeamP::inject (src/beamP.cpp:71-157) of n particles in domain d: the
// reference's two uniforms per particle are Philox draw 0 of counter
// (d << 40 | ctr0 + i); the particles take order keys after every particle
// of (d, sp) (appended to the domain's vector); one placed past the domain's
// cells is a migrant to its owner.
template <bool F32>
__global__ void k_inject(Geom g, SpeciesDev s64, Species32 s32, int sp, long base, long n, int d, double domain_radius,
                         double dl, double velocity, double v_macro_avg, double n_per_macro_avg, uint64_t seed,
                         uint64_t ctr0, const uint64_t *gbase, Migrants mg, StepState *st) {
  const Dom dm = make_dom(g, d);
  const double dr = g.dr, dz = g.dz;
  const uint64_t seq_base = gbase[(long)d * MAX_SPECIES + sp];
  int err = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double rand_r, rand_z;
    const uint64_t ctr = ((uint64_t)d << 40) | (ctr0 + (uint64_t)i);
    uniform2(ctr, 0, seed, rand_r, rand_z);
    const double pos_r = domain_radius * rand_r + dr / 2 + (double)dm.bot_r * dr;
    const double pos_z = dl * rand_z + dz / 2.;
    double norm;
    if (pos_r > dr / 2)
      norm = 2 * PI * pos_r * dr * dz / v_macro_avg;
    else
      norm = PI * (pos_r * pos_r + pos_r * dr + dr * dr / 4.) * dz / v_macro_avg;
    const double w = n_per_macro_avg * norm;
    const long q = base + i;
    const uint64_t seq = seq_base + (uint64_t)i;
    const uint64_t id = (1ull << 62) | ctr;
    if (F32) {
      store_split(s32, q, g, pos_r, pos_z);
      s32.vr[0][q] = s32.vr[1][q] = 0.f;
      s32.vphi[0][q] = s32.vphi[1][q] = 0.f;
      s32.vz[0][q] = s32.vz[1][q] = (float)velocity;
      s32.w[q] = (float)w;
      s32.seq[q] = seq;
      s32.id[q] = id;
    } else {
      s64.r[0][q] = pos_r; s64.r[1][q] = pos_r;
      s64.z[0][q] = pos_z; s64.z[1][q] = pos_z;
      s64.vr[q] = 0;
      s64.vphi[q] = 0;
      s64.vz[q] = velocity;
      s64.w[q] = w;
      s64.seq[q] = seq;
      s64.id[q] = id;
    }
    int d_own;
    if (F32) {
      const uint32_t cell = s32.cell[0][q];
      const int cr = (int)g.div_Nz.div(cell);
      d_own = domain_of(g, cr, (int)(cell - (uint32_t)cr * (uint32_t)g.Nz));
    } else {
      d_own = domain_of(g, cell_number(pos_r, dr), cell_number(pos_z, dz));
    }
    if (d_own != d) append_migrant(mg, st, sp, q, d, d_own, seq, g, &err);
  }
  if (err) atomicOr(&st->err, err);
}

// the domain's next order key follows the injected particles
__global__ void k_advance_gbase(uint64_t *gbase, long idx, long n) { gbase[idx] += (uint64_t)n; }

__global__ void k_count_live(const double *r, long n, unsigned long long *out) {
  unsigned long long c = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) c += r[p] >= 0;
  atomicAdd(out, c);
}

// streaming copy used to calibrate the achievable HBM bandwidth (16 B/lane)
__global__ void __launch_bounds__(256) k_copy(const double2 *__restrict__ a, double2 *__restrict__ b, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) b[t] = a[t];
}

int bench_copy(pcp_ctx *c, long bytes, int iters, double *gbps) {
  long n = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) return -1;
  (void)hipMemsetAsync(a, 0, n * 16, c->stream);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  unsigned grid = (unsigned)std::min<long>((n + 255) / 256, (long)c->n_cu * 64);
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, a, b, n);
  (void)hipEventRecord(e0, c->stream);
  for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, a, b, n);
  (void)hipEventRecord(e1, c->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *gbps = 2.0 * n * 16 * iters / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

void launch_init_plasma(pcp_ctx *c, int sp, long base, long n_new, const long *dev_cell_off, long n_cells,
                        double weight, double vth_r, double vth_phi, double vth_z, uint64_t seed,
                        const long *dev_dom_off) {
  unsigned grid = (unsigned)std::min<long>((n_new + 255) / 256, (long)c->n_cu * 32);
  hipLaunchKernelGGL(k_init_plasma, dim3(grid ? grid : 1), dim3(256), 0, c->stream, c->g, c->sp[sp], base, n_new,
                     dev_cell_off, n_cells, weight, vth_r, vth_phi, vth_z, seed, dev_dom_off);
}

void launch_init_bunch(pcp_ctx *c, int sp, long base, long n_new, double sigma_r, double sigma_z, double z0,
                       double vz, double weight, uint64_t seed) {
  unsigned grid = (unsigned)std::min<long>((n_new + 255) / 256, (long)c->n_cu * 32);
  hipLaunchKernelGGL(k_init_bunch, dim3(grid ? grid : 1), dim3(256), 0, c->stream, c->g, c->sp[sp], base, n_new,
                     sigma_r, sigma_z, z0, vz, weight, seed, c->mig.gbase, sp);
}

void launch_init_reference(pcp_ctx *c, int sp, long base, uint32_t M, const RefInitDom *dev_doms,
                           const RadicalPowers &pw, double temperature, uint64_t seed) {
  const long total = (long)M * c->D;
  unsigned grid = (unsigned)std::min<long>((total + 255) / 256, (long)c->n_cu * 32);
  hipLaunchKernelGGL(k_reset_migrants, dim3(1), dim3(1), 0, c->stream, c->st);
  if (c->s32)
    hipLaunchKernelGGL(k_init_reference<true>, dim3(grid ? grid : 1), dim3(256), 0, c->stream, c->g, SpeciesDev{},
                       c->s32->sp[sp], sp, base, M, dev_doms, pw, temperature, c->sp[sp].m, seed, c->mig, c->st);
  else
    hipLaunchKernelGGL(k_init_reference<false>, dim3(grid ? grid : 1), dim3(256), 0, c->stream, c->g, c->sp[sp],
                       Species32{}, sp, base, M, dev_doms, pw, temperature, c->sp[sp].m, seed, c->mig, c->st);
}

void launch_inject(pcp_ctx *c, int sp, long base, long n, int d, double domain_radius, double dl, double velocity,
                   double v_macro_avg, double n_per_macro_avg, uint64_t seed, uint64_t ctr0, bool migrants) {
  unsigned grid = (unsigned)std::min<long>((n + 255) / 256, (long)c->n_cu * 8);
  if (migrants) hipLaunchKernelGGL(k_reset_migrants, dim3(1), dim3(1), 0, c->stream, c->st);
  if (c->s32)
    hipLaunchKernelGGL(k_inject<true>, dim3(grid ? grid : 1), dim3(256), 0, c->stream, c->g, SpeciesDev{},
                       c->s32->sp[sp], sp, base, n, d, domain_radius, dl, velocity, v_macro_avg, n_per_macro_avg, seed,
                       ctr0, c->mig.gbase, c->mig, c->st);
  else
    hipLaunchKernelGGL(k_inject<false>, dim3(grid ? grid : 1), dim3(256), 0, c->stream, c->g, c->sp[sp], Species32{},
                       sp, base, n, d, domain_radius, dl, velocity, v_macro_avg, n_per_macro_avg, seed, ctr0,
                       c->mig.gbase, c->mig, c->st);
  hipLaunchKernelGGL(k_advance_gbase, dim3(1), dim3(1), 0, c->stream, c->mig.gbase, (long)d * MAX_SPECIES + sp, n);
}

void launch_count_live(pcp_ctx *c, int sp, unsigned long long *dev_out) {
  SpeciesDev &s = c->sp[sp];
  unsigned grid = (unsigned)std::min<long>((s.n + 255) / 256, (long)c->n_cu * 8);
  hipLaunchKernelGGL(k_count_live, dim3(grid ? grid : 1), dim3(256), 0, c->stream, s.r[c->cur], s.n, dev_out);
}

}  // namespace pcp
