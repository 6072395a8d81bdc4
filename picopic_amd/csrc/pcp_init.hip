// Seeded synthetic inputs generated on the device.
//
// Random numbers come from Philox-4x32-10 (pcp_rng.hpp) counted by
// (particle index, draw), so every particle's draws are independent of the
// launch shape.
#include <hip/hip_runtime.h>

#include "pcp_init.hpp"
#include "pcp_physics.hpp"
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

void launch_count_live(pcp_ctx *c, int sp, unsigned long long *dev_out) {
  SpeciesDev &s = c->sp[sp];
  unsigned grid = (unsigned)std::min<long>((s.n + 255) / 256, (long)c->n_cu * 8);
  hipLaunchKernelGGL(k_count_live, dim3(grid ? grid : 1), dim3(256), 0, c->stream, s.r[c->cur], s.n, dev_out);
}

}  // namespace pcp
