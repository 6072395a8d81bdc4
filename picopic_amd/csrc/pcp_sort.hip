// Counting sort of one species by (domain, sort bin, cell), with compaction.
//
// Cell order makes the 64 particles of a wavefront share their field nodes
// (the gather becomes broadcast loads) and their current nodes (the
// wave-aggregated atomics issue one add per node).  The counting and
// scatter passes aggregate their counters per workgroup chunk in LDS.  Removed particles
// (r < 0) are dropped; the tail [n_live, n) is marked removed so that the
// host-side slot count stays valid without a synchronisation.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "pcp_sort.hpp"
#include "pcp_wave.hpp"

namespace pcp {

__global__ void __launch_bounds__(SORT_BLOCK) k_sort_count(Geom g, SpeciesDev s, int cur, uint32_t *key,
                                                           uint32_t *count, uint32_t sentinel) {
  __shared__ SortTable t;
  table_clear(t);
  __syncthreads();
  for (long c0 = (long)blockIdx.x * SORT_CHUNK; c0 < s.n; c0 += (long)gridDim.x * SORT_CHUNK)
    table_count_chunk(t, c0, s.n, sentinel, count, [&](long p) {
      uint32_t k = sentinel;
      const double r = s.r[cur][p], z = s.z[cur][p];
      if (r >= 0) {
        const long cr = (long)floor(r / g.dr), cz = (long)floor(z / g.dz);
        const int I = (int)g.div_nr.div((uint32_t)cr), J = (int)g.div_nz.div((uint32_t)cz);
        const int d = I * g.Dz + J;
        k = sort_cell_key(g, d, cr - (long)I * g.nr, cz - (long)J * g.nz);
      }
      key[p] = k;
      return k;
    });
}

__global__ void __launch_bounds__(SORT_BLOCK) k_sort_scatter(SpeciesDev s, SortBuffers sb, int cur,
                                                             const uint32_t *key, const uint32_t *offset,
                                                             uint32_t *fill, uint32_t sentinel) {
  __shared__ SortTable t;
  const int nxt = cur ^ 1;
  table_clear(t);
  __syncthreads();
  for (long c0 = (long)blockIdx.x * SORT_CHUNK; c0 < s.n; c0 += (long)gridDim.x * SORT_CHUNK)
    table_scatter_chunk(t, c0, s.n, sentinel, key, offset, fill, [&](long p, long q) {
      s.r[nxt][q] = s.r[cur][p];
      s.z[nxt][q] = s.z[cur][p];
      sb.vr[q] = s.vr[p];
      sb.vphi[q] = s.vphi[p];
      sb.vz[q] = s.vz[p];
      sb.w[q] = s.w[p];
      sb.seq[q] = s.seq[p];
      sb.id[q] = s.id[p];
    });
}

// mark [n_live, n) removed in both position buffers
__global__ void k_sort_tail(SpeciesDev s, const uint32_t *n_live_ptr) {
  long n_live = (long)*n_live_ptr;
  for (long p = n_live + (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x) {
    s.r[0][p] = -1;
    s.r[1][p] = -1;
  }
}

__global__ void k_zero_u32(uint32_t *p, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = 0;
}

size_t sort_temp_bytes(long ncells) {
  size_t bytes = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr, (int)(ncells + 1));
  return bytes;
}

int launch_sort(pcp_ctx *c, int sp) {
  SpeciesDev &s = c->sp[sp];
  if (s.n == 0) return 0;
  SortBuffers &sb = c->sortb[sp];
  const long ncells = (long)c->D * c->g.nr * c->g.nz;
  const uint32_t sentinel = (uint32_t)ncells;
  unsigned grid = (unsigned)std::min<long>((s.n + 255) / 256, (long)c->n_cu * 16);
  const unsigned gsort = (unsigned)std::max(1L, std::min((s.n + SORT_CHUNK - 1) / SORT_CHUNK, (long)c->n_cu * 16));
  unsigned gz = (unsigned)std::min<long>((ncells + 256) / 256, (long)c->n_cu * 16);
  hipLaunchKernelGGL(k_zero_u32, dim3(gz), dim3(256), 0, c->stream, c->sort_count, ncells + 1);
  hipLaunchKernelGGL(k_zero_u32, dim3(gz), dim3(256), 0, c->stream, c->sort_fill, ncells + 1);
  hipLaunchKernelGGL(k_sort_count, dim3(gsort), dim3(SORT_BLOCK), 0, c->stream, c->g, s, c->cur, c->sort_key, c->sort_count,
                     sentinel);
  size_t bytes = c->sort_temp_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(c->sort_temp, bytes, c->sort_count, c->sort_offset, (int)(ncells + 1),
                                       c->stream) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_sort_scatter, dim3(gsort), dim3(SORT_BLOCK), 0, c->stream, s, sb, c->cur, c->sort_key, c->sort_offset,
                     c->sort_fill, sentinel);
  // swap the sorted buffers in (this species only; the shared `cur` index
  // keeps its meaning because the two position pointers are exchanged)
  std::swap(s.r[c->cur], s.r[c->cur ^ 1]);
  std::swap(s.z[c->cur], s.z[c->cur ^ 1]);
  std::swap(s.vr, sb.vr);
  std::swap(s.vphi, sb.vphi);
  std::swap(s.vz, sb.vz);
  std::swap(s.w, sb.w);
  std::swap(s.seq, sb.seq);
  std::swap(s.id, sb.id);
  hipLaunchKernelGGL(k_sort_tail, dim3(grid), dim3(256), 0, c->stream, s, c->sort_offset + ncells);
  return 1;
}

}  // namespace pcp
