// C ABI of picopic_amd (include/picopic.h): context life cycle, host <-> device
// transfers, and the step phases.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pcp_init.hpp"
#include "pcp_internal.hpp"
#include "pcp_kernels.hpp"
#include "pcp_sort.hpp"
#include "pcp32.hpp"
#include "pcp_collide.hpp"
#include "pcp_mlp.hpp"
#include "pcp_diag.hpp"

using namespace pcp;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(PCP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

// device scratch released on every return path
template <class T>
struct DevBuf {
  T *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)); }
};

// HIP events destroyed on every return path
struct EventSet {
  std::vector<hipEvent_t> ev;
  EventSet() = default;
  EventSet(const EventSet &) = delete;
  EventSet &operator=(const EventSet &) = delete;
  ~EventSet() {
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  hipError_t create(size_t n) {
    ev.assign(n, nullptr);
    for (auto &e : ev)
      if (hipError_t r = hipEventCreate(&e)) {
        e = nullptr;
        return r;
      }
    return hipSuccess;
  }
};

template <class T>
int dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return fail(PCP_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemset(*p, 0, n * sizeof(T));
  if (e != hipSuccess) return fail(PCP_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
  return 0;
}

// MaxwellSolverYee::set_pml for one domain (maxwellSolverYee.cpp:57-109),
// evaluated on the host once at creation.
void set_pml_host(const pcp_ctx *c, int d, const long pml_size[4], double *sigma) {
  const Geom &g = c->g;
  Dom dm = make_dom(g, d);
  double cell_size_r = g.dr, cell_size_z = g.dz;
  double lenght_sigma_left = cell_size_z * pml_size[1];
  double lenght_sigma_right = cell_size_z * pml_size[3];
  double lenght_sigma_extern = cell_size_r * pml_size[2];
  double radial_cell_offset = dm.bot_r * g.dr;
  double longitudinal_cell_offset = dm.left_z * g.dz;
  double system_radius = g.gsize_r;
  double system_longitude = g.gsize_z;
  double delta_sigma = c->cfg.pml_sigma2 - c->cfg.pml_sigma1;
  auto S = [&](int i, int k) -> double & { return sigma[(long)(i + 2) * g.yr + (k + 2)]; };
  if (pml_size[2] != 0)
    for (int i = 0; i < g.nr; i++)
      for (int k = 0; k < g.nz; k++) {
        unsigned long rel_end = (unsigned long)(g.gca_r - dm.bot_r);
        double dist = (double)(rel_end - (unsigned long)i) * g.dr;
        if (dist < lenght_sigma_extern)
          S(i, k) += c->cfg.pml_sigma1 + delta_sigma / std::pow(lenght_sigma_extern, 2) *
                     std::pow(cell_size_r * (i + 1) + radial_cell_offset - system_radius + lenght_sigma_extern, 2);
      }
  if (pml_size[1] != 0)
    for (int k = 0; k < g.nz; k++)
      for (int i = 0; i < g.nr; i++)
        if (cell_size_z * k + longitudinal_cell_offset < lenght_sigma_left)
          S(i, k) += c->cfg.pml_sigma1 + delta_sigma / std::pow(lenght_sigma_left, 2) *
                     std::pow(lenght_sigma_left - cell_size_z * k - longitudinal_cell_offset, 2);
  if (pml_size[3] != 0)
    for (int k = 0; k < g.nz; k++)
      for (int i = 0; i < g.nr; i++) {
        unsigned long rel_end = (unsigned long)(g.gca_z - dm.left_z);
        double dist = (double)(rel_end - (unsigned long)k) * g.dz;
        if (dist < lenght_sigma_right)
          S(i, k) += c->cfg.pml_sigma1 + delta_sigma / std::pow(lenght_sigma_right, 2) *
                     std::pow(cell_size_z * (k + 1) + longitudinal_cell_offset - system_longitude + lenght_sigma_right, 2);
      }
}

int check_state(pcp_ctx *c) {
  StepState h;
  HIP_TRY(hipMemcpyAsync(&h, c->st, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  int fatal = h.err & (ERR_NAN_POSITION | ERR_BETA_GT_1 | ERR_NAN_FIELD | ERR_MIGRANT_OVERFLOW | ERR_COLL_OVERFLOW);
  if (fatal) {
    char buf[256];
    snprintf(buf, sizeof buf,
             "device reported error bits 0x%x (1 NaN position, 2 beta>1, 4 NaN field, 8 migrant overflow, "
             "32 crowded collision cells (over 256 macro-particles of a species) holding more than 2^22 "
             "macro-particles in one pass)",
             h.err);
    int zero = 0;
    hipMemcpy(&c->st->err, &zero, sizeof zero, hipMemcpyHostToDevice);
    return fail(PCP_E_STATE, buf);
  }
  return 0;
}

double *grid_ptr(pcp_ctx *c, int which, int *ncomp) {
  *ncomp = 3;
  switch (which) {
    case PCP_GRID_E: return c->f.E;
    case PCP_GRID_H: return c->f.H;
    case PCP_GRID_H_AT_ET: return c->f.Hat;
    case PCP_GRID_J: return c->f.J;
    case PCP_GRID_EPS: *ncomp = 1; return c->f.eps;
    case PCP_GRID_SIGMA: *ncomp = 1; return c->f.sigma;
    default: return nullptr;
  }
}

int ensure_capacity(pcp_ctx *c, int sp, long add) {
  if (c->sp[sp].n + add > c->sp[sp].cap) {
    char buf[128];
    snprintf(buf, sizeof buf, "species %d capacity %ld exceeded (%ld + %ld)", sp, c->sp[sp].cap, c->sp[sp].n, add);
    return fail(PCP_E_NOMEM, buf);
  }
  return 0;
}

}  // namespace

// drops the captured step graphs (their kernels hold buffer pointers and
// particle counts)
static void graph_invalidate(pcp_ctx *c);

extern "C" {

const char *pcp_last_error(void) { return g_last_error.c_str(); }

int pcp_create(const pcp_config *cfg, pcp_ctx **out) {
  if (!cfg || !out) return fail(PCP_E_ARG, "null argument");
  *out = nullptr;
  if (cfg->grid_r <= 0 || cfg->grid_z <= 0 || cfg->domains_r <= 0 || cfg->domains_z <= 0)
    return fail(PCP_E_ARG, "grid and domains must be positive");
  if (cfg->grid_r % cfg->domains_r || cfg->grid_z % cfg->domains_z)
    return fail(PCP_E_ARG, "the grid must divide evenly into domains (overlay needs equal domain sizes)");
  if (cfg->n_species < 1 || cfg->n_species > MAX_SPECIES) return fail(PCP_E_ARG, "n_species must be 1..8");
  if (cfg->pusher < 0 || cfg->pusher > PCP_PUSHER_HC) return fail(PCP_E_ARG, "unknown pusher");
  if (cfg->ccs != PCP_CCS_VB && cfg->ccs != PCP_CCS_ZIGZAG) return fail(PCP_E_ARG, "unknown ccs");
  if (cfg->precision != PCP_PREC_FP64 && cfg->precision != PCP_PREC_FP32_SPLIT)
    return fail(PCP_E_ARG, "unknown precision");
  if ((long)cfg->grid_r * cfg->grid_z >= 0x7FFFFFFFL)
    return fail(PCP_E_ARG, "at most 2^31 - 1 cells (32-bit cell indices)");
  if (cfg->domains_r * cfg->domains_z > 4096) return fail(PCP_E_ARG, "at most 4096 domains");
  if (!(cfg->dt > 0)) return fail(PCP_E_ARG, "dt must be positive");

  pcp_ctx *c = new pcp_ctx();
  c->cfg = *cfg;
  Geom &g = c->g;
  g.Dr = (int)cfg->domains_r;
  g.Dz = (int)cfg->domains_z;
  g.nr = (int)(cfg->grid_r / cfg->domains_r);
  g.nz = (int)(cfg->grid_z / cfg->domains_z);
  g.xr = g.nr + 4;
  g.yr = g.nz + 4;
  g.P = (long)g.xr * g.yr;
  if (cfg->precision == PCP_PREC_FP32_SPLIT && 3L * (long)cfg->domains_r * cfg->domains_z * g.P * 4 > 0xFFFFFFFFL) {
    delete c;
    return fail(PCP_E_ARG, "FP32 mode: each field array (3 components of every domain, halos included) must stay "
                           "below 4 GiB (32-bit byte offsets)");
  }
  g.Nr = cfg->grid_r;
  g.Nz = cfg->grid_z;
  // domain geometry exactly as SMB::SMB builds it (src/SMB.cpp:96-103,
  // include/geometry.hpp:52-75)
  g.size_r = cfg->size_r / g.Dr;
  g.size_z = cfg->size_z / g.Dz;
  g.dr = g.size_r / g.nr;
  g.dz = g.size_z / g.nz;
  g.gsize_r = cfg->size_r;
  g.gsize_z = cfg->size_z;
  g.gca_r = (long)(g.gsize_r / g.dr);
  g.gca_z = (long)(g.gsize_z / g.dz);
  g.dt = cfg->dt;
  g.pml_on = cfg->pml_on;
  g.ccs = cfg->ccs;
  g.pusher = cfg->pusher;
  g.bin_r = 1;
  g.bin_z = g.nz;
  g.nbz = 1;
  g.div_bin_r = make_fastdiv(1);
  g.div_bin_z = make_fastdiv((uint32_t)g.nz);
  g.div_nr = make_fastdiv((uint32_t)g.nr);
  g.div_nz = make_fastdiv((uint32_t)g.nz);
  g.div_Dz = make_fastdiv((uint32_t)g.Dz);
  g.div_Nz = make_fastdiv((uint32_t)g.Nz);
  g.inv_vol0 = 1.0 / (PI * g.dz * g.dr * g.dr);
  g.dtr_f = (float)(g.dt / g.dr);
  g.dtz_f = (float)(g.dt / g.dz);
  g.inv_dt_f = (float)(1.0 / g.dt);
  c->D = g.Dr * g.Dz;
  c->cur = 0;

  hipDeviceProp_t prop;
  if (hipGetDevice(&c->device) != hipSuccess || hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
    delete c;
    return fail(PCP_E_HIP, "no HIP device");
  }
  // the allocation-free entry points' fixed device array on this GPU
  {
    static std::mutex m;
    std::lock_guard<std::mutex> lock(m);
    if (continuity_scratch_init(c->device) != 0) {
      delete c;
      return fail(PCP_E_HIP, "hipGetSymbolAddress of the fixed scratch array failed");
    }
  }
  c->n_cu = prop.multiProcessorCount;
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return fail(PCP_E_HIP, "hipStreamCreate failed");
  }
  const bool f32 = cfg->precision == PCP_PREC_FP32_SPLIT;
  long gridn = f32 ? 0 : (long)c->D * 3 * g.P;
  long gp = f32 ? 0 : (long)c->D * g.P;
  int rc = 0;
  if ((rc = dalloc(&c->f.E, gridn)) || (rc = dalloc(&c->f.H, gridn)) || (rc = dalloc(&c->f.Hat, gridn)) ||
      (rc = dalloc(&c->f.J, gridn)) || (rc = dalloc(&c->f.eps, gp)) || (rc = dalloc(&c->f.sigma, gp)) ||
      (rc = dalloc(&c->st, 1)) || (rc = dalloc(&c->diag, DIAG_DOUBLES)) || (rc = dalloc(&c->coll_pass, 1)) ||
      (rc = dalloc(&c->marker, (size_t)c->D * g.P)) ||
      (rc = dalloc(&c->cell_flag, (size_t)g.Nr * g.Nz))) {
    pcp_destroy(c);
    return rc;
  }
  if (f32 && alloc32(c) != 0) {
    pcp_destroy(c);
    return fail(PCP_E_NOMEM, "FP32 split buffers");
  }
  // epsilon = EPSILON0 everywhere (maxwellSolver.hpp:60-62); sigma per PML
  {
    std::vector<double> eps((size_t)c->D * g.P, EPSILON0);
    std::vector<double> sig((size_t)c->D * g.P, 0.0);
    if (cfg->pml_on) {
      long gpml[4] = {0, (long)std::floor(cfg->pml_left * cfg->grid_z), (long)std::floor(cfg->pml_outer * cfg->grid_r),
                      (long)std::floor(cfg->pml_right * cfg->grid_z)};
      for (int d = 0; d < c->D; ++d) {
        Dom dm = make_dom(g, d);
        long top_r = dm.bot_r + g.nr, right_z = dm.left_z + g.nz;
        long psz[4] = {0, 0, 0, 0};
        // SMB.cpp:84-93 (size_t comparisons)
        long gca_r_glob = (long)(cfg->size_r / (cfg->size_r / cfg->grid_r));
        long gca_z_glob = (long)(cfg->size_z / (cfg->size_z / cfg->grid_z));
        if ((unsigned long)(gca_r_glob - top_r) < (unsigned long)gpml[2]) psz[2] = gpml[2];
        if ((unsigned long)(gca_z_glob - right_z) < (unsigned long)gpml[3]) psz[3] = gpml[3];
        if (dm.left_z < gpml[1]) psz[1] = gpml[1];
        set_pml_host(c, d, psz, sig.data() + (size_t)d * g.P);
      }
    }
    bool ok;
    if (f32) {
      std::vector<float> e32(eps.begin(), eps.end()), s32(sig.begin(), sig.end());
      ok = hipMemcpy(c->s32->f.eps, e32.data(), e32.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(c->s32->f.sigma, s32.data(), s32.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    } else {
      ok = hipMemcpy(c->f.eps, eps.data(), eps.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(c->f.sigma, sig.data(), sig.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
      pcp_destroy(c);
      return fail(PCP_E_HIP, "upload of epsilon/sigma failed");
    }
  }
  // species
  long total_cap = 0;
  for (int s = 0; s < cfg->n_species; ++s) {
    SpeciesDev &sp = c->sp[s];
    sp.cap = cfg->species[s].capacity;
    sp.q = cfg->species[s].charge;
    sp.m = cfg->species[s].mass;
    sp.beam = cfg->species[s].is_beam;
    sp.n = 0;
    if (sp.cap < 0 || !(sp.m > 0)) {
      pcp_destroy(c);
      return fail(PCP_E_ARG, "species capacity must be >= 0 and mass > 0");
    }
    total_cap += sp.cap;
    if (f32) continue;
    if ((rc = dalloc(&sp.r[0], sp.cap)) || (rc = dalloc(&sp.r[1], sp.cap)) || (rc = dalloc(&sp.z[0], sp.cap)) ||
        (rc = dalloc(&sp.z[1], sp.cap)) || (rc = dalloc(&sp.vr, sp.cap)) || (rc = dalloc(&sp.vphi, sp.cap)) ||
        (rc = dalloc(&sp.vz, sp.cap)) || (rc = dalloc(&sp.w, sp.cap)) || (rc = dalloc(&sp.seq, sp.cap)) ||
        (rc = dalloc(&sp.id, sp.cap))) {
      pcp_destroy(c);
      return rc;
    }
  }
  // sort scratch: alternate particle arrays and per-cell counters
  long max_cap = 0;
  for (int s = 0; s < cfg->n_species; ++s) {
    SpeciesDev &sp = c->sp[s];
    SortBuffers &b = c->sortb[s];
    max_cap = std::max(max_cap, sp.cap);
    c->sort_every[s] = sp.beam ? 1 : 50;
    if (f32) continue;
    if ((rc = dalloc(&b.vr, sp.cap)) || (rc = dalloc(&b.vphi, sp.cap)) || (rc = dalloc(&b.vz, sp.cap)) ||
        (rc = dalloc(&b.w, sp.cap)) || (rc = dalloc(&b.seq, sp.cap)) || (rc = dalloc(&b.id, sp.cap))) {
      pcp_destroy(c);
      return rc;
    }
  }
  {
    long ncells = (long)c->D * g.nr * g.nz;
    c->sort_key_cap = max_cap;
    c->sort_temp_bytes = sort_temp_bytes(ncells);
    if ((rc = dalloc(&c->sort_key, max_cap)) || (rc = dalloc(&c->sort_count, ncells + 1)) ||
        (rc = dalloc(&c->sort_offset, ncells + 1)) || (rc = dalloc(&c->sort_fill, ncells + 1)) ||
        (rc = dalloc((char **)&c->sort_temp, c->sort_temp_bytes))) {
      pcp_destroy(c);
      return rc;
    }
  }
  // migration bookkeeping
  Migrants &m = c->mig;
  m.cap = cfg->max_migrants > 0 ? cfg->max_migrants : std::max<long>(1 << 16, total_cap / 64);
  m.ngroups = c->D * MAX_SPECIES;
  if ((rc = dalloc(&m.key, m.cap)) || (rc = dalloc(&m.ref, m.cap)) || (rc = dalloc(&m.group, m.cap)) ||
      (rc = dalloc(&m.bucket, m.cap)) || (rc = dalloc(&m.skey, m.cap)) || (rc = dalloc(&m.gcount, m.ngroups)) ||
      (rc = dalloc(&m.gstart, m.ngroups)) || (rc = dalloc(&m.gfill, m.ngroups)) ||
      (rc = dalloc(&m.gtile, m.ngroups + 1)) || (rc = dalloc(&m.gbase, m.ngroups))) {
    pcp_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int pcp_destroy(pcp_ctx *c) {
  if (!c) return 0;
  graph_invalidate(c);
  coll_free(c);
  mlp_free(c);
  hipFree(c->f.E); hipFree(c->f.H); hipFree(c->f.Hat); hipFree(c->f.J);
  hipFree(c->f.eps); hipFree(c->f.sigma); hipFree(c->st); hipFree(c->marker); hipFree(c->cell_flag);
  hipFree(c->diag); hipFree(c->ehist); hipFree(c->probe_buf); hipFree(c->coll_pass);
  hipFree(c->map_n); hipFree(c->map_pw); hipFree(c->map_out);
  for (int s = 0; s < MAX_SPECIES; ++s) hipFree(c->map_pabs[s]);
  for (int s = 0; s < MAX_SPECIES; ++s) {
    SpeciesDev &sp = c->sp[s];
    hipFree(sp.r[0]); hipFree(sp.r[1]); hipFree(sp.z[0]); hipFree(sp.z[1]);
    hipFree(sp.vr); hipFree(sp.vphi); hipFree(sp.vz); hipFree(sp.w); hipFree(sp.seq); hipFree(sp.id);
  }
  for (int s = 0; s < MAX_SPECIES; ++s) {
    SortBuffers &b = c->sortb[s];
    hipFree(b.vr); hipFree(b.vphi); hipFree(b.vz); hipFree(b.w); hipFree(b.seq); hipFree(b.id);
  }
  hipFree(c->sort_key); hipFree(c->sort_count); hipFree(c->sort_offset); hipFree(c->sort_fill);
  hipFree(c->sort_temp);
  Migrants &m = c->mig;
  hipFree(m.key); hipFree(m.ref); hipFree(m.group); hipFree(m.bucket); hipFree(m.skey);
  hipFree(m.gcount); hipFree(m.gstart); hipFree(m.gfill); hipFree(m.gtile); hipFree(m.gbase);
  free32(c);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int pcp_sync(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return check_state(c);
}

void *pcp_stream(pcp_ctx *c) { return c ? (void *)c->stream : nullptr; }

long pcp_grid_points(pcp_ctx *c) { return c ? c->g.P : -1; }

int pcp_upload_species(pcp_ctx *c, int sp, long n, const double *r, const double *z, const double *vr,
                       const double *vphi, const double *vz, const double *w, const uint64_t *id) {
  if (c) graph_invalidate(c);
  if (!c || sp < 0 || sp >= c->cfg.n_species || n < 0) return fail(PCP_E_ARG, "bad argument");
  if (n == 0) return 0;
  if (!r || !z || !vr || !vphi || !vz || !w) return fail(PCP_E_ARG, "null particle array");
  int rc = ensure_capacity(c, sp, n);
  if (rc) return rc;
  const Geom &g = c->g;
  std::vector<uint64_t> gbase(c->mig.ngroups);
  HIP_TRY(hipMemcpy(gbase.data(), c->mig.gbase, gbase.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
  std::vector<uint64_t> seq(n), ids(n);
  for (long p = 0; p < n; ++p) {
    long cr = (long)std::floor(r[p] / g.dr), cz = (long)std::floor(z[p] / g.dz);
    if (!(cr >= 0 && cz >= 0 && cr < g.Nr && cz < g.Nz)) {
      char buf[128];
      snprintf(buf, sizeof buf, "particle %ld at (%g, %g) is outside the simulation area", p, r[p], z[p]);
      return fail(PCP_E_ARG, buf);
    }
    int d = (int)((cr / g.nr) * g.Dz + cz / g.nz);
    seq[p] = gbase[(size_t)d * MAX_SPECIES + sp]++;
    ids[p] = id ? id[p] : (uint64_t)(c->sp[sp].n + p);
  }
  if (c->s32) {
    Species32 &s = c->s32->sp[sp];
    std::vector<uint32_t> cell(n);
    std::vector<float> fr(n), fz(n), fvr(n), fvp(n), fvz(n), fw(n);
    for (long p = 0; p < n; ++p) {
      double ur = r[p] / g.dr, uz = z[p] / g.dz;
      double ir = std::floor(ur), iz = std::floor(uz);
      float a = (float)(ur - ir), b = (float)(uz - iz);
      if (a >= 1.f) { a = 0.f; ir += 1; }
      if (b >= 1.f) { b = 0.f; iz += 1; }
      cell[p] = (uint32_t)((long)ir * g.Nz + (long)iz);
      fr[p] = a; fz[p] = b;
      fvr[p] = (float)vr[p]; fvp[p] = (float)vphi[p]; fvz[p] = (float)vz[p]; fw[p] = (float)w[p];
    }
    size_t off = (size_t)s.n, fb = (size_t)n * sizeof(float);
    for (int k = 0; k < 2; ++k) {
      HIP_TRY(hipMemcpy(s.cell[k] + off, cell.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(s.fr[k] + off, fr.data(), fb, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(s.fz[k] + off, fz.data(), fb, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(s.vr[k] + off, fvr.data(), fb, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(s.vphi[k] + off, fvp.data(), fb, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(s.vz[k] + off, fvz.data(), fb, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(s.w + off, fw.data(), fb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.seq + off, seq.data(), (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.id + off, ids.data(), (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->mig.gbase, gbase.data(), gbase.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    s.n += n;
    c->sp[sp].n = s.n;
    return 0;
  }
  SpeciesDev &s = c->sp[sp];
  size_t off = (size_t)s.n, bytes = (size_t)n * sizeof(double);
  HIP_TRY(hipMemcpy(s.r[0] + off, r, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.r[1] + off, r, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.z[0] + off, z, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.z[1] + off, z, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.vr + off, vr, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.vphi + off, vphi, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.vz + off, vz, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.w + off, w, bytes, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.seq + off, seq.data(), (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s.id + off, ids.data(), (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->mig.gbase, gbase.data(), gbase.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  s.n += n;
  return 0;
}

int pcp_count(pcp_ctx *c, int sp, long *n) {
  if (!c || !n || sp < 0 || sp >= c->cfg.n_species) return fail(PCP_E_ARG, "bad argument");
  if (c->sp[sp].n == 0) { *n = 0; return 0; }
  unsigned long long *d = (unsigned long long *)c->diag, h = 0;
  HIP_TRY(hipMemsetAsync(d, 0, sizeof h, c->stream));
  if (c->s32) launch32_count_live(c, sp, d);
  else launch_count_live(c, sp, d);
  HIP_TRY(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *n = (long)h;
  return 0;
}

int pcp_download_species(pcp_ctx *c, int sp, double *r, double *z, double *vr, double *vphi, double *vz, double *w,
                         uint64_t *id, uint64_t *order, double *r_old, double *z_old) {
  if (!c || sp < 0 || sp >= c->cfg.n_species) return fail(PCP_E_ARG, "bad argument");
  if (c->s32) {
    Species32 &s = c->s32->sp[sp];
    long n = s.n;
    if (n == 0) return 0;
    HIP_TRY(hipStreamSynchronize(c->stream));
    const Geom &g = c->g;
    std::vector<uint32_t> cell(n), cello(n);
    std::vector<float> a(n), b(n), ao(n), bo(n), t(n);
    std::vector<uint64_t> hu(n);
    HIP_TRY(hipMemcpy(cell.data(), s.cell[c->cur], n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(a.data(), s.fr[c->cur], n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(b.data(), s.fz[c->cur], n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cello.data(), s.cell[c->cur ^ 1], n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ao.data(), s.fr[c->cur ^ 1], n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bo.data(), s.fz[c->cur ^ 1], n * 4, hipMemcpyDeviceToHost));
    auto pos = [&](uint32_t cl, float fa, float fb, double &pr, double &pz) {
      long cr = (long)(cl / (uint32_t)g.Nz), cz = (long)(cl % (uint32_t)g.Nz);
      pr = ((double)cr + (double)fa) * g.dr;
      pz = ((double)cz + (double)fb) * g.dz;
    };
    long k = 0;
    for (long p = 0; p < n; ++p) {
      if (cell[p] == DEAD_CELL) continue;
      double pr, pz, por = 0, poz = 0;
      pos(cell[p], a[p], b[p], pr, pz);
      if (cello[p] != DEAD_CELL) pos(cello[p], ao[p], bo[p], por, poz);
      if (r) r[k] = pr;
      if (z) z[k] = pz;
      if (r_old) r_old[k] = por;
      if (z_old) z_old[k] = poz;
      ++k;
    }
    const float *srcs[4] = {s.vr[c->cur], s.vphi[c->cur], s.vz[c->cur], s.w};
    double *dsts[4] = {vr, vphi, vz, w};
    for (int q = 0; q < 4; ++q) {
      if (!dsts[q]) continue;
      HIP_TRY(hipMemcpy(t.data(), srcs[q], n * 4, hipMemcpyDeviceToHost));
      long kk = 0;
      for (long p = 0; p < n; ++p)
        if (cell[p] != DEAD_CELL) dsts[q][kk++] = (double)t[p];
    }
    if (id) {
      HIP_TRY(hipMemcpy(hu.data(), s.id, n * 8, hipMemcpyDeviceToHost));
      long kk = 0;
      for (long p = 0; p < n; ++p)
        if (cell[p] != DEAD_CELL) id[kk++] = hu[p];
    }
    if (order) {
      HIP_TRY(hipMemcpy(hu.data(), s.seq, n * 8, hipMemcpyDeviceToHost));
      long kk = 0;
      for (long p = 0; p < n; ++p)
        if (cell[p] != DEAD_CELL) {
          long cr = (long)(cell[p] / (uint32_t)g.Nz), cz = (long)(cell[p] % (uint32_t)g.Nz);
          uint64_t d = (uint64_t)((cr / g.nr) * g.Dz + cz / g.nz);
          order[kk++] = (d << 40) | hu[p];
        }
    }
    return 0;
  }
  SpeciesDev &s = c->sp[sp];
  long n = s.n;
  if (n == 0) return 0;
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<double> hr(n), hz(n), tmp(n);
  std::vector<uint64_t> hu(n);
  HIP_TRY(hipMemcpy(hr.data(), s.r[c->cur], n * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hz.data(), s.z[c->cur], n * sizeof(double), hipMemcpyDeviceToHost));
  auto emit = [&](double *dst, const std::vector<double> &src) {
    if (!dst) return;
    long k = 0;
    for (long p = 0; p < n; ++p)
      if (hr[p] >= 0) dst[k++] = src[p];
  };
  emit(r, hr);
  emit(z, hz);
  const double *srcs[6] = {s.vr, s.vphi, s.vz, s.w, s.r[c->cur ^ 1], s.z[c->cur ^ 1]};
  double *dsts[6] = {vr, vphi, vz, w, r_old, z_old};
  for (int a = 0; a < 6; ++a) {
    if (!dsts[a]) continue;
    HIP_TRY(hipMemcpy(tmp.data(), srcs[a], n * sizeof(double), hipMemcpyDeviceToHost));
    emit(dsts[a], tmp);
  }
  if (id) {
    HIP_TRY(hipMemcpy(hu.data(), s.id, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    long k = 0;
    for (long p = 0; p < n; ++p)
      if (hr[p] >= 0) id[k++] = hu[p];
  }
  if (order) {
    HIP_TRY(hipMemcpy(hu.data(), s.seq, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    const Geom &g = c->g;
    long k = 0;
    for (long p = 0; p < n; ++p)
      if (hr[p] >= 0) {
        long cr = (long)std::floor(hr[p] / g.dr), cz = (long)std::floor(hz[p] / g.dz);
        uint64_t d = (uint64_t)((cr / g.nr) * g.Dz + cz / g.nz);
        order[k++] = (d << 40) | hu[p];
      }
  }
  return 0;
}

int pcp_init_plasma(pcp_ctx *c, int sp, int n_per_cell, double density_left, double density_right,
                    double temperature_ev, uint64_t seed) {
  if (c) graph_invalidate(c);
  if (!c || sp < 0 || sp >= c->cfg.n_species || n_per_cell <= 0 || !(density_left >= 0) || !(density_right >= 0))
    return fail(PCP_E_ARG, "bad argument");
  const Geom &g = c->g;
  double n_peak = std::max(density_left, density_right);
  if (!(n_peak > 0)) return fail(PCP_E_ARG, "density must be positive");
  // columns: N_col(k) = ppc * Nr * n(z_k) / n_peak, spread over the radial
  // cells in proportion to the ring area (uniform in r^2, equal weights)
  long ncell = (long)c->D * g.nr * g.nz;
  std::vector<long> cnt(ncell), off(ncell + 1), doff(c->D);
  std::vector<uint64_t> gbase(c->mig.ngroups);
  HIP_TRY(hipMemcpy(gbase.data(), c->mig.gbase, gbase.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
  long total = 0;
  for (int d = 0; d < c->D; ++d) {
    Dom dm = make_dom(g, d);
    long dstart = total;
    for (int li = 0; li < g.nr; ++li)
      for (int lk = 0; lk < g.nz; ++lk) {
        long gi = dm.bot_r + li, gk = dm.left_z + lk;
        double zc = (gk + 0.5) * g.dz;
        double dens = density_left + (density_right - density_left) * zc / g.gsize_z;
        double ncol = std::floor((double)n_per_cell * g.Nr * dens / n_peak + 0.5);
        double a0 = std::floor(ncol * (double)gi * gi / ((double)g.Nr * g.Nr));
        double a1 = std::floor(ncol * (double)(gi + 1) * (gi + 1) / ((double)g.Nr * g.Nr));
        long ci = (long)(li * g.nz + lk) + (long)d * g.nr * g.nz;
        cnt[ci] = (long)(a1 - a0);
        total += cnt[ci];
      }
    doff[d] = dstart - (long)gbase[(size_t)d * MAX_SPECIES + sp];
    gbase[(size_t)d * MAX_SPECIES + sp] += total - dstart;
  }
  int rc = ensure_capacity(c, sp, total);
  if (rc) return rc;
  off[0] = 0;
  for (long i = 0; i < ncell; ++i) off[i + 1] = off[i] + cnt[i];
  double weight = n_peak * PI * g.gsize_r * g.gsize_r * g.gsize_z / ((double)n_per_cell * g.Nr * g.Nz);
  double m = c->sp[sp].m;
  double vth = std::sqrt(temperature_ev * EL_CHARGE / m);
  DevBuf<long> d_off, d_doff;
  HIP_TRY(d_off.alloc(ncell + 1));
  HIP_TRY(d_doff.alloc(c->D));
  HIP_TRY(hipMemcpy(d_off.p, off.data(), sizeof(long) * (ncell + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_doff.p, doff.data(), sizeof(long) * c->D, hipMemcpyHostToDevice));
  if (c->s32) launch32_init_plasma(c, sp, c->s32->sp[sp].n, total, d_off.p, ncell, weight, vth, seed, d_doff.p);
  else launch_init_plasma(c, sp, c->sp[sp].n, total, d_off.p, ncell, weight, vth, vth, vth, seed, d_doff.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->mig.gbase, gbase.data(), gbase.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  // the kernel wrote both position buffers
  c->sp[sp].n += total;
  if (c->s32) c->s32->sp[sp].n = c->sp[sp].n;
  return 0;
}

int pcp_init_bunch(pcp_ctx *c, int sp, long n, double n_peak, double sigma_r, double sigma_z, double z0, double gamma,
                   uint64_t seed) {
  if (c) graph_invalidate(c);
  if (!c || sp < 0 || sp >= c->cfg.n_species || n <= 0 || !(gamma >= 1) || !(sigma_r > 0) || !(sigma_z > 0))
    return fail(PCP_E_ARG, "bad argument");
  int rc = ensure_capacity(c, sp, n);
  if (rc) return rc;
  // real particles of a Gaussian bunch: n_peak (2 pi)^(3/2) sigma_r^2 sigma_z
  double real = n_peak * std::pow(2 * PI, 1.5) * sigma_r * sigma_r * sigma_z;
  double weight = real / n;
  double vz = LIGHT_VEL * std::sqrt(1.0 - 1.0 / (gamma * gamma));
  if (c->s32) launch32_init_bunch(c, sp, c->s32->sp[sp].n, n, sigma_r, sigma_z, z0, vz, weight, seed);
  else launch_init_bunch(c, sp, c->sp[sp].n, n, sigma_r, sigma_z, z0, vz, weight, seed);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  // order keys of the bunch are gbase + index: advance every domain's base
  std::vector<uint64_t> gbase(c->mig.ngroups);
  HIP_TRY(hipMemcpy(gbase.data(), c->mig.gbase, gbase.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
  for (int d = 0; d < c->D; ++d) gbase[(size_t)d * MAX_SPECIES + sp] += (uint64_t)n;
  HIP_TRY(hipMemcpy(c->mig.gbase, gbase.data(), gbase.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  c->sp[sp].n += n;
  if (c->s32) c->s32->sp[sp].n = c->sp[sp].n;
  return 0;
}

// grows the migrant list to hold at least `need` entries (its contents are
// per-pass scratch)
static int grow_migrants(pcp_ctx *c, long need) {
  Migrants &m = c->mig;
  if (need <= m.cap) return 0;
  graph_invalidate(c);   // captured steps hold the old buffers
  hipFree(m.key); hipFree(m.ref); hipFree(m.group); hipFree(m.bucket); hipFree(m.skey);
  m.key = nullptr; m.ref = nullptr; m.group = nullptr; m.bucket = nullptr; m.skey = nullptr;
  m.cap = need;
  int rc;
  if ((rc = dalloc(&m.key, m.cap)) || (rc = dalloc(&m.ref, m.cap)) || (rc = dalloc(&m.group, m.cap)) ||
      (rc = dalloc(&m.bucket, m.cap)) || (rc = dalloc(&m.skey, m.cap))) {
    m.cap = 0;
    return rc;
  }
  return 0;
}

int pcp_init_reference(pcp_ctx *c, int sp, long macro_amount, double density_left, double density_right,
                       double temperature_ev, uint64_t seed) {
  if (c) graph_invalidate(c);
  if (!c || sp < 0 || sp >= c->cfg.n_species || macro_amount < 0 || macro_amount > 0xFFFFFFFFL ||
      !(density_left >= 0) || !(density_right >= 0) || !(temperature_ev > 0) || temperature_ev > 3e3)
    return fail(PCP_E_ARG, "bad argument (temperatures above 3e3 reach the sampler's unsupported branches)");
  const Geom &g = c->g;
  // SMB.cpp:125: integer division of the species' macro amount
  const uint32_t M = (uint32_t)macro_amount / (uint32_t)g.Dr / (uint32_t)g.Dz;
  if (M == 0) return 0;
  if (int rc = ensure_capacity(c, sp, (long)M * c->D)) return rc;
  RadicalPowers pw;
  for (int k = 0; k < RADICAL_DIGITS; ++k) {
    // the reference's pow((double)power, (-order)) per digit
    volatile int order = k + 1;
    pw.p13[k] = std::pow(13.0, (double)(-order));
    pw.p11[k] = std::pow(11.0, (double)(-order));
  }
  std::vector<uint64_t> gbase(c->mig.ngroups);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(gbase.data(), c->mig.gbase, gbase.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
  std::vector<RefInitDom> doms(c->D);
  long leavers = 0;
  for (int d = 0; d < c->D; ++d) {
    const Dom dm = make_dom(g, d);
    RefInitDom &P = doms[d];
    const double dr = g.dr, dz = g.dz;
    const uint32_t int_cell = (uint32_t)dm.bot_r, ext_cell = (uint32_t)(dm.bot_r + g.nr);
    const uint32_t left_cell = (uint32_t)dm.left_z, right_cell = (uint32_t)(dm.left_z + g.nz);
    // SMB.cpp:128-130: local densities of the z ramp
    const double drho_by_dz = (density_right - density_left) / g.gsize_z;
    P.dens0 = density_left + drho_by_dz * left_cell * g.dz;
    P.dens1 = density_left + drho_by_dz * right_cell * g.dz;
    // specieP.cpp:138-149: extents, less a cell at the outer and right walls
    P.r_size = (ext_cell - int_cell) * dr;
    P.z_size = (right_cell - left_cell) * dz;
    if (dm.w2) P.r_size -= dr;
    if (dm.w3) P.z_size -= dz;
    P.int_cell = int_cell;
    P.left_cell = left_cell;
    P.seq_base = gbase[(size_t)d * MAX_SPECIES + sp];
    // set_weight (specieP.cpp:280-331): the volume sum in particle order
    double v_sum = 0;
    for (uint32_t k = 0; k < M; ++k) {
      double r, z;
      ref_init_position(dr, dz, P, pw, M, k, r, z);
      v_sum += ref_init_volume(dr, dz, r);
      if (std::floor(r / dr) >= (double)ext_cell || std::floor(z / dz) >= (double)right_cell) ++leavers;
    }
    P.v_avg = v_sum / M;
    double top_r_num = (double)ext_cell, bot_r_num = (double)int_cell;
    if (dm.w2) top_r_num -= 1;
    const double N_total = (P.dens0 + P.dens1) / 2 * PI * dr * dr * (top_r_num * top_r_num - bot_r_num * bot_r_num) *
                           P.z_size;
    P.n_per_macro_avg = N_total / M;
    gbase[(size_t)d * MAX_SPECIES + sp] += M;
  }
  // particles placed past an inner domain edge migrate (with margin for
  // the FP32 split's rounding at a cell boundary)
  if (int rc = grow_migrants(c, leavers + leavers / 8 + 1024)) return rc;
  HIP_TRY(hipMemcpy(c->mig.gbase, gbase.data(), gbase.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  DevBuf<RefInitDom> d_doms;
  HIP_TRY(d_doms.alloc(c->D));
  HIP_TRY(hipMemcpy(d_doms.p, doms.data(), sizeof(RefInitDom) * c->D, hipMemcpyHostToDevice));
  const long base = c->sp[sp].n;
  launch_init_reference(c, sp, base, M, d_doms.p, pw, temperature_ev, seed);
  if (c->s32) launch32_migrant_keys(c);
  else launch_migrant_keys(c);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->sp[sp].n += (long)M * c->D;
  if (c->s32) c->s32->sp[sp].n = c->sp[sp].n;
  return check_state(c);
}

// SMB::inject_beam at the current time (src/SMB.cpp:1308-1322 calls
// BeamP::inject, src/beamP.cpp:71-157, for every domain).  Returns the
// number of particles injected, or a negative status.
static long inject_beams(pcp_ctx *c) {
  const Geom &g = c->g;
  long total = 0;
  for (int sp = 0; sp < c->cfg.n_species; ++sp) {
    BeamHost &b = c->beam[sp];
    if (!b.on) continue;
    const double t = c->t_now;
    for (int d = 0; d < c->D; ++d) {
      BeamDomHost &B = b.dom[d];
      if (B.bunch >= b.bunches_amount) continue;
      const Dom dm = make_dom(g, d);
      const double bunch_start_time =
          b.start_time + (b.bunch_length + b.bunches_distance) / b.velocity * B.bunch;
      const double bunch_finish_time = bunch_start_time + b.bunch_length / b.velocity;
      if (dm.bot_r * g.dr < b.radius && t >= bunch_start_time && t < bunch_finish_time) {
        // the loop `for (unsigned int i = 0; i < macro_per_step_to_inject; ++i)`
        const long n = B.macro_per_step > 0 ? (long)std::ceil(B.macro_per_step) : 0;
        if (n > 0) {
          if (int rc = ensure_capacity(c, sp, n)) return rc;
          const double dr = g.dr, dz = g.dz;
          const double dl = b.velocity * g.dt;
          const double v_macro_avg = PI * b.radius * dr * dz;
          const double N_total = PI * b.radius * b.radius * b.bunch_length * b.bunches_amount * b.density;
          const double n_per_macro_avg = N_total / b.macro_amount;
          // particles past the domain's cells (a bunch wider than the domain)
          const bool spill = B.domain_radius + dr / 2 + dm.bot_r * dr >= (dm.bot_r + g.nr) * dr ||
                             dl + dz / 2. >= g.nz * dz;
          launch_inject(c, sp, c->sp[sp].n, n, d, B.domain_radius, dl, b.velocity, v_macro_avg, n_per_macro_avg,
                        b.seed, B.injected, spill);
          if (spill) {
            if (c->s32) launch32_migrant_keys(c);
            else launch_migrant_keys(c);
          }
          c->sp[sp].n += n;
          if (c->s32) c->s32->sp[sp].n = c->sp[sp].n;
          B.injected += (uint64_t)n;
          total += n;
        }
      }
      // the counter advances in domain (0, 0) only (src/beamP.cpp:146-153)
      if (t > bunch_finish_time && dm.left_z == 0 && dm.bot_r == 0) ++B.bunch;
    }
  }
  if (total) graph_invalidate(c);
  return total;
}

static float *grid_ptr32(pcp_ctx *c, int which, int *ncomp) {
  *ncomp = 3;
  FieldsT<float> &f = c->s32->f;
  switch (which) {
    case PCP_GRID_E: return f.E;
    case PCP_GRID_H: return f.H;
    case PCP_GRID_H_AT_ET: return f.Hat;
    case PCP_GRID_J: return f.J;
    case PCP_GRID_EPS: *ncomp = 1; return f.eps;
    case PCP_GRID_SIGMA: *ncomp = 1; return f.sigma;
    default: return nullptr;
  }
}

int pcp_download_grid(pcp_ctx *c, int which, double *host) {
  if (!c || !host) return fail(PCP_E_ARG, "bad argument");
  int nc;
  if (c->s32) {
    float *p = grid_ptr32(c, which, &nc);
    if (!p) return fail(PCP_E_ARG, "unknown grid");
    HIP_TRY(hipStreamSynchronize(c->stream));
    size_t n = (size_t)c->D * nc * c->g.P;
    std::vector<float> t(n);
    HIP_TRY(hipMemcpy(t.data(), p, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) host[i] = t[i];
    return 0;
  }
  double *p = grid_ptr(c, which, &nc);
  if (!p) return fail(PCP_E_ARG, "unknown grid");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(host, p, sizeof(double) * (size_t)c->D * nc * c->g.P, hipMemcpyDeviceToHost));
  return 0;
}

int pcp_species_map(pcp_ctx *c, int sp, int kind, double *host) {
  if (!c || sp < 0 || sp >= c->cfg.n_species || kind < PCP_MAP_DENSITY || kind > PCP_MAP_MOMENTUM)
    return fail(PCP_E_ARG, "bad argument");
  const size_t DP = (size_t)c->D * c->g.P;
  int rc;
  if (!c->map_n && ((rc = dalloc(&c->map_n, DP)) || (rc = dalloc(&c->map_pw, 3 * DP)) ||
                    (rc = dalloc(&c->map_out, 3 * DP))))
    return rc;
  if (!c->map_pabs[sp]) {
    if ((rc = dalloc(&c->map_pabs[sp], DP))) return rc;
    HIP_TRY(hipMemsetAsync(c->map_pabs[sp], 0, DP * sizeof(double), c->stream));
  }
  if (launch_species_map(c, sp, kind, c->map_n, c->map_pw, c->map_pabs[sp], c->map_out) < 0)
    return fail(PCP_E_HIP, "species map launch failed");
  c->map_ncomp = kind == PCP_MAP_MOMENTUM ? 3 : 1;
  if (host) {
    HIP_TRY(hipMemcpyAsync(host, c->map_out, sizeof(double) * DP * c->map_ncomp, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_probe(pcp_ctx *c, int which, int comp, long r0, long z0, long r1, long z1, double *host) {
  if (!c || !host) return fail(PCP_E_ARG, "bad argument");
  const Geom &g = c->g;
  if (!(r0 >= 0 && z0 >= 0 && r0 <= r1 && z0 <= z1 && r1 <= g.Nr && z1 <= g.Nz))
    return fail(PCP_E_ARG, "probe window outside the grid");
  int nc;
  bool f32 = c->s32 != nullptr;
  const void *grid;
  if (which == PCP_GRID_MAP) {
    if (!c->map_ncomp) return fail(PCP_E_STATE, "no species map computed yet");
    grid = c->map_out;
    nc = c->map_ncomp;
    f32 = false;
  } else {
    grid = f32 ? (const void *)grid_ptr32(c, which, &nc) : (const void *)grid_ptr(c, which, &nc);
  }
  if (!grid) return fail(PCP_E_ARG, "unknown grid");
  if (comp < 0 || comp >= nc) return fail(PCP_E_ARG, "bad component");
  const long n = (r1 - r0) * (z1 - z0);
  if (n == 0) return 0;
  if (n > c->probe_cap) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(c->probe_buf);
    c->probe_buf = nullptr;
    c->probe_cap = 0;
    if (int rc = dalloc(&c->probe_buf, (size_t)n)) return rc;
    c->probe_cap = n;
  }
  launch_probe(c, grid, f32, nc, comp, r0, z0, r1 - r0, z1 - z0, c->probe_buf);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(host, c->probe_buf, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_upload_grid(pcp_ctx *c, int which, const double *host) {
  if (!c || !host) return fail(PCP_E_ARG, "bad argument");
  int nc;
  if (c->s32) {
    float *p = grid_ptr32(c, which, &nc);
    if (!p) return fail(PCP_E_ARG, "unknown grid");
    HIP_TRY(hipStreamSynchronize(c->stream));
    size_t n = (size_t)c->D * nc * c->g.P;
    std::vector<float> t(host, host + n);
    HIP_TRY(hipMemcpy(p, t.data(), n * sizeof(float), hipMemcpyHostToDevice));
    return 0;
  }
  double *p = grid_ptr(c, which, &nc);
  if (!p) return fail(PCP_E_ARG, "unknown grid");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(p, host, sizeof(double) * (size_t)c->D * nc * c->g.P, hipMemcpyHostToDevice));
  return 0;
}

int pcp_yee_e(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (c->s32) launch32_yee_e(c, 0);
  else launch_yee_e(c, 0);
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

int pcp_yee_h(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (c->s32) launch32_yee_h(c);
  else launch_yee_h(c);
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

int pcp_stitch(pcp_ctx *c, int which) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (which != PCP_GRID_E && which != PCP_GRID_H && which != PCP_GRID_H_AT_ET && which != PCP_GRID_J)
    return fail(PCP_E_ARG, "stitch: which must be E, H or J");
  if (c->s32) launch32_stitch(c, which);
  else launch_stitch(c, which);
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

int pcp_push(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (c->s32) return fail(PCP_E_ARG, "the phase entry points run FP64 contexts; FP32 split contexts advance with pcp_step");
  for (int s = 0; s < c->cfg.n_species; ++s) launch_push(c, s);
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

int pcp_move(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (c->s32) return fail(PCP_E_ARG, "the phase entry points run FP64 contexts; FP32 split contexts advance with pcp_step");
  for (int s = 0; s < c->cfg.n_species; ++s) launch_move(c, s);
  HIP_TRY(hipGetLastError());
  c->cur ^= 1;
  return pcp_sync(c);
}

int pcp_migrate(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (c->s32) return fail(PCP_E_ARG, "the phase entry points run FP64 contexts; FP32 split contexts advance with pcp_step");
  launch_step_begin(c);   // resets the migrant counter (and advances the stamp)
  for (int s = 0; s < c->cfg.n_species; ++s) launch_migrate_mark(c, s);
  launch_migrant_keys(c);
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

int pcp_deposit(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (c->s32) return fail(PCP_E_ARG, "the phase entry points run FP64 contexts; FP32 split contexts advance with pcp_step");
  launch_stamp(c);
  launch_fill(c, c->f.J, (long)c->D * 3 * c->g.P, 0.0);   // Domain::reset_current
  for (int s = 0; s < c->cfg.n_species; ++s) launch_mark(c, s, 0);
  for (int s = 0; s < c->cfg.n_species; ++s) launch_deposit_exact(c, s, 0);
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

static int check_coll_args(pcp_ctx *c, int scheme, int se, int si) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (scheme < PCP_COLL_P12 || scheme > PCP_COLL_SURROGATE) return fail(PCP_E_ARG, "unknown collision scheme");
  if (scheme == PCP_COLL_SURROGATE && (!c->mlp || !c->mlp->trained))
    return fail(PCP_E_STATE, "the collision surrogate has no fitted network (pcp_surrogate_fit)");
  if (se < 0 || si < 0 || se >= c->cfg.n_species || si >= c->cfg.n_species || se == si)
    return fail(PCP_E_ARG, "collisions need two distinct species (electrons, ions)");
  return 0;
}

int pcp_collide(pcp_ctx *c, int scheme, int se, int si, uint64_t seed, uint64_t pass) {
  if (int rc = check_coll_args(c, scheme, se, si)) return rc;
  if (launch_collide(c, scheme, se, si, seed, PassNo{nullptr, pass}, c->cur) < 0)
    return fail(PCP_E_HIP, "collision pass failed");
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

static int need_mlp(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (mlp_alloc(c) < 0) return fail(PCP_E_NOMEM, "network buffers");
  return 0;
}

int pcp_mlp_nparams(void) { return MLP_NPARAM; }

int pcp_mlp_init(pcp_ctx *c, uint64_t seed) {
  if (c) graph_invalidate(c);   // captured steps hold the normalisation by value
  if (int rc = need_mlp(c)) return rc;
  if (mlp_init(c, seed) < 0) return fail(PCP_E_HIP, "network initialisation");
  c->mlp->trained = 0;
  return 0;
}

int pcp_mlp_get_params(pcp_ctx *c, float *host) {
  if (int rc = need_mlp(c)) return rc;
  if (!host) return fail(PCP_E_ARG, "null buffer");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(host, c->mlp->params, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int pcp_mlp_set_params(pcp_ctx *c, const float *host) {
  if (c) graph_invalidate(c);   // captured steps hold the normalisation by value
  if (int rc = need_mlp(c)) return rc;
  if (!host) return fail(PCP_E_ARG, "null buffer");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->mlp->params, host, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  c->mlp->inited = 1;
  return 0;
}

static int mlp_batch_buffers(pcp_ctx *c, long n) {
  MlpState *m = c->mlp;
  if (m->bcap >= n) return 0;
  (void)hipFree(m->xb); (void)hipFree(m->tb); (void)hipFree(m->yb);
  m->xb = m->tb = m->yb = nullptr;
  HIP_TRY(hipMalloc(&m->xb, n * MLP_XSTRIDE * sizeof(float)));
  HIP_TRY(hipMalloc(&m->tb, n * MLP_YSTRIDE * sizeof(float)));
  HIP_TRY(hipMalloc(&m->yb, n * MLP_YSTRIDE * sizeof(float)));
  m->bcap = n;
  return 0;
}

int pcp_mlp_forward(pcp_ctx *c, const float *X, long n, float *Y) {
  if (int rc = need_mlp(c)) return rc;
  if (n <= 0 || !X || !Y) return fail(PCP_E_ARG, "bad batch");
  if (int rc = mlp_batch_buffers(c, n)) return rc;
  MlpState *m = c->mlp;
  HIP_TRY(hipMemcpy(m->xb, X, n * MLP_XSTRIDE * sizeof(float), hipMemcpyHostToDevice));
  mlp_forward(c, m->xb, n, m->yb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(Y, m->yb, n * MLP_YSTRIDE * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int pcp_mlp_grad(pcp_ctx *c, const float *X, const float *T, long n, float *grad, float *loss) {
  if (int rc = need_mlp(c)) return rc;
  if (n <= 0 || !X || !T) return fail(PCP_E_ARG, "bad batch");
  if (int rc = mlp_batch_buffers(c, n)) return rc;
  MlpState *m = c->mlp;
  HIP_TRY(hipMemcpy(m->xb, X, n * MLP_XSTRIDE * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m->tb, T, n * MLP_YSTRIDE * sizeof(float), hipMemcpyHostToDevice));
  mlp_grad(c, m->xb, m->tb, n, m->yb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (grad) HIP_TRY(hipMemcpy(grad, m->grads, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) HIP_TRY(hipMemcpy(loss, m->yb, sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int pcp_mlp_adam(pcp_ctx *c, float lr) {
  if (int rc = need_mlp(c)) return rc;
  mlp_adam(c, lr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_surrogate_collect(pcp_ctx *c, int se, int si, uint64_t seed, uint64_t pass, long *rows) {
  if (int rc = check_coll_args(c, PCP_COLL_P12, se, si)) return rc;
  if (int rc = need_mlp(c)) return rc;
  if (mlp_reserve(c, (long)c->D * c->g.nr * c->g.nz) < 0) return fail(PCP_E_NOMEM, "training set");
  if (surrogate_collect(c, se, si, seed, pass) < 0) return fail(PCP_E_HIP, "collection pass failed");
  HIP_TRY(hipGetLastError());
  if (rows) *rows = c->mlp->n;
  return pcp_sync(c);
}

int pcp_surrogate_fit(pcp_ctx *c, int epochs, long batch, float lr, uint64_t seed, float *loss) {
  if (c) graph_invalidate(c);   // captured steps hold the normalisation by value
  if (int rc = need_mlp(c)) return rc;
  if (epochs < 1 || batch < 1) return fail(PCP_E_ARG, "epochs and batch must be positive");
  int rc = mlp_fit(c, epochs, batch, lr, seed, loss);
  if (rc == -2) return fail(PCP_E_STATE, "the training set is too small (pcp_surrogate_collect first)");
  if (rc < 0) return fail(PCP_E_HIP, "training failed");
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

int pcp_surrogate_get_norm(pcp_ctx *c, float *out) {
  if (int rc = need_mlp(c)) return rc;
  MlpState *m = c->mlp;
  for (int k = 0; k < MLP_IN; ++k) { out[k] = m->feat_mean[k]; out[MLP_IN + k] = m->feat_std[k]; }
  for (int k = 0; k < MLP_OUT; ++k) out[2 * MLP_IN + k] = m->targ_scale[k];
  return 0;
}

int pcp_surrogate_set_norm(pcp_ctx *c, const float *in) {
  if (c) graph_invalidate(c);   // captured steps hold the normalisation by value
  if (int rc = need_mlp(c)) return rc;
  MlpState *m = c->mlp;
  for (int k = 0; k < MLP_IN; ++k) {
    if (!(in[MLP_IN + k] > 0)) return fail(PCP_E_ARG, "feature scales must be positive");
    m->feat_mean[k] = in[k];
    m->feat_std[k] = in[MLP_IN + k];
  }
  for (int k = 0; k < MLP_OUT; ++k) m->targ_scale[k] = in[2 * MLP_IN + k];
  m->trained = 1;
  return 0;
}

// relaxation rate of the electrons' anisotropy <dvz^2> - <dvr^2 + dvphi^2> / 2
// (proportional to T_par - T_perp) over `passes` collision passes: minus
// the least-squares slope of its logarithm per pass
static int anisotropy_rate(pcp_ctx *c, int scheme, int se, int si, uint64_t seed, int passes, double *rate) {
  double sx = 0, sy = 0, sxx = 0, sxy = 0;
  for (int k = 0; k <= passes; ++k) {
    double m[7];
    if (int rc = pcp_species_moments(c, se, m)) return rc;
    const double sw = m[0], vr = m[1] / sw, vp = m[2] / sw, vz = m[3] / sw;
    const double a = (m[4] / sw - vz * vz) - 0.5 * (m[5] / sw - vr * vr - vp * vp);
    if (!(a > 0) || !std::isfinite(a))
      return fail(PCP_E_STATE, "calibration needs an electron T_par above T_perp throughout the passes");
    const double x = k, y = std::log(a);
    sx += x; sy += y; sxx += x * x; sxy += x * y;
    if (k < passes)
      if (int rc = pcp_collide(c, scheme, se, si, seed, (uint64_t)k)) return rc;
  }
  const double n = passes + 1;
  *rate = -(n * sxy - sx * sy) / (n * sxx - sx * sx);
  return 0;
}

namespace {
struct SpeciesCopy {
  std::vector<double> r, z, vr, vp, vz, w;
  std::vector<uint64_t> id;
};
}  // namespace

// One run of `passes` passes of `scheme` on a scratch context of c's
// configuration holding copies of species se and si; for the surrogate the
// scratch network is c's with the electron target scales times `fac`.
static int scratch_rate(pcp_ctx *c, const SpeciesCopy cp[2], int scheme, int se, int si, uint64_t seed, int passes,
                        const std::vector<float> &params, const float norm[16], double fac, double *rate) {
  pcp_ctx *t = nullptr;
  if (int rc = pcp_create(&c->cfg, &t)) return rc;
  int rc = 0;
  const int sp[2] = {se, si};
  for (int j = 0; j < 2 && !rc; ++j)
    rc = pcp_upload_species(t, sp[j], (long)cp[j].r.size(), cp[j].r.data(), cp[j].z.data(), cp[j].vr.data(),
                            cp[j].vp.data(), cp[j].vz.data(), cp[j].w.data(), cp[j].id.data());
  if (!rc && scheme == PCP_COLL_SURROGATE) {
    float nm[16];
    std::copy(norm, norm + 16, nm);
    nm[2 * MLP_IN + 0] = (float)(nm[2 * MLP_IN + 0] * fac);
    nm[2 * MLP_IN + 1] = (float)(nm[2 * MLP_IN + 1] * fac);
    rc = pcp_mlp_set_params(t, params.data());
    if (!rc) rc = pcp_surrogate_set_norm(t, nm);
  }
  if (!rc) rc = anisotropy_rate(t, scheme, se, si, seed, passes, rate);
  (void)pcp_destroy(t);
  return rc;
}

int pcp_surrogate_calibrate(pcp_ctx *c, int se, int si, uint64_t seed, int passes, double *factor) {
  if (c) graph_invalidate(c);   // captured steps hold the normalisation by value
  if (int rc = check_coll_args(c, PCP_COLL_SURROGATE, se, si)) return rc;
  if (passes < 2) return fail(PCP_E_ARG, "calibration needs at least 2 passes");
  // copies of the two species as they are now
  SpeciesCopy cp[2];
  const int sp[2] = {se, si};
  for (int j = 0; j < 2; ++j) {
    long n = 0;
    if (int rc = pcp_count(c, sp[j], &n)) return rc;
    if (n == 0) return fail(PCP_E_STATE, "calibration needs particles of both species");
    SpeciesCopy &q = cp[j];
    q.r.resize(n); q.z.resize(n); q.vr.resize(n); q.vp.resize(n); q.vz.resize(n); q.w.resize(n); q.id.resize(n);
    if (int rc = pcp_download_species(c, sp[j], q.r.data(), q.z.data(), q.vr.data(), q.vp.data(), q.vz.data(),
                                      q.w.data(), q.id.data(), nullptr, nullptr, nullptr))
      return rc;
  }
  std::vector<float> params(MLP_NPARAM);
  float norm[16];
  if (int rc = pcp_mlp_get_params(c, params.data())) return rc;
  if (int rc = pcp_surrogate_get_norm(c, norm)) return rc;
  double r_ref = 0, fac = 1.0;
  if (int rc = scratch_rate(c, cp, PCP_COLL_P12, se, si, seed, passes, params, norm, 1.0, &r_ref)) return rc;
  // the rate is close to proportional to the electron output scale: two
  // fixed-point iterations
  for (int it = 0; it < 2; ++it) {
    double r = 0;
    if (int rc = scratch_rate(c, cp, PCP_COLL_SURROGATE, se, si, seed, passes, params, norm, fac, &r)) return rc;
    if (!(r > 0)) return fail(PCP_E_STATE, "the surrogate does not relax the anisotropy: no calibration possible");
    fac *= r_ref / r;
  }
  c->mlp->targ_scale[0] = (float)(c->mlp->targ_scale[0] * fac);
  c->mlp->targ_scale[1] = (float)(c->mlp->targ_scale[1] * fac);
  if (factor) *factor = fac;
  return 0;
}

int pcp_surrogate_fallback_cells(pcp_ctx *c, long *cells) {
  if (!c || !cells) return fail(PCP_E_ARG, "bad argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  *cells = surrogate_fallback_cells(c);
  return 0;
}

int pcp_set_collisions(pcp_ctx *c, int scheme, int se, int si, uint64_t seed) {
  if (!c) return fail(PCP_E_ARG, "null context");
  graph_invalidate(c);
  if (scheme < 0) {
    c->coll_scheme = -1;
    return 0;
  }
  if (int rc = check_coll_args(c, scheme, se, si)) return rc;
  // the phase's buffers exist before any step is captured
  if (coll_prepare(c, se, si) < 0) return fail(PCP_E_NOMEM, "collision buffers");
  if (scheme == COLL_SURROGATE && c->s32 && fsurr_prepare(c, se, si) < 0)
    return fail(PCP_E_NOMEM, "surrogate collision buffers");
  c->coll_scheme = scheme;
  c->coll_se = se;
  c->coll_si = si;
  c->coll_seed = seed;
  // passes count from 0 on the device
  launch_set_u64(c, c->coll_pass, 0);
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // extern "C"

// the kernels of one step after the sorts: solve_maxvell, advance_particles
// (with the collision phase between push and move), solve_current
// (src/PiCoPiC.cpp:299-345); flips c->cur
template <class Mark>
static int step_body(pcp_ctx *c, Mark mark) {
  const bool f32 = c->s32 != nullptr;
  launch_step_begin(c);
  if (f32) {
    launch32_yee_e(c, 1);
    launch32_stitch(c, PCP_GRID_E);
    launch32_yee_h(c);
    launch32_stitch(c, PCP_GRID_H);
  } else {
    launch_yee_e(c, 1);
    launch_stitch(c, PCP_GRID_E);
    launch_yee_h(c);
    launch_stitch(c, PCP_GRID_H);
  }
  if (c->elog) launch_field_energy(c, c->ehist, true);
  mark(2);
  const bool coll = c->coll_scheme >= 0;
  // FP32 with the surrogate: the fused phase (moments in the push, per-cell
  // rescalings applied by the move kernel)
  const bool fused = coll && f32 && c->coll_scheme == COLL_SURROGATE;
  const float *cpar = fused ? fsurr_par(c) : nullptr;
  if (fused) {
    double *mom = fsurr_mom(c);
    for (int s = 0; s < c->cfg.n_species; ++s)
      launch32_push_mom(c, s, mom, s == c->coll_se ? 0 : s == c->coll_si ? 1 : -1);
    mark(3);
    if (fsurr_cells(c, c->coll_se, c->coll_si, c->coll_seed, PassNo{c->coll_pass, 0}) < 0)
      return fail(PCP_E_HIP, "surrogate collision phase failed");
    launch_inc_u64(c, c->coll_pass);
  } else if (coll) {
    // SMB::advance_particles: push, collide (before the move), move
    for (int s = 0; s < c->cfg.n_species; ++s) {
      if (f32) launch32_push(c, s);
      else launch_push(c, s);
    }
    mark(3);
    if (launch_collide(c, c->coll_scheme, c->coll_se, c->coll_si, c->coll_seed, PassNo{c->coll_pass, 0},
                       c->cur ^ 1) < 0)
      return fail(PCP_E_HIP, "collision pass failed");
    launch_inc_u64(c, c->coll_pass);
  } else {
    mark(3);   // the push runs in the particle kernel
  }
  mark(4);
  if (f32) launch32_step_batch(c, !coll, cpar, c->coll_se, c->coll_si);
  else
    for (int s = 0; s < c->cfg.n_species; ++s) launch_step_fused(c, s, !coll);
  c->cur ^= 1;
  mark(5);
  if (f32) launch32_migrant_keys(c);
  else launch_migrant_keys(c);
  // strict axial motions reset J_r nodes: take back the contributions of
  // the particles that precede the resetting one (device-side no-op when no
  // reset happened)
  if (c->g.ccs == PCP_CCS_VB) {
    if (f32) launch32_fix(c, !coll, cpar, c->coll_se, c->coll_si);
    else
      for (int s = 0; s < c->cfg.n_species; ++s) launch_fix(c, s);
  }
  if (f32) launch32_fold_j(c);
  mark(6);
  if (f32) launch32_stitch(c, PCP_GRID_J);
  else launch_stitch(c, PCP_GRID_J);
  // FP64 kinetic energies of the log (FP32 sums them in the particle kernel)
  if (c->elog && !f32)
    for (int s = 0; s < c->cfg.n_species; ++s) launch_kinetic_energy_log(c, s);
  mark(7);
  return 0;
}

static void graph_invalidate(pcp_ctx *c) {
  for (auto &e : c->graphs)
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
  c->graphs.clear();
}

constexpr size_t GRAPH_CACHE = 8;

// what a captured step depends on besides the device state: `cur`, and
// each species' count and buffer pointers (exchanged by the sorts and by
// the velocity-slot exchange of collision steps)
static std::vector<uintptr_t> step_key(const pcp_ctx *c) {
  std::vector<uintptr_t> k;
  k.push_back((uintptr_t)c->cur);
  for (int s = 0; s < c->cfg.n_species; ++s) {
    if (c->s32) {
      const Species32 &q = c->s32->sp[s];
      const void *p[] = {q.cell[0], q.cell[1], q.fr[0], q.fr[1], q.fz[0], q.fz[1], q.vr[0], q.vr[1],
                         q.vphi[0], q.vphi[1], q.vz[0], q.vz[1], q.w, q.seq, q.id};
      k.push_back((uintptr_t)q.n);
      for (const void *x : p) k.push_back((uintptr_t)x);
    } else {
      const SpeciesDev &q = c->sp[s];
      const void *p[] = {q.r[0], q.r[1], q.z[0], q.z[1], q.vr, q.vphi, q.vz, q.w, q.seq, q.id};
      k.push_back((uintptr_t)q.n);
      for (const void *x : p) k.push_back((uintptr_t)x);
    }
  }
  return k;
}

// the host-side effects of step_body, which a graph replay does not run:
// the velocity-slot exchange of FP32 collision steps and the flip of `cur`
static void step_host_effects(pcp_ctx *c) {
  if (c->s32 && c->coll_scheme >= 0)
    for (int s = 0; s < c->cfg.n_species; ++s) {
      Species32 &q = c->s32->sp[s];
      std::swap(q.vr[0], q.vr[1]);
      std::swap(q.vphi[0], q.vphi[1]);
      std::swap(q.vz[0], q.vz[1]);
    }
  c->cur ^= 1;
}

// the executable graph of the current state: a cached one, or a new capture
// of step_body (the host state it changes is restored, the caller applies
// step_host_effects after launching)
static int step_graph(pcp_ctx *c, hipGraphExec_t *out) {
  std::vector<uintptr_t> key = step_key(c);
  for (auto &e : c->graphs)
    if (e.key == key) {
      e.used = ++c->graph_tick;
      *out = e.exec;
      return 0;
    }
  const int cur = c->cur;
  std::vector<Species32> saved32;
  if (c->s32) saved32.assign(c->s32->sp, c->s32->sp + c->cfg.n_species);
  hipGraph_t graph;
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  int rc = step_body(c, [](int) {});
  hipError_t e = hipStreamEndCapture(c->stream, &graph);
  c->cur = cur;   // nothing ran during the capture
  if (c->s32) std::copy(saved32.begin(), saved32.end(), c->s32->sp);
  if (rc) return rc;
  if (e != hipSuccess) return fail(PCP_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(PCP_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  if (c->graphs.size() >= GRAPH_CACHE) {
    auto lru = std::min_element(c->graphs.begin(), c->graphs.end(),
                                [](const pcp_ctx::StepGraph &a, const pcp_ctx::StepGraph &b) { return a.used < b.used; });
    (void)hipGraphExecDestroy(lru->exec);
    c->graphs.erase(lru);
  }
  c->graphs.push_back(pcp_ctx::StepGraph{std::move(key), exec, ++c->graph_tick});
  ++c->graph_captures;
  *out = exec;
  return 0;
}

extern "C" {

int pcp_collision_passes(pcp_ctx *c, uint64_t *passes) {
  if (!c || !passes) return fail(PCP_E_ARG, "bad argument");
  HIP_TRY(hipMemcpyAsync(passes, c->coll_pass, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_graph_stats(pcp_ctx *c, long *launches, long *captures) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (launches) *launches = c->graph_launches;
  if (captures) *captures = c->graph_captures;
  return 0;
}

int pcp_set_graph(pcp_ctx *c, int on) {
  if (!c) return fail(PCP_E_ARG, "null context");
  c->graph_on = on != 0;
  if (!c->graph_on) graph_invalidate(c);
  return 0;
}

int pcp_set_compensated_current(pcp_ctx *c, int on) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (!c->s32) return fail(PCP_E_ARG, "compensated deposition is a mode of FP32 split contexts");
  graph_invalidate(c);
  if (set_compensated32(c, on != 0)) return fail(PCP_E_HIP, "cannot allocate the compensation array");
  return 0;
}

int pcp_set_beam(pcp_ctx *c, int sp, long macro_amount, double start_time, double radius, double density,
                 int bunches_amount, double bunch_length, double bunches_distance, double velocity, uint64_t seed) {
  if (!c || sp < 0 || sp >= c->cfg.n_species || macro_amount <= 0 || macro_amount > 0xFFFFFFFFL ||
      bunches_amount <= 0 || !(velocity > 0) || !(radius > 0) || !(bunch_length > 0) || !(density >= 0))
    return fail(PCP_E_ARG, "bad argument");
  if (!c->sp[sp].beam) return fail(PCP_E_ARG, "the species was not created as a beam");
  const Geom &g = c->g;
  BeamHost &b = c->beam[sp];
  b.on = 1;
  b.macro_amount = (unsigned int)macro_amount;
  b.bunches_amount = (unsigned int)bunches_amount;
  b.start_time = start_time;
  b.radius = radius;
  b.density = density;
  b.bunch_length = bunch_length;
  b.bunches_distance = bunches_distance;
  b.velocity = velocity;
  b.seed = seed;
  b.dom.assign(c->D, BeamDomHost{});
  // BeamP::BeamP in every domain (src/beamP.cpp:40-66)
  for (int d = 0; d < c->D; ++d) {
    const Dom dm = make_dom(g, d);
    BeamDomHost &B = b.dom[d];
    const long top_r = dm.bot_r + g.nr;
    if (dm.left_z > 0)
      B.domain_radius = 0;
    else if (radius > top_r * g.dr)
      B.domain_radius = g.size_r;
    else if (radius > dm.bot_r * g.dr)
      B.domain_radius = radius - dm.bot_r * g.dr;
    else
      B.domain_radius = 0;
    const double macro_normalizer = B.domain_radius / radius;
    const double bunch_macro_amount = (double)(b.macro_amount / b.bunches_amount) * macro_normalizer;
    B.macro_per_step = bunch_macro_amount * velocity * g.dt / bunch_length;
  }
  return 0;
}

int pcp_inject(pcp_ctx *c) {
  if (!c) return fail(PCP_E_ARG, "null context");
  long n = inject_beams(c);
  if (n < 0) return (int)n;
  HIP_TRY(hipGetLastError());
  return 0;
}

int pcp_set_time(pcp_ctx *c, double t) {
  if (!c) return fail(PCP_E_ARG, "null context");
  c->t_now = t;
  return 0;
}

double pcp_time(pcp_ctx *c) { return c ? c->t_now : 0.0; }

int pcp_step(pcp_ctx *c, int n_steps) {
  if (!c || n_steps < 0) return fail(PCP_E_ARG, "bad argument");
  // optional per-phase device timing: events on the context stream
  // sort: the step's particle sorts; fields: Yee + stitches; push: the push
  // of a step with collisions (with the cell moments of the surrogate);
  // collide: its collision phase; particles: the fused particle kernel(s)
  constexpr int NPH = 7;
  static const char *names[NPH] = {"sort", "fields", "push", "collide", "particles", "fixup", "stitch_j"};
  EventSet evs;
  if (c->timing) HIP_TRY(evs.create((size_t)n_steps * (NPH + 1)));
  std::vector<hipEvent_t> &ev = evs.ev;
  const bool f32 = c->s32 != nullptr;
  for (int t = 0; t < n_steps; ++t) {
    auto mark = [&](int k) {
      if (c->timing) (void)hipEventRecord(ev[(size_t)t * (NPH + 1) + k], c->stream);
    };
    // SMB::inject_beam first (src/PiCoPiC.cpp:320), outside any graph
    if (long ni = inject_beams(c); ni < 0) return (int)ni;
    mark(0);
    // the sorts exchange particle buffers; they run outside the step graph,
    // whose cache is keyed by the buffers
    for (int s = 0; s < c->cfg.n_species; ++s)
      if (c->sort_every[s] > 0 && c->steps_done % c->sort_every[s] == 0)
        if ((f32 ? launch32_sort(c, s) : launch_sort(c, s)) < 0) return fail(PCP_E_HIP, "sort failed");
    mark(1);
    if (c->graph_on) {
      // one graph launch for the step; with timing on, events 0 and NPH
      // bracket the sorts and the launch, and the phases inside the graph
      // are not separated (they read as zero)
      hipGraphExec_t exec = nullptr;
      if (int rc = step_graph(c, &exec)) return rc;
      HIP_TRY(hipGraphLaunch(exec, c->stream));
      mark(NPH);
      ++c->graph_launches;
      step_host_effects(c);
    } else {
      if (int rc = step_body(c, mark)) return rc;
    }
    ++c->steps_done;
    c->t_now += c->g.dt;
  }
  HIP_TRY(hipGetLastError());
  int rc = pcp_sync(c);
  if (c->timing && n_steps > 0) {
    double acc[NPH] = {}, tot = 0;
    for (int t = 0; t < n_steps; ++t) {
      float ms = 0;
      const size_t e0 = (size_t)t * (NPH + 1);
      if (c->graph_on) {
        (void)hipEventElapsedTime(&ms, ev[e0], ev[e0 + NPH]);
        tot += ms;
        continue;
      }
      for (int k = 0; k < NPH; ++k) {
        (void)hipEventElapsedTime(&ms, ev[e0 + k], ev[e0 + k + 1]);
        acc[k] += ms;
        tot += ms;
      }
    }
    c->ktimes.clear();
    for (int k = 0; k < NPH; ++k) c->ktimes.emplace_back(names[k], acc[k] / n_steps);
    c->ktimes.emplace_back("step", tot / n_steps);
  }
  return rc;
}

int pcp_sort(pcp_ctx *c, int sp) {
  if (c) graph_invalidate(c);
  if (!c || sp < 0 || sp >= c->cfg.n_species) return fail(PCP_E_ARG, "bad argument");
  if ((c->s32 ? launch32_sort(c, sp) : launch_sort(c, sp)) < 0) return fail(PCP_E_HIP, "sort failed");
  HIP_TRY(hipGetLastError());
  return pcp_sync(c);
}

int pcp_set_sort_interval(pcp_ctx *c, int sp, int every) {
  if (!c || sp < 0 || sp >= c->cfg.n_species || every < 0) return fail(PCP_E_ARG, "bad argument");
  c->sort_every[sp] = every;
  return 0;
}

int pcp_set_sort_bin(pcp_ctx *c, int bin_r, int bin_z) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (bin_r < 1 || bin_z < 1 || c->g.nr % bin_r || c->g.nz % bin_z)
    return fail(PCP_E_ARG, "sort bin must divide the domain's cells (nr, nz)");
  graph_invalidate(c);
  c->g.bin_r = bin_r;
  c->g.bin_z = bin_z;
  c->g.nbz = c->g.nz / bin_z;
  c->g.div_bin_r = make_fastdiv((uint32_t)bin_r);
  c->g.div_bin_z = make_fastdiv((uint32_t)bin_z);
  return 0;
}

int pcp_set_launch(pcp_ctx *c, int block, int blocks_per_cu) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if ((block != 128 && block != 256 && block != 512) || blocks_per_cu < 0 || blocks_per_cu > 32)
    return fail(PCP_E_ARG, "block must be 128, 256 or 512 and blocks_per_cu in [0, 32]");
  graph_invalidate(c);
  c->tune.block = block;
  c->tune.blocks_per_cu = blocks_per_cu;
  return 0;
}

int pcp_get_tuning(pcp_ctx *c, pcp_tuning *out) {
  if (!c || !out) return fail(PCP_E_ARG, "bad argument");
  out->bin_r = c->g.bin_r;
  out->bin_z = c->g.bin_z;
  out->block = c->tune.block;
  out->blocks_per_cu = c->tune.blocks_per_cu;
  out->sort_every = 0;
  for (int s = 0; s < c->cfg.n_species; ++s)
    if (!c->cfg.species[s].is_beam) out->sort_every = std::max(out->sort_every, c->sort_every[s]);
  out->ms_per_step = 0;
  return 0;
}

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// sort every species, then time `steps` steps with no sort among them
int tune_trial(pcp_ctx *c, int steps, double *ms) {
  for (int s = 0; s < c->cfg.n_species; ++s)
    if (int rc = pcp_sort(c, s)) return rc;
  const double t0 = now_ms();
  if (int rc = pcp_step(c, steps)) return rc;
  *ms = (now_ms() - t0) / steps;
  return 0;
}

// Times each of n candidates (set(i) applies candidate i) twice, in forward
// and then in reverse order, and picks the least sum: the simulation advances
// with every trial and its step time drifts, which the two orders cancel to
// first order (one forward pass favoured whichever candidates came first).
extern "C++" {
template <class Set>
int tune_pick(pcp_ctx *c, int steps, int n, Set set, int *best_i, double *best_ms) {
  std::vector<double> score((size_t)n, 0.0);
  for (int pass = 0; pass < 2; ++pass)
    for (int j = 0; j < n; ++j) {
      const int i = pass ? n - 1 - j : j;
      if (int rc = set(i)) return rc;
      double ms;
      if (int rc = tune_trial(c, steps, &ms)) return rc;
      score[(size_t)i] += ms;
    }
  *best_i = (int)(std::min_element(score.begin(), score.end()) - score.begin());
  *best_ms = score[(size_t)*best_i] / 2;
  return 0;
}
}  // extern "C++"

}  // namespace

// The tuner times candidates on the live state, so the simulation advances
// by every step it takes (about 80 x steps + 24 in all).
int pcp_autotune(pcp_ctx *c, int steps, pcp_tuning *out) {
  if (!c || steps < 1 || steps > 1000) return fail(PCP_E_ARG, "bad argument");
  int saved[MAX_SPECIES];
  for (int s = 0; s < c->cfg.n_species; ++s) {
    saved[s] = c->sort_every[s];
    c->sort_every[s] = 0;
  }
  const int timing = c->timing;
  c->timing = 0;
  auto restore = [&](int rc) {
    for (int s = 0; s < c->cfg.n_species; ++s) c->sort_every[s] = saved[s];
    c->timing = timing;
    return rc;
  };
  const Geom &g = c->g;
  // 1. sort bin, at the default launch shape
  std::vector<std::pair<int, int>> bins = {{1, g.nz}};
  for (int br : {1, 2, 4, 8, 16})
    for (int bz : {4, 8, 16, 32, 64})
      if (br * bz >= 8 && br * bz <= 512 && g.nr % br == 0 && g.nz % bz == 0 && bz < g.nz) bins.push_back({br, bz});
  double best = 1e300;
  int bi = 0;
  if (int rc = tune_pick(c, steps, (int)bins.size(),
                         [&](int i) { return pcp_set_sort_bin(c, bins[(size_t)i].first, bins[(size_t)i].second); },
                         &bi, &best))
    return restore(rc);
  if (int rc = pcp_set_sort_bin(c, bins[(size_t)bi].first, bins[(size_t)bi].second)) return restore(rc);
  // 2. launch shape of the batched FP32 kernel
  if (c->s32) {
    static const int shapes[9][2] = {{128, 0}, {128, 1}, {128, 2}, {256, 0}, {256, 1},
                                     {256, 2}, {512, 0}, {512, 1}, {512, 2}};
    int li = 0;
    if (int rc = tune_pick(c, steps, 9, [&](int i) { return pcp_set_launch(c, shapes[i][0], shapes[i][1]); }, &li,
                           &best))
      return restore(rc);
    if (int rc = pcp_set_launch(c, shapes[li][0], shapes[li][1])) return restore(rc);
  }
  // 3. sort interval of the plasma species: the step time t(j) j steps
  // after a sort grows as the particles leave cell order; with t(j) ~ a + b j
  // and sort time t_s, an interval k costs t_s / k + a + b (k - 1) / 2 per
  // step, least at k = sqrt(2 t_s / b)
  // beams keep their own interval while the plasma order ages
  for (int s = 0; s < c->cfg.n_species; ++s)
    if (c->cfg.species[s].is_beam) c->sort_every[s] = saved[s];
  double t_sort = 0;
  {
    if (int rc = pcp_sync(c)) return restore(rc);
    const double t0 = now_ms();
    for (int s = 0; s < c->cfg.n_species; ++s)
      if (!c->cfg.species[s].is_beam)
        if (int rc = pcp_sort(c, s)) return restore(rc);
    t_sort = now_ms() - t0;
  }
  const int J = 24;
  double sj = 0, st = 0, sjj = 0, sjt = 0;
  for (int j = 0; j < J; ++j) {
    const double t0 = now_ms();
    if (int rc = pcp_step(c, 1)) return restore(rc);
    const double t = now_ms() - t0;
    sj += j, st += t, sjj += (double)j * j, sjt += j * t;
  }
  const double b = (J * sjt - sj * st) / (J * sjj - sj * sj);
  int every = 200;
  if (b > 0) every = (int)std::lround(std::sqrt(2 * t_sort / b));
  every = std::max(1, std::min(200, every));
  for (int s = 0; s < c->cfg.n_species; ++s) saved[s] = c->cfg.species[s].is_beam ? saved[s] : every;
  restore(0);
  if (out) {
    pcp_get_tuning(c, out);
    out->ms_per_step = best;
  }
  return 0;
}

int pcp_get_field_e(pcp_ctx *c, int d, double r, double z, double *out) {
  if (!c || !out || d < 0 || d >= c->D) return fail(PCP_E_ARG, "bad argument");
  {
    Dom dm = make_dom(c->g, d);
    double r0 = dm.bot_r * c->g.dr, z0 = dm.left_z * c->g.dz;
    if (!(r >= r0 && r < r0 + c->g.nr * c->g.dr && z >= z0 && z < z0 + c->g.nz * c->g.dz))
      return fail(PCP_E_ARG, "point outside domain");
  }
  double *dev = c->diag;
  if (c->s32) launch32_gather_point(c, d, r, z, 0, dev);
  else launch_gather_point(c, d, r, z, 0, dev);
  HIP_TRY(hipMemcpyAsync(out, dev, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_get_field_h(pcp_ctx *c, int d, double r, double z, double *out) {
  if (!c || !out || d < 0 || d >= c->D) return fail(PCP_E_ARG, "bad argument");
  {
    Dom dm = make_dom(c->g, d);
    double r0 = dm.bot_r * c->g.dr, z0 = dm.left_z * c->g.dz;
    if (!(r >= r0 && r < r0 + c->g.nr * c->g.dr && z >= z0 && z < z0 + c->g.nz * c->g.dz))
      return fail(PCP_E_ARG, "point outside domain");
  }
  double *dev = c->diag;
  if (c->s32) launch32_gather_point(c, d, r, z, 1, dev);
  else launch_gather_point(c, d, r, z, 1, dev);
  HIP_TRY(hipMemcpyAsync(out, dev, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_field_energy(pcp_ctx *c, double *out) {
  if (!c || !out) return fail(PCP_E_ARG, "bad argument");
  double *d = c->diag;
  HIP_TRY(hipMemsetAsync(d, 0, sizeof(double), c->stream));
  launch_field_energy(c, d);
  HIP_TRY(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_kinetic_energy(pcp_ctx *c, int sp, double *out) {
  if (!c || !out || sp < 0 || sp >= c->cfg.n_species) return fail(PCP_E_ARG, "bad argument");
  double *d = c->diag;
  HIP_TRY(hipMemsetAsync(d, 0, sizeof(double), c->stream));
  launch_kinetic_energy(c, sp, d);
  HIP_TRY(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_species_moments(pcp_ctx *c, int sp, double *out) {
  if (!c || !out || sp < 0 || sp >= c->cfg.n_species) return fail(PCP_E_ARG, "bad argument");
  double *d = c->diag;
  HIP_TRY(hipMemsetAsync(d, 0, 7 * sizeof(double), c->stream));
  launch_species_moments(c, sp, d);
  HIP_TRY(hipMemcpyAsync(out, d, 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int pcp_continuity_residual(pcp_ctx *c, double *out) {
  if (!c || !out) return fail(PCP_E_ARG, "bad argument");
  double *d = c->diag;
  if (launch_continuity_residual(c, d) < 0) return fail(PCP_E_HIP, "continuity residual");
  HIP_TRY(hipMemcpy(out, d, 3 * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pcp_continuity_residual_noalloc(pcp_ctx *c, double *out) {
  if (!c || !out) return fail(PCP_E_ARG, "bad argument");
  const int rc = launch_continuity_residual_noalloc(c, c->diag);
  if (rc == -2) {
    char buf[200];
    snprintf(buf, sizeof buf,
             "grid of %ld x %ld nodes (halo included) exceeds the fixed scratch array of %ld doubles; use "
             "pcp_continuity_residual",
             (long)(c->g.Nr + 4), (long)(c->g.Nz + 4), CONTINUITY_SCRATCH_DOUBLES);
    return fail(PCP_E_ARG, buf);
  }
  if (rc == -3) return fail(PCP_E_STATE, "fixed scratch array not set up on this device");
  if (rc < 0) return fail(PCP_E_HIP, "continuity residual");
  HIP_TRY(hipMemcpy(out, c->diag, 3 * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pcp_copy_bandwidth(pcp_ctx *c, long bytes, int iters, double *gbps) {
  if (!c || !gbps || bytes < 16 || iters < 1) return fail(PCP_E_ARG, "bad argument");
  if (bench_copy(c, bytes, iters, gbps) != 0) return fail(PCP_E_NOMEM, "copy buffers");
  return 0;
}

int pcp_set_energy_log(pcp_ctx *c, int on) {
  if (!c) return fail(PCP_E_ARG, "null context");
  if (on && !c->ehist) {
    int rc = dalloc(&c->ehist, (size_t)EHIST * EHIST_COLS);
    if (rc) return rc;
  }
  graph_invalidate(c);
  c->elog = on != 0;
  return 0;
}

int pcp_energy_log(pcp_ctx *c, long max_rows, double *out, long *rows) {
  if (!c || !out || !rows || max_rows < 0) return fail(PCP_E_ARG, "bad argument");
  *rows = 0;
  if (!c->ehist) return 0;
  std::vector<double> h((size_t)EHIST * EHIST_COLS);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(h.data(), c->ehist, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  // rows in step order, the most recent max_rows of them
  std::vector<long> order;
  for (long k = 0; k < EHIST; ++k)
    if (h[(size_t)k * EHIST_COLS] > 0) order.push_back(k);
  std::sort(order.begin(), order.end(),
            [&](long a, long b) { return h[(size_t)a * EHIST_COLS] < h[(size_t)b * EHIST_COLS]; });
  const long n = std::min<long>((long)order.size(), max_rows);
  const long first = (long)order.size() - n;
  for (long k = 0; k < n; ++k)
    for (int j = 0; j < EHIST_COLS; ++j)
      out[k * EHIST_COLS + j] = h[(size_t)order[first + k] * EHIST_COLS + j];
  *rows = n;
  return 0;
}

int pcp_set_timing(pcp_ctx *c, int on) {
  if (!c) return fail(PCP_E_ARG, "null context");
  c->timing = on;
  return 0;
}

double pcp_kernel_ms(pcp_ctx *c, const char *name) {
  if (!c || !name) return -1;
  for (auto &kv : c->ktimes)
    if (kv.first == name) return kv.second;
  return -1;
}

}  // extern "C"
