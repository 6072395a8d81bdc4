// FP32 split-precision particle loop (see pcp32.hpp for the storage format).
//
// The algorithm is the reference's (same staggering, the same CELL_VOLUME
// factors and the same quirks), rewritten in cell units so that FP32 keeps
// its precision: r = (i + f) dr with integer i and FP32 f, every ring-volume
// weight becomes a product of small exact factors, and the radial move is
// tracked as an increment of f.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>

#include "pcp32.hpp"
#include "pcp_rng.hpp"
#include "pcp_kernels.hpp"
#include "pcp_physics.hpp"
#include "pcp_wave.hpp"

namespace pcp {

// ---------------------------------------------------------------------------
// gather in cell units
//
// For a component whose radial node is i = floor(u - 1/2) (E_r, H_z, H_phi)
// the reference weights are PI dz1 (r2^2 - r1^2) / CELL_VOLUME(i + 1) and
// PI dz1 (r3^2 - r2^2) / CELL_VOLUME(i + 3) with r1 = r - dr/2, r2 = (i+1) dr,
// r3 = r + dr/2 and CELL_VOLUME(i + m) = PI dz dr^2 (2 i + m + 1).  With
// e = u - 1/2 - i in [0, 1) they become
//   a1 = (1 - e)(2i + 1 + e) / (2i + 2),   a2 = e (2i + 2 + e) / (2i + 4).
// For node i = floor(u) (E_z, E_phi, H_r), e = u - i, r2 = (i + 1/2) dr and
// the volumes CELL_VOLUME(i), CELL_VOLUME(i + 2):
//   a1 = (1 - e)(2i + e) / (2i + 1),       a2 = e (2i + 1 + e) / (2i + 3).
// Axial weights are 1 - e_z and e_z for the node below and above.
// Near the axis (i <= 1) the radial factors are evaluated in FP64.
// ---------------------------------------------------------------------------
// FP32-mode arithmetic uses the hardware reciprocal, reciprocal square
// root and square root (v_rcp_f32, v_rsq_f32, v_sqrt_f32; each within one
// ulp), where the correctly rounded forms cost 8-12 instructions each.
__device__ __forceinline__ float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsqrt_f(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sqrt_f(float x) { return __builtin_amdgcn_sqrtf(x); }

struct W2 {
  int i;
  float a1, a2;
};

__device__ __forceinline__ W2 rad_half(int cr, float fr) {
  W2 w;
  const bool lo = fr < 0.5f;
  const float e = lo ? fr + 0.5f : fr - 0.5f;
  w.i = cr - (lo ? 1 : 0);
  if (w.i <= 1) {
    double ed = e, id = w.i;
    w.a1 = (float)((1.0 - ed) * (2 * id + 1 + ed) / (2 * id + 2));
    w.a2 = (float)(ed * (2 * id + 2 + ed) / (2 * id + 4));
  } else {
    float fi = (float)w.i;
    w.a1 = (1.f - e) * (2.f * fi + 1.f + e) * rcp_f(2.f * fi + 2.f);
    w.a2 = e * (2.f * fi + 2.f + e) * rcp_f(2.f * fi + 4.f);
  }
  return w;
}

__device__ __forceinline__ W2 rad_int(int cr, float fr) {
  W2 w;
  const float e = fr;
  w.i = cr;
  if (w.i <= 1) {
    double ed = e, id = w.i;
    w.a1 = (float)((1.0 - ed) * (2 * id + ed) / (2 * id + 1));
    w.a2 = (float)(ed * (2 * id + 1 + ed) / (2 * id + 3));
  } else {
    float fi = (float)w.i;
    w.a1 = (1.f - e) * (2.f * fi + e) * rcp_f(2.f * fi + 1.f);
    w.a2 = e * (2.f * fi + 1.f + e) * rcp_f(2.f * fi + 3.f);
  }
  return w;
}

struct Z2 {
  int k;
  float w1, w2;
};
__device__ __forceinline__ Z2 z_int(int cz, float fz) { return Z2{cz, 1.f - fz, fz}; }
__device__ __forceinline__ Z2 z_half(int cz, float fz) {
  const bool lo = fz < 0.5f;
  const float e = lo ? fz + 0.5f : fz - 0.5f;
  return Z2{cz - (lo ? 1 : 0), 1.f - e, e};
}

struct GridView32 {
  const float *G;   // component 0 of the domain
  long P;
  int yr;
  long bot_r, left_z;
  __device__ __forceinline__ float at(int c, int i, int k) const {
    return G[c * P + (long)(i - bot_r + 2) * yr + (k - left_z + 2)];
  }
};

// E (maxwellSolver.hpp:211-316) in cell units
__device__ __forceinline__ void gather32_e(const GridView32 &E, int cr, int cz, float fr, float fz, float out[3]) {
#pragma clang fp contract(fast)
  W2 ra = rad_half(cr, fr), ri = rad_int(cr, fr);
  Z2 zi = z_int(cz, fz), zh = z_half(cz, fz);
  float c0 = 0.f, c1 = 0.f, c2 = 0.f;
  // E_r: node (floor(u - 1/2), floor(w))
  c0 += E.at(0, ra.i, zi.k) * (ra.a1 * zi.w1);
  c0 += E.at(0, ra.i + 1, zi.k) * (ra.a2 * zi.w1);
  c0 += E.at(0, ra.i, zi.k + 1) * (ra.a1 * zi.w2);
  c0 += E.at(0, ra.i + 1, zi.k + 1) * (ra.a2 * zi.w2);
  // E_z: node (floor(u), floor(w - 1/2))
  c2 += E.at(2, ri.i, zh.k) * (ri.a1 * zh.w1);
  c2 += E.at(2, ri.i + 1, zh.k) * (ri.a2 * zh.w1);
  c2 += E.at(2, ri.i, zh.k + 1) * (ri.a1 * zh.w2);
  c2 += E.at(2, ri.i + 1, zh.k + 1) * (ri.a2 * zh.w2);
  // E_phi: node (floor(u), floor(w)); the 4th term goes to E_z (quirk Q2)
  c1 += E.at(1, ri.i, zi.k) * (ri.a1 * zi.w1);
  c1 += E.at(1, ri.i + 1, zi.k) * (ri.a2 * zi.w1);
  c1 += E.at(1, ri.i, zi.k + 1) * (ri.a1 * zi.w2);
  c2 += E.at(1, ri.i + 1, zi.k + 1) * (ri.a2 * zi.w2);
  out[0] = c0; out[1] = c1; out[2] = c2;
}

// H (maxwellSolver.hpp:101-209) in cell units
__device__ __forceinline__ void gather32_h(const GridView32 &H, int cr, int cz, float fr, float fz, float out[3]) {
#pragma clang fp contract(fast)
  W2 ra = rad_half(cr, fr), ri = rad_int(cr, fr);
  Z2 zi = z_int(cz, fz), zh = z_half(cz, fz);
  float c0 = 0.f, c1 = 0.f, c2 = 0.f;
  // H_z: node (floor(u - 1/2), floor(w))
  c2 += H.at(2, ra.i, zi.k) * (ra.a1 * zi.w1);
  c2 += H.at(2, ra.i + 1, zi.k) * (ra.a2 * zi.w1);
  c2 += H.at(2, ra.i, zi.k + 1) * (ra.a1 * zi.w2);
  c2 += H.at(2, ra.i + 1, zi.k + 1) * (ra.a2 * zi.w2);
  // H_r: node (floor(u), floor(w - 1/2))
  c0 += H.at(0, ri.i, zh.k) * (ri.a1 * zh.w1);
  c0 += H.at(0, ri.i + 1, zh.k) * (ri.a2 * zh.w1);
  c0 += H.at(0, ri.i, zh.k + 1) * (ri.a1 * zh.w2);
  c0 += H.at(0, ri.i + 1, zh.k + 1) * (ri.a2 * zh.w2);
  // H_phi: node (floor(u - 1/2), floor(w - 1/2))
  c1 += H.at(1, ra.i, zh.k) * (ra.a1 * zh.w1);
  c1 += H.at(1, ra.i + 1, zh.k) * (ra.a2 * zh.w1);
  c1 += H.at(1, ra.i, zh.k + 1) * (ra.a1 * zh.w2);
  c1 += H.at(1, ra.i + 1, zh.k + 1) * (ra.a2 * zh.w2);
  out[0] = c0; out[1] = c1; out[2] = c2;
}

// ---------------------------------------------------------------------------
// pushers in FP32 (same algebra as pcp_physics.hpp::push_velocity)
// ---------------------------------------------------------------------------
struct PushConst {
  float c2m;     // q dt / (2 m)
  float c2m_mu;  // q dt / (2 m) * MAGN_CONST
  float inv_c2;  // 1 / c^2
};

// the species of one batched step launch; end[k] = block trips of species
// 0..k
struct Batch32 {
  Species32 s[MAX_SPECIES];
  PushConst K[MAX_SPECIES];
  long end[MAX_SPECIES];
  int sp[MAX_SPECIES];
  int nsp;
  uint64_t *marker;    // strict-motion reset markers [D][P]
  uint32_t *cflag;     // cells near reset nodes [Nr][Nz]
  // collision surrogate of the step (null: none): per global cell, for the
  // electron (slot 0) and ion (slot 1) species, the mean velocity V and the
  // factors (perp, par) of v -> V + a (v - V) (SURR_PAR_STRIDE floats per cell)
  const float *cpar;
  int cslot[MAX_SPECIES];   // -1: the species does not collide
};

__device__ __forceinline__ void apply_cell_rescale(const float *cpar, int cslot, uint32_t cell, float v[3]) {
  const float *P = cpar + (size_t)cell * SURR_PAR_STRIDE + cslot * 8;
  const float Vr = P[0], Vp = P[1], Vz = P[2], ap = P[3], az = P[4];
  v[0] = Vr + ap * (v[0] - Vr);
  v[1] = Vp + ap * (v[1] - Vp);
  v[2] = Vz + az * (v[2] - Vz);
}

__device__ __forceinline__ float len2f(const float v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

__device__ __forceinline__ void push32(int pusher, const PushConst &K, float e[3], float b[3], float v[3], int *err) {
  if (pusher <= PCP_PUSHER_BORIS_ADAPTIVE) {
    const bool rel = pusher != PCP_PUSHER_BORIS;
    for (int i = 0; i < 3; ++i) { e[i] *= K.c2m; b[i] *= K.c2m_mu; }
    if (rel) {
      float beta = len2f(v) * K.inv_c2;
      if (beta > 1.f) *err |= ERR_BETA_GT_1;
      float gamma = rsqrt_f(1.f - beta);
      for (int i = 0; i < 3; ++i) v[i] *= gamma;
    }
    for (int i = 0; i < 3; ++i) v[i] += e[i];
    if (rel) {
      float gi = rsqrt_f(1.f + len2f(v) * K.inv_c2);
      for (int i = 0; i < 3; ++i) b[i] *= gi;
    }
    float const2 = 2.f * rcp_f(1.f + len2f(b));
    float t0 = v[0], t1 = v[1], t2 = v[2];
    v[0] = t0 + const2 * ((t1 - t0 * b[2] + t2 * b[0]) * b[2] - (t2 + t0 * b[1] - t1 * b[0]) * b[1]);
    v[1] = t1 + const2 * (-(t0 + t1 * b[2] - t2 * b[1]) * b[2] + (t2 + t0 * b[1] - t1 * b[0]) * b[0]);
    v[2] = t2 + const2 * ((t0 + t1 * b[2] - t2 * b[1]) * b[1] - (t1 - t0 * b[2] + t2 * b[0]) * b[0]);
    for (int i = 0; i < 3; ++i) v[i] += e[i];
    if (rel) {
      float gi = rsqrt_f(1.f + len2f(v) * K.inv_c2);
      for (int i = 0; i < 3; ++i) v[i] *= gi;
    }
  } else if (pusher == PCP_PUSHER_VAY) {
    float beta = len2f(v) * K.inv_c2;
    if (beta > 1.f) *err |= ERR_BETA_GT_1;
    float gamma = rsqrt_f(1.f - beta);
    for (int i = 0; i < 3; ++i) v[i] *= gamma;
    float up[3];
    for (int i = 0; i < 3; ++i) { up[i] = v[i] + e[i] * (2.f * K.c2m); b[i] *= K.c2m_mu; }
    float gi = rsqrt_f(1.f + len2f(v) * K.inv_c2);
    up[0] += gi * (v[1] * b[2] - v[2] * b[1]);
    up[1] += gi * (v[2] * b[0] - v[0] * b[2]);
    up[2] += gi * (v[0] * b[1] - v[1] * b[0]);
    float alpha = 1.f + len2f(up);   // unit mix of the reference (quirk Q4)
    float B2 = len2f(b);
    float s = alpha - B2;
    float ub = up[0] * b[0] + up[1] * b[1] + up[2] * b[2];
    alpha = rsqrt_f(0.5f * (s + sqrt_f(s * s + 4.f * (B2 + ub * ub))));
    for (int i = 0; i < 3; ++i) b[i] *= alpha;
    s = rcp_f(1.f + len2f(b));
    alpha = up[0] * b[0] + up[1] * b[1] + up[2] * b[2];
    float psm[3];
    psm[0] = s * (up[0] + alpha * b[0] + b[2] * up[1] - b[1] * up[2]);
    psm[1] = s * (up[1] + alpha * b[1] + b[0] * up[2] - b[2] * up[0]);
    psm[2] = s * (up[2] + alpha * b[2] + b[1] * up[0] - b[0] * up[1]);
    gi = rsqrt_f(1.f + len2f(psm) * K.inv_c2);
    for (int i = 0; i < 3; ++i) v[i] = psm[i] * gi;
  } else {
    float beta = len2f(v) * K.inv_c2;
    if (beta > 1.f) *err |= ERR_BETA_GT_1;
    float gamma = rsqrt_f(1.f - beta);
    float um[3], psm[3], b2[3], bc[3], up[3];
    for (int i = 0; i < 3; ++i) {
      v[i] *= gamma;
      psm[i] = e[i] * K.c2m;
      um[i] = v[i] + psm[i];
      b[i] *= K.c2m_mu;
    }
    float gfm2 = 1.f + len2f(um);   // quirk Q5
    float B2 = len2f(b);
    float bu = b[0] * um[0] + b[1] * um[1] + b[2] * um[2];
    float gb = gfm2 - B2;
    gamma = rsqrt_f(0.5f * (gb + sqrt_f(gb * gb + 4.f * (B2 + bu * bu))));
    for (int i = 0; i < 3; ++i) { b[i] *= gamma; b2[i] = b[i] * b[i]; }
    bc[0] = b[0] * b[1]; bc[1] = b[1] * b[2]; bc[2] = b[2] * b[0];
    float inv_det_B = rcp_f(1.f + b2[0] + b2[1] + b2[0]);   // b2[0] twice (pusherHC.cpp:94)
    up[0] = ((1.f + b2[0] - b2[1] - b2[2]) * um[0] + 2.f * (bc[0] + b[2]) * um[1] + 2.f * (bc[2] - b[2]) * um[2]) * inv_det_B;
    up[1] = (2.f * (bc[0] - b[2]) * um[0] + (1.f - b2[0] + b2[1] - b2[2]) * um[1] + 2.f * (bc[1] + b[0]) * um[2]) * inv_det_B;
    up[2] = (2.f * (bc[2] + b[1]) * um[0] + 2.f * (bc[1] - b[0]) * um[1] + (1.f - b2[0] - b2[1] + b2[2]) * um[2]) * inv_det_B;
    for (int i = 0; i < 3; ++i) psm[i] += up[i];
    float gi = rsqrt_f(1.f + len2f(psm) * K.inv_c2);
    for (int i = 0; i < 3; ++i) v[i] = psm[i] * gi;
  }
}

// ---------------------------------------------------------------------------
// move in split form: mover_cylindrical + back_position_to_rz track the new
// radius as r_old + dr_x + Y^2 / (X + sqrt(X^2 + Y^2)) (X = r + v_r dt,
// Y = v_phi dt), an increment of the in-cell offset; particles that cross a
// wall are reflected in FP64 (reflect_plasma / reflect_beam).
// ---------------------------------------------------------------------------
struct Split {
  int cr, cz;
  float fr, fz;
};

__device__ __forceinline__ void renorm(int &c, float &f) {
  float fl = floorf(f);
  c += (int)fl;
  f -= fl;
  if (f >= 1.f) { f -= 1.f; c += 1; }
}

// Returns false when the particle left the simulation area.  (rd, zd) is the
// unrounded end point in metres -- the old split position plus the FP32
// displacement, summed in FP64 -- from which the current is deposited: the
// stored offset rounds displacements below 2^-24 of a cell away, and a
// rounded-away radial step would turn an ordinary move into the reference's
// strict axial motion (which resets J_r nodes).
// EXACT_END = false (ZigZag, which deposits from the split positions): the
// FP64 end point is only formed for a wall reflection, from the FP32 offset.
template <bool EXACT_END = true>
__device__ __forceinline__ bool move32(const Geom &g, const Dom &o, int beam, float dtr, float dtz, Split &p, float v[3],
                                       double &rd, double &zd) {
  const Split old = p;
  float du = v[0] * dtr, dph = v[1] * dtr, dw = v[2] * dtz;
  float xoff = p.fr + du;
  float X = (float)p.cr + xoff;
  float Y = dph;
  float sn = 0.f;
  double ur = (double)p.cr + (double)p.fr + (double)du;   // unrounded radius, cell units
  if (Y != 0.f) {
    float R = sqrt_f(X * X + Y * Y);
    if (X > 0.f) {
      float inc = Y * Y * rcp_f(X + R);
      xoff += inc;
      ur += (double)inc;
    } else {
      xoff = R - (float)p.cr;
      ur = (double)R;
    }
    sn = Y * rcp_f(R);
  } else if (X < 0.f) {
    xoff = -X - (float)p.cr;   // |X|: the cylinder radius of a particle that crossed the axis
    ur = -ur;
  }
  p.fr = xoff;
  p.fz = p.fz + dw;
  if (EXACT_END) {
    rd = ur * g.dr;
    zd = ((double)old.cz + (double)old.fz + (double)dw) * g.dz;
  }
  // walls (local cell units): r < 1/2 or > nr - 1/2, z < 1/2 or > nz - 1/2
  const float lr = (float)(p.cr - (int)o.bot_r) + p.fr, lz = (float)(p.cz - (int)o.left_z) + p.fz;
  const bool hit = beam ? (o.w0 && lr < 0.5f)
                        : ((o.w0 && lr < 0.5f) || (o.w2 && lr > g.nr - 0.5f) || (o.w1 && lz < 0.5f) ||
                           (o.w3 && lz > g.nz - 0.5f));
  if (hit) {
    if (!EXACT_END) {
      rd = ((double)p.cr + (double)p.fr) * g.dr;
      zd = ((double)p.cz + (double)p.fz) * g.dz;
    }
    double por = ((double)old.cr + (double)old.fr) * g.dr, poz = ((double)old.cz + (double)old.fz) * g.dz;
    double vr = v[0], vz = v[2];
    if (beam) reflect_beam(g, o, rd, vr);
    else reflect_plasma(g, o, rd, zd, por, poz, vr, vz);
    v[0] = (float)vr;
    v[2] = (float)vz;
    double qr = rd / g.dr, qz = zd / g.dz;
    double fcr = floor(qr), fcz = floor(qz);
    p.cr = (int)fcr; p.fr = (float)(qr - fcr);
    p.cz = (int)fcz; p.fz = (float)(qz - fcz);
  }
  renorm(p.cr, p.fr);
  renorm(p.cz, p.fz);
  // back_velocity_to_rz
  float cs = sqrt_f(1.f - sn * sn);
  float vr = v[0], vp = v[1];
  v[0] = cs * vr - sn * vp;
  v[1] = sn * vr + cs * vp;
  // a non-finite position leaves the area, as in the FP64 mover
  return isfinite(p.fr) && isfinite(p.fz) && p.cr >= 0 && p.cz >= 0 && p.cr < g.Nr && p.cz < g.Nz;
}

// ZigZag (currentZigZag.cpp:22-136) in FP32, positions relative to the old
// cell: uo = f_old, un = (i_n - i_o) + f_new (likewise in z)
__device__ __forceinline__ ZZTermsT<float> zigzag32(const Geom &g, const Dom &dn, const Split &o,
                                         const Split &n, float vphi, float qw, float inv_dt) {
  const float dr = (float)g.dr, dz = (float)g.dz;
  const int di = n.cr - o.cr, dk = n.cz - o.cz;
  const float uo = o.fr, un = (float)di + n.fr;
  const float wo = o.fz, wn = (float)dk + n.fz;
  // relay points (currentZigZag.hpp:28-30) relative to the old cell
  const float rr = fminf(fminf(0.f, (float)di) + 1.f, fmaxf(fmaxf(0.f, (float)di), 0.5f * (uo + un)));
  const float rz = fminf(fminf(0.f, (float)dk) + 1.f, fmaxf(fmaxf(0.f, (float)dk), 0.5f * (wo + wn)));
  const int i_middle = o.cr + (int)floorf(0.5f * (uo + un));
  const float one_over_volume = (float)g.inv_vol0 * rcp_f((float)(2 * i_middle + 1));
  const float charge_over_dt = qw * inv_dt;
  const float F_phi = qw * vphi;
  const float F_r1 = charge_over_dt * dr * (rr - uo);
  const float F_r2 = charge_over_dt * dr * (un - rr);
  const float F_z1 = charge_over_dt * dz * (rz - wo);
  const float F_z2 = charge_over_dt * dz * (wn - rz);
  // weights in metres (quirk Q11)
  const float W_r1 = dr * (0.5f * (uo + rr));
  const float W_r2 = dr * (0.5f * (un + rr) - (float)di);
  const float W_z1 = dz * (0.5f * (wo + rz));
  const float W_z2 = dz * (0.5f * (wn + rz) - (float)dk);
  const float mWr1 = 1.f - W_r1, mWr2 = 1.f - W_r2, mWz1 = 1.f - W_z1, mWz2 = 1.f - W_z2;
  const int io = o.cr - (int)dn.bot_r, ko = o.cz - (int)dn.left_z;
  const int in = n.cr - (int)dn.bot_r, kn = n.cz - (int)dn.left_z;
  const float A = one_over_volume;
  ZZTermsT<float> Z;
  Z.io = io; Z.ko = ko; Z.in = in; Z.kn = kn;
  // the reference's two azimuthal halves of each J_r and J_z term (weights
  // 1 - W_phi and W_phi) land on the same r-z node; their sum is the term
  // without the azimuthal weight
  const float Ar1 = A * F_r1, Ar2 = A * F_r2, Az1 = A * F_z1, Az2 = A * F_z2, Aphi = A * F_phi;
  Z.v[0] = Ar1 * mWz1;
  Z.v[1] = Ar1 * W_z1;
  Z.v[2] = Ar2 * mWz2;
  Z.v[3] = Ar2 * W_z2;
  Z.v[4] = Aphi * mWr2 * mWz2;
  Z.v[5] = Aphi * mWr2 * W_z1;   // uses W_z1 (currentZigZag.cpp:104)
  Z.v[6] = Aphi * W_r2 * mWz2;
  Z.v[7] = Aphi * W_r2 * W_z2;
  Z.v[8] = Az1 * mWr1;
  Z.v[9] = Az1 * W_r1;
  Z.v[10] = Az2 * mWr2;
  Z.v[11] = Az2 * W_r2;
  return Z;
}

// (cr, cz >= 0: cells of live particles)
__device__ __forceinline__ int domain_of32(const Geom &g, int cr, int cz) {
  return (int)g.div_nr.div((uint32_t)cr) * g.Dz + (int)g.div_nz.div((uint32_t)cz);
}

// domain of a cell and its description; D1: the context has one domain
// (the domain is 0, its walls are the area's)
template <bool D1>
__device__ __forceinline__ int domain_of32t(const Geom &g, int cr, int cz) {
  return D1 ? 0 : domain_of32(g, cr, cz);
}
template <bool D1>
__device__ __forceinline__ Dom make_dom_t(const Geom &g, int d) {
  if (!D1) return make_dom(g, d);
  Dom o;
  o.I = o.J = 0;
  o.bot_r = o.left_z = 0;
  o.w0 = o.w1 = o.w2 = o.w3 = 1;
  return o;
}

__device__ __forceinline__ Split unpack(uint32_t cell, float fr, float fz, const Geom &g) {
  Split s;
  s.cr = (int)g.div_Nz.div(cell);
  s.cz = (int)(cell - (uint32_t)s.cr * (uint32_t)g.Nz);
  s.fr = fr;
  s.fz = fz;
  return s;
}

__device__ __forceinline__ uint64_t order_key32(int sp, uint64_t seq) { return ((uint64_t)sp << 36) | seq; }


// ---------------------------------------------------------------------------
// the fused FP32 step kernel
// ---------------------------------------------------------------------------
// Gather, push and move of live particle p: positions from slot `from`,
// velocities from slot `vfrom`; do_push = false moves with the stored
// velocity (steps whose collision phase sits between push and move).  The
// fix-up pass calls it again on the pre-step slots and gets bit-identical
// (n, v, rd, zd).
// With !do_push and a surrogate collision phase (cpar, cslot >= 0), the
// stored (pushed) velocity first takes its cell's rescaling.
template <bool EXACT_END = true, bool D1 = false>
__device__ __forceinline__ bool advance32(const Geom &g, const FieldsT<float> &f, const Species32 &s, long p,
                                          uint32_t cell, int from, int vfrom, bool do_push, const PushConst &K,
                                          float dtr, float dtz, Split &o, Split &n, float v[3], double &rd,
                                          double &zd, int &d_old, int *err, const float *cpar = nullptr,
                                          int cslot = -1) {
  o = unpack(cell, s.fr[from][p], s.fz[from][p], g);
  d_old = domain_of32t<D1>(g, o.cr, o.cz);
  const Dom dm = make_dom_t<D1>(g, d_old);
  v[0] = s.vr[vfrom][p]; v[1] = s.vphi[vfrom][p]; v[2] = s.vz[vfrom][p];
  if (!do_push && cpar && cslot >= 0) apply_cell_rescale(cpar, cslot, cell, v);
  if (do_push) {
    GridView32 EV{f.E + (long)d_old * 3 * g.P, g.P, g.yr, dm.bot_r, dm.left_z};
    GridView32 HV{f.H + (long)d_old * 3 * g.P, g.P, g.yr, dm.bot_r, dm.left_z};
    float e[3], h[3];
    gather32_e(EV, o.cr, o.cz, o.fr, o.fz, e);
    gather32_h(HV, o.cr, o.cz, o.fr, o.fz, h);
    push32(g.pusher, K, e, h, v, err);
  }
  n = o;
  return move32<EXACT_END>(g, dm, s.beam, dtr, dtz, n, v, rd, zd);
}

// gather + push only (steps with a collision phase): velocities of slot cur
// pushed into slot cur ^ 1
__global__ void __launch_bounds__(256) k32_push(Geom g, FieldsT<float> f, Species32 s, int cur, StepState *st,
                                                PushConst K) {
  int err = 0;
  const int nxt = cur ^ 1;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x) {
    uint32_t cell = s.cell[cur][p];
    if (cell == DEAD_CELL) continue;
    Split o = unpack(cell, s.fr[cur][p], s.fz[cur][p], g);
    const int d = domain_of32(g, o.cr, o.cz);
    const Dom dm = make_dom(g, d);
    GridView32 EV{f.E + (long)d * 3 * g.P, g.P, g.yr, dm.bot_r, dm.left_z};
    GridView32 HV{f.H + (long)d * 3 * g.P, g.P, g.yr, dm.bot_r, dm.left_z};
    float e[3], h[3];
    gather32_e(EV, o.cr, o.cz, o.fr, o.fz, e);
    gather32_h(HV, o.cr, o.cz, o.fr, o.fz, h);
    float v[3] = {s.vr[cur][p], s.vphi[cur][p], s.vz[cur][p]};
    push32(g.pusher, K, e, h, v, &err);
    s.vr[nxt][p] = v[0]; s.vphi[nxt][p] = v[1]; s.vz[nxt][p] = v[2];
  }
  if (err) atomicOr(&st->err, err);
}

// Gather + push of a step with the fused surrogate collision phase: the
// pushed velocities go to slot cur ^ 1, and for a colliding species (slot >=
// 0) their per-cell moments are added to mom[cell][slot][0..6]: count,
// sum w, sum w v (r, phi, z), sum w v_z^2, sum w (v_r^2 + v_phi^2) -- the
// sums behind the surrogate's temperatures.  A wavefront takes 64
// consecutive (cell-ordered) particles; the lanes of a cell form runs whose
// sums (DPP run sums in FP32) leave one FP64 atomic per run and moment.
__global__ void __launch_bounds__(256) k32_push_mom(Geom g, FieldsT<float> f, Species32 s, int cur, StepState *st,
                                                    PushConst K, double *mom, int slot) {
  int err = 0;
  const int nxt = cur ^ 1;
  const int lane = threadIdx.x & 63;
  const long waves = (s.n + 63) / 64;
  for (long wv = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); wv < waves;
       wv += (long)gridDim.x * (blockDim.x >> 6)) {
    const long p = wv * 64 + lane;
    uint32_t key = DEAD_CELL;
    float v[3] = {0.f, 0.f, 0.f}, w = 0.f;
    if (p < s.n) {
      const uint32_t cell = s.cell[cur][p];
      if (cell != DEAD_CELL) {
        Split o = unpack(cell, s.fr[cur][p], s.fz[cur][p], g);
        const int d = domain_of32(g, o.cr, o.cz);
        const Dom dm = make_dom(g, d);
        GridView32 EV{f.E + (long)d * 3 * g.P, g.P, g.yr, dm.bot_r, dm.left_z};
        GridView32 HV{f.H + (long)d * 3 * g.P, g.P, g.yr, dm.bot_r, dm.left_z};
        float e[3], h[3];
        gather32_e(EV, o.cr, o.cz, o.fr, o.fz, e);
        gather32_h(HV, o.cr, o.cz, o.fr, o.fz, h);
        v[0] = s.vr[cur][p]; v[1] = s.vphi[cur][p]; v[2] = s.vz[cur][p];
        push32(g.pusher, K, e, h, v, &err);
        s.vr[nxt][p] = v[0]; s.vphi[nxt][p] = v[1]; s.vz[nxt][p] = v[2];
        key = cell;
        w = s.w[p];
      }
    }
    if (slot >= 0) {
      const Runs R = runs_of(key);
      const bool live = key != DEAD_CELL;
      float q[7] = {live ? 1.f : 0.f, w, w * v[0], w * v[1], w * v[2], w * v[2] * v[2],
                    w * (v[0] * v[0] + v[1] * v[1])};
#pragma unroll
      for (int k = 0; k < 7; ++k) q[k] = run_sum_f(q[k], R);
      // one atomic instruction per run: lane k adds moment k of the run's
      // cell, so the instruction's seven lanes hit one cache line (seven
      // instructions with one lane per run each measured 2.3 ms more per
      // full-form step: the L2 takes same-line lanes of an instruction
      // together)
      unsigned long long tails = __ballot(R.tail && live);
      while (tails) {
        const int t = __ffsll((long long)tails) - 1;
        tails &= tails - 1ull;
        const uint32_t kt = (uint32_t)__builtin_amdgcn_readlane((int)key, t);
        float mine = 0.f;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
          const float qk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q[k]), t));
          if (lane == k) mine = qk;
        }
        if (lane < 7) atomicAdd(mom + (size_t)kt * SURR_PAR_STRIDE + slot * 8 + lane, (double)mine);
      }
    }
  }
  if (err) atomicOr(&st->err, err);
}

// domain whose grid receives the current of a particle ending at (rd, zd)
__device__ __forceinline__ int deposit_domain32(const Geom &g, double rd, double zd) {
  int cr = min(max(cell_number(rd, g.dr), 0), (int)g.Nr - 1);
  int cz = min(max(cell_number(zd, g.dz), 0), (int)g.Nz - 1);
  return domain_of32(g, cr, cz);
}

// the deposit of one lane's particle (ZigZag in FP32 relative to the old
// cell; Villasenor-Buneman segments split in FP64 from the old split
// position and the unrounded end point, their terms in the FP32 cell-local
// form except on the axis cell and for long steps); every lane of the
// wavefront calls it.  COMP: compensated run sums and atomics, the
// rounding errors into the companion array at element offset jc from J.
template <int CCS, bool COMP>
__device__ __forceinline__ void deposit32(bool live, const Geom &g, const Dom &dn, const Split &o,
                                          const Split &n, float vphi, float qw, float inv_dt, double rd, double zd,
                                          double p_charge, const FieldsT<float> &f, int d_dep, StepState *st,
                                          int *err, uint64_t *marker, uint32_t *cflag, const uint64_t *seq, long p,
                                          int sp, long jc) {
  if (CCS == PCP_CCS_ZIGZAG) {
    zigzag_runs<COMP>(f.J, (long)d_dep * 3 * g.P, g.P, g.yr, g.nr, g.nz, zigzag32(g, dn, o, n, vphi, qw, inv_dt), live,
                      err, jc);
  } else {
    const double r = rd, z = zd;
    const double ro = ((double)o.cr + (double)o.fr) * g.dr, zo = ((double)o.cz + (double)o.fz) * g.dz;
    AzTerms A = vb_azimuthal_terms_local(g.nr, g.nz, g.dr, g.dz, dn.bot_r, dn.left_z, r, z, ro, zo, (double)vphi,
                                         p_charge);
    SegRecorder R;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      R.rn[q] = 1.5 * g.dr; R.zn[q] = 1.25 * g.dz; R.ro[q] = 1.25 * g.dr; R.zo[q] = 1.5 * g.dz;
      R.i[q] = 1; R.k[q] = 1;
    }
    if (live) vb_rz_dispatch(R, g.dr, g.dz, r, z, ro, zo, err);
    if (live && R.is_strict) {
      const ResetMark M{marker + (long)d_dep * g.P, cflag, ((st->step & 0xFFFFFF) << 40) | order_key32(sp, seq[p]),
                        (uint32_t)st->step, dn.bot_r, dn.left_z, g.Nr, g.Nz, g.yr, g.nr, g.nz, &st->need_exact};
      ResetSink<float> A1{f.J + (long)d_dep * 3 * g.P, g.P, g.yr, g.nr, g.nz, err, M, COMP ? jc : 0};
      vb_strict(A1, g.dr, g.dz, g.dt, dn.bot_r, dn.left_z, r, z, ro, zo, p_charge);
    }
    vb_runs<COMP>(f.J, (long)d_dep * 3 * g.P, g.P, g.yr, g, dn, live, A, R, p_charge, err, jc);
  }
}

// PUSH: velocities slot cur -> pushed, moved -> slot cur ^ 1.
// !PUSH: pushed (and collided) velocities of slot cur ^ 1 -> moved -> slot
// cur; the caller then exchanges the two velocity slots.
// One wavefront trip of the fused step: particle p of this lane (the lanes
// of a wavefront take consecutive particles; all 64 lanes take part, live
// or not, since the deposit's run sums move data between lanes).
// Returns the lane's kinetic energy w m v^2 / (s (1 + s)), s = sqrt(1 - v^2/c^2)
// (= w m c^2 (gamma - 1) without cancellation) after the move when `ke` is
// set, else 0.
template <int CCS, bool PUSH, bool D1, bool COMP>
__device__ __forceinline__ float step_trip32(const Geom &g, const FieldsT<float> &f, const Species32 &s, int sp,
                                             const PushConst &K, long p, int cur, StepState *st, const Migrants &mg,
                                             uint64_t *marker, uint32_t *cflag, int &err, const float *cpar,
                                             int cslot, bool ke, long jc) {
  const int nxt = cur ^ 1;
  const float dtr = g.dtr_f, dtz = g.dtz_f;
  const float inv_dt = g.inv_dt_f;
  bool live = false;
  Split o{1, 1, 0.5f, 0.5f}, n{1, 1, 0.5f, 0.5f};
  float v[3] = {0.f, 0.f, 0.f};
  double rd = 1.5 * g.dr, zd = 1.5 * g.dz;
  float w = 0.f;
  int d_new = 0;
  if (p < s.n) {
    uint32_t cell = s.cell[cur][p];
    int d_old = 0;
    const int vin = PUSH ? cur : nxt, vout = PUSH ? nxt : cur;
    if (cell != DEAD_CELL &&
        advance32<CCS != PCP_CCS_ZIGZAG, D1>(g, f, s, p, cell, cur, vin, PUSH, K, dtr, dtz, o, n, v, rd, zd, d_old,
                                             &err, PUSH ? nullptr : cpar, cslot)) {
      s.vr[vout][p] = v[0]; s.vphi[vout][p] = v[1]; s.vz[vout][p] = v[2];
      s.cell[nxt][p] = (uint32_t)n.cr * (uint32_t)g.Nz + (uint32_t)n.cz;
      s.fr[nxt][p] = n.fr;
      s.fz[nxt][p] = n.fz;
      d_new = domain_of32t<D1>(g, n.cr, n.cz);
      if (!D1 && d_new != d_old) append_migrant(mg, st, sp, p, d_old, d_new, s.seq[p], g, &err);
      w = s.w[p];
      live = true;
    } else {
      s.cell[nxt][p] = DEAD_CELL;
    }
  }
  if (!live) { o = n = Split{1, 1, 0.5f, 0.5f}; d_new = 0; v[1] = 0.f; w = 0.f; rd = 1.5 * g.dr; zd = 1.5 * g.dz; }
  float kin = 0.f;
  if (ke) {
    const float b2 = len2f(v) * K.inv_c2;
    const float sq = sqrt_f(1.f - b2);
    kin = live ? w * (float)s.m * len2f(v) * rcp_f(sq * (1.f + sq)) : 0.f;
  }
  const float qw = (float)s.q * w;
  // deposit domain: ZigZag deposits in the particle's new domain, VB in the
  // domain of its FP64 end point
  const int d_dep = CCS == PCP_CCS_ZIGZAG || D1 ? d_new : (live ? deposit_domain32(g, rd, zd) : 0);
  const Dom dn = make_dom_t<D1>(g, d_dep);
  deposit32<CCS, COMP>(live, g, dn, o, n, v[1], qw, inv_dt, rd, zd, s.q * (double)w, f, d_dep, st, &err, marker,
                       cflag, s.seq, p, sp, jc);
  return kin;
}

// The fused step of every species in one persistent launch.  A trip is one
// wavefront of 64 consecutive particles of one species; each wavefront
// takes `chunk` consecutive trips per increment of a counter in StepState
// until the trips of all species are handed out, so that trips of uneven
// cost (dead slots, domains of different density, strict motions, walls)
// balance over the resident wavefronts.  The wavefronts never synchronise:
// a block-wide ticket and its barrier made every wavefront wait for its
// outstanding J atomics once per trip (1.6x slower on the wakefield case),
// and a counter increment per trip serialises on the counter.
struct BatchArgs32 {
  Geom g;
  FieldsT<float> f;
  Batch32 b;
  StepState *st;
  Migrants mg;
  double *ehist;   // energy log (null: off); kinetic energies summed per species
  long jc;         // compensated current: element offset of Jc from J (0: off)
  int cur, chunk;
};
typedef __attribute__((address_space(4))) const BatchArgs32 KernargBatch32;

// The arguments are one struct, read through the kernel-argument segment
// pointer made opaque on every trip: otherwise the compiler keeps the
// geometry and the species tables in scalar registers across the trip loop,
// runs out of them and spills into vector lanes, whose reloads are VALU
// work (34 spilled registers and 101 lane moves in the loop before; none
// after, 82 instead of 107 VGPRs, 7 % faster on the wakefield case).  The
// ZigZag kernel then fits 80 VGPRs without spilling: six wavefronts per
// SIMD instead of five (4 % faster); the VB kernel fits 128 instead of 130:
// four instead of three (9 % faster on the batched case).  The compensated
// deposit (COMP) carries twice the run-sum state and is held to four.
template <int CCS, bool PUSH, int BLOCK, bool D1, bool COMP>
__global__ void __launch_bounds__(BLOCK)
    __attribute__((amdgpu_waves_per_eu(CCS == PCP_CCS_ZIGZAG && !COMP ? 6 : 4))) k32_step_batch(BatchArgs32 args) {
  (void)args;
  int err = 0;
  KernargBatch32 *A = (KernargBatch32 *)__builtin_amdgcn_kernarg_segment_ptr();
  const long total = A->b.end[A->b.nsp - 1];   // trips of 64 particles
  const int lane = threadIdx.x & 63;
  const bool ke = A->ehist != nullptr;
  // kinetic energy of the trips since the last flush, all of species ke_sp
  double ke_acc = 0.0;
  int ke_sp = -1;
  auto flush = [&]() {
    const double tot = wave_sum(ke_acc);
    if (lane == 0 && ke_sp >= 0) atomicAdd(ehist_row(A->ehist, A->st->step) + 2 + A->b.sp[ke_sp], tot);
    ke_acc = 0.0;
  };
  for (;;) {
    KernargBatch32 *a = A;
    asm volatile("" : "+s"(a));
    unsigned tk = 0;
    if (lane == 0) tk = atomicAdd(&a->st->queue, 1u);
    const long t0 = (long)(unsigned)__builtin_amdgcn_readfirstlane((int)tk) * a->chunk;
    if (t0 >= total) break;
    const long t1 = t0 + a->chunk < total ? t0 + a->chunk : total;
    for (long t = t0; t < t1; ++t) {
      KernargBatch32 *q = A;
      asm volatile("" : "+s"(q));
      const BatchArgs32 &x = *(const BatchArgs32 *)q;
      int k = 0;
      while (k + 1 < x.b.nsp && t >= x.b.end[k]) ++k;
      const long first = k ? x.b.end[k - 1] : 0;
      if (ke && k != ke_sp) {
        flush();
        ke_sp = k;
      }
      const float kin = step_trip32<CCS, PUSH, D1, COMP>(x.g, x.f, x.b.s[k], x.b.sp[k], x.b.K[k],
                                                         (t - first) * 64 + lane, x.cur, x.st, x.mg, x.b.marker,
                                                         x.b.cflag, err, x.b.cpar, x.b.cslot[k], ke, x.jc);
      if (ke) ke_acc += (double)kin;
    }
  }
  if (ke) flush();
  if (err) atomicOr(&A->st->err, err);
}

// Fix-up pass of a step whose strict motions reset J_r nodes: particles
// whose pre-step cell is flagged redo their gather/push/move from the
// pre-step slots (bit-identical to the step) and take back their J_r
// contributions at nodes that a later particle in the reference order reset.
// Two launches: a lean screen over every particle (the pre-step cells, 16
// per thread in four 16-byte loads, and their flags) lists the flagged ones,
// and the redo runs over that list.  The redo's deposit code needs many
// registers; screening inside it ran at its low occupancy, latency-bound
// (0.69 ms per step on 1.5e8 particles, against ~0.1 ms for the screen).
// A list longer than its capacity makes the redo screen every particle
// itself, as a single launch did.
constexpr uint32_t FIX_LIST_CAP = 1u << 20;

__global__ void __launch_bounds__(256) k32_fix_screen(Batch32 b, int cur, StepState *st, uint64_t *list,
                                                      unsigned long long *count) {
  if (!st->need_exact) return;
  const uint32_t fval = (uint32_t)st->step;
  const long total = b.end[b.nsp - 1] * 16;   // groups of 16 particles (b.end counts 256)
  const int lane = threadIdx.x & 63;
  for (long t0 = (long)blockIdx.x * blockDim.x; t0 < total; t0 += (long)gridDim.x * blockDim.x) {
    const long t = t0 + threadIdx.x;
    uint32_t hits = 0;   // bit j: particle p0 + j sits in a flagged cell
    int k = 0;
    long p0 = 0;
    if (t < total) {
      while (k + 1 < b.nsp && t >= b.end[k] * 16) ++k;
      const Species32 &s = b.s[k];
      p0 = (t - (k ? b.end[k - 1] * 16 : 0)) * 16;
      uint32_t co[16];
      const uint32_t *cell = s.cell[cur ^ 1];
      if (p0 + 15 < s.n) {
        // 16-byte aligned: allocations are, and p0 is a multiple of 16
        const uint4 *q = reinterpret_cast<const uint4 *>(cell + p0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint4 x = q[j];
          co[4 * j] = x.x; co[4 * j + 1] = x.y; co[4 * j + 2] = x.z; co[4 * j + 3] = x.w;
        }
      } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) co[j] = p0 + j < s.n ? cell[p0 + j] : DEAD_CELL;
      }
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (co[j] != DEAD_CELL && b.cflag[co[j]] == fval) hits |= 1u << j;
    }
    // one counter atomic per wavefront that found any
    const int nh = __popc(hits);
    if (__ballot(nh != 0) == 0ull) continue;
    int incl = nh;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    unsigned long long base = 0;
    if (lane == 63) base = atomicAdd(count, (unsigned long long)incl);
    base = __shfl(base, 63, 64);
    unsigned long long slot = base + (unsigned long long)(incl - nh);
    for (int j = 0; j < 16; ++j)
      if (hits >> j & 1u) {
        if (slot < FIX_LIST_CAP) list[slot] = ((uint64_t)k << 40) | (uint64_t)(p0 + j);
        ++slot;
      }
  }
}

// the redo of flagged particle p of batch species k
__device__ __forceinline__ void fix_particle32(const Geom &g, const FieldsT<float> &f, const Batch32 &b, int k, long p,
                                               uint32_t co, int cur, int do_push, uint64_t stamp, long jc, int &err) {
  const Species32 &s = b.s[k];
  if (s.cell[cur][p] == DEAD_CELL) return;
  Split o, n;
  float v[3];
  double r, z;
  int d_old;
  if (!advance32(g, f, s, p, co, cur ^ 1, cur ^ 1, do_push != 0, b.K[k], g.dtr_f, g.dtz_f, o, n, v, r, z, d_old,
                 &err, do_push ? nullptr : b.cpar, b.cslot[k]))
    return;
  const double ro = ((double)o.cr + (double)o.fr) * g.dr, zo = ((double)o.cz + (double)o.fz) * g.dz;
  const int d = deposit_domain32(g, r, z);
  const Dom dm = make_dom(g, d);
  FixSink<float> S{f.J + (long)d * 3 * g.P, b.marker + (long)d * g.P, g.yr, g.nr, g.nz, stamp,
                   order_key32(b.sp[k], s.seq[p]), jc};
  vb_rz(S, g.dr, g.dz, g.dt, dm.bot_r, dm.left_z, r, z, ro, zo, s.q * (double)s.w[p], &err);
}

__global__ void __launch_bounds__(256) k32_fix(Geom g, FieldsT<float> f, Batch32 b, int cur, StepState *st,
                                               int do_push, long jc, const uint64_t *list,
                                               const unsigned long long *count) {
  if (!st->need_exact) return;
  const uint64_t stamp = st->step & 0xFFFFFF;
  const uint32_t fval = (uint32_t)st->step;
  int err = 0;
  const unsigned long long nl = *count;
  if (nl <= FIX_LIST_CAP) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)nl; i += (long)gridDim.x * blockDim.x) {
      const uint64_t e = list[i];
      const int k = (int)(e >> 40);
      const long p = (long)(e & ((1ull << 40) - 1));
      fix_particle32(g, f, b, k, p, b.s[k].cell[cur ^ 1][p], cur, do_push, stamp, jc, err);
    }
  } else {
    const long total = b.end[b.nsp - 1] * 64;   // groups of 4 particles (b.end counts 256)
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
      int k = 0;
      while (k + 1 < b.nsp && t >= b.end[k] * 64) ++k;
      const Species32 &s = b.s[k];
      const long p0 = (t - (k ? b.end[k - 1] * 64 : 0)) * 4;
      for (int j = 0; j < 4; ++j) {
        const long p = p0 + j;
        if (p >= s.n) break;
        const uint32_t co = s.cell[cur ^ 1][p];
        if (co != DEAD_CELL && b.cflag[co] == fval) fix_particle32(g, f, b, k, p, co, cur, do_push, stamp, jc, err);
      }
    }
  }
  if (err) atomicOr(&st->err, err);
}

// ---------------------------------------------------------------------------
// sort by (domain, cell) and compaction
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t sort_key32(const Geom &g, uint32_t cell) {
  const int cr = (int)g.div_Nz.div(cell), cz = (int)(cell - (uint32_t)cr * (uint32_t)g.Nz);
  const int d = domain_of32(g, cr, cz);
  const int I = (int)g.div_Dz.div((uint32_t)d), J = d - I * g.Dz;
  return sort_cell_key(g, d, cr - (long)I * g.nr, cz - (long)J * g.nz);
}

__global__ void __launch_bounds__(SORT_BLOCK) k32_sort_count(Geom g, Species32 s, int cur, uint32_t *key,
                                                             uint32_t *count, uint32_t sentinel) {
  __shared__ SortTable t;
  table_clear(t);
  __syncthreads();
  for (long c0 = (long)blockIdx.x * SORT_CHUNK; c0 < s.n; c0 += (long)gridDim.x * SORT_CHUNK)
    table_count_chunk(t, c0, s.n, sentinel, count, [&](long p) {
      const uint32_t c = s.cell[cur][p];
      const uint32_t k = c != DEAD_CELL ? sort_key32(g, c) : sentinel;
      key[p] = k;
      return k;
    });
}

__global__ void __launch_bounds__(SORT_BLOCK) k32_sort_scatter(Species32 s, SortBuffers32 sb, int cur,
                                                               const uint32_t *key, const uint32_t *offset,
                                                               uint32_t *fill, uint32_t sentinel) {
  __shared__ SortTable t;
  const int nxt = cur ^ 1;
  table_clear(t);
  __syncthreads();
  for (long c0 = (long)blockIdx.x * SORT_CHUNK; c0 < s.n; c0 += (long)gridDim.x * SORT_CHUNK)
    table_scatter_chunk(t, c0, s.n, sentinel, key, offset, fill, [&](long p, long q) {
      s.cell[nxt][q] = s.cell[cur][p];
      s.fr[nxt][q] = s.fr[cur][p];
      s.fz[nxt][q] = s.fz[cur][p];
      s.vr[nxt][q] = s.vr[cur][p];
      s.vphi[nxt][q] = s.vphi[cur][p];
      s.vz[nxt][q] = s.vz[cur][p];
      sb.w[q] = s.w[p];
      sb.seq[q] = s.seq[p];
      sb.id[q] = s.id[p];
    });
}

__global__ void k32_sort_tail(Species32 s, const uint32_t *n_live_ptr) {
  long n_live = (long)*n_live_ptr;
  for (long p = n_live + (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x) {
    s.cell[0][p] = DEAD_CELL;
    s.cell[1][p] = DEAD_CELL;
  }
}

__global__ void k32_zero_u32(uint32_t *p, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = 0;
}

__global__ void k32_count_live(const uint32_t *cell, long n, unsigned long long *out) {
  unsigned long long c = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
    c += cell[p] != DEAD_CELL;
  atomicAdd(out, c);
}

// ---------------------------------------------------------------------------
// synthetic inputs in split form (same sampling as pcp_init.hip)
// ---------------------------------------------------------------------------


__global__ void k32_init_plasma(Geom g, Species32 s, long base, long n_new, const long *cell_off, long n_cells,
                                double weight, double vth, uint64_t seed, const long *dom_off) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n_new; p += (long)gridDim.x * blockDim.x) {
    long lo = 0, hi = n_cells - 1;
    while (lo < hi) {
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
    if (r < r0) r = r0;
    if (r >= r1) r = nextafter(r1, 0.0);
    if (z >= (gk + 1) * g.dz) z = nextafter((gk + 1) * g.dz, 0.0);
    long q = base + p;
    store_split(s, q, g, r, z);
    s.vr[0][q] = s.vr[1][q] = (float)(n0 * vth);
    s.vphi[0][q] = s.vphi[1][q] = (float)(n1 * vth);
    s.vz[0][q] = s.vz[1][q] = (float)(n2 * vth);
    s.w[q] = (float)weight;
    s.seq[q] = (uint64_t)(p - dom_off[d]);
    s.id[q] = (uint64_t)q;
  }
}

__global__ void k32_init_bunch(Geom g, Species32 s, long base, long n_new, double sigma_r, double sigma_z, double z0,
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
    store_split(s, q, g, r, z);
    s.vr[0][q] = s.vr[1][q] = 0.f;
    s.vphi[0][q] = s.vphi[1][q] = 0.f;
    s.vz[0][q] = s.vz[1][q] = (float)vz;
    s.w[q] = (float)weight;
    int d = (int)((long)floor(r / g.dr) / g.nr) * g.Dz + (int)((long)floor(z / g.dz) / g.nz);
    s.seq[q] = gbase[(long)d * MAX_SPECIES + sp] + (uint64_t)p;
    s.id[q] = (uint64_t)q;
  }
}

__global__ void k32_gather_point(Geom g, FieldsT<float> f, int d, double r, double z, int which, double *out) {
  Dom dm = make_dom(g, d);
  double ur = r / g.dr, uz = z / g.dz;
  int cr = (int)floor(ur), cz = (int)floor(uz);
  float fr = (float)(ur - cr), fz = (float)(uz - cz);
  GridView32 V{(which == 0 ? f.E : f.H) + (long)d * 3 * g.P, g.P, g.yr, dm.bot_r, dm.left_z};
  float c[3];
  if (which == 0) gather32_e(V, cr, cz, fr, fz, c);
  else gather32_h(V, cr, cz, fr, fz, c);
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
}

// ---------------------------------------------------------------------------
// FP64 energy reductions (diagnostics)
// ---------------------------------------------------------------------------
// out: the sum's destination; with st, out is the energy log and the sum
// goes to column `col` of the current step's row
__device__ __forceinline__ double *log_dest(double *out, const StepState *st, int col) {
  return st ? ehist_row(out, st->step) + col : out;
}

template <class T>
__global__ void k_field_energy(Geom g, FieldsT<T> f, double *out, const StepState *st) {
  double acc = 0;
  long total = (long)g.Dr * g.Dz * g.nr * g.nz;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    long cells = (long)g.nr * g.nz;
    int d = (int)(t / cells);
    long c = t - (long)d * cells;
    int i = (int)(c / g.nz), k = (int)(c % g.nz);
    long o = (long)(i + 2) * g.yr + (k + 2);
    Dom dm = make_dom(g, d);
    double vol = PI * g.dz * g.dr * g.dr * (2.0 * (double)(i + dm.bot_r) + 1.0);
    double e2 = 0, h2 = 0;
    for (int cc = 0; cc < 3; ++cc) {
      double e = (double)f.E[((long)d * 3 + cc) * g.P + o];
      double h = (double)f.H[((long)d * 3 + cc) * g.P + o];
      e2 += e * e;
      h2 += h * h;
    }
    acc += 0.5 * (EPSILON0 * e2 + MAGN_CONST * h2) * vol;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(log_dest(out, st, 1), red[0]);
}

template <class V>
__device__ __forceinline__ double kin(double m, double w, V vr, V vp, V vz) {
  double v2 = (double)vr * vr + (double)vp * vp + (double)vz * vz;
  double b2 = v2 / LIGHT_VEL_POW_2;
  // m c^2 (gamma - 1) = m v^2 / (sqrt(1 - b2) (1 + sqrt(1 - b2))), cancellation free
  double s = sqrt(1.0 - b2);
  return w * m * v2 / (s * (1.0 + s));
}

__global__ void k_kinetic_energy64(SpeciesDev s, int cur, double *out, const StepState *st, int col) {
  double acc = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x)
    if (s.r[cur][p] >= 0) acc += kin(s.m, s.w[p], s.vr[p], s.vphi[p], s.vz[p]);
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(log_dest(out, st, col), red[0]);
}

__global__ void k_kinetic_energy32(Species32 s, int cur, double *out) {
  double acc = 0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < s.n; p += (long)gridDim.x * blockDim.x)
    if (s.cell[cur][p] != DEAD_CELL) acc += kin(s.m, (double)s.w[p], s.vr[cur][p], s.vphi[cur][p], s.vz[cur][p]);
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for32(long n, int block) { return (unsigned)((n + block - 1) / block); }
static inline unsigned pgrid32(pcp_ctx *c, long n, int block) {
  long want = (n + block - 1) / block;
  long cap = (long)c->n_cu * 16;
  if (want > cap) want = cap;
  return (unsigned)(want < 1 ? 1 : want);
}

template <class T>
int dalloc32(T **p, size_t n) {
  if (n == 0) n = 1;
  if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) return -1;
  if (hipMemset(*p, 0, n * sizeof(T)) != hipSuccess) return -1;
  return 0;
}

int alloc32(pcp_ctx *c) {
  State32 *S = new State32();
  c->s32 = S;
  const Geom &g = c->g;
  long gridn = (long)c->D * 3 * g.P;
  if (dalloc32(&S->f.E, gridn) || dalloc32(&S->f.H, gridn) || dalloc32(&S->f.Hat, gridn) ||
      dalloc32(&S->f.J, gridn) || dalloc32(&S->f.eps, (size_t)c->D * g.P) || dalloc32(&S->f.sigma, (size_t)c->D * g.P) ||
      dalloc32(&S->fix_list, FIX_LIST_CAP) || dalloc32(&S->fix_count, 1))
    return -1;
  for (int s = 0; s < c->cfg.n_species; ++s) {
    Species32 &sp = S->sp[s];
    SortBuffers32 &b = S->sb[s];
    sp.cap = c->cfg.species[s].capacity;
    sp.q = c->cfg.species[s].charge;
    sp.m = c->cfg.species[s].mass;
    sp.beam = c->cfg.species[s].is_beam;
    sp.n = 0;
    if (dalloc32(&sp.cell[0], sp.cap) || dalloc32(&sp.cell[1], sp.cap) || dalloc32(&sp.fr[0], sp.cap) ||
        dalloc32(&sp.fr[1], sp.cap) || dalloc32(&sp.fz[0], sp.cap) || dalloc32(&sp.fz[1], sp.cap) ||
        dalloc32(&sp.vr[0], sp.cap) || dalloc32(&sp.vphi[0], sp.cap) || dalloc32(&sp.vz[0], sp.cap) ||
        dalloc32(&sp.vr[1], sp.cap) || dalloc32(&sp.vphi[1], sp.cap) || dalloc32(&sp.vz[1], sp.cap) ||
        dalloc32(&sp.w, sp.cap) || dalloc32(&sp.seq, sp.cap) || dalloc32(&sp.id, sp.cap) ||
        dalloc32(&b.w, sp.cap) ||
        dalloc32(&b.seq, sp.cap) || dalloc32(&b.id, sp.cap))
      return -1;
  }
  return 0;
}

void free32(pcp_ctx *c) {
  State32 *S = c->s32;
  if (!S) return;
  (void)hipFree(S->f.E); (void)hipFree(S->f.H); (void)hipFree(S->f.Hat); (void)hipFree(S->f.J); (void)hipFree(S->Jc);
  (void)hipFree(S->fix_list); (void)hipFree(S->fix_count);
  (void)hipFree(S->f.eps); (void)hipFree(S->f.sigma);
  for (int s = 0; s < MAX_SPECIES; ++s) {
    Species32 &sp = S->sp[s];
    SortBuffers32 &b = S->sb[s];
    (void)hipFree(sp.cell[0]); (void)hipFree(sp.cell[1]); (void)hipFree(sp.fr[0]); (void)hipFree(sp.fr[1]);
    (void)hipFree(sp.fz[0]); (void)hipFree(sp.fz[1]); (void)hipFree(sp.vr[0]); (void)hipFree(sp.vphi[0]);
    (void)hipFree(sp.vz[0]); (void)hipFree(sp.vr[1]); (void)hipFree(sp.vphi[1]); (void)hipFree(sp.vz[1]);
    (void)hipFree(sp.w); (void)hipFree(sp.seq); (void)hipFree(sp.id);
    (void)hipFree(b.w);
    (void)hipFree(b.seq); (void)hipFree(b.id);
  }
  delete S;
  c->s32 = nullptr;
}

void launch32_yee_e(pcp_ctx *c, int clear_j) { launch_yee_e_t<float>(c, c->s32->f, clear_j); }
void launch32_yee_h(pcp_ctx *c) { launch_yee_h_t<float>(c, c->s32->f); }
void launch32_stitch(pcp_ctx *c, int which) { launch_stitch_t<float>(c, c->s32->f, which); }

static PushConst push_const(const Species32 &s, double dt) {
  PushConst K;
  double c2m = s.q * dt / (2 * s.m);
  K.c2m = (float)c2m;
  K.c2m_mu = (float)(c2m * MAGN_CONST);
  K.inv_c2 = (float)(1.0 / LIGHT_VEL_POW_2);
  return K;
}

void launch32_push(pcp_ctx *c, int sp) {
  Species32 &s = c->s32->sp[sp];
  if (!s.n) return;
  hipLaunchKernelGGL(k32_push, dim3(pgrid32(c, s.n, 256)), dim3(256), 0, c->stream, c->g, c->s32->f, s, c->cur, c->st,
                     push_const(s, c->g.dt));
}

void launch32_push_mom(pcp_ctx *c, int sp, double *mom, int slot) {
  Species32 &s = c->s32->sp[sp];
  if (!s.n) return;
  const long waves = (s.n + 63) / 64;
  const unsigned grid = (unsigned)std::max(1L, std::min((waves + 3) / 4, (long)c->n_cu * 32));
  hipLaunchKernelGGL(k32_push_mom, dim3(grid), dim3(256), 0, c->stream, c->g, c->s32->f, s, c->cur, c->st,
                     push_const(s, c->g.dt), mom, slot);
}

// element offset of the current's compensation array from J (0: the
// deposit is not compensated)
static long comp_offset(const pcp_ctx *c) { return c->s32->Jc ? (long)(c->s32->Jc - c->s32->f.J) : 0; }

// J += Jc and Jc = 0 over every node: the compensated deposit's rounding
// errors folded into the current once per step, before the stitch
__global__ void __launch_bounds__(256) k32_fold_j(float *J, float *Jc, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float e = Jc[i];
    if (e != 0.f) {
      J[i] += e;
      Jc[i] = 0.f;
    }
  }
}

void launch32_fold_j(pcp_ctx *c) {
  if (!c->s32->Jc) return;
  const long n = (long)c->D * 3 * c->g.P;
  const unsigned grid = (unsigned)std::max(1L, std::min((n + 255) / 256, (long)c->n_cu * 16));
  hipLaunchKernelGGL(k32_fold_j, dim3(grid), dim3(256), 0, c->stream, c->s32->f.J, c->s32->Jc, n);
}

int set_compensated32(pcp_ctx *c, bool on) {
  if (on == (c->s32->Jc != nullptr)) return 0;
  if (!on) {
    (void)hipFree(c->s32->Jc);
    c->s32->Jc = nullptr;
    return 0;
  }
  const size_t bytes = (size_t)c->D * 3 * c->g.P * sizeof(float);
  float *p = nullptr;
  if (hipMalloc((void **)&p, bytes) != hipSuccess) return -1;
  if (hipMemsetAsync(p, 0, bytes, c->stream) != hipSuccess) {
    (void)hipFree(p);
    return -1;
  }
  c->s32->Jc = p;
  return 0;
}

// consecutive wavefront trips per counter increment of the batched step
constexpr int TRIPS_PER_TICKET = 16;

template <int CCS, bool PUSH, int BLOCK, bool D1, bool COMP = false>
static void launch_batch_t(pcp_ctx *c, const Batch32 &b) {
  static int occ = 0;   // resident blocks per CU of this instance
  if (!occ &&
      (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k32_step_batch<CCS, PUSH, BLOCK, D1, COMP>, BLOCK, 0) !=
           hipSuccess ||
       occ < 1))
    occ = 1;
  const int per_cu = c->tune.blocks_per_cu > 0 ? c->tune.blocks_per_cu : occ;
  const int chunk = TRIPS_PER_TICKET;
  const long waves = (b.end[b.nsp - 1] + chunk - 1) / chunk;
  const long grid = std::max(1L, std::min((waves + BLOCK / 64 - 1) / (BLOCK / 64), (long)c->n_cu * per_cu));
  BatchArgs32 args{c->g, c->s32->f, b, c->st, c->mig, c->elog ? c->ehist : nullptr, COMP ? comp_offset(c) : 0,
                   c->cur, chunk};
  hipLaunchKernelGGL((k32_step_batch<CCS, PUSH, BLOCK, D1, COMP>), dim3((unsigned)grid), dim3(BLOCK), 0, c->stream,
                     args);
}

template <int CCS>
static void launch_batch_ccs(pcp_ctx *c, const Batch32 &b, bool push, int block) {
  // the compensated deposit runs at one launch shape
  if (c->s32->Jc) {
    if (c->D != 1) {
      if (push) launch_batch_t<CCS, true, 256, false, true>(c, b);
      else launch_batch_t<CCS, false, 256, false, true>(c, b);
    } else {
      if (push) launch_batch_t<CCS, true, 256, true, true>(c, b);
      else launch_batch_t<CCS, false, 256, true, true>(c, b);
    }
    return;
  }
  if (c->D != 1) {
    if (!push) launch_batch_t<CCS, false, 256, false>(c, b);
    else if (block == 128) launch_batch_t<CCS, true, 128, false>(c, b);
    else if (block == 512) launch_batch_t<CCS, true, 512, false>(c, b);
    else launch_batch_t<CCS, true, 256, false>(c, b);
    return;
  }
  if (!push) launch_batch_t<CCS, false, 256, true>(c, b);
  else if (block == 128) launch_batch_t<CCS, true, 128, true>(c, b);
  else if (block == 512) launch_batch_t<CCS, true, 512, true>(c, b);
  else launch_batch_t<CCS, true, 256, true>(c, b);
}

// the collision slots of the species (electrons 0, ions 1, others -1) when
// the step's surrogate rescaling `cpar` is given
static void batch_coll(pcp_ctx *c, Batch32 &b, const float *cpar, int se, int si) {
  b.cpar = cpar;
  for (int k = 0; k < b.nsp; ++k) b.cslot[k] = !cpar ? -1 : b.sp[k] == se ? 0 : b.sp[k] == si ? 1 : -1;
}

void launch32_step_batch(pcp_ctx *c, bool push, const float *cpar, int se, int si) {
  Batch32 b{};
  const int block = push ? c->tune.block : 256;
  long acc = 0;   // trips of 64 particles
  for (int sp = 0; sp < c->cfg.n_species; ++sp) {
    const Species32 &s = c->s32->sp[sp];
    if (!s.n) continue;
    acc += (s.n + 63) / 64;
    b.s[b.nsp] = s;
    b.K[b.nsp] = push_const(s, c->g.dt);
    b.end[b.nsp] = acc;
    b.sp[b.nsp] = sp;
    ++b.nsp;
  }
  b.marker = c->marker;
  b.cflag = c->cell_flag;
  batch_coll(c, b, cpar, se, si);
  if (b.nsp) {
    if (c->g.ccs == PCP_CCS_ZIGZAG) launch_batch_ccs<PCP_CCS_ZIGZAG>(c, b, push, block);
    else launch_batch_ccs<PCP_CCS_VB>(c, b, push, block);
  }
  if (!push)
    for (int sp = 0; sp < c->cfg.n_species; ++sp) {
      Species32 &s = c->s32->sp[sp];
      std::swap(s.vr[0], s.vr[1]);
      std::swap(s.vphi[0], s.vphi[1]);
      std::swap(s.vz[0], s.vz[1]);
    }
}

void launch32_fix(pcp_ctx *c, bool push, const float *cpar, int se, int si) {
  if (c->g.ccs != PCP_CCS_VB) return;
  Batch32 b{};
  long acc = 0;
  for (int sp = 0; sp < c->cfg.n_species; ++sp) {
    const Species32 &s = c->s32->sp[sp];
    if (!s.n) continue;
    acc += (s.n + 255) / 256;
    b.s[b.nsp] = s;
    b.K[b.nsp] = push_const(s, c->g.dt);
    b.end[b.nsp] = acc;
    b.sp[b.nsp] = sp;
    ++b.nsp;
  }
  if (!b.nsp) return;
  b.marker = c->marker;
  b.cflag = c->cell_flag;
  batch_coll(c, b, cpar, se, si);
  // the screen: 256-particle units, 16 threads per unit; then the redo
  launch_set_u64(c, (uint64_t *)c->s32->fix_count, 0);
  const unsigned gs = (unsigned)std::max(1L, std::min((acc * 16 + 255) / 256, (long)c->n_cu * 32));
  hipLaunchKernelGGL(k32_fix_screen, dim3(gs), dim3(256), 0, c->stream, b, c->cur, c->st, c->s32->fix_list,
                     c->s32->fix_count);
  const unsigned grid = (unsigned)std::max(1L, std::min((acc + 3) / 4, (long)c->n_cu * 16));
  hipLaunchKernelGGL(k32_fix, dim3(grid), dim3(256), 0, c->stream, c->g, c->s32->f, b, c->cur, c->st, push ? 1 : 0,
                     comp_offset(c), c->s32->fix_list, c->s32->fix_count);
}

void launch32_migrant_keys(pcp_ctx *c) {
  if (c->D == 1) return;
  SeqTable t;
  for (int i = 0; i < MAX_SPECIES; ++i) t.p[i] = c->s32->sp[i].seq;
  launch_migrant_keys_t(c, t);
}

int launch32_sort(pcp_ctx *c, int sp) {
  Species32 &s = c->s32->sp[sp];
  if (s.n == 0) return 0;
  SortBuffers32 &sb = c->s32->sb[sp];
  const long ncells = (long)c->D * c->g.nr * c->g.nz;
  const uint32_t sentinel = (uint32_t)ncells;
  unsigned grid = pgrid32(c, s.n, 256);
  const unsigned gsort = (unsigned)std::max(1L, std::min((s.n + SORT_CHUNK - 1) / SORT_CHUNK, (long)c->n_cu * 16));
  unsigned gz = (unsigned)std::min<long>((ncells + 256) / 256, (long)c->n_cu * 16);
  hipLaunchKernelGGL(k32_zero_u32, dim3(gz), dim3(256), 0, c->stream, c->sort_count, ncells + 1);
  hipLaunchKernelGGL(k32_zero_u32, dim3(gz), dim3(256), 0, c->stream, c->sort_fill, ncells + 1);
  hipLaunchKernelGGL(k32_sort_count, dim3(gsort), dim3(SORT_BLOCK), 0, c->stream, c->g, s, c->cur, c->sort_key, c->sort_count,
                     sentinel);
  size_t bytes = c->sort_temp_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(c->sort_temp, bytes, c->sort_count, c->sort_offset, (int)(ncells + 1),
                                       c->stream) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(k32_sort_scatter, dim3(gsort), dim3(SORT_BLOCK), 0, c->stream, s, sb, c->cur, c->sort_key,
                     c->sort_offset, c->sort_fill, sentinel);
  std::swap(s.cell[c->cur], s.cell[c->cur ^ 1]);
  std::swap(s.fr[c->cur], s.fr[c->cur ^ 1]);
  std::swap(s.fz[c->cur], s.fz[c->cur ^ 1]);
  std::swap(s.vr[c->cur], s.vr[c->cur ^ 1]);
  std::swap(s.vphi[c->cur], s.vphi[c->cur ^ 1]);
  std::swap(s.vz[c->cur], s.vz[c->cur ^ 1]);
  std::swap(s.w, sb.w);
  std::swap(s.seq, sb.seq);
  std::swap(s.id, sb.id);
  hipLaunchKernelGGL(k32_sort_tail, dim3(grid), dim3(256), 0, c->stream, s, c->sort_offset + ncells);
  return 1;
}

void launch32_count_live(pcp_ctx *c, int sp, unsigned long long *dev_out) {
  Species32 &s = c->s32->sp[sp];
  hipLaunchKernelGGL(k32_count_live, dim3(pgrid32(c, s.n, 256)), dim3(256), 0, c->stream, s.cell[c->cur], s.n, dev_out);
}

void launch32_init_plasma(pcp_ctx *c, int sp, long base, long n_new, const long *dev_cell_off, long n_cells,
                          double weight, double vth, uint64_t seed, const long *dev_dom_off) {
  unsigned grid = (unsigned)std::min<long>((n_new + 255) / 256, (long)c->n_cu * 32);
  hipLaunchKernelGGL(k32_init_plasma, dim3(grid ? grid : 1), dim3(256), 0, c->stream, c->g, c->s32->sp[sp], base,
                     n_new, dev_cell_off, n_cells, weight, vth, seed, dev_dom_off);
}

void launch32_init_bunch(pcp_ctx *c, int sp, long base, long n_new, double sigma_r, double sigma_z, double z0,
                         double vz, double weight, uint64_t seed) {
  unsigned grid = (unsigned)std::min<long>((n_new + 255) / 256, (long)c->n_cu * 32);
  hipLaunchKernelGGL(k32_init_bunch, dim3(grid ? grid : 1), dim3(256), 0, c->stream, c->g, c->s32->sp[sp], base,
                     n_new, sigma_r, sigma_z, z0, vz, weight, seed, c->mig.gbase, sp);
}

void launch32_gather_point(pcp_ctx *c, int d, double r, double z, int which, double *dev_out) {
  hipLaunchKernelGGL(k32_gather_point, dim3(1), dim3(1), 0, c->stream, c->g, c->s32->f, d, r, z, which, dev_out);
}

void launch_field_energy(pcp_ctx *c, double *dev_out, bool to_log) {
  long total = (long)c->D * c->g.nr * c->g.nz;
  unsigned grid = (unsigned)std::min<long>((total + 255) / 256, (long)c->n_cu * 8);
  const StepState *st = to_log ? c->st : nullptr;
  if (c->s32)
    hipLaunchKernelGGL(k_field_energy<float>, dim3(grid), dim3(256), 0, c->stream, c->g, c->s32->f, dev_out, st);
  else
    hipLaunchKernelGGL(k_field_energy<double>, dim3(grid), dim3(256), 0, c->stream, c->g, c->f, dev_out, st);
}

// velocity moments of a species: sum w, sum w v (3), sum w v_z^2,
// sum w (v_r^2 + v_phi^2), live count
template <class S, class F>
__global__ void k_species_moments(S s, long n, F live_at, double *out) {
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    double w, v[3];
    if (!live_at(s, p, w, v)) continue;
    a[0] += w;
    a[1] += w * v[0];
    a[2] += w * v[1];
    a[3] += w * v[2];
    a[4] += w * v[2] * v[2];
    a[5] += w * (v[0] * v[0] + v[1] * v[1]);
    a[6] += 1.0;
  }
  for (int k = 0; k < 7; ++k) {
    double v = a[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&out[k], v);
  }
}

struct Live64 {
  int cur;
  __device__ bool operator()(const SpeciesDev &s, long p, double &w, double v[3]) const {
    if (!(s.r[cur][p] >= 0)) return false;
    w = s.w[p]; v[0] = s.vr[p]; v[1] = s.vphi[p]; v[2] = s.vz[p];
    return true;
  }
};
struct Live32 {
  int cur;
  __device__ bool operator()(const Species32 &s, long p, double &w, double v[3]) const {
    if (s.cell[cur][p] == DEAD_CELL) return false;
    w = s.w[p]; v[0] = s.vr[cur][p]; v[1] = s.vphi[cur][p]; v[2] = s.vz[cur][p];
    return true;
  }
};

void launch_species_moments(pcp_ctx *c, int sp, double *dev_out) {
  if (c->s32) {
    Species32 &s = c->s32->sp[sp];
    if (!s.n) return;
    hipLaunchKernelGGL((k_species_moments<Species32, Live32>), dim3(pgrid32(c, s.n, 256)), dim3(256), 0, c->stream,
                       s, s.n, Live32{c->cur}, dev_out);
  } else {
    SpeciesDev &s = c->sp[sp];
    if (!s.n) return;
    hipLaunchKernelGGL((k_species_moments<SpeciesDev, Live64>), dim3(pgrid32(c, s.n, 256)), dim3(256), 0, c->stream,
                       s, s.n, Live64{c->cur}, dev_out);
  }
}

void launch_kinetic_energy(pcp_ctx *c, int sp, double *dev_out) {
  if (c->s32) {
    Species32 &s = c->s32->sp[sp];
    if (!s.n) return;
    hipLaunchKernelGGL(k_kinetic_energy32, dim3(pgrid32(c, s.n, 256)), dim3(256), 0, c->stream, s, c->cur, dev_out);
  } else {
    SpeciesDev &s = c->sp[sp];
    if (!s.n) return;
    hipLaunchKernelGGL(k_kinetic_energy64, dim3(pgrid32(c, s.n, 256)), dim3(256), 0, c->stream, s, c->cur, dev_out,
                       nullptr, 0);
  }
}

void launch_kinetic_energy_log(pcp_ctx *c, int sp) {
  SpeciesDev &s = c->sp[sp];
  if (!s.n || c->s32) return;
  hipLaunchKernelGGL(k_kinetic_energy64, dim3(pgrid32(c, s.n, 256)), dim3(256), 0, c->stream, s, c->cur, c->ehist,
                     c->st, 2 + sp);
}

}  // namespace pcp
