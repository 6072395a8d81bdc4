// The reference's initial plasma distribution, per particle (device side):
// the "flat" placement of SpecieP::rectangular_random_placement (radical
// inverses of the particle's index), set_weight's r-proportional weights and
// thermal_velocity_distribution's Maxwell-Juettner energies
// (src/specieP.cpp:127-189, 280-331, 346-384;
// src/math/maxwellJuettner.cpp:182-255).
//
// The reference inverts tabulated CDFs (linear interpolation in log space)
// drawn from a time-seeded mt19937; here each inverse CDF is solved to double
// precision by bisection and the uniforms come from Philox (counter
// domain << 40 | index in the domain, key = seed; proposal j uses draw j, the
// velocity directions draws 1000 and 1001), so a run is reproducible and
// the energy law is the reference's without its table error:
//   T < 0.1:  E = T invF(U), invF the inverse CDF of Gamma(3/2), with the
//             reference's tail branches kept: log(-log U) > 2 gives
//             (4/3) sqrt(pi) U^(2/3) (its THREE_OVER_4 constant is 4/3),
//             log(-log U) < -19 gives 1;
//   T >= 0.1: gamma = T x, x solving g(x) = -log(1 - U) + g(1/T) with
//             g(x) = x - log(1 + x + x^2/2) (Gamma(3) truncated to
//             gamma >= 1), accepted when a second uniform is below
//             sqrt(1 - 1/gamma^2); E = gamma - 1 (the temperature is taken in
//             units of m c^2, as the reference takes it).
#pragma once

#include "pcp_internal.hpp"
#include "pcp_rng.hpp"

namespace pcp {

// host-computed pow(b, -k) of the radical inverses (the reference calls
// pow per digit; the host's libm values keep the placements bit-identical)
constexpr int RADICAL_DIGITS = 24;
struct RadicalPowers {
  double p13[RADICAL_DIGITS], p11[RADICAL_DIGITS];
};

// per-domain parameters of the placement and the weights
struct RefInitDom {
  double r_size, z_size;        // placement extents after the wall reduction
  double dens0, dens1;          // local left / right density
  double v_avg;                 // set_weight's mean macro volume (host sum)
  double n_per_macro_avg;
  uint32_t int_cell, left_cell;
  uint64_t seq_base;            // order key of the domain's first particle
};

// math::random::random_reverse (src/math/rand.cpp:60-76) in base BASE,
// the digit powers taken from the table
template <int BASE>
__host__ __device__ __forceinline__ double radical_inverse_t(uint32_t vel, const double *pw) {
  int int_vel = (int)vel;
  double r = 0;
  int order = 1;
  while (int_vel >= 1) {
    const double ost = (double)(int_vel % BASE);
    r = r + ost * pw[order - 1];
    int_vel = (int)(((double)int_vel - ost) / BASE);
    order = order + 1;
  }
  return r;
}

// inverse of the Gamma(3/2) CDF P(3/2, x) = erf(s) - 2 s exp(-s^2)/sqrt(pi),
// s = sqrt(x): bisection on s in [0, 6], the complement above the median
__device__ __forceinline__ double gamma32_quantile(double U) {
  const double two_over_sqrt_pi = 1.1283791670955126;
  double lo = 0.0, hi = 6.0;
  for (int it = 0; it < 80; ++it) {
    const double s = 0.5 * (lo + hi);
    double f;
    if (U < 0.5)
      f = (erf(s) - two_over_sqrt_pi * s * exp(-s * s)) - U;
    else
      f = (1.0 - U) - (erfc(s) + two_over_sqrt_pi * s * exp(-s * s));
    if (f < 0) lo = s; else hi = s;
  }
  const double s = 0.5 * (lo + hi);
  return s * s;
}

__device__ __forceinline__ double mj_g(double x) { return x - log1p(x + 0.5 * x * x); }

// inverse of g on [x_lo, x_lo + 200] (g increasing)
__device__ __forceinline__ double mj_g_inverse(double t, double x_lo) {
  double lo = x_lo, hi = x_lo + 200.0;
  for (int it = 0; it < 100; ++it) {
    const double x = 0.5 * (lo + hi);
    if (mj_g(x) < t) lo = x; else hi = x;
  }
  return 0.5 * (lo + hi);
}

// one energy of maxwellJuettner(); NaN when 1000 proposals were rejected
__device__ __forceinline__ double mj_energy(double temperature, uint64_t seed, uint64_t ctr) {
  if (temperature < 0.1) {
    double U, u1;
    uniform2(ctr, 0, seed, U, u1);
    const double lnlnU = log(-log(U));
    double invF;
    if (lnlnU > 2.)
      invF = (4. / 3.) * sqrt(PI) * pow(U, 2. / 3.);
    else if (lnlnU < -19.)
      invF = 1.;
    else
      invF = gamma32_quantile(U);
    return temperature * invF;
  }
  const double invT = 1. / temperature;
  const double g0 = mj_g(invT);
  for (uint32_t j = 0; j < 1000; ++j) {
    double U, V;
    uniform2(ctr, j, seed, U, V);
    const double x = mj_g_inverse(-log(1. - U) + g0, invT);
    const double gamma = temperature * x;
    if (V < sqrt(1. - 1. / (gamma * gamma))) return gamma - 1.;
  }
  return __builtin_nan("");
}

// thermal_velocity_distribution for one particle (specieP.cpp:346-384)
__device__ __forceinline__ void thermal_velocity(double energy, double mass, uint64_t seed, uint64_t ctr,
                                                 double v[3]) {
  const double two_over_mass = 2 * EL_CHARGE / mass;
  const double mc_inv = EL_CHARGE / (mass * LIGHT_VEL);
  const double norm = 0.7071067811865475;
  double therm_vel_cmp = energy * mc_inv;
  if (therm_vel_cmp > REL_LIMIT) {
    const double gamma_inv = pow(1.0 + therm_vel_cmp * therm_vel_cmp / LIGHT_VEL_POW_2, -0.5);
    therm_vel_cmp *= gamma_inv;
  } else {
    therm_vel_cmp = sqrt(energy * two_over_mass);
  }
  therm_vel_cmp *= norm;
  double u[4];
  uniform2(ctr, 1000, seed, u[0], u[1]);
  uniform2(ctr, 1001, seed, u[2], u[3]);
  for (int k = 0; k < 3; ++k) v[k] = (u[k] * 2.0 + -1.0) * therm_vel_cmp;
}

// Position of particle k of a domain placing M particles (host and device:
// the host sums set_weight's volumes over the same positions).
__host__ __device__ __forceinline__ void ref_init_position(double dr, double dz, const RefInitDom &P,
                                                           const RadicalPowers &pw, uint32_t M, uint32_t k,
                                                           double &r, double &z) {
  double rand_r = radical_inverse_t<13>(k, pw.p13);
  double rand_z = radical_inverse_t<11>(M - 1 - k, pw.p11);
  if (rand_r == 0) rand_r = MNZL;
  if (rand_z == 0) rand_z = MNZL;
  r = P.r_size * rand_r;
  if (P.dens0 == P.dens1) {
    z = P.z_size * rand_z;
  } else {
    const double dn = P.dens1 - P.dens0, dl = P.dens0;
    z = P.z_size / dn * (sqrt(dl * dl + rand_z * (2 * dl * dn + dn * dn)) - dl);
  }
  r += P.int_cell * dr;
  r += dr / 2.;
  z += P.left_cell * dz;
  z += dz / 2.;
}

// set_weight's per-particle volume 2 pi r dr dz (pi r^2 dz at r <= dr/2)
__host__ __device__ __forceinline__ double ref_init_volume(double dr, double dz, double r) {
  if (r > dr / 2) return 2 * PI * r * dr * dz;
  return PI * r * r * dz;
}

// The state of particle k of domain d: position, weight and velocity.
// Returns false when the energy draw did not accept.
__device__ __forceinline__ bool ref_init_particle(const Geom &g, const RefInitDom &P, const RadicalPowers &pw,
                                                  uint32_t M, uint32_t k, int d, double temperature, double mass,
                                                  uint64_t seed, double &r, double &z, double v[3], double &w) {
  const double dr = g.dr, dz = g.dz;
  ref_init_position(dr, dz, P, pw, M, k, r, z);
  double norm;
  if (r > dr / 2)
    norm = 2 * PI * r * dr * dz / P.v_avg;
  else
    norm = PI * r * r * dz / P.v_avg;
  w = P.n_per_macro_avg * norm;
  const uint64_t ctr = ((uint64_t)d << 40) | k;
  const double e = mj_energy(temperature, seed, ctr);
  thermal_velocity(e, mass, seed, ctr, v);
  return e == e;
}

}  // namespace pcp
