/*
 * CPU oracle for the PiCoPiC per-step particle loop (test infrastructure only).
 *
 * A plain-C restatement of the reference's FP64 algorithm: field gather,
 * Boris / Vay / Higuera-Cary pushers, cylindrical mover + reflection + rz
 * rotation + cell binding, Villasenor-Buneman and ZigZag current deposition,
 * the Yee E/H update with PML, the domain-edge overlay stitching, the
 * runaway collector (particle migration between domains) and the P12 binary
 * collision operator.  Every function cites the reference file:line it
 * follows; arithmetic is written in the same order as the reference so that
 * an IEEE build (-ffp-contract=off, no fast-math) reproduces it.
 *
 * This library is test infrastructure: only tests, the smoke check and the
 * CPU baseline leg of the benchmark load it, to check the device results.
 * It is never the product path.
 *
 * The parity status of each function (which known-answer vector, closed
 * form, invariant or brute-force check pins it, or "parity unpinned") is
 * listed at the top of picopic_oracle.c.
 */
#ifndef PICOPIC_ORACLE_H
#define PICOPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/constant.hpp:28-58 */
#define ORC_PI 3.1415926535897932
#define ORC_EPSILON0 8.85E-12
#define ORC_EL_MASS 9.1E-31
#define ORC_EL_CHARGE 1.6E-19
#define ORC_LIGHT_VEL 3.0E8
#define ORC_MAGN_CONST 1.26E-6
#define ORC_MNZL 1e-15
#define ORC_PLANK_CONST 6.6261e-34
/* configure.ac:224 */
#define ORC_REL_LIMIT 5e7

enum { ORC_PUSHER_BORIS = 0, ORC_PUSHER_BORIS_REL = 1, ORC_PUSHER_BORIS_ADAPTIVE = 2,
       ORC_PUSHER_VAY = 3, ORC_PUSHER_HC = 4 };
enum { ORC_CCS_VB = 0, ORC_CCS_ZIGZAG = 1 };
enum { ORC_GRID_E = 0, ORC_GRID_H = 1, ORC_GRID_H_AT_ET = 2, ORC_GRID_J = 3,
       ORC_GRID_EPS = 4, ORC_GRID_SIGMA = 5 };

/* reference include/geometry.hpp:27-100 (per-domain Geometry) */
typedef struct {
  double size[2];
  double global_size[2];
  double cell_size[2];
  long cell_dims[4];          /* bot_r, left_z, top_r, right_z */
  long cell_amount[2];
  long global_cell_amount[2];
  int walls[4];               /* r=0, z=0, r=R, z=L */
  long pml_size[4];
  double pml_sigma[2];
} orc_geometry;

/* reference include/algo/grid.hpp:24-58: halo grid, x = r rows, y = z contiguous */
typedef struct {
  int o_s;
  int x_real, y_real;
  int x_size, y_size;
  double *g;                  /* g[i * y_real + j] */
} orc_grid;

/* reference include/specieP.hpp:74-119 (Particle), plus a stable id */
typedef struct {
  double pos_r, pos_phi, pos_z;
  double pos_old_r, pos_old_phi, pos_old_z;
  double vel_r, vel_phi, vel_z;
  double weight;
  double sin;
  long cell_r, cell_z;
  uint64_t id;
} orc_particle;

/* global configuration (the JSON geometry/time/particles blocks) */
typedef struct {
  double size[2];             /* radius, longitude [m] */
  long grid[2];               /* cells r, z */
  long domains[2];            /* domains r, z */
  double dt;
  int pusher;                 /* ORC_PUSHER_* */
  int ccs;                    /* ORC_CCS_* */
  int pml_on;                 /* ENABLE_PML */
  double pml_left, pml_right, pml_outer;   /* fractions of the grid */
  double pml_sigma1, pml_sigma2;
  int n_species;
  double charge[8];           /* SI (already multiplied by EL_CHARGE) */
  double mass[8];             /* SI */
  int is_beam[8];             /* BeamP::reflect instead of SpecieP::reflect */
} orc_config;

typedef struct orc_smb orc_smb;

/* ---- SMB-level API (reference src/SMB.cpp) ---- */
orc_smb *orc_smb_create(const orc_config *cfg);
void orc_smb_destroy(orc_smb *s);
int orc_smb_n_domains(const orc_smb *s);
void orc_smb_domain_geometry(const orc_smb *s, int d, orc_geometry *out);
/* cells per domain including halo: (Nr_d + 4) x (Nz_d + 4) */
long orc_smb_grid_points(const orc_smb *s);

/* Place particles: each goes to the domain that owns its cell; order within a
 * domain follows input order (reference Domain::distribute + bind). */
int orc_smb_add_particles(orc_smb *s, int sp, long n, const double *r, const double *z,
                          const double *vr, const double *vphi, const double *vz,
                          const double *w, const uint64_t *id);
long orc_smb_count(const orc_smb *s, int sp);
/* Domain-major, vector order within a domain. dom may be NULL. */
void orc_smb_get_particles(const orc_smb *s, int sp, double *r, double *z, double *vr,
                           double *vphi, double *vz, double *w, uint64_t *id, int *dom,
                           double *r_old, double *z_old);

/* whole halo grids, domain-major: [domain][comp][x_real][y_real] */
void orc_smb_get_grid(const orc_smb *s, int which, double *out);
void orc_smb_set_grid(orc_smb *s, int which, const double *in);

/* step phases (reference src/SMB.cpp:1231-1306, src/domain.cpp:118-204) */
enum { ORC_PH_YEE_E = 0, ORC_PH_OVERLAY_E, ORC_PH_YEE_H, ORC_PH_OVERLAY_H,
       ORC_PH_PUSH, ORC_PH_COLLIDE, ORC_PH_MOVE, ORC_PH_MIGRATE, ORC_PH_RESET_J,
       ORC_PH_DEPOSIT, ORC_PH_OVERLAY_J, ORC_PH_COUNT };
void orc_smb_phase(orc_smb *s, int phase);
/* beam injection + solve_maxvell + advance_particles + solve_current, n
 * times; the clock (TimeSim::current) advances by dt after each */
void orc_smb_step(orc_smb *s, int n);
double orc_smb_time(const orc_smb *s);
void orc_smb_set_time(orc_smb *s, double t);
/* BeamP::BeamP for species sp in every domain (src/beamP.cpp:24-66); the
 * injection (BeamP::inject, src/beamP.cpp:71-157) then runs at the start of
 * every step.  Injected particle ids: 1 << 62 | domain << 40 | count. */
int orc_smb_set_beam(orc_smb *s, int sp, long macro_amount, double start_time, double radius, double density,
                     long bunches_amount, double bunch_length, double bunches_distance, double velocity,
                     uint64_t seed);
/* SMB::inject_beam at the current time */
void orc_smb_inject(orc_smb *s);
void orc_set_threads(int n);

/* ---- unit-level entry points used by the known-answer tests ---- */
void orc_grid_overlay_x(orc_grid *a, orc_grid *b);
void orc_grid_overlay_y(orc_grid *a, orc_grid *b);
void orc_grid_overlay_xy(orc_grid *a, orc_grid *b);
void orc_grid_overlay_set(orc_grid *a, double v);

/* gather on domain d of an SMB */
void orc_smb_get_field_e(const orc_smb *s, int d, double r, double z, double out[3]);
void orc_smb_get_field_h(const orc_smb *s, int d, double r, double z, double out[3]);
/* one particle velocity update with given (already gathered) fields */
void orc_push_velocity(int pusher, double charge, double mass, double dt,
                       const double e[3], const double h[3], double v[3]);
double orc_lorenz_factor(double sq_velocity);
double orc_lorenz_factor_inv(double sq_velocity);

/* ---- collisions (reference src/collisions.cpp, src/collisions/collisionsP12.cpp) ----
 * Random draws are taken from a caller-supplied stream so that a GPU
 * implementation fed the same stream can be compared pair by pair. */
typedef struct {
  const double *u;  /* uniform [0,1) draws, consumed in order */
  long n, pos;
} orc_rng_stream;
/* P12 on one domain for the electron species se and ion species si, cells
 * visited r-major exactly like CollisionsP12::collide; the per-cell particle
 * order is the given vector order (the random shuffle is the caller's). */
void orc_smb_collide_p12(orc_smb *s, int d, int se, int si, orc_rng_stream *rng);
void orc_p12_collide_single(double m_a, double m_b, double q_a, double q_b,
                            double *va, double wa, double *vb, double wb,
                            double dens_a, double dens_b, double debye, double dt,
                            orc_rng_stream *rng);
double orc_debye_length(double n_e, double n_i, double t_e, double t_i);

/* collision schemes: configure.ac --with-coulomb-collisions (P12 default) */
enum { ORC_COLL_P12 = 0, ORC_COLL_TA77S = 1, ORC_COLL_SK98 = 2 };
/* TA77S and SK98 collide_single (collisionsTA77S.cpp:31-219,
 * collisionsSK98.cpp:29-244); stream draws: normal deviate from two
 * uniforms (Box-Muller), Phi, and (TA77S) dzeta */
void orc_ta77s_collide_single(double m_a, double m_b, double q_a, double q_b,
                              double *va, double wa, double *vb, double wb,
                              double dens_a, double dens_b, double debye, double dt,
                              orc_rng_stream *rng);
void orc_sk98_collide_single(double m_a, double m_b, double q_a, double q_b,
                             double *va, double wa, double *vb, double wb,
                             double dens_a, double dens_b, double debye, double dt,
                             orc_rng_stream *rng);
/* Philox-4x32-10 (Salmon et al., SC11) as the device library computes it:
 * counter (ctr, draw, 0x5eed), key = seed */
void orc_philox4x32(uint64_t ctr, uint32_t draw, uint64_t seed, uint32_t out[4]);
void orc_philox4x32_raw(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* One collision pass of `scheme` on domain d with the counter-based draws of
 * the device library: each cell's particles are shuffled by sorting on a
 * Philox key of (particle id, step) (ties by id) -- the reference's
 * time-seeded random_shuffle has no reproducible permutation -- and pair q of
 * global cell c takes the four uniforms orc_coll_uniforms(seed, c, q, step). */
void orc_smb_collide(orc_smb *s, int d, int scheme, int se, int si, uint64_t seed, uint64_t step);
void orc_coll_uniforms(uint64_t seed, uint64_t cell, uint32_t q, uint64_t step, double u[4]);
uint32_t orc_coll_key(uint64_t seed, uint64_t id, uint64_t step);
/* enable the collision phase of orc_smb_step (scheme -1: off) */
void orc_smb_set_collisions(orc_smb *s, int scheme, int se, int si, uint64_t seed);
double orc_coulomb_logarithm(double m_a, double m_b, double debye, double v_rel);

/* ---- initial distribution (src/specieP.cpp:79-189, 280-331, 346-384;
 * src/math/maxwellJuettner.cpp:182-255): flat placement by radical inverses,
 * r-proportional weights, Maxwell-Juettner thermal velocities with Philox
 * draws (counter domain << 40 | index, key seed).  Returns -2 when an energy
 * draw did not accept. */
int orc_smb_init_reference(orc_smb *s, int sp, long macro_amount, double density_left, double density_right,
                           double temperature, uint64_t seed, uint64_t id_base);
double orc_random_reverse(double vel, int power);
double orc_gamma32_quantile(double U);
double orc_mj_g(double x);
double orc_mj_g_inverse(double t, double x_lo);
double orc_mj_energy(double temperature, uint64_t seed, uint64_t ctr);

/* species maps of the probes (src/specieP.cpp:665-839, include/algo/weighter.hpp:27-132)
 * with OutController's pre-dump overlay (src/outController.cpp:409-491);
 * out: [D][ncomp][Nr_d+4][Nz_d+4], ncomp 3 for the momentum map.  The map
 * grids of a species belong to the last SMB that asked for them. */
enum { ORC_MAP_DENSITY_ = 0, ORC_MAP_TEMPERATURE_ = 1, ORC_MAP_ENERGY_ = 2, ORC_MAP_MOMENTUM_ = 3 };
void orc_smb_species_map(orc_smb *s, int sp, int kind, double *out);

#ifdef __cplusplus
}
#endif
#endif
