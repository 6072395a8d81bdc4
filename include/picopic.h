/*
 * picopic_amd -- MI355X (gfx950) implementation of the PiCoPiC per-step
 * particle loop behind a C ABI.
 *
 * One context (`pcp_ctx`) holds one shared-memory block (SMB) of the
 * reference: an r x z grid of equal domains with halo-2 field grids, and the
 * particles of every species, all resident in device memory.  The caller owns
 * host buffers; the context owns all device memory.  Calls are ordered on the
 * context's stream; a context is driven by one host thread.
 *
 * Errors: every entry point returns 0 on success and a negative PCP_E* code
 * otherwise; pcp_last_error() gives the message of the calling thread's last
 * failure.  Nothing throws across this boundary.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * cosmonaut-ok/PiCoPiC).
 */
#ifndef PICOPIC_H
#define PICOPIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCP_OK 0
#define PCP_E_ARG -1      /* bad argument */
#define PCP_E_HIP -2      /* HIP runtime error */
#define PCP_E_NOMEM -3    /* capacity exceeded or device allocation failed */
#define PCP_E_STATE -4    /* a kernel reported a FATAL condition (NaN, beta > 1, ...) */

/* pushers: configure.ac:96-98,226-239 (--with-pusher) */
enum { PCP_PUSHER_BORIS = 0, PCP_PUSHER_BORIS_RELATIVISTIC = 1, PCP_PUSHER_BORIS_ADAPTIVE = 2,
       PCP_PUSHER_VAY = 3, PCP_PUSHER_HC = 4 };
/* charge conservation schemes: configure.ac:108-110,268-275 (--with-ccs) */
enum { PCP_CCS_VB = 0, PCP_CCS_ZIGZAG = 1 };
/* grids: MaxwellSolver::field_e/field_h (include/maxwellSolver.hpp:39-46),
 * MaxwellSolverYee::field_h_at_et, Current::current, epsilon, sigma */
enum { PCP_GRID_E = 0, PCP_GRID_H = 1, PCP_GRID_H_AT_ET = 2, PCP_GRID_J = 3,
       PCP_GRID_EPS = 4, PCP_GRID_SIGMA = 5,
       PCP_GRID_MAP = 6 /* the last map of pcp_species_map (probes only) */ };
/* species maps of the probes (SpecieP::calc_*, src/specieP.cpp:665-839) */
enum { PCP_MAP_DENSITY = 0, PCP_MAP_TEMPERATURE = 1, PCP_MAP_ENERGY = 2, PCP_MAP_MOMENTUM = 3 };
/* binary collision schemes: configure.ac --with-coulomb-collisions
 * (CollisionsP12 is the reference default) */
enum { PCP_COLL_P12 = 0, PCP_COLL_TA77S = 1, PCP_COLL_SK98 = 2,
       /* no reference counterpart: the fitted per-cell network (below) on
        * the cells it covers, P12 on the others */
       PCP_COLL_SURROGATE = 3 };
/* precision of the particle/field storage, fixed at creation.  FP32_SPLIT
 * stores each position as a packed cell index plus FP32 in-cell offsets and
 * keeps fields, velocities and J in FP32; such contexts run pcp_step and the
 * field entry points (the push/move/migrate/deposit phase entry points are
 * FP64 only). */
enum { PCP_PREC_FP64 = 0, PCP_PREC_FP32_SPLIT = 1 };

typedef struct {
  double charge;        /* SI [C]; SpecieP::charge = Z * EL_CHARGE (src/specieP.cpp:40) */
  double mass;          /* SI [kg]; SpecieP::mass = A * EL_MASS (src/specieP.cpp:41) */
  int is_beam;          /* reflect like BeamP::reflect (src/beamP.cpp:159-177) */
  long capacity;        /* maximum number of macro-particles held */
} pcp_species_desc;

/* The JSON geometry/time blocks (src/cfg.cpp:379-439) plus the compile-time
 * switches of configure.ac as run-time fields. */
typedef struct {
  double size_r, size_z;         /* geometry.size.radius / longitude [m] */
  long grid_r, grid_z;           /* geometry.grid.radius / longitude [cells] */
  long domains_r, domains_z;     /* geometry.domains_amount; must divide the grid */
  double dt;                     /* time.step [s] */
  int pusher;                    /* PCP_PUSHER_* */
  int ccs;                       /* PCP_CCS_* */
  int pml_on;                    /* ENABLE_PML */
  double pml_left, pml_right, pml_outer;   /* geometry.pml.*_wall (fractions) */
  double pml_sigma1, pml_sigma2;           /* geometry.pml.sigma_1/2 */
  int precision;                 /* PCP_PREC_* */
  int n_species;
  pcp_species_desc species[8];
  long max_migrants;             /* per-step capacity of the migration list (0: default) */
} pcp_config;

typedef struct pcp_ctx pcp_ctx;

const char *pcp_last_error(void);

/* SMB::SMB + Domain::Domain + solver constructors (src/SMB.cpp:28-192,
 * src/domain.cpp:26-84, src/maxwellSolver/maxwellSolverYee.cpp:22-109,
 * include/current.hpp:45-53). Fields start at zero, epsilon = EPSILON0,
 * sigma from set_pml. */
int pcp_create(const pcp_config *cfg, pcp_ctx **out);
int pcp_destroy(pcp_ctx *ctx);
int pcp_sync(pcp_ctx *ctx);

/* Domain::distribute + bind (src/domain.cpp:86-94): append particles to
 * species `sp`.  Each particle joins the domain owning its cell; within a
 * domain the input order is the reference's vector order.  `id` (optional)
 * is carried through unchanged. */
int pcp_upload_species(pcp_ctx *ctx, int sp, long n, const double *r, const double *z,
                       const double *vr, const double *vphi, const double *vz,
                       const double *w, const uint64_t *id);
/* number of live particles of species sp */
int pcp_count(pcp_ctx *ctx, int sp, long *n);
/* live particles, in device order; any pointer may be NULL. `order` receives
 * (domain << 40 | vector position key) -- sorting by it gives the reference's
 * domain-major vector order.  r_old/z_old are the positions before the last
 * move. */
int pcp_download_species(pcp_ctx *ctx, int sp, double *r, double *z, double *vr, double *vphi,
                         double *vz, double *w, uint64_t *id, uint64_t *order, double *r_old,
                         double *z_old);
/* Synthetic seeded inputs generated on the device (counter-based RNG), in
 * cell order: n_per_cell particles per cell with r^2 and z uniform inside the
 * cell, weight density * cell volume / n_per_cell, Maxwellian velocities of
 * temperature_ev.  density_right != density_left gives the linear z ramp of
 * SMB::SMB (src/SMB.cpp:134-136), rounding the per-cell count. */
int pcp_init_plasma(pcp_ctx *ctx, int sp, int n_per_cell, double density_left,
                    double density_right, double temperature_ev, uint64_t seed);
/* Gaussian bunch (sigma_r, sigma_z) centred at z0 on the axis, moving along z
 * with Lorentz factor gamma, peak density n_peak, n macro-particles. */
int pcp_init_bunch(pcp_ctx *ctx, int sp, long n, double n_peak, double sigma_r, double sigma_z,
                   double z0, double gamma, uint64_t seed);
/* The reference's own initial plasma: SMB::SMB + Domain::distribute for one
 * plasma species (src/SMB.cpp:123-137, src/specieP.cpp:79-189 "flat"
 * placement, 280-331 set_weight, 346-384 thermal velocities,
 * src/math/maxwellJuettner.cpp:182-255).  Every domain places
 * macro_amount / Dr / Dz particles at radical-inverse positions (bases 13 and
 * 11, as math::random::random_reverse), uniform in r with weights
 * proportional to r, and gives them Maxwell-Juettner thermal velocities of
 * temperature_ev (the reference's units); draws are Philox with `seed` in
 * place of the reference's time-seeded mt19937.  Particles a domain places
 * past an inner domain edge join the owning domain as the first runaway
 * collector would move them.  Replaces SpecieP::fullyfill_spatial_distribution
 * + velocity_distribution (include/specieP.hpp:225-229). */
int pcp_init_reference(pcp_ctx *ctx, int sp, long macro_amount, double density_left,
                       double density_right, double temperature_ev, uint64_t seed);
/* One probe frame (OutWriter::operator(), src/outWriter.cpp:49-100, as
 * OutController maps a probe onto the domains, src/outController.cpp:49-309):
 * component comp (0 r, 1 phi, 2 z) of grid `which` over the global cells
 * [r0, r1) x [z0, z1), each value read from the domain that owns the cell,
 * row-major (r outer) into host_out ((r1 - r0) x (z1 - z0) doubles).  The
 * reference's "rec" probe is this window; "col" z is [0, Nr) x [z, z + 1),
 * "row" r is [r, r + 1) x [0, Nz), "dot" is one cell.  Probes read
 * E (PCP_GRID_E), H (PCP_GRID_H_AT_ET, as the reference's writers do) and
 * J (PCP_GRID_J).  The window is gathered on the device. */
int pcp_probe(pcp_ctx *ctx, int which, int comp, long r0, long z0, long r1, long z1, double *host_out);

/* A species map as the reference computes it before a probe dump:
 * weight_cylindrical of every particle into its domain (include/algo/
 * weighter.hpp:27-132), SpecieP::calc_density / calc_temperature /
 * calc_energy / calc_momentum per domain (src/specieP.cpp:665-839) and
 * OutController's overlay (src/outController.cpp:409-491), which sums the
 * r- and z-neighbours' halo deposits into the density and leaves the other
 * maps as each domain computed them.  Density in m^-3, temperature and
 * energy in eV, momentum (3 components) in kg m/s per particle.  As in the
 * reference, the temperature map uses the species' momentum_abs map as the
 * last energy map left it (zero before any).  host_out ([D][ncomp][Nr_d+4]
 * [Nz_d+4], interior nodes defined) may be NULL: the map stays on the device
 * for pcp_probe(ctx, PCP_GRID_MAP, comp, ...).  Replaces the calc_* calls of
 * OutController::operator() (src/outController.cpp:380-408). */
int pcp_species_map(pcp_ctx *ctx, int sp, int kind, double *host_out);

/* A particle beam of bunches injected over time: BeamP::BeamP for beam species
 * sp in every domain (src/beamP.cpp:24-66).  From then on pcp_step injects at
 * the start of every step as SMB::inject_beam does (src/SMB.cpp:1308-1322,
 * BeamP::inject src/beamP.cpp:71-157): while the context time lies in a
 * bunch's window [start, start + length / velocity), each domain at z = 0
 * whose radial range starts below `radius` appends ceil(macro_per_step)
 * particles (macro_per_step = macro_amount / bunches_amount (integer
 * division) x its share of the radius x velocity dt / length) at r uniform
 * over its share, z uniform over [dz/2, dz/2 + velocity dt), v = (0, 0,
 * velocity), weights proportional to r.  The reference's uniform() draws
 * are Philox draw 0 of counter (domain << 40 | injected count), key seed.
 * Injected particle ids: 1 << 62 | domain << 40 | count.  The species must
 * have been created with is_beam = 1 and enough capacity.  Replaces
 * BeamP's constructor and SpecieP::inject (include/beamP.hpp:43-62). */
int pcp_set_beam(pcp_ctx *ctx, int sp, long macro_amount, double start_time, double radius,
                 double density, int bunches_amount, double bunch_length, double bunches_distance,
                 double velocity, uint64_t seed);
/* SMB::inject_beam at the current context time (for callers that drive the
 * phases themselves; pcp_step injects by itself). */
int pcp_inject(pcp_ctx *ctx);
/* TimeSim::current of the context: 0 at creation, + dt after every step of
 * pcp_step (src/PiCoPiC.cpp:343 advances it after each iteration). */
int pcp_set_time(pcp_ctx *ctx, double t);
double pcp_time(pcp_ctx *ctx);

/* Whole halo grids, domain-major [domain][component][Nr_d+4][Nz_d+4]
 * (one component for EPS and SIGMA), as Grid<double> stores them
 * (include/algo/grid.hpp:24-58). */
int pcp_download_grid(pcp_ctx *ctx, int which, double *host);
int pcp_upload_grid(pcp_ctx *ctx, int which, const double *host);
long pcp_grid_points(pcp_ctx *ctx);

/* MaxwellSolverYee::calc_field_e / calc_field_h for every domain
 * (src/maxwellSolver/maxwellSolverYee.cpp:111-257) */
int pcp_yee_e(pcp_ctx *ctx);
int pcp_yee_h(pcp_ctx *ctx);
/* SMB::field_e_overlay / field_h_overlay / current_overlay
 * (src/SMB.cpp:566-596, 784-823, 1010-1044); which = E, H (stitches H and
 * H_at_et) or J.  After the J stitch only interior J values are defined. */
int pcp_stitch(pcp_ctx *ctx, int which);
/* Pusher::operator() for every domain (include/pusher.hpp:58) */
int pcp_push(pcp_ctx *ctx);
/* SMB::advance_particles after the push: dump_position_to_old, mover,
 * back_position_to_rz, reflect, back_velocity_to_rz, bind_cell_numbers
 * (src/SMB.cpp:1290-1303, src/specieP.cpp:482-663) */
int pcp_move(pcp_ctx *ctx);
/* SMB::particles_runaway_collector (src/SMB.cpp:194-344): removal and the
 * reference's vector order of migrating particles */
int pcp_migrate(pcp_ctx *ctx);
/* Domain::reset_current + Current::current_distribution
 * (src/domain.cpp:135-139, include/current.hpp:53) */
int pcp_deposit(pcp_ctx *ctx);
/* n whole steps: solve_maxvell, advance_particles, solve_current
 * (src/PiCoPiC.cpp:299-345) with push + move + deposit fused */
int pcp_step(pcp_ctx *ctx, int n_steps);

/* Collisions::operator() (include/collisions.hpp:49; P12
 * src/collisions/collisionsP12.cpp:571-579, TA77S
 * src/collisions/collisionsTA77S.cpp:568-576, SK98
 * src/collisions/collisionsSK98.cpp:441-450) for every domain: one pass
 * between electron species sp_el and ion species sp_ion on the current
 * velocities.  random_sort is an ordering of each cell's particles by a
 * Philox key of (particle id, pass) -- the reference shuffles with a
 * time-seeded std::random_shuffle -- and pair q of global cell c takes the
 * uniforms Philox(seed; c << 24 | q, 2 pass + 1).  Cells of any size: those
 * over 256 macro-particles of a species are staged in global memory, at most
 * 2^22 macro-particles in all such cells of one pass (PCP_E_STATE beyond). */
int pcp_collide(pcp_ctx *ctx, int scheme, int sp_el, int sp_ion, uint64_t seed, uint64_t pass);
/* Enable (scheme >= 0) or disable (scheme < 0) the collision phase of
 * pcp_step, run between the push and the move like Domain::collide in
 * SMB::advance_particles (src/SMB.cpp:1287-1289); passes count from 0.
 * PCP_COLL_SURROGATE in an FP32_SPLIT context runs fused into the step: the
 * push kernel sums each cell's velocity moments, the network evaluates every
 * cell, and the move kernel applies each cell's conserving rescale to the
 * colliding species as it reads the velocities; cells the network does not
 * cover take the P12 pass as in pcp_collide. */
int pcp_set_collisions(pcp_ctx *ctx, int scheme, int sp_el, int sp_ion, uint64_t seed);

/* ---- collision surrogate: no counterpart in the reference ----
 * A per-cell network (6 inputs: log10 of n_e, n_i, Te_par, Te_perp, Ti and
 * dt; two tanh layers of 64; 4 outputs: relative changes of Te_par,
 * Te_perp, Ti_par, Ti_perp per pass) evaluated and trained with BF16 MFMA
 * kernels.  Its output rescales each species' thermal velocities about the
 * species' mean velocity (z: par, r and phi: perp); a common factor then
 * restores the cell's kinetic energy, so every species' momentum and the
 * cell's energy are conserved.  Cells with fewer than 8 macro-particles of a
 * species, degenerate moments or features beyond 4 standard deviations of
 * the training set take the P12 pass instead (PCP_COLL_SURROGATE).
 *
 * Raw network access (rows of 8 floats in, first 6 used; rows of 4 out):
 * the parameter block is W1[64][32] b1[64] W2[64][64] b2[64] W3[16][64]
 * b3[16] (padding zero). */
int pcp_mlp_nparams(void);
int pcp_mlp_init(pcp_ctx *ctx, uint64_t seed);
int pcp_mlp_get_params(pcp_ctx *ctx, float *host);
int pcp_mlp_set_params(pcp_ctx *ctx, const float *host);
int pcp_mlp_forward(pcp_ctx *ctx, const float *x, long n, float *y);
/* gradient of sum (y - t)^2 / n over the batch; *loss = that sum */
int pcp_mlp_grad(pcp_ctx *ctx, const float *x, const float *t, long n, float *grad, float *loss);
/* one Adam step (beta 0.9 / 0.999) with the last gradient */
int pcp_mlp_adam(pcp_ctx *ctx, float lr);
/* one P12 pass (as pcp_collide) whose collided cells append a training row
 * (raw features, relative temperature changes) to the context's training
 * set; *rows = rows held */
int pcp_surrogate_collect(pcp_ctx *ctx, int sp_el, int sp_ion, uint64_t seed, uint64_t pass, long *rows);
/* normalise the training set and train: `epochs` shuffled passes of
 * mini-batches with Adam; *loss = mean batch loss of the last pass */
int pcp_surrogate_fit(pcp_ctx *ctx, int epochs, long batch, float lr, uint64_t seed, float *loss);
/* normalisation of a fitted network: mean[6], std[6], target scale[4] */
int pcp_surrogate_get_norm(pcp_ctx *ctx, float *out16);
int pcp_surrogate_set_norm(pcp_ctx *ctx, const float *in16);
/* cells of the last surrogate pass that took the P12 fallback */
int pcp_surrogate_fallback_cells(pcp_ctx *ctx, long *cells);
/* A posteriori calibration of a fitted network (no reference counterpart):
 * the relaxation rate of the electron anisotropy over `passes` passes is
 * measured for P12 and for the surrogate on scratch contexts of this
 * context's configuration holding copies of its two species (as much
 * device memory again), and the electron target scales are multiplied by
 * the factor (*factor) that makes the surrogate's rate P12's (two
 * fixed-point iterations).  The context's own particles are not changed.
 * The plasma must keep T_par > T_perp for the electrons over the passes. */
int pcp_surrogate_calibrate(pcp_ctx *ctx, int sp_el, int sp_ion, uint64_t seed, int passes, double *factor);

/* Replay each step from a captured hipGraph (no reference counterpart):
 * one graph launch per step holds the field update, the push, the
 * collision phase (its pass number lives on the device and advances inside
 * the graph), the move and deposit, the migration keys, the fix-up and the
 * current overlay.  The beam injection and the sorts run before it as
 * ordinary launches.  Graphs are cached by buffer state (the particle
 * buffers, counts and buffer parity), so the sorts' buffer exchanges cycle
 * through a few captured graphs; uploads, initialisation, tuning and
 * collision or network changes drop them.  With pcp_set_timing, a step's
 * time is its sorts plus its graph launch, and the phases inside the graph
 * read as zero. */
int pcp_set_graph(pcp_ctx *ctx, int on);

/* Compensated current deposition of an FP32 split context (no reference
 * counterpart; the reference sums J in FP64, currentVB.cpp:731-735,
 * currentZigZag.cpp:22-136).  On: each wavefront's run sums carry a
 * Kahan-style compensation term, and each FP32 atomic on a J node returns
 * the node's previous value, from which the exact rounding error of that
 * atomic is recovered and added, with the run's compensation, to a
 * companion array of J's layout; the step folds the companion into J before
 * the current overlay.  The deposited J then carries one FP32 rounding per
 * node instead of one per atomic and scan step.  Costs a second array of
 * J's size and about twice the deposit's run-sum arithmetic.  Returns
 * PCP_E_ARG for FP64 contexts. */
int pcp_set_compensated_current(pcp_ctx *ctx, int on);
/* graph launches and captures since the context was created (no reference
 * counterpart) */
int pcp_graph_stats(pcp_ctx *ctx, long *launches, long *captures);
/* collision passes the step has run since pcp_set_collisions: the device
 * counter that numbers the passes' draws (no reference counterpart) */
int pcp_collision_passes(pcp_ctx *ctx, uint64_t *passes);

/* Counting sort of species sp by (domain, sort bin, cell), dropping removed particles
 * (no reference counterpart: the reference keeps per-domain vectors).  The
 * reference order keys travel with the particles, so results do not depend
 * on the storage order.  pcp_step sorts species sp before every `every`-th
 * step (0: never; defaults: beams 1, plasma 50). */
int pcp_sort(pcp_ctx *ctx, int sp);
int pcp_set_sort_interval(pcp_ctx *ctx, int sp, int every);

/* Launch tuning (no reference counterpart: the reference's OpenMP loops
 * have no launch shape).  The sort bin orders the cells of a domain in
 * bin_r x bin_z tiles (each must divide the domain's cell count in r and
 * z; default 1 x nz, plain row-major), so that the particles of one
 * workgroup trip cover a compact patch of field and current nodes.  The
 * launch shape is the workgroup size of the FP32 particle kernel (128, 256
 * or 512) and its resident workgroups per CU (0: as many as fit).
 * pcp_autotune times the candidates on the live state (the simulation
 * advances by the steps it takes: about 80 x steps + 24), keeps the fastest
 * bin and shape, and sets the plasma species' sort interval from the
 * measured sort cost and the growth of the step time after a sort. */
typedef struct pcp_tuning {
  int bin_r, bin_z;
  int block, blocks_per_cu;
  int sort_every;        /* largest sort interval of the non-beam species */
  double ms_per_step;    /* best trial step time (pcp_autotune) */
} pcp_tuning;
int pcp_set_sort_bin(pcp_ctx *ctx, int bin_r, int bin_z);
int pcp_set_launch(pcp_ctx *ctx, int block, int blocks_per_cu);
int pcp_get_tuning(pcp_ctx *ctx, pcp_tuning *out);
int pcp_autotune(pcp_ctx *ctx, int steps, pcp_tuning *out);

/* field gather at one point of domain d (include/maxwellSolver.hpp:101-316),
 * evaluated on the device; out[3] = (r, phi, z) components */
int pcp_get_field_e(pcp_ctx *ctx, int d, double r, double z, double *out);
int pcp_get_field_h(pcp_ctx *ctx, int d, double r, double z, double *out);

/* FP64 reductions over the device state (diagnostics; the reference computes
 * such sums on the host from its maps, src/specieP.cpp:665-839):
 * field energy  sum over interior nodes of (EPSILON0 E^2 + MAGN_CONST H^2)/2
 *               times the ring volume PI dz dr^2 (2 i + 1)   [J]
 * kinetic energy of species sp: sum of w m c^2 (gamma - 1)    [J] */
int pcp_field_energy(pcp_ctx *ctx, double *out);
int pcp_kinetic_energy(pcp_ctx *ctx, int sp, double *out);
/* velocity moments of species sp (the sums behind SpecieP::calc_temperature,
 * src/specieP.cpp:725-839, taken over the whole block): out[7] = sum w,
 * sum w v_r, sum w v_phi, sum w v_z, sum w v_z^2, sum w (v_r^2 + v_phi^2),
 * number of live macro-particles */
int pcp_species_moments(pcp_ctx *ctx, int sp, double *out);
/* Discrete continuity residual of the last step (no reference counterpart;
 * it checks currentVB.cpp's charge conservation on the device): for every
 * node (i dr, k dz), 0 <= i <= Nr-2, 1 <= k <= Nz-2, the change of the
 * charge the particles' ring clouds ([r -+ dr/2] x [z -+ dz/2]) put in the
 * node's control volume plus dt times the current leaving it through J_r
 * (cylinder r = (i+1/2) dr) and J_z (annulus at z = (k+1/2) dz, pi dr^2/4 at
 * the axis).  out[3] = max |residual|, max |charge change|, largest |q w|. */
int pcp_continuity_residual(pcp_ctx *ctx, double *out);
/* pcp_continuity_residual without allocating: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-node charge array lives in a fixed device array of 10 x 2^20 doubles
 * (80 MiB, one per GPU, held until the process ends), so the grid's
 * (Nr + 4) x (Nz + 4) nodes may not exceed 10,485,760 (PCP_E_ARG otherwise).
 * Every call on one GPU shares the array: calls that overlap on one GPU
 * race.  Same results as pcp_continuity_residual. */
int pcp_continuity_residual_noalloc(pcp_ctx *ctx, double *out);

/* In-step FP64 energy log (no reference counterpart; the reference sums
 * energies on the host from its output maps, src/specieP.cpp:665-839).
 * While on, every pcp_step step records, on the device, one row of
 * 2 + 8 doubles: the step number, the field energy (as pcp_field_energy,
 * after the step's E and H updates) and the kinetic energy of each species
 * (as pcp_kinetic_energy, after the move).  FP32 contexts sum the kinetic
 * energies inside the fused particle kernel.  The log keeps the last 4096
 * steps; pcp_energy_log copies the most recent min(max_rows, logged) rows,
 * oldest first, into out (rows x 10 doubles) and sets *rows. */
int pcp_set_energy_log(pcp_ctx *ctx, int on);
int pcp_energy_log(pcp_ctx *ctx, long max_rows, double *out, long *rows);

/* achievable HBM bandwidth of a plain device copy of `bytes` bytes (read +
 * write counted), averaged over iters launches [GB/s]; a calibration aid for
 * the roofline figures, no reference counterpart */
int pcp_copy_bandwidth(pcp_ctx *ctx, long bytes, int iters, double *gbps);

/* hipStream_t the context enqueues on (as void*) */
void *pcp_stream(pcp_ctx *ctx);
/* average device time [ms] per step of the named phase over the last
 * pcp_step call when timing is enabled with pcp_set_timing(ctx, 1): "sort",
 * "fields", "push" (the push of a step with collisions), "collide",
 * "particles" (the fused move + deposit launch), "fixup", "stitch_j" and
 * "step" (their sum; with graph mode only "step" is measured); -1 if
 * unknown */
int pcp_set_timing(pcp_ctx *ctx, int on);
double pcp_kernel_ms(pcp_ctx *ctx, const char *name);

#ifdef __cplusplus
}
#endif
#endif
