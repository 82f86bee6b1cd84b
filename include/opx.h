/*
 * opx.h -- C ABI of the MI355X (gfx950) real-time lattice Yang-Mills leapfrog
 * and its colored-particle (CPIC) coupling.
 *
 * The ABI replaces OpenPixi's FieldSolver / Interpolation / FieldMeasurements
 * seams for the FastTYMSolver path. Every entry point is a plain C function on
 * pointers and sizes; nothing throws across it. Reference citations are to
 * pixi/src/main/java/org/openpixi/pixi/ in the reference repository.
 *
 * Conventions
 *   - Return value: OPX_OK (0) or a negative OPX_E_* code; opx_last_error()
 *     returns the message of the last failure on a context.
 *   - Host buffers use the reference's per-cell element order (AoS, cell-major,
 *     direction next, then element components):
 *       links   [cell][dir][G]   G = 4 (SU(2): e0..e3) or 18 (SU(3): 9 real, 9 imaginary parts, row-major)
 *       algebra [cell][dir][A]   A = 3 (SU(2): v0..v2) or 9 (SU(3): Hermitian storage
 *                                 of math/SU3AlgebraElement.java:3-11)
 *       charge  [cell][A]
 *     cell = ((x0*N1)+x1)*N2+x2 (physics/grid/Grid.java:757-770).
 *     The device keeps each field in blocks of 64 consecutive cells, a block
 *     holding every component (dir-major, then element) of its cells, one
 *     component of the 64 cells after another; FP64 SU(3) fields hold the
 *     components in pairs (components 2k and 2k+1 of a cell side by side).
 *     Fields are in the context's precision; SU(3) algebra is kept as 8
 *     Gell-Mann coefficients.
 *     OPX_FP32 SU(3) contexts store the first two rows of each link (12 reals)
 *     and rebuild the third as conj(row0 x row1), on the device and on download.
 *   - One host thread per context; calls are ordered on the context's HIP
 *     stream. Entry points that return host data synchronize that stream.
 */
#ifndef OPX_H
#define OPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct opx_ctx opx_ctx;

enum {
    OPX_OK = 0,
    OPX_E_ARG = -1,          /* invalid argument (null pointer, size mismatch, bad index) */
    OPX_E_UNSUPPORTED = -2,  /* unsupported number of colours / dimensions / precision */
    OPX_E_DEVICE = -3,       /* HIP runtime error (allocation, launch, copy) */
    OPX_E_STATE = -4,        /* call not valid in the current state (e.g. no particles) */
    OPX_E_NUMERIC = -5       /* health monitor tripped (NaN / unitarity defect) */
};

enum { OPX_FP64 = 0, OPX_FP32 = 1 };

enum {
    OPX_FIELD_U = 0,      /* links at t - dt/2            (Cell.U,     physics/grid/Cell.java:26) */
    OPX_FIELD_UNEXT = 1,  /* links at t + dt/2            (Cell.Unext, physics/grid/Cell.java:29) */
    OPX_FIELD_E = 2,      /* electric field at t          (Cell.E,     physics/grid/Cell.java:20) */
    OPX_FIELD_J = 3,      /* current at t + dt/2          (Cell.J,     physics/grid/Cell.java:14) */
    OPX_FIELD_RHO = 4     /* charge density at t          (Cell.rho,   physics/grid/Cell.java:17) */
};

/* indices into opx_measure's output */
enum {
    OPX_M_E2_0 = 0, OPX_M_E2_1 = 1, OPX_M_E2_2 = 2,   /* FieldMeasurements.calculateEsquared(grid, i) */
    OPX_M_B2_0 = 3, OPX_M_B2_1 = 4, OPX_M_B2_2 = 5,   /* FieldMeasurements.calculateBsquared(grid, i) */
    OPX_M_GAUSS = 6,                                  /* FieldMeasurements.calculateGaussConstraint  */
    OPX_M_RHO2 = 7,                                   /* FieldMeasurements.calculateTotalChargeSquared */
    OPX_M_COUNT = 8
};

/* ---- context ------------------------------------------------------------
 * Replaces the state held by Grid(Settings) (physics/grid/Grid.java:443-473):
 * colours, cells per dimension, lattice spacings a_i, time step a_t and
 * coupling g. Links start as identities, fields at zero, all cells active
 * and evaluatable (physics/grid/Cell.java:49-74). num_dims must be 3.
 * precision: OPX_FP64, or OPX_FP32 (links, E, J, rho in float; measurements,
 * reunitarization in double). */
int opx_create(int num_colors, int num_dims, const int *num_cells, const double *spacing,
               double time_step, double coupling, int precision, opx_ctx **out);
void opx_destroy(opx_ctx *ctx);
const char *opx_last_error(const opx_ctx *ctx);
int opx_get_info(const opx_ctx *ctx, int64_t *num_cells_total, int *num_colors, int *precision);
/* Launch on a caller-owned hipStream_t (NULL restores the context's own stream). */
int opx_set_stream(opx_ctx *ctx, void *hip_stream);
int opx_sync(opx_ctx *ctx);

/* ---- data at the seam ---------------------------------------------------
 * Grid.setU/setUnext/setE/addJ/setRho and getters, whole lattice at once
 * (physics/grid/Grid.java:126-265). n_values must equal the full buffer size. */
int opx_upload_field(opx_ctx *ctx, int field, const double *host, int64_t n_values);
int opx_download_field(opx_ctx *ctx, int field, double *host, int64_t n_values);
/* Links U = exp(i sum_a c_a t_a) from Gell-Mann/Pauli coefficients [cell][dir][nc*nc-1]
 * (AlgebraElement.getLink, math/SU2AlgebraElement.java:140-152,
 * math/SU3AlgebraElement.java:407-417), computed on the device. */
int opx_set_links_from_coefficients(opx_ctx *ctx, int field, const double *coeffs, int64_t n_values);
/* E from generator coefficients [cell][dir][nc*nc-1] (AlgebraElement.set(j, v)). */
int opx_set_efield_from_coefficients(opx_ctx *ctx, const double *coeffs, int64_t n_values);

/* Grid.setActiveRegion / setEvaluationRegion (physics/grid/Grid.java:992-1020):
 * inclusive boxes in cell coordinates. */
int opx_set_active_region(opx_ctx *ctx, const int *lo, const int *hi);
int opx_set_evaluation_region(opx_ctx *ctx, const int *lo, const int *hi);

/* ---- FieldSolver seam (FastTYMSolver) -----------------------------------
 * opx_step:        FastTYMSolver.step -> CombinedUpdate for every active cell
 *                  (physics/fields/FastTYMSolver.java:25-32,40-62):
 *                  E_i += dt proj(U_i S_i) - g a_i a_t J_i;  Unext_i = exp(-i dt E_i) U_i
 * opx_step_links:  FastTYMSolver.stepLinks -> UpdateLinks (:35-38,65-84)
 * opx_store_fields: Grid.storeFields, U <-> Unext (physics/grid/Grid.java:591-593)
 * opx_reset_current / opx_reset_charge: Grid.resetCurrent / resetCharge (:570-586) */
int opx_step(opx_ctx *ctx, double dt);
int opx_step_links(opx_ctx *ctx, double dt);
/* Field solver used by opx_step, opx_leapfrog and opx_simulation_step:
 *   OPX_SOLVER_FAST      FastTYMSolver (default);
 *   OPX_SOLVER_IMPLICIT  ImplicitTYMSolver(iterations, damping)
 *                        (physics/fields/ImplicitTYMSolver.java:18-58), the solver of the
 *                        "temporal implicit cgc ngp" and "temporal optimized implicit cgc ngp"
 *                        simulation types (physics/Settings.java:560-577): the explicit update,
 *                        then `iterations` passes that rebuild E from plaquettes averaged over
 *                        the past and future links (implicit in the transverse plane and along
 *                        the beam direction x0), each followed by the damping
 *                        Unext = U^damping Unext^(1 - damping) (GroupElement.pow).
 *                        With damping 0 or 1 the powers are not evaluated: x^0 is the
 *                        identity and x^1 is x up to the logarithm's rounding (~1e-16). */
enum { OPX_SOLVER_FAST = 0, OPX_SOLVER_IMPLICIT = 1 };
int opx_set_field_solver(opx_ctx *ctx, int solver, int iterations, double damping);
/* GroupElement.pow(x) = getLink(x getAlgebraElement()) (math/SU2GroupElement.java:234-271,
 * math/SU3GroupElement.java:239-520) on the device for `count` elements in reference storage
 * (4 or 18 reals each), rounded to the given precision and raised in double precision, as the
 * damping of the implicit solver evaluates it in every precision. */
int opx_group_pow(int num_colors, int precision, int count, const double *elements, double x, double *out);
int opx_store_fields(opx_ctx *ctx);
int opx_reset_current(opx_ctx *ctx);
int opx_reset_charge(opx_ctx *ctx);

/* Kernel form of the fused update (no counterpart in the reference):
 *   OPX_KERNEL_LINK   one thread per link, neighbour links from global memory
 *                     (any lattice, any active region);
 *   OPX_KERNEL_STREAM tiles of (x1,x2) columns marching along x0, each plaquette
 *                     computed once and exchanged through LDS;
 *   OPX_KERNEL_PLAQ   two kernels: proj of the three plaquettes of every cell into a
 *                     scratch buffer (3 algebra elements per cell), then one thread per
 *                     link builds the force from four of them and two adjoint actions;
 *   OPX_KERNEL_TILE   SU(2) only: tiles of (x1,x2) cells marching along x0 with the
 *                     links of three planes staged in LDS; OPX_OPT_CHUNKS sets the
 *                     number of x0 chunks (0 = derived from the CU count) and
 *                     OPX_OPT_TILE the store policy of E and Unext (0 non-temporal,
 *                     1 cached);
 *   OPX_KERNEL_PLANE  a block updates the three links of every cell of a tile of one
 *                     x0 plane, the plane's links (tile plus a one-cell halo) staged in
 *                     LDS and the x0 +- 1 neighbours read from global memory;
 *                     OPX_OPT_TILE the tile shape (0: 8x16, 1: 8x8, 2: 4x16 cells
 *                     over (x1,x2)), one thread per link;
 *   OPX_KERNEL_CELL   one thread per cell updates its three links, the two upper staples of
 *                     a plane sharing one product (33 link reads and 21 staple products per
 *                     cell instead of 39 and 24); OPX_OPT_TILE the row tile the blocks sweep
 *                     along x0 (0: 32, 1: 16, 2: 8 rows of x1)
 *   OPX_KERNEL_AUTO   LINK until opx_autotune picks a form (default).
 * OPX_OPT_TILE selects the STREAM tile shape (0: 16x16, 1: 8x32, 2: 16x32,
 * 3: 8x64, 4: 8x16 threads over (x1,x2)); OPX_OPT_CHUNKS the number of x0
 * chunks (0 = derived from the CU count). opx_autotune times every candidate
 * on scratch copies of E and Unext (the lattice state is not changed) over reps
 * launches, re-times the three fastest in four interleaved rounds of max(reps, 8)
 * launches with no host synchronization between them, keeps the one with the
 * fastest round and returns its time per launch in *best_ms. */
/* OPX_OPT_GLOBAL_POINTERS 1: the per-link, cell and plaquette forms read their fields through
 * plain global pointers, as they do by themselves for fields of 4 GiB or more (0, the default:
 * buffer instructions below 4 GiB); for testing that path on small lattices. */
/* OPX_OPT_REUNIT_EVERY N > 0: every N-th Simulation step of opx_simulation_step (counted by
 * opx_simulation_steps_done) reprojects the new links (Unext) onto the group in double precision,
 * as opx_leapfrog's reunit_every does: inside the update where the kernel form can, else by a pass
 * after it. 0 (the default): never. No counterpart in the reference, whose double-precision links
 * are not reprojected; for the FP32 form. */
enum { OPX_OPT_KERNEL = 0, OPX_OPT_TILE = 1, OPX_OPT_CHUNKS = 2, OPX_OPT_GLOBAL_POINTERS = 3,
       OPX_OPT_REUNIT_EVERY = 4 };
enum { OPX_KERNEL_AUTO = 0, OPX_KERNEL_LINK = 1, OPX_KERNEL_STREAM = 2, OPX_KERNEL_PLAQ = 3, OPX_KERNEL_TILE = 4,
       OPX_KERNEL_PLANE = 5, OPX_KERNEL_CELL = 6 };
int opx_set_option(opx_ctx *ctx, int option, int value);
int opx_get_option(const opx_ctx *ctx, int option, int *value);
int opx_autotune(opx_ctx *ctx, int reps, double *best_ms);

/* Reprojects every link of U and Unext onto the group in double precision
 * (SU(2): quaternion normalisation; SU(3): Gram-Schmidt on rows 1-2, row 3 =
 * conj(row1 x row2)). No counterpart in the reference (which never
 * reunitarizes); used by the FP32 form every few steps. */
int opx_reunitarize(opx_ctx *ctx);

/* Projects E onto Gauss's law with respect to the links U (and rho):
 * E' = E - D lambda with D^T D lambda = G, where G is the constraint of
 * Grid.getGaussConstraint (physics/grid/Grid.java:944-968) and D its covariant
 * difference; solved by conjugate gradients in double precision until
 * |r| <= tol |r0| or max_iter iterations. Returns the iteration count and the
 * mean |G|^2 before and after (of the stored field, in its precision). The
 * reference's iterative GaussConstraintRestoration
 * (diagnostics/methods/GaussConstraintRestoration.java:79-163) relaxes towards
 * the same constraint surface; this solves for the nearest point on it. */
int opx_gauss_project(opx_ctx *ctx, double tol, int max_iter, int *iterations, double *gauss_before,
                      double *gauss_after);

/* `steps` leapfrog steps of pure Yang-Mills: for each, storeFields then step
 * (physics/Simulation.java:389-395 with no particles). Every `reunit_every`-th
 * step (counted over the context's life; 0 = never) reprojects the new links
 * (Unext) onto the group in double precision, as opx_reunitarize does: inside
 * the update kernel in the cell form, as a separate pass in the others. */
int opx_leapfrog(opx_ctx *ctx, int steps, int reunit_every);

/* ---- observables seam (FieldMeasurements, double precision) -------------
 * out[OPX_M_COUNT]: mean E_i^2/(g a_i)^2, mean 1/2(B_i^2(U)+B_i^2(Unext))/(g a_i)^2,
 * mean |G|^2, sqrt(sum rho^2)/N over evaluatable cells
 * (physics/measurements/FieldMeasurements.java:68-230). Sums are compensated
 * (Kahan/two-sum) and reduced in a fixed order. */
int opx_measure(opx_ctx *ctx, double *out);
/* ProjectedEnergyDensity along `direction` (diagnostics/methods/ProjectedEnergyDensity.java:130-170):
 * four arrays of num_cells[direction] doubles: transverse electric/magnetic,
 * longitudinal electric/magnetic energy per slice. */
int opx_projected_energy_density(opx_ctx *ctx, int direction, double *eT_el, double *eT_mag,
                                 double *eL_el, double *eL_mag);
/* opx_projected_energy_density without any device allocation. No counterpart in
 * the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * per-slice sums go to a static device buffer of 4 x 4096 doubles (128 KiB, one
 * per GPU, held until the process ends) instead of the context's own buffer, so
 * `direction` may have at most 4096 slices (more: OPX_E_ARG). Results are those
 * of opx_projected_energy_density bit for bit. Every call on a GPU shares that
 * GPU's buffer: calls that overlap on one GPU race. The context's GPU must be
 * current on the first call. */
int opx_projected_energy_density_noalloc(opx_ctx *ctx, int direction, double *eT_el, double *eT_mag,
                                         double *eL_el, double *eL_mag);
/* Longitudinal Poynting vector along `direction` (PoyntingComputation,
 * diagnostics/methods/ProjectedEnergyDensity.java:178-262): two arrays of
 * num_cells[direction] doubles, spatially and time averaged and only time averaged,
 * summed over every cell of each slice. SU(2) only: returns OPX_E_UNSUPPORTED for
 * SU(3), where the reference's SU3AlgebraElement.mult(AlgebraElement) throws. */
int opx_projected_poynting(opx_ctx *ctx, int direction, double *averaged, double *time_averaged);

/* ---- particle seam: CGC super particles (light-cone NGP) ----------------
 * Replaces CGCSuperParticleInterpolationNGP (physics/grid/CGCSuperParticleInterpolationNGP.java:19-93)
 * over a list of CGCSuperParticle (physics/particles/CGCSuperParticle.java:16-46).
 * Arrays have `count` entries; charges holds sum(num) algebra elements in
 * reference storage, particle after particle. The list order is kept. */
int opx_superparticles_upload(opx_ctx *ctx, int count, const int *orientation, const int *num,
                              const int *index_offset, const int *per_plane, const int *sub_lattice_shift,
                              const int *particles_per_cell, const double *charges, int64_t n_values);
int opx_superparticles_download_charges(opx_ctx *ctx, double *charges, int64_t n_values);
/* interpolateToGrid at integer clock t: J_0 += +-Q a/dt and parallel transport by Unext_0 */
int opx_interpolate_to_grid(opx_ctx *ctx, int t);
/* interpolateChargedensity at integer clock t: rho += Q at the NGP offset */
int opx_interpolate_chargedensity(opx_ctx *ctx, int t);
/* the list's size: super particles and charges in total (CGC super particles of the
 * simulation's particle list, Simulation.particles) */
int opx_superparticles_count(const opx_ctx *ctx, int *count, int64_t *total_charges);
/* the list's descriptors, `count` entries each (the constructor arguments of CGCSuperParticle,
 * physics/particles/CGCSuperParticle.java:33-46) */
int opx_superparticles_download_descriptors(opx_ctx *ctx, int count, int *orientation, int *num,
                                            int *index_offset, int *per_plane, int *sub_lattice_shift,
                                            int *particles_per_cell);

/* ---- CGC initial condition (beam along x0) ------------------------------
 * Replaces CGCInitialCondition.applyInitialCondition (physics/initial/CGC/CGCInitialCondition.java:55-93)
 * for direction 0 with an MVModel charge density, LightConePoissonSolverImprovedFull and
 * LightConeNGPSuperParticleCreator. One call of each, in this order, per nucleus. The charge
 * density and the Gauss violation of the last solve stay on the device between the calls
 * (double precision); every axis may hold at most 4096 cells. Entry points that take host
 * buffers synchronize the context's stream. */
/* MVModel.initialize (physics/initial/CGC/MVModel.java:101-157): noise holds the unit Gaussian
 * draws, (colours^2 - 1) * cells values, component after component and cells in index order
 * within a component (the order the reference takes them from java.util.Random.nextGaussian).
 * uv_cutoff = ultravioletCutoffTransverse, coherence_length = longitudinalCoherenceLength,
 * ir_coefficient = infraredCoefficient. rho_out (cells * A reals, reference storage) may be NULL. */
int opx_cgc_mv_charge(opx_ctx *ctx, double location, double width, double mu, double uv_cutoff,
                      double coherence_length, double ir_coefficient, const double *noise, int64_t n_noise,
                      double *rho_out, int64_t n_rho);
/* as opx_cgc_mv_charge with the noise drawn on the device from java.util.Random(seed): the
 * values new Random(seed).nextGaussian() returns, in order (MVModel with useSeed). */
int opx_cgc_mv_charge_seeded(opx_ctx *ctx, double location, double width, double mu, double uv_cutoff,
                             double coherence_length, double ir_coefficient, int64_t seed, double *rho_out,
                             int64_t n_rho);
/* the first `count` values of new java.util.Random(seed).nextGaussian() (the generator and polar
 * method the JDK specifies, with StrictMath.log), computed on the device into a host buffer */
int opx_java_gaussians(int64_t seed, int64_t count, double *out);
/* an arbitrary charge density for the next solve (IInitialChargeDensity.getChargeDensity),
 * cells * A reals in reference storage */
int opx_cgc_set_charge(opx_ctx *ctx, const double *rho, int64_t n_values);
/* LightConePoissonSolverImprovedFull.initialize + solve (physics/initial/CGC/
 * LightConePoissonSolverImprovedFull.java:34-263): transverse Poisson equation per x0 sheet,
 * path-ordered Wilson lines V at t = -at/2 and +at/2, transverse links U and Unext rotated by
 * them, E = -proj(Unext U^+)/at in every direction, and the Gauss violation of the fields the
 * Wilson lines alone produce (getGaussViolation). orientation is +1 or -1; gauss_out (cells * A
 * reals, may be NULL) receives that violation. */
int opx_cgc_lightcone_solve(opx_ctx *ctx, int orientation, double *gauss_out, int64_t n_values);
/* LightConeNGPSuperParticleCreator.initialize (physics/initial/CGC/
 * LightConeNGPSuperParticleCreator.java:76-245) from the last solve's Gauss violation: the block
 * of planes whose largest |G| exceeds 1e-9 of the largest, min(num_threads, block) subdivisions
 * of round(a_0/at) super particles each, their charges refined to 2nd then 4th order along the
 * beam (100 sweeps each), APPENDED to the context's super-particle list (Simulation.particles.add).
 * added, z_start and block_width (each may be NULL) report the particles added and the block. */
int opx_cgc_create_superparticles(opx_ctx *ctx, int orientation, int num_threads, int *added, int *z_start,
                                  int *block_width);
/* the charge density of the next solve (MVModel.getChargeDensity) or the Gauss violation of the
 * last one (LightConePoissonSolverImprovedFull.getGaussViolation), cells * A reals */
enum { OPX_CGC_CHARGE = 0, OPX_CGC_GAUSS = 1 };
int opx_cgc_download(opx_ctx *ctx, int which, double *out, int64_t n_values);
/* WilsonLineObservables (physics/initial/CGC/WilsonLineObservables.java:36-131) on the Wilson
 * lines the last solve left on the transverse plane behind the nucleus (getV: x0 = 0 for
 * orientation +1, N0 - 1 for -1). Tadpole: the mean of Re tr V. Dipole: tr(V_x V_y^+) averaged
 * over the pairs in each of num_bins = (int)sqrt(N1 N2) / 2 bins of periodic transverse distance,
 * bin_width = min(N1 a1, N2 a2) / 2 / num_bins; corr and counts hold `capacity` >= num_bins
 * entries (both NULL: only num_bins and bin_width are returned). An empty bin reads 0/0. */
int opx_cgc_wilson_tadpole(opx_ctx *ctx, double *value);
int opx_cgc_wilson_dipole(opx_ctx *ctx, int capacity, int *num_bins, double *bin_width, double *corr,
                          int64_t *counts);
/* LightConeNGPParticleCreator.initialize (physics/initial/CGC/LightConeNGPParticleCreator.java:82-234)
 * from the last solve's Gauss violation, for the non-optimized CGC types: round(a_0/at) particles
 * per cell of the planes z_start..z_end (the block of the super-particle creator, its last plane
 * included), charges refined along every transverse cell's list, APPENDED to the context's
 * per-particle list (opx_particles_*) with Q1 = Q0 for every particle of the list. added,
 * z_start and z_end (each may be NULL) report the particles added and the planes. */
int opx_cgc_create_particles(opx_ctx *ctx, int orientation, int64_t *added, int *z_start, int *z_end);
/* frees the charge density and Gauss violation (IInitialChargeDensity.clear) */
int opx_cgc_clear(opx_ctx *ctx);

/* ---- particle seam: per-particle CGC NGP particles ----------------------
 * Replaces CGCParticleInterpolationNGP (physics/grid/CGCParticleInterpolationNGP.java:15-104)
 * and CGCParticleSolver (physics/movement/solver/CGCParticleSolver.java:15-36) over a list of
 * CGCParticle moving along grid axis `direction`. pos0 and vel hold 3 doubles per
 * particle, charges one algebra element per particle in reference storage (Q0).
 * Positions stay in double in every precision: the nearest grid point
 * (int) rint(pos / a) (physics/util/GridFunctions.java:67-73) is bit-exact. */
int opx_particles_upload(opx_ctx *ctx, int count, const int *direction, const double *pos0, const double *vel,
                         const double *charges, int64_t n_values);
/* pos0, pos1 (3 per particle) and Q0, Q1 (reference storage); any output may be NULL */
int opx_particles_download(opx_ctx *ctx, double *pos0, double *pos1, double *Q0, double *Q1, int64_t count);
/* the number of particles in the context's per-particle list */
int opx_particles_count(const opx_ctx *ctx, int64_t *count);
/* interpolateChargedensity: rho(ngp(pos0)) += Q0 */
int opx_particles_interpolate_chargedensity(opx_ctx *ctx);
/* updatePosition (pos1 = pos0 + v dt), interpolateToParticle (the crossed link of Unext)
 * and updateCharge (Q1 = U^dagger Q0 U on a crossing, else Q0 <-> Q1) */
int opx_particles_push(opx_ctx *ctx, double dt);
/* interpolateToGrid: on a crossing J_d += Q0 a_d/dt at the old cell (v > 0) or -Q1 a_d/dt at the new cell */
int opx_particles_interpolate_to_grid(opx_ctx *ctx);
/* YangMillsParticle.reassignValues (physics/particles/YangMillsParticle.java:117-125): pos0 <-> pos1, Q0 <-> Q1 */
int opx_particles_reassign(opx_ctx *ctx);

/* ---- batched events ------------------------------------------------------
 * Many independent lattices (events) of uneven sizes advanced together: one
 * persistent kernel per step runs the fused update of every event (a fixed
 * grid striding over per-event work items), optionally replayed from a
 * captured hipGraph. No counterpart in the reference (where ensembles are
 * independent jobs); each event follows FastTYMSolver exactly as a single
 * context does.
 * opx_batch_event returns a context view of one event that borrows the batch
 * memory: use it for opx_upload_field / opx_download_field /
 * opx_set_links_from_coefficients / opx_measure / regions; stepping a view is
 * rejected (OPX_E_STATE) and opx_destroy of a view is left to the batch.
 * dims holds 3 cells per event; spacing, time step and coupling are shared. */
typedef struct opx_batch opx_batch;
enum {
    OPX_BATCH_OPT_GRID = 0,   /* workgroups of the persistent kernel */
    OPX_BATCH_OPT_FORM = 1    /* 0: per-link kernel; 1, 2: LDS-tile kernel with non-temporal / cached
                                 stores of E and Unext (SU(2) only); 3: cell kernel (the three links
                                 of a cell per thread, as OPX_KERNEL_CELL) */
};
int opx_batch_create(int num_colors, int count, const int *dims, const double *spacing, double time_step,
                     double coupling, int precision, opx_batch **out);
void opx_batch_destroy(opx_batch *batch);
const char *opx_batch_last_error(const opx_batch *batch);
opx_ctx *opx_batch_event(opx_batch *batch, int event);
int opx_batch_set_stream(opx_batch *batch, void *hip_stream);
int opx_batch_set_option(opx_batch *batch, int option, int value);
int opx_batch_get_option(const opx_batch *batch, int option, int *value);
int opx_batch_get_info(const opx_batch *batch, int *count, int64_t *total_cells, int *grid, int *steps_done);
/* stepLinks on every event */
int opx_batch_step_links(opx_batch *batch, double dt);
/* `steps` x { storeFields, fused update } on every event, reunitarizing after
 * every reunit_every-th step; use_graph != 0 replays captured periods of steps */
int opx_batch_leapfrog(opx_batch *batch, int steps, int reunit_every, int use_graph);
/* out[event * OPX_M_COUNT + k]: opx_measure of each event */
int opx_batch_measure(opx_batch *batch, double *out);
int opx_batch_sync(opx_batch *batch);
/* times persistent-grid sizes (scratch copies of E and Unext) and keeps the fastest */
int opx_batch_autotune(opx_batch *batch, int reps, double *best_ms);

/* ---- Simulation driver (physics/Simulation.java:318-452) -----------------
 * opx_simulation_initialize: stepLinks, rho, J deposit at t = 0.
 * opx_simulation_step: n times { clock++, storeFields, step, resetCharge,
 * interpolateChargedensity, resetCurrent, interpolateToGrid }. Nothing inside a step reads
 * the charge density, so it is gathered on the last of the n steps only; the state after the
 * call is the same. Diagnostics that read rho between steps call with n = 1.
 * opx_simulation_steps_done returns the integer clock. */
int opx_simulation_initialize(opx_ctx *ctx);
int opx_simulation_step(opx_ctx *ctx, int n);
int opx_simulation_steps_done(const opx_ctx *ctx);

/* Health monitor: counts non-finite values and links with |U U^dagger - 1|
 * above tol (a stricter relative of UnitarityTester, below). */
int opx_check_health(opx_ctx *ctx, double tol, int64_t *non_finite, int64_t *non_unitary);

/* UnitarityTester.calculate (diagnostics/methods/UnitarityTester.java:31-46,
 * GroupElement.getRealTrace): the links (cell, dir) where |Re tr U| or |Re tr Unext|
 * is not below num_colors + tol (the reference uses tol = 1e-8; NaN counts as not
 * below). *non_unitary receives their number; cells (up to max_cells entries, may
 * be NULL when max_cells is 0) the cell index of each, in the reference's loop order
 * (cell, then direction; a cell appears once per flagged direction). */
int opx_unitarity_test(opx_ctx *ctx, double tol, int64_t *non_unitary, int64_t *cells, int64_t max_cells);

/* GaussConstraintRestoration.iterateRestorationAlgorithm
 * (diagnostics/methods/GaussConstraintRestoration.java:79-163, after hep-ph/9603384):
 * the violation C(x) = sum_j [Unext_j(x) U_j(x)^+ - U_j(x-j)^+ Unext_j(x-j)] / (a_t a)
 * with total (1/N) sum_x |proj C(x) - rho(x)|^2; each iteration backs up E and Unext,
 * applies E_j(x) += -a gamma/2 proj(U_j(x) C(x+j) U_j(x)^+ - C(x)) on every cell, then
 * stepLinks (Unext = exp(-i a_t E) U on the active cells). It stops when the total is
 * below absolute_value (*reason 0), when a correction made it larger, which is then
 * undone (1), after max_iterations corrections (2), or at once when max_iterations
 * <= 0 (3). *iterations: corrections kept; violation_before/after: the total before
 * and for the state left behind. Evenly spaced lattices only (the reference's
 * Grid.getLatticeSpacing throws otherwise): OPX_E_UNSUPPORTED. In FP32 contexts the
 * violation and the correction are computed in double and E is stored in float. */
int opx_gauss_restore(opx_ctx *ctx, double gamma, int max_iterations, double absolute_value, int *iterations,
                      int *reason, double *violation_before, double *violation_after);

/* ---- SU(2) adjoint of the leapfrog (differentiation; no counterpart in the reference) ----
 *
 * Treats the context's current U as the initial links and E as the initial
 * electric field E0, runs the forward map of opx_step_links followed by
 * opx_leapfrog(ctx, steps, reunit_every) (FastTYMSolver.java:40-84), and
 * evaluates the loss
 *   L = lattice average of e_T - e_L
 *     = (1/N_dir) sum_k [eT_el + eT_mag - eL_el - eL_mag](k)
 * of opx_projected_energy_density(ctx, direction, ...) on the final state
 * (ProjectedEnergyDensity.java:132-160). When grad is not NULL it then returns
 * dL/dE0 (n_values = num_cells * 3 * 3, reference per-cell order) from a
 * hand-written reverse sweep: (U, E) is checkpointed every checkpoint_every
 * steps and each segment is recomputed before it is swept backwards. The
 * context is left in the final forward state. SU(2), pure Yang-Mills, every
 * cell active and evaluated; in FP32 contexts the sweep runs in float.
 * With grad == NULL only the forward and the loss run (finite differences). */
int opx_adjoint_gradient(opx_ctx *ctx, int steps, int reunit_every, int checkpoint_every, int direction,
                         double *loss, double *grad, int64_t n_values);

#ifdef __cplusplus
}
#endif
#endif
