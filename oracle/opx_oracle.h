/*
 * opx_oracle.h -- FP64 CPU restatement of OpenPixi's temporal-gauge Yang-Mills
 * leapfrog (FastTYMSolver), its measurements and its NGP colored-particle coupling.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * path: only tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load it. The product library never links or calls it.
 *
 * Storage follows the reference object layout, flattened per cell:
 *   links  U[(cell*3 + dir)*G + c]   G = 4  (SU(2): e0..e3, U = e0 + i e_k sigma_k)
 *                                     G = 18 (SU(3): 9 real parts row-major, then 9 imaginary parts)
 *   algebra E[(cell*3 + dir)*A + c]  A = 3  (SU(2): generator coefficients v_a, t_a = sigma_a/2)
 *                                     A = 9  (SU(3): Hermitian storage
 *                                               [v0, v1-i v3, v2-i v6; v1+i v3, v4, v5-i v7; v2+i v6, v5+i v7, v8])
 *   cell index ((x0*N1)+x1)*N2+x2, x0 slowest.
 * References: pixi/src/main/java/org/openpixi/pixi/physics/grid/Cell.java:14-29,
 *             math/SU2GroupElement.java:14, math/SU3GroupElement.java:25,
 *             math/SU3AlgebraElement.java:3-11, physics/grid/Grid.java:525-545.
 *
 * Parity status of each group of functions is stated in oracle/README (tests/ pin them).
 */
#ifndef OPX_ORACLE_H
#define OPX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- element algebra (nc = 2 or 3) ---- */
int  orc_group_size(int nc);     /* 4 or 18 */
int  orc_alg_size(int nc);       /* 3 or 9  */
int  orc_alg_dim(int nc);        /* 3 or 8 generator coefficients */
void orc_group_mult(int nc, const double *a, const double *b, double *out);
void orc_group_adj(int nc, const double *a, double *out);
void orc_group_proj(int nc, const double *a, double *alg_out);
void orc_alg_get_link(int nc, const double *alg, double *grp_out);      /* exp(i sum v_a t_a) */
void orc_alg_act(int nc, const double *alg, const double *g, double *alg_out); /* g X g^dagger */
double orc_alg_square(int nc, const double *alg);
double orc_alg_get(int nc, const double *alg, int j);                  /* generator coefficient j */
void orc_alg_set(int nc, double *alg, int j, double value);            /* reference incremental set */
void orc_alg_from_coeffs(int nc, const double *coeffs, double *alg);   /* zero element then set(j) */
/* batch helpers: count elements, coefficients packed nc*nc-1 per element */
void orc_links_from_coeffs(int nc, int64_t count, const double *coeffs, double *links);
void orc_algs_from_coeffs(int nc, int64_t count, const double *coeffs, double *algs);

/* ---- lattice state ---- */
typedef struct orc_grid {
    int nc;
    int dims[3];
    double as[3];
    double at;
    double g;
    int64_t n;          /* number of cells */
    int64_t cum[4];     /* cummulatedCellCount */
    int gsz, asz;
    double *U, *Unext, *E, *J, *rho;
    unsigned char *active, *evaluatable;
    int nthreads;       /* OpenMP threads used by the cell loops (1 = sequential) */
} orc_grid;

orc_grid *orc_grid_create(int nc, const int dims[3], const double as[3], double at, double g);
void orc_grid_destroy(orc_grid *gr);
void orc_grid_set_threads(orc_grid *gr, int nthreads);
double *orc_grid_field(orc_grid *gr, int which);   /* 0 U, 1 Unext, 2 E, 3 J, 4 rho */
unsigned char *orc_grid_mask(orc_grid *gr, int which); /* 0 active, 1 evaluatable */
int64_t orc_shift(const orc_grid *gr, int64_t index, int dir, int orientation);
void orc_set_active_region(orc_grid *gr, const int lo[3], const int hi[3]);
void orc_set_evaluation_region(orc_grid *gr, const int lo[3], const int hi[3]);

/* ---- FieldSolver seam (FastTYMSolver) ---- */
void orc_staple_sum(const orc_grid *gr, int64_t index, int d, double *S_out);
void orc_step(orc_grid *gr, double at);         /* CombinedUpdate over all cells */
void orc_step_links(orc_grid *gr, double at);   /* UpdateLinks over all cells */
/* ImplicitTYMSolver.step (opx_oracle_implicit.c) and the group log / power it uses */
void orc_implicit_step(orc_grid *gr, double at, int iterations, double damping);
void orc_group_log(int nc, const double *u, double *v);
void orc_group_pow(int nc, const double *u, double x, double *o);
/* CGC initial condition, beam along x0 (opx_oracle_cgc.c) */
void orc_regulate_charge_gaussian(const int dims[3], double uvt, double long_width, double ir, double aT,
                                  double aL, double *rho);
void orc_poisson_2d(int n1, int n2, double aT, const double *rho, double *phi);
void orc_mv_charge(const orc_grid *gr, double location, double width, double mu, double uvt, double long_width,
                   double ir, const double *noise, double *rho);
void orc_wilson_segment(const orc_grid *gr, double z1, double z2, int M, const double *P1, const double *P2, double *W);
void orc_lightcone_solve(orc_grid *gr, const double *rho, int orientation, double *gauss);
/* as orc_lightcone_solve; VT (N1 N2 group elements, may be NULL) receives the Wilson lines at -at/2
 * on the plane behind the nucleus (x0 = 0 for orientation +1, N0 - 1 for -1) */
void orc_lightcone_solve_vt(orc_grid *gr, const double *rho, int orientation, double *gauss, double *VT);
void orc_lightcone_particles(const orc_grid *gr, const double *gauss, int orientation, int z_start, int z_end, int ppc,
                             double *pos0, double *pos1, double *vel, double *charges);
double orc_wilson_tadpole(int nc, long T, const double *VT);
void orc_wilson_dipole(int nc, int n1, int n2, double a1, double a2, const double *VT, int nbins, double *corr,
                       long *counts);
void orc_lightcone_block(const orc_grid *gr, const double *gauss, int *z_start, int *block_width);
/* one list's refinement sweeps (order 2 or 4), as both light-cone creators apply them */
void orc_refine_list(double *q, int nn, int A, int ppc, int skip, int order, int sweeps);
void orc_lightcone_superparticles(const orc_grid *gr, const double *gauss, int orientation, int z_start,
                                  int block_width, int subdivisions, int ppc, int *sp_orientation, int *sp_num,
                                  int *sp_index_offset, int *sp_per_plane, int *sp_shift, int *sp_ppc, double *charges);
void orc_store_fields(orc_grid *gr);            /* U <-> Unext */
void orc_reset_current(orc_grid *gr);
void orc_reset_charge(orc_grid *gr);

/* ---- measurements ---- */
void orc_plaquette(const orc_grid *gr, int64_t index, int d1, int d2, int o1, int o2, int time_index, double *out);
void orc_get_B(const orc_grid *gr, int64_t index, int dir, int time_index, double *alg_out);
void orc_gauss_constraint(const orc_grid *gr, int64_t index, double *alg_out);
double orc_gauss_constraint_squared(const orc_grid *gr, int64_t index);
/* out: [E2_0,E2_1,E2_2,B2_0,B2_1,B2_2,gauss_mean_sq,total_charge_squared] (FieldMeasurements) */
void orc_measure(const orc_grid *gr, double *out);
/* ProjectedEnergyDensity per longitudinal slice: out arrays of length dims[dir] */
void orc_projected_energy_density(const orc_grid *gr, int dir, double *eT_el, double *eT_mag,
                                  double *eL_el, double *eL_mag);
int orc_projected_poynting(const orc_grid *gr, int dir, double *avg, double *tavg);
void orc_gauge_transform(orc_grid *gr, const double *g_per_cell);

/* ---- colored particles (CGC super particles, light-cone NGP) ---- */
typedef struct orc_superparticles {
    int count;
    int *orientation, *num, *index_offset, *per_plane, *shift, *ppc;
    int64_t *q_offset;      /* start of each particle's charges in q (in algebra elements) */
    double *q;              /* charges, algebra storage (A reals each) */
    int64_t total_charges;
} orc_superparticles;

int orc_sp_current_offset(int orientation, int index_offset, int per_plane, int shift, int ppc, int t);
int orc_sp_ngp_offset(int orientation, int index_offset, int per_plane, int shift, int ppc, int t);
int orc_sp_needs_update(int orientation, int shift, int ppc, int t);
void orc_sp_interpolate_to_grid(orc_grid *gr, orc_superparticles *sp, int t);
void orc_sp_interpolate_chargedensity(orc_grid *gr, const orc_superparticles *sp, int t);

/* per-particle CGC NGP particles (CGCParticle + CGCParticleSolver + CGCParticleInterpolationNGP) */
typedef struct orc_particles {
    int count;
    int *direction;
    double *pos0, *pos1, *vel;   /* 3 per particle */
    double *Q0, *Q1;             /* algebra storage */
    double *Ulink;               /* group storage, per particle */
    unsigned char *update_charge;
} orc_particles;
int orc_nearest_grid_point(double pos, double as);
void orc_p_update_positions(orc_particles *p, double dt);
void orc_p_interpolate_to_particle(const orc_grid *gr, orc_particles *p);
void orc_p_update_charges(orc_particles *p);
void orc_p_interpolate_to_grid(orc_grid *gr, const orc_particles *p);
void orc_p_interpolate_chargedensity(orc_grid *gr, const orc_particles *p);
void orc_p_reassign(orc_particles *p);

/* ---- Simulation driver (Simulation.initialize / step) ---- */
/* particles may be NULL. steps_done is the integer clock (totalSimulationSteps). */
void orc_sim_initialize(orc_grid *gr, orc_superparticles *sp, orc_particles *p, double at);
void orc_sim_step(orc_grid *gr, orc_superparticles *sp, orc_particles *p, double at, int *steps_done);

/* ---- SU(2) adjoint of the leapfrog (opx_oracle_adjoint.c) ---- */
void orc_adj_exp_vjp(const double *w, const double *Mbar, double *wbar);
void orc_adj_force_vjp(const orc_grid *gr, const double *U, const double *Ebar, double scale, double *Ubar);
double orc_adj_bsq(const orc_grid *gr, const double *U, int dir, double *Ubar);
double orc_adj_loss(const orc_grid *gr, const double *Uprev, const double *Ulast, const double *E, int dir);
int orc_adjoint_gradient(orc_grid *gr, int steps, int dir, double *loss, double *grad);

/* ---- GaussConstraintRestoration and UnitarityTester (opx_oracle_gauss.c) ---- */
/* gv: n group elements; returns the mean violation |proj gv - rho|^2 */
double orc_restore_violation(const orc_grid *gr, double *gv);
void orc_restore_correct(orc_grid *gr, const double *gv, double gamma);
/* iterateRestorationAlgorithm: returns the corrections kept (-1: uneven spacings); reason 0 = the
 * absolute value was reached, 1 = a correction made the violation worse and was undone, 2 = the
 * maximum number of iterations, 3 = no iteration ran; history (max_iterations + 1 doubles, may be
 * NULL) receives the violation before and after each kept correction */
int orc_gauss_restore(orc_grid *gr, double gamma, int max_iterations, double absolute_value, int *reason,
                      double *before, double *after, double *history);
int64_t orc_unitarity_flags(const orc_grid *gr, double tol, unsigned char *flags);

/* ---- timing helper for the CPU baseline ---- */
double orc_time_steps(orc_grid *gr, int steps, int warmup);   /* seconds for `steps` leapfrog steps */

#ifdef __cplusplus
}
#endif
#endif
