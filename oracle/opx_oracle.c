/*
 * opx_oracle.c -- FP64 CPU restatement of OpenPixi's FastTYMSolver path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Built with -O2 -ffp-contract=off so
 * that every double operation rounds once, as Java evaluates doubles. Term
 * order follows the reference's algebra where it is a plain formula
 * (SU(2) quaternion product, staple sums, E-update order); the SU(3) 3x3
 * product is an ordinary row-by-column complex product (the reference's
 * Mathematica-generated term order is not reproduced, so SU(3) results agree
 * with the reference to rounding, not bit for bit).
 *
 * Citations are to pixi/src/main/java/org/openpixi/pixi/ in the reference.
 */
#include "opx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== */
/* SU(2): U = e0 + i e_k sigma_k; algebra v_a with t_a = sigma_a / 2      */
/* ===================================================================== */

/* math/SU2GroupElement.java:206-217 (mult) */
static void su2_mult(const double *a, const double *b, double *o)
{
    double r0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    double r1 = a[0] * b[1] + a[1] * b[0] - a[2] * b[3] + a[3] * b[2];
    double r2 = a[0] * b[2] + a[2] * b[0] - a[3] * b[1] + a[1] * b[3];
    double r3 = a[0] * b[3] + a[3] * b[0] - a[1] * b[2] + a[2] * b[1];
    o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3;
}

/* math/SU2GroupElement.java:147-153 (adj) */
static void su2_adj(const double *a, double *o)
{
    o[0] = a[0]; o[1] = -a[1]; o[2] = -a[2]; o[3] = -a[3];
}

/* math/SU2GroupElement.java:254-263 (proj: u_a = 2 e_a) */
static void su2_proj(const double *a, double *v)
{
    v[0] = 2 * a[1]; v[1] = 2 * a[2]; v[2] = 2 * a[3];
}

/* math/SU2AlgebraElement.java:140-152 (getLink = exp(i v_a sigma_a / 2)) */
static void su2_get_link(const double *v, double *o)
{
    double sum = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double mod = sqrt(sum) / 2;
    double sinfakt = (mod < 1.E-20) ? 0.0 : 0.5 / mod * sin(mod);
    o[0] = cos(mod);
    o[1] = v[0] * sinfakt;
    o[2] = v[1] * sinfakt;
    o[3] = v[2] * sinfakt;
}

/* math/SU2AlgebraElement.java:159-172 (act: g X g^dagger) */
static void su2_act(const double *v, const double *g, double *o)
{
    double xm[4] = {0.0, v[0] / 2, v[1] / 2, v[2] / 2};
    double gd[4], t[4], r[4];
    su2_adj(g, gd);
    su2_mult(xm, gd, t);
    su2_mult(g, t, r);
    su2_proj(r, o);
}

/* ===================================================================== */
/* SU(3): 18 reals (9 real parts row-major, 9 imaginary parts row-major)  */
/* algebra: 9-real Hermitian storage (see header)                         */
/* ===================================================================== */

static void su3_mult(const double *a, const double *b, double *o)
{
    double t[18];
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 3; k++) {
                double ar = a[3 * r + k], ai = a[9 + 3 * r + k];
                double br = b[3 * k + c], bi = b[9 + 3 * k + c];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            t[3 * r + c] = re;
            t[9 + 3 * r + c] = im;
        }
    }
    memcpy(o, t, sizeof t);
}

/* math/SU3GroupElement.java:101-131 (adj = conjugate transpose) */
static void su3_adj(const double *a, double *o)
{
    double t[18];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            t[3 * r + c] = a[3 * c + r];
            t[9 + 3 * r + c] = -a[9 + 3 * c + r];
        }
    memcpy(o, t, sizeof t);
}

/* math/SU3GroupElement.java:494-505: traceless part of (M - M^dagger)/(2i),
 * returned in Hermitian storage. */
static void su3_proj(const double *m, double *v)
{
    double d0 = m[9], d1 = m[13], d2 = m[17];
    double tr = (d0 + d1 + d2) / 3;
    v[0] = d0 - tr;
    v[4] = d1 - tr;
    v[8] = d2 - tr;
    v[1] = (m[10] + m[12]) / 2;   /* Re H01: (Im M01 + Im M10)/2 */
    v[3] = (m[1] - m[3]) / 2;     /* -Im H01: (Re M01 - Re M10)/2 */
    v[2] = (m[11] + m[15]) / 2;
    v[6] = (m[2] - m[6]) / 2;
    v[5] = (m[14] + m[16]) / 2;
    v[7] = (m[5] - m[7]) / 2;
}

/* Hermitian storage -> full matrix H (group storage) */
static void su3_herm_to_matrix(const double *v, double *h)
{
    double t[18] = {v[0], v[1], v[2],
                    v[1], v[4], v[5],
                    v[2], v[5], v[8],
                    0.0, -v[3], -v[6],
                    v[3], 0.0, -v[7],
                    v[6], v[7], 0.0};
    memcpy(h, t, sizeof t);
}

/* math/SU3AlgebraElement.java:144-168 (act: g X g^dagger, then Hermitian part) */
static void su3_act(const double *v, const double *g, double *o)
{
    double h[18], gd[18], t[18], r[18];
    su3_herm_to_matrix(v, h);
    su3_adj(g, gd);
    su3_mult(h, gd, t);
    su3_mult(g, t, r);
    o[0] = r[0];
    o[4] = r[4];
    o[8] = r[8];
    o[1] = (r[1] + r[3]) / 2;
    o[2] = (r[2] + r[6]) / 2;
    o[5] = (r[5] + r[7]) / 2;
    o[3] = (r[12] - r[10]) / 2;
    o[6] = (r[15] - r[11]) / 2;
    o[7] = (r[16] - r[14]) / 2;
}

/* math/SU3AlgebraElement.java:174-176 (square = 2 tr H^2 = sum_a c_a^2) */
static double su3_square(const double *v)
{
    return 2 * (v[0] * v[0] + v[4] * v[4] + v[8] * v[8] +
                2 * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3] + v[5] * v[5] + v[6] * v[6] + v[7] * v[7]));
}

/* math/SU3AlgebraElement.java:389-400: sum_{k<=n} (iH)^k / k!, accumulated term by term */
static void su3_exp_taylor(const double *v, int n, double *o)
{
    double res[18] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double term[18];
    double ih[18];
    memcpy(term, res, sizeof term);
    /* iH = -Im H + i Re H */
    double h[18];
    su3_herm_to_matrix(v, h);
    for (int k = 0; k < 9; k++) {
        ih[k] = -h[9 + k];
        ih[9 + k] = h[k];
    }
    for (int i = 1; i <= n; i++) {
        su3_mult(term, ih, term);
        double f = 1.0 / i;
        for (int k = 0; k < 18; k++) term[k] *= f;
        for (int k = 0; k < 18; k++) res[k] += term[k];
    }
    memcpy(o, res, sizeof res);
}

/* exp(iH) through the spectral decomposition of the traceless Hermitian H:
 * eigenvalues from the trigonometric solution of l^3 - p l - det H = 0,
 * eigenvectors as cross products of two rows of H - l. The branch structure
 * (Taylor-15 for tr H^2 <= 1, Taylor-25 when two eigenvalues nearly coincide)
 * follows math/SU3AlgebraElement.java:240-309,407-417. */
static void su3_get_link(const double *v, double *o)
{
    double trsq = su3_square(v) / 2;
    if (trsq <= 1.0) {
        su3_exp_taylor(v, 15, o);
        return;
    }
    double h[18];
    su3_herm_to_matrix(v, h);
    /* det H (real for Hermitian H) */
    double dre, dim;
    {
        /* det = sum over permutations; complex arithmetic */
        double ar[9], ai[9];
        for (int k = 0; k < 9; k++) { ar[k] = h[k]; ai[k] = h[9 + k]; }
#define CM(xr, xi, yr, yi, zr, zi) do { double _r = (xr) * (yr) - (xi) * (yi); double _i = (xr) * (yi) + (xi) * (yr); zr = _r; zi = _i; } while (0)
        double m1r, m1i, m2r, m2i, c0r, c0i, c1r, c1i, c2r, c2i;
        /* cofactor expansion along row 0 */
        CM(ar[4], ai[4], ar[8], ai[8], m1r, m1i);
        CM(ar[5], ai[5], ar[7], ai[7], m2r, m2i);
        CM(ar[0], ai[0], m1r - m2r, m1i - m2i, c0r, c0i);
        CM(ar[3], ai[3], ar[8], ai[8], m1r, m1i);
        CM(ar[5], ai[5], ar[6], ai[6], m2r, m2i);
        CM(ar[1], ai[1], m1r - m2r, m1i - m2i, c1r, c1i);
        CM(ar[3], ai[3], ar[7], ai[7], m1r, m1i);
        CM(ar[4], ai[4], ar[6], ai[6], m2r, m2i);
        CM(ar[2], ai[2], m1r - m2r, m1i - m2i, c2r, c2i);
#undef CM
        dre = c0r - c1r + c2r;
        dim = c0i - c1i + c2i;
        (void)dim;
    }
    double p = trsq / 2;                 /* l^3 - p l - det = 0 */
    double s = sqrt(p / 3);
    double cth = (dre / 2) / (s * s * s);
    if (cth > 1) cth = 1;
    if (cth < -1) cth = -1;
    double th = acos(cth);
    double lam[3];
    for (int i = 0; i < 3; i++) lam[i] = 2 * s * cos((th + 2 * M_PI * i) / 3);
    if (fabs(1 - lam[0] / lam[1]) < 1.E-2 || fabs(1 - lam[0] / lam[2]) < 1.E-2 ||
        fabs(1 - lam[1] / lam[2]) < 1.E-2) {
        su3_exp_taylor(v, 25, o);
        return;
    }
    double res[18];
    memset(res, 0, sizeof res);
    for (int e = 0; e < 3; e++) {
        /* rows of M = H - lam I */
        double mr[9], mi[9];
        for (int k = 0; k < 9; k++) { mr[k] = h[k]; mi[k] = h[9 + k]; }
        mr[0] -= lam[e]; mr[4] -= lam[e]; mr[8] -= lam[e];
        /* three pairwise cross products, keep the largest */
        double best = -1, wr[3] = {0, 0, 0}, wi[3] = {0, 0, 0};
        for (int pr = 0; pr < 3; pr++) {
            int a = pr, b = (pr + 1) % 3;
            double cr[3], ci[3];
            for (int k = 0; k < 3; k++) {
                int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
                double xr = mr[3 * a + k1], xi = mi[3 * a + k1];
                double yr = mr[3 * b + k2], yi = mi[3 * b + k2];
                double ur = mr[3 * a + k2], ui = mi[3 * a + k2];
                double zr = mr[3 * b + k1], zi = mi[3 * b + k1];
                cr[k] = (xr * yr - xi * yi) - (ur * zr - ui * zi);
                ci[k] = (xr * yi + xi * yr) - (ur * zi + ui * zr);
            }
            double nn = 0;
            for (int k = 0; k < 3; k++) nn += cr[k] * cr[k] + ci[k] * ci[k];
            if (nn > best) {
                best = nn;
                memcpy(wr, cr, sizeof wr);
                memcpy(wi, ci, sizeof wi);
            }
        }
        double nrm = sqrt(best);
        for (int k = 0; k < 3; k++) { wr[k] /= nrm; wi[k] /= nrm; }
        double er = cos(lam[e]), ei = sin(lam[e]);
        /* res += e^{i lam} w w^dagger */
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                double pr_ = wr[r] * wr[c] + wi[r] * wi[c];
                double pi_ = wi[r] * wr[c] - wr[r] * wi[c];
                res[3 * r + c] += er * pr_ - ei * pi_;
                res[9 + 3 * r + c] += er * pi_ + ei * pr_;
            }
    }
    memcpy(o, res, sizeof res);
}

/* math/SU3AlgebraElement.java:117-129 */
static double su3_get(const double *v, int j)
{
    switch (j) {
    case 0: return 2 * v[1];
    case 1: return 2 * v[3];
    case 2: return v[0] - v[4];
    case 3: return 2 * v[2];
    case 4: return 2 * v[6];
    case 5: return 2 * v[5];
    case 6: return 2 * v[7];
    case 7: return (v[0] + v[4] - 2 * v[8]) / sqrt(3.0);
    default: return 0.0;
    }
}

/* math/SU3AlgebraElement.java:93-115 (incremental set of one coefficient) */
static void su3_set(double *v, int j, double value)
{
    double diff = (value - su3_get(v, j)) / 2;
    switch (j) {
    case 0: v[1] += diff; break;
    case 1: v[3] += diff; break;
    case 2: v[0] += diff; v[4] -= diff; break;
    case 3: v[2] += diff; break;
    case 4: v[6] += diff; break;
    case 5: v[5] += diff; break;
    case 6: v[7] += diff; break;
    case 7: v[0] += diff / sqrt(3.0); v[4] += diff / sqrt(3.0); v[8] -= 2 * diff / sqrt(3.0); break;
    default: break;
    }
}

/* ===================================================================== */
/* dispatch                                                               */
/* ===================================================================== */

int orc_group_size(int nc) { return nc == 2 ? 4 : 18; }
int orc_alg_size(int nc) { return nc == 2 ? 3 : 9; }
int orc_alg_dim(int nc) { return nc * nc - 1; }

void orc_group_mult(int nc, const double *a, const double *b, double *o)
{
    if (nc == 2) su2_mult(a, b, o); else su3_mult(a, b, o);
}
void orc_group_adj(int nc, const double *a, double *o)
{
    if (nc == 2) su2_adj(a, o); else su3_adj(a, o);
}
void orc_group_proj(int nc, const double *a, double *v)
{
    if (nc == 2) su2_proj(a, v); else su3_proj(a, v);
}
void orc_alg_get_link(int nc, const double *v, double *o)
{
    if (nc == 2) su2_get_link(v, o); else su3_get_link(v, o);
}
void orc_alg_act(int nc, const double *v, const double *g, double *o)
{
    if (nc == 2) su2_act(v, g, o); else su3_act(v, g, o);
}
double orc_alg_square(int nc, const double *v)
{
    if (nc == 2) return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    return su3_square(v);
}
double orc_alg_get(int nc, const double *v, int j)
{
    return nc == 2 ? v[j] : su3_get(v, j);
}
void orc_alg_set(int nc, double *v, int j, double value)
{
    if (nc == 2) v[j] = value; else su3_set(v, j, value);
}
void orc_alg_from_coeffs(int nc, const double *c, double *v)
{
    int a = orc_alg_size(nc), d = orc_alg_dim(nc);
    for (int k = 0; k < a; k++) v[k] = 0.0;
    for (int j = 0; j < d; j++) orc_alg_set(nc, v, j, c[j]);
}

void orc_links_from_coeffs(int nc, int64_t count, const double *c, double *links)
{
    int d = orc_alg_dim(nc), G = orc_group_size(nc), A = orc_alg_size(nc);
    for (int64_t i = 0; i < count; i++) {
        double v[9];
        orc_alg_from_coeffs(nc, c + i * d, v);
        orc_alg_get_link(nc, v, links + i * G);
        (void)A;
    }
}

void orc_algs_from_coeffs(int nc, int64_t count, const double *c, double *algs)
{
    int d = orc_alg_dim(nc), A = orc_alg_size(nc);
    for (int64_t i = 0; i < count; i++) orc_alg_from_coeffs(nc, c + i * d, algs + i * A);
}

static void grp_identity(int nc, double *o)
{
    if (nc == 2) {
        o[0] = 1; o[1] = o[2] = o[3] = 0;
    } else {
        for (int k = 0; k < 18; k++) o[k] = 0.0;
        o[0] = o[4] = o[8] = 1.0;
    }
}

/* ===================================================================== */
/* grid                                                                   */
/* ===================================================================== */

orc_grid *orc_grid_create(int nc, const int dims[3], const double as[3], double at, double g)
{
    if (nc != 2 && nc != 3) return NULL;
    orc_grid *gr = (orc_grid *)calloc(1, sizeof(orc_grid));
    gr->nc = nc;
    gr->n = 1;
    for (int i = 0; i < 3; i++) {
        gr->dims[i] = dims[i];
        gr->as[i] = as[i];
        gr->n *= dims[i];
    }
    /* physics/grid/Grid.java:534-538 */
    gr->cum[3] = 1;
    for (int i = 2; i >= 0; i--) gr->cum[i] = gr->cum[i + 1] * dims[i];
    gr->at = at;
    gr->g = g;
    gr->gsz = orc_group_size(nc);
    gr->asz = orc_alg_size(nc);
    gr->U = (double *)calloc((size_t)gr->n * 3 * gr->gsz, sizeof(double));
    gr->Unext = (double *)calloc((size_t)gr->n * 3 * gr->gsz, sizeof(double));
    gr->E = (double *)calloc((size_t)gr->n * 3 * gr->asz, sizeof(double));
    gr->J = (double *)calloc((size_t)gr->n * 3 * gr->asz, sizeof(double));
    gr->rho = (double *)calloc((size_t)gr->n * gr->asz, sizeof(double));
    gr->active = (unsigned char *)malloc((size_t)gr->n);
    gr->evaluatable = (unsigned char *)malloc((size_t)gr->n);
    memset(gr->active, 1, (size_t)gr->n);
    memset(gr->evaluatable, 1, (size_t)gr->n);
    /* physics/grid/Cell.java:63-68: links start as identities */
    for (int64_t c = 0; c < gr->n * 3; c++) {
        grp_identity(nc, gr->U + c * gr->gsz);
        grp_identity(nc, gr->Unext + c * gr->gsz);
    }
    gr->nthreads = 1;
    return gr;
}

void orc_grid_destroy(orc_grid *gr)
{
    if (!gr) return;
    free(gr->U); free(gr->Unext); free(gr->E); free(gr->J); free(gr->rho);
    free(gr->active); free(gr->evaluatable);
    free(gr);
}

void orc_grid_set_threads(orc_grid *gr, int nthreads) { gr->nthreads = nthreads < 1 ? 1 : nthreads; }

double *orc_grid_field(orc_grid *gr, int which)
{
    switch (which) {
    case 0: return gr->U;
    case 1: return gr->Unext;
    case 2: return gr->E;
    case 3: return gr->J;
    case 4: return gr->rho;
    default: return NULL;
    }
}

unsigned char *orc_grid_mask(orc_grid *gr, int which) { return which == 0 ? gr->active : gr->evaluatable; }

/* physics/grid/Grid.java:802-823 */
int64_t orc_shift(const orc_grid *gr, int64_t index, int dir, int orientation)
{
    int64_t result = index;
    int64_t within = (index / gr->cum[dir + 1]) % gr->dims[dir];
    if (orientation > 0) {
        if (within == gr->dims[dir] - 1) result -= gr->cum[dir];
        result += gr->cum[dir + 1];
    } else if (orientation < 0) {
        if (within == 0) result += gr->cum[dir];
        result -= gr->cum[dir + 1];
    }
    return result;
}

static void cell_pos(const orc_grid *gr, int64_t index, int pos[3])
{
    for (int i = 2; i >= 0; i--) {
        pos[i] = (int)(index % gr->dims[i]);
        index /= gr->dims[i];
    }
}

/* physics/grid/Grid.java:992-1020 (set*Region: inclusive box) */
static void set_region(orc_grid *gr, unsigned char *mask, const int lo[3], const int hi[3])
{
    for (int64_t i = 0; i < gr->n; i++) {
        int p[3];
        cell_pos(gr, i, p);
        unsigned char in = 1;
        for (int j = 0; j < 3; j++)
            if (p[j] < lo[j] || hi[j] < p[j]) { in = 0; break; }
        mask[i] = in;
    }
}
void orc_set_active_region(orc_grid *gr, const int lo[3], const int hi[3]) { set_region(gr, gr->active, lo, hi); }
void orc_set_evaluation_region(orc_grid *gr, const int lo[3], const int hi[3]) { set_region(gr, gr->evaluatable, lo, hi); }

#define GU(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * (gr)->gsz)
#define GA(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * (gr)->asz)

/* physics/grid/Grid.java:675-696 */
void orc_staple_sum(const orc_grid *gr, int64_t index, int d, double *S)
{
    int nc = gr->nc, G = gr->gsz;
    for (int k = 0; k < G; k++) S[k] = 0.0;
    int64_t ci1 = orc_shift(gr, index, d, 1);
    for (int i = 0; i < 3; i++) {
        if (i == d) continue;
        int64_t ci2 = orc_shift(gr, index, i, 1);
        int64_t ci3 = orc_shift(gr, ci1, i, -1);
        int64_t ci4 = orc_shift(gr, index, i, -1);
        double a[18], b[18], U1[18], U2[18];
        /* U1 = U_i(x+d) U_d(x+i)^dagger U_i(x)^dagger */
        orc_group_adj(nc, GU(gr, gr->U, ci2, d), a);
        orc_group_mult(nc, GU(gr, gr->U, ci1, i), a, U1);
        orc_group_adj(nc, GU(gr, gr->U, index, i), a);
        orc_group_mult(nc, U1, a, U1);
        /* U2 = (U_d(x-i) U_i(x+d-i))^dagger U_i(x-i) */
        orc_group_mult(nc, GU(gr, gr->U, ci4, d), GU(gr, gr->U, ci3, i), b);
        orc_group_adj(nc, b, U2);
        orc_group_mult(nc, U2, GU(gr, gr->U, ci4, i), U2);
        double area = 1.0 / pow(gr->as[i], 2);
        for (int k = 0; k < G; k++) U1[k] += U2[k];
        for (int k = 0; k < G; k++) U1[k] *= area;
        for (int k = 0; k < G; k++) S[k] += U1[k];
    }
}

/* physics/fields/FastTYMSolver.java:40-62 (CombinedUpdate.execute) */
static void combined_update(orc_grid *gr, int64_t index, double at, const double *unit_factor)
{
    if (!gr->active[index]) return;
    int nc = gr->nc, A = gr->asz;
    for (int i = 0; i < 3; i++) {
        double S[18], temp[18], pr[9], V[18];
        orc_staple_sum(gr, index, i, S);
        orc_group_mult(nc, GU(gr, gr->U, index, i), S, temp);
        orc_group_proj(nc, temp, pr);
        double *E = GA(gr, gr->E, index, i);
        for (int k = 0; k < A; k++) pr[k] = pr[k] * at;
        for (int k = 0; k < A; k++) E[k] += pr[k];
        const double *J = GA(gr, gr->J, index, i);
        for (int k = 0; k < A; k++) E[k] += J[k] * unit_factor[i];
        double w[9];
        for (int k = 0; k < A; k++) w[k] = E[k] * (-at);
        orc_alg_get_link(nc, w, V);
        orc_group_mult(nc, V, GU(gr, gr->U, index, i), V);
        memcpy(GU(gr, gr->Unext, index, i), V, sizeof(double) * gr->gsz);
    }
}

/* physics/fields/FastTYMSolver.java:25-32 */
void orc_step(orc_grid *gr, double at)
{
    double uf[3];
    for (int i = 0; i < 3; i++) uf[i] = -(gr->g * gr->as[i]) * gr->at;
    int64_t n = gr->n;
#pragma omp parallel for schedule(static) num_threads(gr->nthreads) if (gr->nthreads > 1)
    for (int64_t c = 0; c < n; c++) combined_update(gr, c, at, uf);
}

/* physics/fields/FastTYMSolver.java:65-84 */
void orc_step_links(orc_grid *gr, double at)
{
    int nc = gr->nc, A = gr->asz;
    int64_t n = gr->n;
#pragma omp parallel for schedule(static) num_threads(gr->nthreads) if (gr->nthreads > 1)
    for (int64_t c = 0; c < n; c++) {
        if (!gr->active[c]) continue;
        for (int k = 0; k < 3; k++) {
            double w[9], V[18];
            const double *E = GA(gr, gr->E, c, k);
            for (int m = 0; m < A; m++) w[m] = E[m] * (-at);
            orc_alg_get_link(nc, w, V);
            orc_group_mult(nc, V, GU(gr, gr->U, c, k), V);
            memcpy(GU(gr, gr->Unext, c, k), V, sizeof(double) * gr->gsz);
        }
    }
}

/* physics/grid/Cell.java:173-177, applied to every cell */
void orc_store_fields(orc_grid *gr)
{
    double *t = gr->U;
    gr->U = gr->Unext;
    gr->Unext = t;
}

void orc_reset_current(orc_grid *gr) { memset(gr->J, 0, sizeof(double) * (size_t)gr->n * 3 * gr->asz); }
void orc_reset_charge(orc_grid *gr) { memset(gr->rho, 0, sizeof(double) * (size_t)gr->n * gr->asz); }

/* ===================================================================== */
/* measurements                                                           */
/* ===================================================================== */

static const double *link_at(const orc_grid *gr, int time_index, int64_t idx, int d)
{
    return GU(gr, time_index == 0 ? gr->U : gr->Unext, idx, d);
}

/* physics/grid/Grid.java:712-727 */
static void get_link(const orc_grid *gr, int64_t index, int dir, int orientation, int time_index, double *o)
{
    if (orientation < 0) {
        orc_group_adj(gr->nc, link_at(gr, time_index, orc_shift(gr, index, dir, orientation), dir), o);
    } else {
        memcpy(o, link_at(gr, time_index, index, dir), sizeof(double) * gr->gsz);
    }
}

/* physics/grid/Grid.java:609-642 */
void orc_plaquette(const orc_grid *gr, int64_t index, int d1, int d2, int o1, int o2, int time_index, double *out)
{
    int64_t x1 = index;
    int64_t x2 = orc_shift(gr, x1, d1, o1);
    int64_t x3 = orc_shift(gr, x2, d2, o2);
    int64_t x4 = orc_shift(gr, x3, d1, -o1);
    double U1[18], U2[18], U3[18], U4[18], U[18];
    get_link(gr, x1, d1, o1, time_index, U1);
    get_link(gr, x2, d2, o2, time_index, U2);
    get_link(gr, x3, d1, -o1, time_index, U3);
    get_link(gr, x4, d2, -o2, time_index, U4);
    grp_identity(gr->nc, U);
    orc_group_mult(gr->nc, U, U1, U);
    orc_group_mult(gr->nc, U, U2, U);
    orc_group_mult(gr->nc, U, U3, U);
    orc_group_mult(gr->nc, U, U4, U);
    memcpy(out, U, sizeof(double) * gr->gsz);
}

/* physics/grid/Grid.java:877-900 */
void orc_get_B(const orc_grid *gr, int64_t index, int dir, int time_index, double *o)
{
    int j = 0, k = 0;
    switch (dir) {
    case 0: j = 1; k = 2; break;
    case 1: j = 2; k = 0; break;
    case 2: j = 0; k = 1; break;
    }
    double P[18];
    orc_plaquette(gr, index, j, k, 1, 1, time_index, P);
    orc_group_proj(gr->nc, P, o);
    double f = gr->as[dir] / (gr->as[j] * gr->as[k]);
    for (int m = 0; m < gr->asz; m++) o[m] = o[m] * f;
}

/* physics/grid/Grid.java:944-968 */
void orc_gauss_constraint(const orc_grid *gr, int64_t index, double *gauss)
{
    int nc = gr->nc, A = gr->asz;
    for (int m = 0; m < A; m++) gauss[m] = 0.0;
    for (int i = 0; i < 3; i++) {
        double unit_factor = gr->g * gr->as[i];
        int64_t sh = orc_shift(gr, index, i, -1);
        double E[9], L[18];
        get_link(gr, index, i, -1, 0, L);
        orc_alg_act(nc, GA(gr, gr->E, sh, i), L, E);
        for (int m = 0; m < A; m++) E[m] *= -1.0;
        const double *Ex = GA(gr, gr->E, index, i);
        for (int m = 0; m < A; m++) E[m] += Ex[m];
        double f = 1.0 / (gr->as[i] * unit_factor);
        for (int m = 0; m < A; m++) E[m] *= f;
        for (int m = 0; m < A; m++) gauss[m] += E[m];
    }
    const double *r = gr->rho + index * A;
    for (int m = 0; m < A; m++) gauss[m] = gauss[m] - r[m];
}

double orc_gauss_constraint_squared(const orc_grid *gr, int64_t index)
{
    double g[9];
    orc_gauss_constraint(gr, index, g);
    return orc_alg_square(gr->nc, g);
}

/* physics/measurements/FieldMeasurements.java:68-230, summed sequentially in cell order */
void orc_measure(const orc_grid *gr, double *out)
{
    int nc = gr->nc;
    double ev[3] = {0, 0, 0}, bv[3] = {0, 0, 0}, uf[3];
    double gsum = 0.0, qsum = 0.0;
    for (int i = 0; i < 3; i++) uf[i] = pow(gr->g * gr->as[i], -2);
    for (int64_t c = 0; c < gr->n; c++) {
        if (!gr->evaluatable[c]) continue;
        for (int i = 0; i < 3; i++) ev[i] += orc_alg_square(nc, GA(gr, gr->E, c, i)) * uf[i];
        for (int i = 0; i < 3; i++) {
            double b0[9], b1[9];
            orc_get_B(gr, c, i, 0, b0);
            orc_get_B(gr, c, i, 1, b1);
            bv[i] += 0.5 * (orc_alg_square(nc, b0) + orc_alg_square(nc, b1)) * uf[i];
        }
        gsum += orc_gauss_constraint_squared(gr, c);
        qsum += orc_alg_square(nc, gr->rho + c * gr->asz);
    }
    double n = (double)gr->n;
    for (int i = 0; i < 3; i++) {
        out[i] = ev[i] / n;
        out[3 + i] = bv[i] / n;
    }
    out[6] = gsum / n;
    out[7] = sqrt(qsum) / n;
}

/* diagnostics/methods/ProjectedEnergyDensity.java:130-170 */
void orc_projected_energy_density(const orc_grid *gr, int dir, double *eT_el, double *eT_mag,
                                  double *eL_el, double *eL_mag)
{
    int nc = gr->nc;
    int ncell = gr->dims[dir];
    double area = (double)gr->n / (double)gr->dims[dir];
    double uf[3];
    for (int i = 0; i < 3; i++) uf[i] = pow(gr->g * gr->as[i], -2) / area;
    for (int k = 0; k < ncell; k++) eT_el[k] = eT_mag[k] = eL_el[k] = eL_mag[k] = 0.0;
    for (int64_t c = 0; c < gr->n; c++) {
        if (!gr->evaluatable[c]) continue;
        int p[3];
        cell_pos(gr, c, p);
        int proj = p[dir];
        double tel = 0, lel = 0, tmag = 0, lmag = 0;
        for (int j = 0; j < 3; j++) {
            double el = 0.5 * orc_alg_square(nc, GA(gr, gr->E, c, j)) * uf[j];
            double b0[9], b1[9];
            orc_get_B(gr, c, j, 0, b0);
            orc_get_B(gr, c, j, 1, b1);
            double mag = 0.25 * (orc_alg_square(nc, b0) + orc_alg_square(nc, b1)) * uf[j];
            if (j == dir) { lel += el; lmag += mag; }
            else { tel += el; tmag += mag; }
        }
        eT_el[proj] += tel;
        eT_mag[proj] += tmag;
        eL_el[proj] += lel;
        eL_mag[proj] += lmag;
    }
}

/* diagnostics/methods/ProjectedEnergyDensity.java:178-262 (PoyntingComputation), summed over
 * every cell in cell order. The longitudinal Poynting vector, spatially and time averaged
 * (avg) and only time averaged (tavg). SU(2) only: SU3AlgebraElement.mult(AlgebraElement)
 * throws in the reference (math/SU3AlgebraElement.java:196-201); returns -1 for SU(3). */
int orc_projected_poynting(const orc_grid *gr, int dir, double *avg, double *tavg)
{
    if (gr->nc != 2) return -1;
    const int A = gr->asz, G = gr->gsz;
    const int ncell = gr->dims[dir];
    const double area = (double)gr->n / (double)gr->dims[dir];
    double ufF[3], ufEB[3];
    for (int i = 0; i < 3; i++) {
        ufF[i] = 1.0 / ((gr->as[i] * gr->as[dir]) * gr->g * sqrt(area));
        ufEB[i] = 1.0 / ((gr->g * gr->as[i]) * sqrt(area));
    }
    for (int k = 0; k < ncell; k++) avg[k] = tavg[k] = 0.0;
    for (int64_t c = 0; c < gr->n; c++) {
        double local_avg = 0.0;
        for (int i = 0; i < 3; i++) {
            if (i == dir) continue;
            /* plaquettes around the (dir, i) face, both time levels */
            double FG[4] = {0, 0, 0, 0}, P[4];
            for (int t = 0; t < 2; t++) {
                orc_plaquette(gr, c, dir, i, 1, 1, t, P);
                for (int m = 0; m < G; m++) FG[m] += P[m];
                orc_plaquette(gr, c, i, dir, -1, 1, t, P);
                for (int m = 0; m < G; m++) FG[m] += P[m];
                orc_plaquette(gr, c, i, dir, 1, -1, t, P);
                for (int m = 0; m < G; m++) FG[m] += P[m];
                orc_plaquette(gr, c, dir, i, -1, -1, t, P);
                for (int m = 0; m < G; m++) FG[m] += P[m];
            }
            double F[3];
            orc_group_proj(2, FG, F);
            for (int m = 0; m < A; m++) F[m] *= ufF[i] / 8.0;
            /* E_i averaged over x and x - i (transported to x) */
            int64_t sh = orc_shift(gr, c, i, -1);
            double L[4], E2[3], E[3];
            get_link(gr, c, i, -1, 0, L);
            orc_alg_act(2, GA(gr, gr->E, sh, i), L, E2);
            const double *E1 = GA(gr, gr->E, c, i);
            for (int m = 0; m < A; m++) E[m] = (E1[m] + E2[m]) * (0.5 * ufEB[i]);
            double dot = 0.0;
            for (int m = 0; m < A; m++) dot += E[m] * F[m];
            local_avg += dot;
        }
        const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
        double E1[3], E2[3], B1[3], B2[3], b0[3], b1[3];
        for (int m = 0; m < A; m++) {
            E1[m] = GA(gr, gr->E, c, d1)[m] * ufEB[d1];
            E2[m] = GA(gr, gr->E, c, d2)[m] * ufEB[d2];
        }
        orc_get_B(gr, c, d1, 0, b0);
        orc_get_B(gr, c, d1, 1, b1);
        for (int m = 0; m < A; m++) B1[m] = (b0[m] + b1[m]) * (ufEB[d1] * 0.5);
        orc_get_B(gr, c, d2, 0, b0);
        orc_get_B(gr, c, d2, 1, b1);
        for (int m = 0; m < A; m++) B2[m] = (b0[m] + b1[m]) * (ufEB[d2] * 0.5);
        double e1b2 = 0.0, e2b1 = 0.0;
        for (int m = 0; m < A; m++) {
            e1b2 += E1[m] * B2[m];
            e2b1 += E2[m] * B1[m];
        }
        int p[3];
        cell_pos(gr, c, p);
        avg[p[dir]] += local_avg;
        tavg[p[dir]] += e1b2 - e2b1;
    }
    return 0;
}

/* physics/gauge/GaugeTransformation.java:56-86: U -> g(x) U g(x+i)^dagger on U and Unext,
 * E -> proj(g(x) linearized(E) g(x)^dagger). g_per_cell holds one group element per cell. */
void orc_gauge_transform(orc_grid *gr, const double *gpc)
{
    int nc = gr->nc, G = gr->gsz, A = gr->asz;
    for (int64_t c = 0; c < gr->n; c++) {
        for (int d = 0; d < 3; d++) {
            int64_t s = orc_shift(gr, c, d, 1);
            double gd[18], t[18], lin[18], gdx[18];
            orc_group_adj(nc, gpc + s * G, gd);
            orc_group_mult(nc, gpc + c * G, GU(gr, gr->U, c, d), t);
            orc_group_mult(nc, t, gd, GU(gr, gr->U, c, d));
            orc_group_mult(nc, gpc + c * G, GU(gr, gr->Unext, c, d), t);
            orc_group_mult(nc, t, gd, GU(gr, gr->Unext, c, d));
            /* linearized link: SU2AlgebraElement.java:133-138, SU3AlgebraElement.java:402-405 */
            double *E = GA(gr, gr->E, c, d);
            if (nc == 2) {
                double sum = (E[0] * E[0] + E[1] * E[1] + E[2] * E[2]) / 4;
                lin[0] = sqrt(1.0 - sum); lin[1] = E[0] / 2; lin[2] = E[1] / 2; lin[3] = E[2] / 2;
            } else {
                double v[18] = {1, E[3], E[6], -E[3], 1, E[7], -E[6], -E[7], 1,
                                E[0], E[1], E[2], E[1], E[4], E[5], E[2], E[5], E[8]};
                memcpy(lin, v, sizeof v);
            }
            orc_group_adj(nc, gpc + c * G, gdx);
            orc_group_mult(nc, gpc + c * G, lin, t);
            orc_group_mult(nc, t, gdx, t);
            orc_group_proj(nc, t, E);
            (void)A;
        }
    }
}

/* ===================================================================== */
/* CGC super particles                                                    */
/* ===================================================================== */

static int floor_div(int a, int b)
{
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}

/* physics/particles/CGCSuperParticle.java:71-77 */
int orc_sp_current_offset(int orientation, int index_offset, int per_plane, int shift, int ppc, int t)
{
    if (orientation > 0) return index_offset + ((t + shift) / ppc) * per_plane;
    /* floor((-t + shift) / (1.0 * ppc)) on integers */
    return index_offset + floor_div(-t + shift, ppc) * per_plane;
}

/* physics/particles/CGCSuperParticle.java:85-93 */
int orc_sp_ngp_offset(int orientation, int index_offset, int per_plane, int shift, int ppc, int t)
{
    int cur = orc_sp_current_offset(orientation, index_offset, per_plane, shift, ppc, t);
    if (orientation > 0) return cur + ((((t + shift) % ppc) < ppc / 2) ? 0 : per_plane);
    return cur + ((((t - shift + ppc + ppc / 2 - 1) % ppc) < ppc / 2) ? 0 : per_plane);
}

/* physics/particles/CGCSuperParticle.java:100-105 */
int orc_sp_needs_update(int orientation, int shift, int ppc, int t)
{
    if (orientation > 0) return 0 == (t + shift + ppc / 2 + 1) % ppc;
    return 0 == (t - shift + ppc / 2) % ppc;
}

/* physics/grid/CGCSuperParticleInterpolationNGP.java:19-73 (useOffset = false) */
void orc_sp_interpolate_to_grid(orc_grid *gr, orc_superparticles *sp, int t)
{
    int nc = gr->nc, A = gr->asz;
    double at = gr->at, as = gr->as[0];
    int64_t total = gr->n;
    for (int p = 0; p < sp->count; p++) {
        if (!orc_sp_needs_update(sp->orientation[p], sp->shift[p], sp->ppc[p], t)) continue;
        int off = orc_sp_current_offset(sp->orientation[p], sp->index_offset[p], sp->per_plane[p], sp->shift[p], sp->ppc[p], t);
        int quickfix = (sp->ppc[p] == 1) ? sp->per_plane[p] : 0;
        int imin = (off < 0) ? -off : 0;
        int64_t lim = (int64_t)off + sp->num[p];
        if (lim > total - quickfix) lim = total - quickfix;
        int64_t imax = lim - off;
        if (imax < 0) imax = 0;
        int64_t ireg = imax - imin;
        double *Q = sp->q + sp->q_offset[p] * A;
        for (int64_t i = 0; i < ireg; i++) {
            int64_t j = i + imin;
            int64_t index = off + j;
            double *q = Q + j * A;
            double Jv[9], U[18], Ud[18], nq[9];
            const double *Un = GU(gr, gr->Unext, index, 0);
            if (sp->orientation[p] > 0) {
                for (int m = 0; m < A; m++) Jv[m] = q[m] * (as / at);
                double *Jc = GA(gr, gr->J, index + quickfix, 0);
                for (int m = 0; m < A; m++) Jc[m] += Jv[m];
                orc_group_adj(nc, Un, Ud);
                orc_alg_act(nc, q, Ud, nq);
                memcpy(q, nq, sizeof(double) * A);
            } else {
                memcpy(U, Un, sizeof(double) * gr->gsz);
                orc_alg_act(nc, q, U, nq);
                memcpy(q, nq, sizeof(double) * A);
                for (int m = 0; m < A; m++) Jv[m] = q[m] * (-as / at);
                double *Jc = GA(gr, gr->J, index, 0);
                for (int m = 0; m < A; m++) Jc[m] += Jv[m];
            }
        }
    }
}

/* physics/grid/CGCSuperParticleInterpolationNGP.java:75-93 */
void orc_sp_interpolate_chargedensity(orc_grid *gr, const orc_superparticles *sp, int t)
{
    int A = gr->asz;
    int64_t total = gr->n;
    for (int p = 0; p < sp->count; p++) {
        int off = orc_sp_ngp_offset(sp->orientation[p], sp->index_offset[p], sp->per_plane[p], sp->shift[p], sp->ppc[p], t);
        int imin = (off < 0) ? -off : 0;
        int64_t lim = (int64_t)off + sp->num[p];
        if (lim > total) lim = total;
        int64_t imax = lim - off;
        if (imax < 0) imax = 0;
        int64_t ireg = imax - imin;
        const double *Q = sp->q + sp->q_offset[p] * A;
        for (int64_t i = 0; i < ireg; i++) {
            int64_t j = i + imin;
            double *r = gr->rho + (off + j) * A;
            for (int m = 0; m < A; m++) r[m] += Q[j * A + m];
        }
    }
}

/* ===================================================================== */
/* per-particle CGC NGP                                                   */
/* ===================================================================== */

/* physics/util/GridFunctions.java:67-73: (int) Math.rint(pos / as) */
int orc_nearest_grid_point(double pos, double as) { return (int)rint(pos / as); }

/* physics/grid/Grid.java getCellIndex with periodic coordinates */
static int64_t cell_index_periodic(const orc_grid *gr, const int c[3])
{
    int64_t idx = 0;
    for (int i = 0; i < 3; i++) {
        int p = ((c[i] % gr->dims[i]) + gr->dims[i]) % gr->dims[i];
        idx = (i == 0) ? p : idx * gr->dims[i] + p;
    }
    return idx;
}

/* physics/movement/solver/CGCParticleSolver.java:15-21 */
void orc_p_update_positions(orc_particles *p, double dt)
{
    for (int k = 0; k < p->count; k++)
        for (int i = 0; i < 3; i++) p->pos1[3 * k + i] = p->pos0[3 * k + i] + p->vel[3 * k + i] * dt;
}

/* physics/grid/CGCParticleInterpolationNGP.java:70-104 */
void orc_p_interpolate_to_particle(const orc_grid *gr, orc_particles *p)
{
    int G = gr->gsz;
    for (int k = 0; k < p->count; k++) {
        int d = p->direction[k];
        int no[3], nn[3];
        for (int i = 0; i < 3; i++) {
            no[i] = orc_nearest_grid_point(p->pos0[3 * k + i], gr->as[i]);
            nn[i] = orc_nearest_grid_point(p->pos1[3 * k + i], gr->as[i]);
        }
        if (no[d] == nn[d]) continue;
        int64_t io = cell_index_periodic(gr, no), in = cell_index_periodic(gr, nn);
        if (p->vel[3 * k + d] > 0) {
            memcpy(p->Ulink + k * G, GU(gr, gr->Unext, io, d), sizeof(double) * G);
        } else {
            orc_group_adj(gr->nc, GU(gr, gr->Unext, in, d), p->Ulink + k * G);
        }
        p->update_charge[k] = 1;
    }
}

/* physics/movement/solver/CGCParticleSolver.java:24-36; nc is deduced from the
 * algebra size stored with the particles (3 -> SU(2), 9 -> SU(3)). */
static int g_p_nc = 2;
void orc_p_update_charges(orc_particles *p)
{
    int nc = g_p_nc, A = orc_alg_size(nc), G = orc_group_size(nc);
    for (int k = 0; k < p->count; k++) {
        if (p->update_charge[k]) {
            double ud[18];
            orc_group_adj(nc, p->Ulink + k * G, ud);
            orc_alg_act(nc, p->Q0 + k * A, ud, p->Q1 + k * A);
            p->update_charge[k] = 0;
        } else {
            double t[9];
            memcpy(t, p->Q1 + k * A, sizeof(double) * A);
            memcpy(p->Q1 + k * A, p->Q0 + k * A, sizeof(double) * A);
            memcpy(p->Q0 + k * A, t, sizeof(double) * A);
        }
    }
}

/* physics/grid/CGCParticleInterpolationNGP.java:15-50 */
void orc_p_interpolate_to_grid(orc_grid *gr, const orc_particles *p)
{
    int A = gr->asz;
    for (int k = 0; k < p->count; k++) {
        int d = p->direction[k];
        int no[3], nn[3];
        for (int i = 0; i < 3; i++) {
            no[i] = orc_nearest_grid_point(p->pos0[3 * k + i], gr->as[i]);
            nn[i] = orc_nearest_grid_point(p->pos1[3 * k + i], gr->as[i]);
        }
        if (no[d] == nn[d]) continue;
        int64_t io = cell_index_periodic(gr, no), in = cell_index_periodic(gr, nn);
        if (p->vel[3 * k + d] > 0) {
            double f = gr->as[d] / gr->at;
            double *Jc = GA(gr, gr->J, io, d);
            for (int m = 0; m < A; m++) Jc[m] += p->Q0[k * A + m] * f;
        } else {
            double f = -gr->as[d] / gr->at;
            double *Jc = GA(gr, gr->J, in, d);
            for (int m = 0; m < A; m++) Jc[m] += p->Q1[k * A + m] * f;
        }
    }
}

/* physics/grid/CGCParticleInterpolationNGP.java:52-68 */
void orc_p_interpolate_chargedensity(orc_grid *gr, const orc_particles *p)
{
    int A = gr->asz;
    for (int k = 0; k < p->count; k++) {
        int no[3];
        for (int i = 0; i < 3; i++) no[i] = orc_nearest_grid_point(p->pos0[3 * k + i], gr->as[i]);
        int64_t io = cell_index_periodic(gr, no);
        for (int m = 0; m < A; m++) gr->rho[io * A + m] += p->Q0[k * A + m];
    }
}

/* physics/particles/YangMillsParticle.java:117-125 */
void orc_p_reassign(orc_particles *p)
{
    int A = orc_alg_size(g_p_nc);
    for (int k = 0; k < p->count; k++) {
        for (int i = 0; i < 3; i++) {
            double t = p->pos0[3 * k + i];
            p->pos0[3 * k + i] = p->pos1[3 * k + i];
            p->pos1[3 * k + i] = t;
        }
        for (int m = 0; m < A; m++) {
            double t = p->Q0[k * A + m];
            p->Q0[k * A + m] = p->Q1[k * A + m];
            p->Q1[k * A + m] = t;
        }
    }
}

/* ===================================================================== */
/* Simulation driver                                                      */
/* ===================================================================== */

/* physics/Simulation.java:318-352 */
void orc_sim_initialize(orc_grid *gr, orc_superparticles *sp, orc_particles *p, double at)
{
    g_p_nc = gr->nc;
    orc_step_links(gr, at);
    orc_reset_charge(gr);
    if (sp) orc_sp_interpolate_chargedensity(gr, sp, 0);
    if (p) orc_p_interpolate_chargedensity(gr, p);
    if (p) orc_p_update_positions(p, at);
    if (p) orc_p_interpolate_to_particle(gr, p);
    if (p) orc_p_update_charges(p);
    orc_reset_current(gr);
    if (sp) orc_sp_interpolate_to_grid(gr, sp, 0);
    if (p) orc_p_interpolate_to_grid(gr, p);
}

/* physics/Simulation.java:373-435 (diagnostics and external currents omitted) */
void orc_sim_step(orc_grid *gr, orc_superparticles *sp, orc_particles *p, double at, int *steps_done)
{
    g_p_nc = gr->nc;
    (*steps_done)++;
    int t = *steps_done;
    if (p) orc_p_reassign(p);
    orc_store_fields(gr);
    orc_step(gr, at);
    orc_reset_charge(gr);
    if (sp) orc_sp_interpolate_chargedensity(gr, sp, t);
    if (p) orc_p_interpolate_chargedensity(gr, p);
    if (p) orc_p_update_positions(p, at);
    if (p) orc_p_interpolate_to_particle(gr, p);
    if (p) orc_p_update_charges(p);
    orc_reset_current(gr);
    if (sp) orc_sp_interpolate_to_grid(gr, sp, t);
    if (p) orc_p_interpolate_to_grid(gr, p);
}

/* ===================================================================== */
/* CPU baseline timing                                                    */
/* ===================================================================== */

double orc_time_steps(orc_grid *gr, int steps, int warmup)
{
    for (int s = 0; s < warmup; s++) {
        orc_store_fields(gr);
        orc_step(gr, gr->at);
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int s = 0; s < steps; s++) {
        orc_store_fields(gr);
        orc_step(gr, gr->at);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
