/*
 * opx_oracle_implicit.c -- FP64 CPU restatement of OpenPixi's ImplicitTYMSolver
 * (physics/fields/ImplicitTYMSolver.java:35-263) and of the group logarithm and
 * power it uses for damping (math/SU2GroupElement.java:234-271,
 * math/SU3GroupElement.java:239-520).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device implicit solver. Nothing in
 * the product path links or calls this file.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "opx_oracle.h"

#define GU(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * (gr)->gsz)
#define GA(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * (gr)->asz)

/* ===================================================================== */
/* logarithm and power                                                    */
/* ===================================================================== */

/* SU2GroupElement.getAlgebraElement (math/SU2GroupElement.java:234-251):
 * v = 2 asin|e| / |e| e, or proj = 2 e when |e| < 1e-15 */
static void su2_log(const double *u, double *v)
{
    double norm = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    norm = sqrt(norm);
    if (norm < 1.E-15) {
        v[0] = 2 * u[1]; v[1] = 2 * u[2]; v[2] = 2 * u[3];
        return;
    }
    double f = 2.0 * asin(norm) / norm;
    v[0] = f * u[1]; v[1] = f * u[2]; v[2] = f * u[3];
}

/* hermiticize (math/SU3GroupElement.java:455-472): a (nearly) Hermitian matrix in group
 * storage -> 9-real algebra storage, off-diagonal pairs averaged */
static void su3_hermiticize(const double *m, double *v)
{
    v[0] = m[0];
    v[4] = m[4];
    v[8] = m[8];
    v[1] = (m[1] + m[3]) / 2;
    v[2] = (m[2] + m[6]) / 2;
    v[5] = (m[5] + m[7]) / 2;
    v[3] = (m[12] - m[10]) / 2;
    v[6] = (m[15] - m[11]) / 2;
    v[7] = (m[16] - m[14]) / 2;
}

/* complex 3x3 product in group storage (real parts then imaginary parts, row-major) */
static void cmat_mul(const double *a, const double *b, double *o)
{
    double t[18];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 3; k++) {
                const double ar = a[3 * r + k], ai = a[9 + 3 * r + k];
                const double br = b[3 * k + c], bi = b[9 + 3 * k + c];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            t[3 * r + c] = re;
            t[9 + 3 * r + c] = im;
        }
    memcpy(o, t, sizeof t);
}

/* log(U) by the series of log(1 + X), X = U - 1, n terms (algebraElementTaylorSeries,
 * math/SU3GroupElement.java:425-449): term_m = term_{m-1} (1 - U), term_0 = -1,
 * log U = sum_m term_m / m; the algebra element is -i log U. */
static void su3_log_series(const double *u, int n, double *v)
{
    double term[18] = {0}, mult[18], sum[18] = {0};
    term[0] = term[4] = term[8] = -1.0;
    for (int k = 0; k < 18; k++) mult[k] = -u[k];
    mult[0] += 1.0; mult[4] += 1.0; mult[8] += 1.0;
    for (int m = 1; m <= n; m++) {
        cmat_mul(term, mult, term);
        const double w = 1.0 / m;
        for (int k = 0; k < 18; k++) sum[k] += term[k] * w;
    }
    double h[18];
    for (int k = 0; k < 9; k++) {
        h[k] = sum[k + 9];
        h[k + 9] = -sum[k];
    }
    su3_hermiticize(h, v);
}

/* normalize (math/SU3GroupElement.java:239-253): unit length unless |x| < 1e-12 */
static int normalize6(double *x)
{
    double norm = 0.0;
    for (int k = 0; k < 6; k++) norm += x[k] * x[k];
    norm = sqrt(norm);
    if (fabs(norm) < 1.E-12) return 0;
    for (int k = 0; k < 6; k++) x[k] /= norm;
    return 1;
}

/* SU3GroupElement.getAlgebraElement (math/SU3GroupElement.java:260-453,476-485):
 * H with U = exp(i H), 9-real storage. Eigenvalues from the depressed characteristic
 * cubic (det U = 1 assumed), eigenvectors as a non-vanishing column of
 * (U - l_j)(U - l_k), phases made traceless by shifting the first; the series form
 * is used near the identity (Re tr U >= 3 - 1e-2, 15 terms) and for nearly
 * degenerate eigenvalues (50 terms). */
static void su3_log(const double *u, double *v)
{
    const double trRe = u[0] + u[4] + u[8], trIm = u[9] + u[13] + u[17];
    if (trRe >= 3 - 1.E-2) {
        su3_log_series(u, 15, v);
        return;
    }
    double u2[18];
    cmat_mul(u, u, u2);
    const double sqRe = u2[0] + u2[4] + u2[8], sqIm = u2[9] + u2[13] + u2[17];
    /* l = X + tr/3 with X^3 + p X + q = 0 */
    const double pRe = (trRe * trRe - trIm * trIm) / 6 - sqRe / 2;
    const double pIm = trRe * trIm / 3 - sqIm / 2;
    const double qRe = -1 - (trRe * sqRe - trIm * sqIm) / 6 + 5 * trRe * (trRe * trRe - 3 * trIm * trIm) / 54;
    const double qIm = -(trRe * sqIm + trIm * sqRe) / 6 + 5 * trIm * (3 * trRe * trRe - trIm * trIm) / 54;
    /* W^3 = (-q + sqrt(q^2 + 4 p^3 / 27)) / 2 with the real part of the radicand */
    double disc = qRe * qRe - qIm * qIm + 4 * pRe * (pRe * pRe - 3 * pIm * pIm) / 27;
    if (disc < 0) disc = 0;
    const double rad = sqrt(disc);
    const double wRe = (-qRe + rad) / 2, wIm = -qIm / 2;
    const double th = atan2(wIm, wRe);
    const double r = pow(wRe * wRe + wIm * wIm, 1. / 6);
    double lRe[3], lIm[3];
    for (int i = 0; i < 3; i++) {
        const double t = (th + 2 * M_PI * i) / 3;
        if (fabs(r) == 0) {
            lRe[i] = trRe / 3;
            lIm[i] = trIm / 3;
        } else {
            const double c = cos(t), s = sin(t);
            lRe[i] = r * c - (pRe * c + pIm * s) / (3 * r) + trRe / 3;
            lIm[i] = r * s + (pRe * s - pIm * c) / (3 * r) + trIm / 3;
        }
    }
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 3; k++) {
        const int a = pairs[k][0], b = pairs[k][1];
        if (fabs(1 - lRe[a] / lRe[b]) < 1.E-4 && fabs(1 - lIm[a] / lIm[b]) < 1.E-4) {
            su3_log_series(u, 50, v);
            return;
        }
    }
    /* eigenvectors: x_i = a column of (U - l_j)(U - l_k), [re0, re1, re2, im0, im1, im2] */
    double vec[3][6];
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        double A[18], B[18], M[18];
        memcpy(A, u, sizeof A);
        memcpy(B, u, sizeof B);
        for (int q = 0; q < 3; q++) {
            A[4 * q] -= lRe[j]; A[9 + 4 * q] -= lIm[j];
            B[4 * q] -= lRe[k]; B[9 + 4 * q] -= lIm[k];
        }
        cmat_mul(A, B, M);
        int ok = 0;
        for (int col = 0; col < 3 && !ok; col++) {
            for (int q = 0; q < 3; q++) {
                vec[i][q] = M[3 * q + col];
                vec[i][3 + q] = M[9 + 3 * q + col];
            }
            ok = normalize6(vec[i]);
        }
        if (!ok) {
            for (int q = 0; q < 6; q++) vec[i][q] = 0.0;
            vec[i][i] = 1.0;
        }
    }
    double ph[3];
    for (int i = 0; i < 3; i++) ph[i] = atan2(lIm[i], lRe[i]);
    const double sum = ph[0] + ph[1] + ph[2];
    ph[0] -= sum;
    /* H = sum_i ph_i x_i x_i^dagger = (columns x_i ph_i) (rows x_i^dagger) */
    double L[18], R[18], H[18];
    for (int rr = 0; rr < 3; rr++)
        for (int i = 0; i < 3; i++) {
            L[3 * rr + i] = vec[i][rr] * ph[i];
            L[9 + 3 * rr + i] = vec[i][3 + rr] * ph[i];
            R[3 * i + rr] = vec[i][rr];
            R[9 + 3 * i + rr] = -vec[i][3 + rr];
        }
    cmat_mul(L, R, H);
    su3_hermiticize(H, v);
}

void orc_group_log(int nc, const double *u, double *v)
{
    if (nc == 2) su2_log(u, v);
    else su3_log(u, v);
}

/* GroupElement.pow(x) = getAlgebraElement().mult(x).getLink() */
void orc_group_pow(int nc, const double *u, double x, double *o)
{
    double v[9];
    orc_group_log(nc, u, v);
    const int A = nc == 2 ? 3 : 9;
    for (int k = 0; k < A; k++) v[k] = v[k] * x;
    orc_alg_get_link(nc, v, o);
}

/* ===================================================================== */
/* ImplicitTYMSolver.step                                                 */
/* ===================================================================== */

/* the link of direction dir at `index` at time -1 (Unext: the links before U),
 * 0 (U) or +1 (the implicit estimate iU) (ImplicitStep.getU, ImplicitTYMSolver.java:203-215) */
static const double *link_t(const orc_grid *gr, const double *iU, int64_t index, int dir, int time)
{
    if (time < 0) return GU(gr, gr->Unext, index, dir);
    if (time == 0) return GU(gr, gr->U, index, dir);
    return iU + (index * 3 + dir) * gr->gsz;
}

/* ImplicitStep.getPlaquetteSum (ImplicitTYMSolver.java:148-193), beam direction 0:
 * sum_{i != d} a_i^-2 U_d(x) [U_i(x+d) U_d(x+i)^+ U_i(x)^+ + (U_d(x-i) U_i(x+d-i))^+ U_i(x-i)]
 * with the links of the transverse plane at `time`, or only the beam-direction links. */
static void plaquette_sum(const orc_grid *gr, const double *iU, int64_t index, int d, int time, double *S)
{
    const int nc = gr->nc, G = gr->gsz, beam = 0;
    for (int k = 0; k < G; k++) S[k] = 0.0;
    const int64_t ci1 = orc_shift(gr, index, d, 1);
    for (int i = 0; i < 3; i++) {
        if (i == d) continue;
        int td = 0, ti = 0;
        if (i != beam && d != beam) {
            td = time;
            ti = time;
        } else if (d == beam) {
            td = time;
        } else if (i == beam) {
            ti = time;
        }
        const int64_t ci2 = orc_shift(gr, index, i, 1);
        const int64_t ci3 = orc_shift(gr, ci1, i, -1);
        const int64_t ci4 = orc_shift(gr, index, i, -1);
        double a[18], b[18], U1[18], U2[18];
        orc_group_adj(nc, link_t(gr, iU, ci2, d, td), a);
        orc_group_mult(nc, link_t(gr, iU, ci1, i, ti), a, U1);
        orc_group_adj(nc, link_t(gr, iU, index, i, ti), a);
        orc_group_mult(nc, U1, a, U1);
        orc_group_mult(nc, link_t(gr, iU, ci4, d, td), link_t(gr, iU, ci3, i, ti), b);
        orc_group_adj(nc, b, U2);
        orc_group_mult(nc, U2, link_t(gr, iU, ci4, i, ti), U2);
        const double area = 1.0 / pow(gr->as[i], 2);
        for (int k = 0; k < G; k++) U1[k] += U2[k];
        orc_group_mult(nc, link_t(gr, iU, index, d, td), U1, U1);
        for (int k = 0; k < G; k++) U1[k] *= area;
        for (int k = 0; k < G; k++) S[k] += U1[k];
    }
}

/* ImplicitTYMSolver.step (ImplicitTYMSolver.java:35-58): ImplicitBegin, then
 * `iterations` times (swap, ImplicitStep, ImplicitDamping), then ImplicitEnd. The
 * implicit grid is a full copy (U, Unext, E) of the grid at the start of the step. */
void orc_implicit_step(orc_grid *gr, double at, int iterations, double damping)
{
    const int nc = gr->nc, G = gr->gsz, A = gr->asz;
    const int64_t n = gr->n;
    const size_t lsz = sizeof(double) * (size_t)n * 3 * G, asz = sizeof(double) * (size_t)n * 3 * A;
    double *iU = (double *)malloc(lsz), *iUn = (double *)malloc(lsz), *iE = (double *)malloc(asz);
    memcpy(iU, gr->U, lsz);
    memcpy(iUn, gr->Unext, lsz);
    memcpy(iE, gr->E, asz);
    double uf[3];
    for (int i = 0; i < 3; i++) uf[i] = -(gr->g * gr->as[i]) * gr->at;
    /* ImplicitBegin (:61-80) */
    for (int64_t c = 0; c < n; c++) {
        if (!gr->active[c]) continue;
        for (int i = 0; i < 3; i++) {
            double S[18], temp[18], pr[9], w[9], V[18];
            orc_staple_sum(gr, c, i, S);
            orc_group_mult(nc, GU(gr, gr->U, c, i), S, temp);
            orc_group_proj(nc, temp, pr);
            double *E = iE + (c * 3 + i) * A;
            for (int k = 0; k < A; k++) E[k] += pr[k] * at;
            const double *J = GA(gr, gr->J, c, i);
            for (int k = 0; k < A; k++) E[k] += J[k] * uf[i];
            for (int k = 0; k < A; k++) w[k] = E[k] * (-at);
            orc_alg_get_link(nc, w, V);
            orc_group_mult(nc, V, GU(gr, gr->U, c, i), V);
            memcpy(iUn + (c * 3 + i) * G, V, sizeof(double) * G);
        }
    }
    for (int it = 0; it < iterations; it++) {
        double *t = iU;
        iU = iUn;
        iUn = t;
        /* ImplicitStep (:96-121) */
        for (int64_t c = 0; c < n; c++) {
            if (!gr->active[c]) continue;
            for (int i = 0; i < 3; i++) {
                double S[18], pr[9], w[9], V[18];
                double *E = iE + (c * 3 + i) * A;
                memcpy(E, GA(gr, gr->E, c, i), sizeof(double) * A);
                plaquette_sum(gr, iU, c, i, +1, S);
                orc_group_proj(nc, S, pr);
                for (int k = 0; k < A; k++) E[k] += pr[k] * (at * 0.5);
                plaquette_sum(gr, iU, c, i, -1, S);
                orc_group_proj(nc, S, pr);
                for (int k = 0; k < A; k++) E[k] += pr[k] * (at * 0.5);
                const double *J = GA(gr, gr->J, c, i);
                for (int k = 0; k < A; k++) E[k] += J[k] * uf[i];
                for (int k = 0; k < A; k++) w[k] = E[k] * (-at);
                orc_alg_get_link(nc, w, V);
                orc_group_mult(nc, V, GU(gr, gr->U, c, i), V);
                memcpy(iUn + (c * 3 + i) * G, V, sizeof(double) * G);
            }
        }
        /* ImplicitDamping (:218-231): Unext = U^damping Unext^(1 - damping) */
        for (int64_t c = 0; c < n; c++) {
            if (!gr->active[c]) continue;
            for (int i = 0; i < 3; i++) {
                double U1[18], U2[18];
                orc_group_pow(nc, iUn + (c * 3 + i) * G, 1 - damping, U1);
                orc_group_pow(nc, iU + (c * 3 + i) * G, damping, U2);
                orc_group_mult(nc, U2, U1, iUn + (c * 3 + i) * G);
            }
        }
    }
    /* ImplicitEnd (:236-250) */
    for (int64_t c = 0; c < n; c++) {
        if (!gr->active[c]) continue;
        for (int i = 0; i < 3; i++) {
            memcpy(GA(gr, gr->E, c, i), iE + (c * 3 + i) * A, sizeof(double) * A);
            memcpy(GU(gr, gr->Unext, c, i), iUn + (c * 3 + i) * G, sizeof(double) * G);
        }
    }
    free(iU);
    free(iUn);
    free(iE);
}
