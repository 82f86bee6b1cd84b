/*
 * opx_oracle_gauss.c -- FP64 CPU restatement of OpenPixi's GaussConstraintRestoration and
 * UnitarityTester diagnostics.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device implementation; the product library
 * never links or calls it.
 *
 * Restated from pixi/src/main/java/org/openpixi/pixi/diagnostics/methods/
 *   GaussConstraintRestoration.java:79-163 (iterateRestorationAlgorithm, computeGaussViolation,
 *   applyCorrection, backupGrid/restoreGrid) and UnitarityTester.java:31-46, in the reference's
 *   matrix form: the violation gv is the full matrix sum C * 1/(a_t a) and every projection is
 *   taken of the matrix expression, as the reference does.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "opx_oracle.h"

#define GUX(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * (gr)->gsz)
#define GAX(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * (gr)->asz)

/* componentwise linear algebra of group storage (both storages are linear in the matrix) */
static void gx_axpy(int gsz, double a, const double *x, double *y)
{
    for (int k = 0; k < gsz; k++) y[k] += a * x[k];
}

/* GaussConstraintRestoration.java:117-135 */
double orc_restore_violation(const orc_grid *gr, double *gv)
{
    const int nc = gr->nc, G = gr->gsz, A = gr->asz;
    const double factor = 1.0 / (gr->at * gr->as[0]);
    double total = 0.0;
    for (int64_t i = 0; i < gr->n; i++) {
        double C[18] = {0}, C1[18], C2[18], t[18], v[9];
        for (int j = 0; j < 3; j++) {
            const int64_t k = orc_shift(gr, i, j, -1);
            orc_group_adj(nc, GUX(gr, gr->U, i, j), t);
            orc_group_mult(nc, GUX(gr, gr->Unext, i, j), t, C1);          /* Unext(i,j) U(i,j)^+ */
            orc_group_adj(nc, GUX(gr, gr->U, k, j), t);
            orc_group_mult(nc, t, GUX(gr, gr->Unext, k, j), C2);          /* U(k,j)^+ Unext(k,j) */
            gx_axpy(G, 1.0, C1, C);
            gx_axpy(G, -1.0, C2, C);
        }
        double *g = gv + i * G;
        for (int m = 0; m < G; m++) g[m] = C[m] * factor;
        orc_group_proj(nc, g, v);
        for (int m = 0; m < A; m++) v[m] -= gr->rho[i * A + m];
        total += orc_alg_square(nc, v);
    }
    return total / (double)gr->n;
}

/* GaussConstraintRestoration.java:142-158: every cell and direction, then updateLinks */
void orc_restore_correct(orc_grid *gr, const double *gv, double gamma)
{
    const int nc = gr->nc, G = gr->gsz, A = gr->asz;
    const double factor2 = -0.5 * gr->as[0] * gamma;
    for (int64_t i = 0; i < gr->n; i++) {
        for (int j = 0; j < 3; j++) {
            double t[18], C1[18], ua[18], v[9];
            const int64_t k = orc_shift(gr, i, j, 1);
            const double *u = GUX(gr, gr->U, i, j);
            orc_group_mult(nc, u, gv + k * G, t);
            orc_group_adj(nc, u, ua);
            orc_group_mult(nc, t, ua, C1);                                 /* U gv(k) U^+ */
            gx_axpy(G, -1.0, gv + i * G, C1);                              /* - gv(i) */
            orc_group_proj(nc, C1, v);
            double *e = GAX(gr, gr->E, i, j);
            for (int m = 0; m < A; m++) e[m] += v[m] * factor2;
        }
    }
    orc_step_links(gr, gr->at);
}

/* GaussConstraintRestoration.java:79-112 */
int orc_gauss_restore(orc_grid *gr, double gamma, int max_iterations, double absolute_value, int *reason,
                      double *before, double *after, double *history)
{
    if (gr->as[0] != gr->as[1] || gr->as[0] != gr->as[2]) return -1;   /* Grid.getLatticeSpacing throws */
    const size_t gsize = (size_t)gr->n * gr->gsz, esize = (size_t)gr->n * 3 * gr->asz;
    const size_t lsize = (size_t)gr->n * 3 * gr->gsz;
    double *gv = (double *)malloc(sizeof(double) * gsize);
    double *E0 = (double *)malloc(sizeof(double) * esize);
    double *Un0 = (double *)malloc(sizeof(double) * lsize);
    double total = orc_restore_violation(gr, gv);
    double old = total;
    if (before) *before = total;
    if (history) history[0] = total;
    int done = 0, why = 3;
    for (int i = 0; i < max_iterations; i++) {
        if (total < absolute_value) {
            why = 0;
            break;
        }
        memcpy(E0, gr->E, sizeof(double) * esize);          /* backupGrid: E and Unext change */
        memcpy(Un0, gr->Unext, sizeof(double) * lsize);
        orc_restore_correct(gr, gv, gamma);
        total = orc_restore_violation(gr, gv);
        if (old - total < 0) {
            memcpy(gr->E, E0, sizeof(double) * esize);      /* restoreGrid */
            memcpy(gr->Unext, Un0, sizeof(double) * lsize);
            why = 1;
            break;
        }
        old = total;
        done = i + 1;
        if (history) history[done] = total;
        if (i == max_iterations - 1) why = 2;
    }
    if (after) *after = old;
    if (reason) *reason = why;
    free(gv);
    free(E0);
    free(Un0);
    return done;
}

/* UnitarityTester.java:31-46 with getRealTrace; flags[i*3 + j] = 1 where U or Unext fails */
int64_t orc_unitarity_flags(const orc_grid *gr, double tol, unsigned char *flags)
{
    int64_t count = 0;
    for (int64_t i = 0; i < gr->n; i++) {
        for (int j = 0; j < 3; j++) {
            int bad = 0;
            for (int t = 0; t < 2; t++) {
                const double *u = GUX(gr, t ? gr->Unext : gr->U, i, j);
                const double tr = gr->nc == 2 ? 2.0 * u[0] : u[0] + u[4] + u[8];
                if (!(fabs(tr) < gr->nc + tol)) bad = 1;
            }
            flags[i * 3 + j] = (unsigned char)bad;
            count += bad;
        }
    }
    return count;
}
