/*
 * opx_oracle_adjoint.c -- FP64 CPU adjoint (reverse-mode derivative) of the
 * SU(2) temporal-gauge leapfrog in opx_oracle.c.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). The reference has no differentiation;
 * this file differentiates the forward map that the reference defines
 * (FastTYMSolver.java:40-84 combined update and link step; the measurement of
 * ProjectedEnergyDensity.java:132-160) and nothing here follows reference code.
 *
 * Forward map (pure Yang-Mills, every cell active, J = 0), with E0 the initial
 * electric field and U_init the initial links:
 *   U^0     = exp(-at E0) U_init                                   (stepLinks)
 *   E^{n+1} = E^n + at F(U^n),  F_l(U) = proj(U_l S_l(U))          (fused step)
 *   U^{n+1} = exp(-at E^{n+1}) U^n
 * Loss after N steps, the lattice average of the transverse-minus-longitudinal
 * energy density for longitudinal direction `dir`:
 *   L = (1/n) sum_x sum_j s_j (g a_j)^-2 [ |E^N_j|^2 / 2 + (|B_j(U^{N-1})|^2 + |B_j(U^N)|^2) / 4 ]
 * with s_j = -1 for j == dir and +1 otherwise.
 *
 * Every function treats an SU(2) element as a real 4-vector (e0, e1, e2, e3)
 * and returns the transposed Jacobian of the real map; the inner product is the
 * Euclidean one on R^4 (= Re tr(A^dagger B) / 2 for the 2x2 matrices), so for
 * a product r = x y the cotangents are x_bar = r_bar y^dagger and
 * y_bar = x^dagger r_bar, with ^dagger the quaternion conjugate.
 *
 * Each plaquette p = (x, mu < nu) with loop factors a = U_mu(x), b = U_nu(x+mu),
 * c = U_mu(x+nu)^dagger, d = U_nu(x)^dagger enters the force of its four links
 * through three cyclic rotations of the loop:
 *   F_mu(x)     +=  w_nu proj(abcd)      F_nu(x)    += -w_mu proj(abcd)
 *   F_nu(x+mu)  +=  w_mu proj(bcda)      F_mu(x+nu) += -w_nu proj(dabc)
 * (w_k = a_k^-2), so <Ebar, F> = sum_p <C0, proj abcd> + <C1, proj bcda> + <C3, proj dabc>
 * with C0 = w_nu Ebar_a - w_mu Ebar_d, C1 = w_mu Ebar_b, C3 = -w_nu Ebar_c.
 */
#include "opx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GU(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * 4)
#define GA(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * 3)

static void qmul(const double *a, const double *b, double *o)
{
    double r[4];
    orc_group_mult(2, a, b, r);
    memcpy(o, r, sizeof r);
}

static void qconj(const double *a, double *o)
{
    o[0] = a[0];
    o[1] = -a[1];
    o[2] = -a[2];
    o[3] = -a[3];
}

static void qmul3(const double *a, const double *b, const double *c, double *o)
{
    double t[4];
    qmul(a, b, t);
    qmul(t, c, o);
}

/* transpose of proj (proj(q) = 2 (q1, q2, q3)) */
static void lift(const double *C, double *q)
{
    q[0] = 0.0;
    q[1] = 2.0 * C[0];
    q[2] = 2.0 * C[1];
    q[3] = 2.0 * C[2];
}

/* exp(w) = (cos(r/2), w sin(r/2)/r), r = |w| (SU2AlgebraElement.getLink); wbar = J^T Mbar */
void orc_adj_exp_vjp(const double *w, const double *Mbar, double *wbar)
{
    double r2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    double r = sqrt(r2);
    double f, fp_over_r, s_half;   /* f = sin(r/2)/r, (df/dr)/r */
    if (r < 0.1) {   /* Taylor series of sin(r/2)/r; the closed form cancels for small r */
        f = 0.5 - r2 / 48.0 + r2 * r2 / 3840.0 - r2 * r2 * r2 / 645120.0;
        fp_over_r = -1.0 / 24.0 + r2 / 960.0 - r2 * r2 / 107520.0 + r2 * r2 * r2 / 23224320.0;
    } else {
        f = sin(0.5 * r) / r;
        fp_over_r = (0.5 * r * cos(0.5 * r) - sin(0.5 * r)) / (r2 * r);
    }
    s_half = f; /* d cos(r/2) / dw = -(1/2) sin(r/2) w / r = -(f/2) w */
    double dot = w[0] * Mbar[1] + w[1] * Mbar[2] + w[2] * Mbar[3];
    for (int k = 0; k < 3; k++) wbar[k] = -0.5 * s_half * w[k] * Mbar[0] + f * Mbar[k + 1] + fp_over_r * dot * w[k];
}

/* cotangents of the four loop factors for sum_k <C_k, proj R_k>, R0 = abcd, R1 = bcda, R3 = dabc */
static void plaquette_cotangents(const double *a, const double *b, const double *c, const double *d, const double *C0,
                                 const double *C1, const double *C3, double *ga, double *gb, double *gc, double *gd)
{
    double c0[4], c1[4], c3[4];
    lift(C0, c0);
    lift(C1, c1);
    lift(C3, c3);
    double ad[4], bd[4], cd_[4], dd[4];
    qconj(a, ad);
    qconj(b, bd);
    qconj(c, cd_);
    qconj(d, dd);
    double t1[4], t2[4], t3[4], x[4];
    /* factor a: c0 (bcd)^+ + (bcd)^+ c1 + d^+ c3 (bc)^+ */
    double bcd_d[4];
    qmul3(dd, cd_, bd, bcd_d);
    qmul(c0, bcd_d, t1);
    qmul(bcd_d, c1, t2);
    qmul(cd_, bd, x); /* (bc)^+ */
    qmul3(dd, c3, x, t3);
    for (int k = 0; k < 4; k++) ga[k] = t1[k] + t2[k] + t3[k];
    /* factor b: a^+ c0 (cd)^+ + c1 (cda)^+ + (da)^+ c3 c^+ */
    qmul(dd, cd_, x); /* (cd)^+ */
    qmul3(ad, c0, x, t1);
    double cda_d[4];
    qmul3(ad, dd, cd_, cda_d);
    qmul(c1, cda_d, t2);
    qmul(ad, dd, x); /* (da)^+ */
    qmul3(x, c3, cd_, t3);
    for (int k = 0; k < 4; k++) gb[k] = t1[k] + t2[k] + t3[k];
    /* factor c: (ab)^+ c0 d^+ + b^+ c1 (da)^+ + (dab)^+ c3 */
    qmul(bd, ad, x); /* (ab)^+ */
    qmul3(x, c0, dd, t1);
    qmul(ad, dd, x); /* (da)^+ */
    qmul3(bd, c1, x, t2);
    double dab_d[4];
    qmul3(bd, ad, dd, dab_d);
    qmul(dab_d, c3, t3);
    for (int k = 0; k < 4; k++) gc[k] = t1[k] + t2[k] + t3[k];
    /* factor d: (abc)^+ c0 + (bc)^+ c1 a^+ + c3 (abc)^+ */
    double abc_d[4];
    qmul3(cd_, bd, ad, abc_d);
    qmul(abc_d, c0, t1);
    qmul(cd_, bd, x); /* (bc)^+ */
    qmul3(x, c1, ad, t2);
    qmul(c3, abc_d, t3);
    for (int k = 0; k < 4; k++) gd[k] = t1[k] + t2[k] + t3[k];
}

/* Ubar += scale * (dF/dU)^T Ebar, scattered plaquette by plaquette */
void orc_adj_force_vjp(const orc_grid *gr, const double *U, const double *Ebar, double scale, double *Ubar)
{
    for (int64_t x = 0; x < gr->n; x++) {
        for (int mu = 0; mu < 3; mu++) {
            for (int nu = mu + 1; nu < 3; nu++) {
                int64_t xm = orc_shift(gr, x, mu, 1), xn = orc_shift(gr, x, nu, 1);
                double wmu = 1.0 / (gr->as[mu] * gr->as[mu]), wnu = 1.0 / (gr->as[nu] * gr->as[nu]);
                double a[4], b[4], c[4], d[4];
                memcpy(a, GU(gr, U, x, mu), sizeof a);
                memcpy(b, GU(gr, U, xm, nu), sizeof b);
                qconj(GU(gr, U, xn, mu), c);
                qconj(GU(gr, U, x, nu), d);
                const double *Ea = GA(gr, Ebar, x, mu), *Eb = GA(gr, Ebar, xm, nu);
                const double *Ec = GA(gr, Ebar, xn, mu), *Ed = GA(gr, Ebar, x, nu);
                double C0[3], C1[3], C3[3];
                for (int k = 0; k < 3; k++) {
                    C0[k] = wnu * Ea[k] - wmu * Ed[k];
                    C1[k] = wmu * Eb[k];
                    C3[k] = -wnu * Ec[k];
                }
                double ga[4], gb[4], gc[4], gd[4];
                plaquette_cotangents(a, b, c, d, C0, C1, C3, ga, gb, gc, gd);
                double gcc[4], gdc[4];
                qconj(gc, gcc);
                qconj(gd, gdc);
                for (int k = 0; k < 4; k++) {
                    GU(gr, Ubar, x, mu)[k] += scale * ga[k];
                    GU(gr, Ubar, xm, nu)[k] += scale * gb[k];
                    GU(gr, Ubar, xn, mu)[k] += scale * gcc[k];
                    GU(gr, Ubar, x, nu)[k] += scale * gdc[k];
                }
            }
        }
    }
}

/* the magnetic part of L for one link field: sum_x sum_j s_j (g a_j)^-2 / (4n) |B_j(U)(x)|^2;
 * B_j = proj(plaquette in the plane (j+1, j+2)) a_j / (a_{j+1} a_{j+2}) (Grid.java:877-900).
 * Returns the value; adds its gradient to Ubar when Ubar is not NULL. */
double orc_adj_bsq(const orc_grid *gr, const double *U, int dir, double *Ubar)
{
    double total = 0.0;
    for (int64_t x = 0; x < gr->n; x++) {
        for (int mu = 0; mu < 3; mu++) {
            for (int nu = mu + 1; nu < 3; nu++) {
                int j = 3 - mu - nu;
                double s = j == dir ? -1.0 : 1.0;
                double fj = gr->as[j] / (gr->as[mu] * gr->as[nu]);
                double coef = s * fj * fj / (4.0 * (double)gr->n * gr->g * gr->g * gr->as[j] * gr->as[j]);
                int64_t xm = orc_shift(gr, x, mu, 1), xn = orc_shift(gr, x, nu, 1);
                double a[4], b[4], c[4], d[4], P[4], pr[3];
                memcpy(a, GU(gr, U, x, mu), sizeof a);
                memcpy(b, GU(gr, U, xm, nu), sizeof b);
                qconj(GU(gr, U, xn, mu), c);
                qconj(GU(gr, U, x, nu), d);
                qmul3(a, b, c, P);
                qmul(P, d, P);
                orc_group_proj(2, P, pr);
                total += coef * (pr[0] * pr[0] + pr[1] * pr[1] + pr[2] * pr[2]);
                if (!Ubar) continue;
                /* d/dR0 of coef |proj R0|^2 = lift(2 coef proj R0) */
                double C0[3] = {2 * coef * pr[0], 2 * coef * pr[1], 2 * coef * pr[2]}, Z[3] = {0, 0, 0};
                double ga[4], gb[4], gc[4], gd[4], gcc[4], gdc[4];
                plaquette_cotangents(a, b, c, d, C0, Z, Z, ga, gb, gc, gd);
                qconj(gc, gcc);
                qconj(gd, gdc);
                for (int k = 0; k < 4; k++) {
                    GU(gr, Ubar, x, mu)[k] += ga[k];
                    GU(gr, Ubar, xm, nu)[k] += gb[k];
                    GU(gr, Ubar, xn, mu)[k] += gcc[k];
                    GU(gr, Ubar, x, nu)[k] += gdc[k];
                }
            }
        }
    }
    return total;
}

/* L for (E^N, U^{N-1}, U^N) */
double orc_adj_loss(const orc_grid *gr, const double *Uprev, const double *Ulast, const double *E, int dir)
{
    double el = 0.0;
    for (int64_t x = 0; x < gr->n; x++)
        for (int j = 0; j < 3; j++) {
            const double *e = GA(gr, E, x, j);
            double s = j == dir ? -1.0 : 1.0;
            el += s * 0.5 * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) / (gr->g * gr->g * gr->as[j] * gr->as[j]);
        }
    el /= (double)gr->n;
    return el + orc_adj_bsq(gr, Uprev, dir, NULL) + orc_adj_bsq(gr, Ulast, dir, NULL);
}

/* one link: Ubar <- M^dagger Ubar, Ebar += -at J_exp^T (Ubar U^dagger), M = exp(-at E) */
static void link_step_vjp(const double *E, const double *U, double at, double *Ubar, double *Ebar)
{
    double w[3] = {-at * E[0], -at * E[1], -at * E[2]}, M[4], Md[4], Ud[4], Mbar[4], wbar[3], t[4];
    orc_alg_get_link(2, w, M);
    qconj(U, Ud);
    qmul(Ubar, Ud, Mbar);
    orc_adj_exp_vjp(w, Mbar, wbar);
    for (int k = 0; k < 3; k++) Ebar[k] += -at * wbar[k];
    qconj(M, Md);
    qmul(Md, Ubar, t);
    memcpy(Ubar, t, sizeof t);
}

/* Gradient of L after `steps` steps with respect to the initial E (gr->E), the
 * initial links being gr->U. Stores every state. On return gr holds the final
 * forward state (U = U^{N-1}, Unext = U^N, E = E^N); grad has n*3*3 values. */
int orc_adjoint_gradient(orc_grid *gr, int steps, int dir, double *loss, double *grad)
{
    if (gr->nc != 2 || steps < 1) return -1;
    const size_t nl = (size_t)gr->n * 3, ub = nl * 4 * sizeof(double), eb = nl * 3 * sizeof(double);
    double *Uinit = malloc(ub), *Uh = malloc(ub * (size_t)(steps + 1)), *Eh = malloc(eb * (size_t)(steps + 1));
    double *Ubar = calloc(nl * 4, sizeof(double)), *Ebar = calloc(nl * 3, sizeof(double));
    if (!Uinit || !Uh || !Eh || !Ubar || !Ebar) {
        free(Uinit); free(Uh); free(Eh); free(Ubar); free(Ebar);
        return -2;
    }
    const double at = gr->at;
    memcpy(Uinit, gr->U, ub);
    memcpy(Eh, gr->E, eb);                 /* E^0 */
    orc_step_links(gr, at);
    memcpy(Uh, gr->Unext, ub);             /* U^0 */
    for (int n = 0; n < steps; n++) {
        orc_store_fields(gr);
        orc_step(gr, at);
        memcpy(Uh + (size_t)(n + 1) * nl * 4, gr->Unext, ub);
        memcpy(Eh + (size_t)(n + 1) * nl * 3, gr->E, eb);
    }
    const double *UN = Uh + (size_t)steps * nl * 4, *UN1 = Uh + (size_t)(steps - 1) * nl * 4;
    const double *EN = Eh + (size_t)steps * nl * 3;
    *loss = orc_adj_loss(gr, UN1, UN, EN, dir);
    for (int64_t x = 0; x < gr->n; x++)
        for (int j = 0; j < 3; j++) {
            double s = j == dir ? -1.0 : 1.0;
            double f = s / ((double)gr->n * gr->g * gr->g * gr->as[j] * gr->as[j]);
            for (int k = 0; k < 3; k++) GA(gr, Ebar, x, j)[k] = f * GA(gr, EN, x, j)[k];
        }
    orc_adj_bsq(gr, UN, dir, Ubar);
    for (int n = steps - 1; n >= 0; n--) {
        const double *Un = Uh + (size_t)n * nl * 4, *En1 = Eh + (size_t)(n + 1) * nl * 3;
        for (size_t l = 0; l < nl; l++) link_step_vjp(En1 + l * 3, Un + l * 4, at, Ubar + l * 4, Ebar + l * 3);
        orc_adj_force_vjp(gr, Un, Ebar, at, Ubar);
        if (n == steps - 1) orc_adj_bsq(gr, Un, dir, Ubar);
    }
    for (size_t l = 0; l < nl; l++) link_step_vjp(Eh + l * 3, Uinit + l * 4, at, Ubar + l * 4, Ebar + l * 3);
    memcpy(grad, Ebar, eb);
    free(Uinit); free(Uh); free(Eh); free(Ubar); free(Ebar);
    return 0;
}
