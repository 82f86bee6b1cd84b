/*
 * opx_oracle_cgc.c -- FP64 CPU restatement of OpenPixi's CGC initial condition for the
 * optimized CPIC simulation types (physics/initial/CGC/CGCInitialCondition.java:55-93):
 * the MV-model charge density (MVModel.java:101-157) with its Fourier-space regulation
 * (FourierFunctions.java:86-128), the light-cone Poisson solver with path-ordered Wilson
 * lines (LightConePoissonSolverImprovedFull.java:40-291) and the super-particle creator with
 * charge refinement (LightConeNGPSuperParticleCreator.java:76-307).
 *
 * The beam direction is x0: the super-particle creator indexes cells z-major and shifts
 * along x0. The reference draws its charges from java.util.Random and transforms them with
 * JTransforms; here the unit Gaussian noise is an input and the discrete Fourier transforms
 * are this file's own (radix 2, or direct sums for other lengths).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device CGC initialization. Nothing in the
 * product path links or calls this file.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "opx_oracle.h"

#define GU(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * (gr)->gsz)
#define GA(gr, buf, idx, d) ((buf) + ((idx) * 3 + (d)) * (gr)->asz)

/* ===================================================================== */
/* discrete Fourier transforms                                            */
/* ===================================================================== */

/* X_k = sum_n x_n exp(sign 2 pi i k n / N) along one axis of a strided complex array */
static void dft_axis(double *re, double *im, int n, long stride, int sign)
{
    double *tr = (double *)malloc(sizeof(double) * n), *ti = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; k++) {
        tr[k] = re[k * stride];
        ti[k] = im[k * stride];
    }
    if ((n & (n - 1)) == 0) {
        /* iterative radix-2: bit reversal, then butterflies */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t = tr[i]; tr[i] = tr[j]; tr[j] = t;
                t = ti[i]; ti[i] = ti[j]; ti[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            const double ang = sign * 2 * M_PI / len;
            for (int i = 0; i < n; i += len)
                for (int j = 0; j < len / 2; j++) {
                    const double wr = cos(ang * j), wi = sin(ang * j);
                    const double ur = tr[i + j], ui = ti[i + j];
                    const double vr = tr[i + j + len / 2] * wr - ti[i + j + len / 2] * wi;
                    const double vi = tr[i + j + len / 2] * wi + ti[i + j + len / 2] * wr;
                    tr[i + j] = ur + vr; ti[i + j] = ui + vi;
                    tr[i + j + len / 2] = ur - vr; ti[i + j + len / 2] = ui - vi;
                }
        }
        for (int k = 0; k < n; k++) {
            re[k * stride] = tr[k];
            im[k * stride] = ti[k];
        }
    } else {
        for (int k = 0; k < n; k++) {
            double sr = 0, si = 0;
            for (int m = 0; m < n; m++) {
                const double a = sign * 2 * M_PI * (double)((long)k * m % n) / n;
                sr += tr[m] * cos(a) - ti[m] * sin(a);
                si += tr[m] * sin(a) + ti[m] * cos(a);
            }
            re[k * stride] = sr;
            im[k * stride] = si;
        }
    }
    free(tr);
    free(ti);
}

/* d-dimensional transform of a row-major complex array; inverse scaled by 1/N */
static void dft_nd(double *re, double *im, const int *dims, int nd, int sign)
{
    long total = 1;
    for (int a = 0; a < nd; a++) total *= dims[a];
    for (int a = 0; a < nd; a++) {
        long stride = 1;
        for (int b = a + 1; b < nd; b++) stride *= dims[b];
        const long outer = total / (stride * dims[a]);
        for (long o = 0; o < outer; o++)
            for (long s = 0; s < stride; s++) {
                const long base = o * stride * dims[a] + s;
                dft_axis(re + base, im + base, dims[a], stride, sign);
            }
    }
    if (sign > 0)
        for (long k = 0; k < total; k++) {
            re[k] /= total;
            im[k] /= total;
        }
}

/* FourierFunctions.computeEffectiveTransverseMomentumSquared / computeTransverseMomentumSquared /
 * computeLatticeMomentum1D (FourierFunctions.java:182-235), for the transverse plane (x1, x2) */
static double kt_eff2(int p1, int p2, int n1, int n2, double aT)
{
    double m = 2.0 * 2;
    m -= 2.0 * cos((2.0 * M_PI * p1) / n1);
    m -= 2.0 * cos((2.0 * M_PI * p2) / n2);
    return m / (aT * aT);
}
static double kt2(int p1, int p2, int n1, int n2, double aT)
{
    const double twopi = 2.0 * M_PI;
    double s = 0.0;
    const int p[2] = {p1, p2}, n[2] = {n1, n2};
    for (int i = 0; i < 2; i++) {
        const double c = p[i] < n[i] / 2 ? twopi * p[i] / (aT * n[i]) : twopi * (n[i] - p[i]) / (aT * n[i]);
        s += c * c;
    }
    return s;
}
static double k_long(int pos, int n, double aL)
{
    const double delta = pos / ((double)n);
    if (delta < 0.5) return 2.0 * delta * M_PI / aL;
    return 2.0 * (delta - 1.0) * M_PI / aL;
}

/* FourierFunctions.regulateChargeDensityGaussian (FourierFunctions.java:86-128), beam along x0 */
void orc_regulate_charge_gaussian(const int dims[3], double uvt, double long_width, double ir, double aT,
                                  double aL, double *rho)
{
    const long n = (long)dims[0] * dims[1] * dims[2];
    double *re = (double *)malloc(sizeof(double) * n), *im = (double *)calloc(n, sizeof(double));
    memcpy(re, rho, sizeof(double) * n);
    dft_nd(re, im, dims, 3, -1);
    re[0] = 0.0;
    im[0] = 0.0;
    for (long c = 0; c < n; c++) {
        const int z = (int)(c / ((long)dims[1] * dims[2]));
        const int p1 = (int)((c / dims[2]) % dims[1]), p2 = (int)(c % dims[2]);
        const double kte = kt_eff2(p1, p2, dims[1], dims[2], aT), kt = kt2(p1, p2, dims[1], dims[2], aT);
        const double kl = fabs(k_long(z, dims[0], aL));
        if (kt <= uvt * uvt && kt > 0) {
            const double reg = kte / (kte + ir * ir);
            const double lreg = exp(-0.25 * kl * kl * long_width * long_width);
            re[c] *= reg * lreg;
            im[c] *= reg * lreg;
        } else {
            re[c] = 0.0;
            im[c] = 0.0;
        }
    }
    dft_nd(re, im, dims, 3, +1);
    memcpy(rho, re, sizeof(double) * n);
    free(re);
    free(im);
}

/* FourierFunctions.solvePoisson2D (FourierFunctions.java:137-160) for one x0 sheet */
void orc_poisson_2d(int n1, int n2, double aT, const double *rho, double *phi)
{
    const long n = (long)n1 * n2;
    double *re = (double *)malloc(sizeof(double) * n), *im = (double *)calloc(n, sizeof(double));
    memcpy(re, rho, sizeof(double) * n);
    const int dims[2] = {n1, n2};
    dft_nd(re, im, dims, 2, -1);
    re[0] = 0.0;
    im[0] = 0.0;
    for (long c = 1; c < n; c++) {
        const double k = kt_eff2((int)(c / n2), (int)(c % n2), n1, n2, aT);
        re[c] *= 1.0 / k;
        im[c] *= 1.0 / k;
    }
    dft_nd(re, im, dims, 2, +1);
    memcpy(phi, re, sizeof(double) * n);
    free(re);
    free(im);
}

/* ===================================================================== */
/* MV model                                                               */
/* ===================================================================== */

/* MVModel.initialize (MVModel.java:101-157): noise[j * n + cell] is the unit Gaussian drawn for
 * component j (components outer, cells inner, as the reference draws them); rho receives
 * algebra storage per cell (zero outside the box around `location` and outside the active region) */
void orc_mv_charge(const orc_grid *gr, double location, double width, double mu, double uvt, double long_width,
                   double ir, const double *noise, double *rho)
{
    const int nc = gr->nc, A = gr->asz, ncomp = nc * nc - 1;
    const long n = gr->n;
    const double aL = gr->as[0], aT = gr->as[1];
    memset(rho, 0, sizeof(double) * n * A);
    double *tmp = (double *)malloc(sizeof(double) * n);
    /* commons-math Gaussian(location, width).value: norm exp(-(x - mean)^2 / (2 sigma^2)) */
    const double norm = 1.0 / (width * sqrt(2 * M_PI)), i2s2 = 0.5 / (width * width);
    const double rcw = mu * gr->g / sqrt(aL * aT * aT);
    const double box = gr->dims[0] * aL;
    const double zmin = fmax(location - box / 2.0, 0.0), zmax = fmin(location + box / 2.0, box);
    const int lmin = (int)floor(zmin / aL), lmax = (int)ceil(zmax / aL);
    for (int j = 0; j < ncomp; j++) {
        for (long c = 0; c < n; c++) {
            const int z = (int)(c / ((long)gr->dims[1] * gr->dims[2]));
            const double xm = z * aL - location;
            const double profile = sqrt(norm * exp(-xm * xm * i2s2));
            tmp[c] = noise[j * n + c] * rcw * profile;
        }
        orc_regulate_charge_gaussian(gr->dims, uvt, long_width, ir, aT, aL, tmp);
        for (long c = 0; c < n; c++) {
            const int z = (int)(c / ((long)gr->dims[1] * gr->dims[2]));
            if (lmin < z && z < lmax && gr->active[c]) orc_alg_set(nc, rho + c * A, j, tmp[c]);
        }
    }
    free(tmp);
}

/* ===================================================================== */
/* light-cone Poisson solver                                              */
/* ===================================================================== */

/* W(z1, z2, M, P1, P2) (LightConePoissonSolverImprovedFull.java:271-291): the path-ordered
 * product over M midpoints of the linear interpolation between P1 and P2 */
void orc_wilson_segment(const orc_grid *gr, double z1, double z2, int M, const double *P1, const double *P2, double *W)
{
    const int nc = gr->nc, A = gr->asz, G = gr->gsz;
    double id[18] = {0};
    if (nc == 2) id[0] = 1.0;
    else id[0] = id[4] = id[8] = 1.0;
    memcpy(W, id, sizeof(double) * G);
    const double dz = fabs(z2 - z1) / ((double)M) * gr->as[0];
    const double gf = -gr->g * dz;
    const double ui = floor((z1 + z2) * 0.5 + 0.5);   /* Math.round */
    const double u1 = z1 - ui, u2 = z2 - ui;
    double PM[9], PD[9], P[9], L[18];
    for (int k = 0; k < A; k++) {
        PM[k] = (P1[k] + P2[k]) * 0.5;
        PD[k] = P2[k] - P1[k];
    }
    for (int m = 0; m < M; m++) {
        const double u = u1 + (u2 - u1) * (m + 0.5) / ((double)M);
        for (int k = 0; k < A; k++) P[k] = (PM[k] + PD[k] * u) * gf;
        orc_alg_get_link(nc, P, L);
        orc_group_mult(nc, W, L, W);
    }
}

/* LightConePoissonSolverImprovedFull.solve (:40-263), beam along x0 */
void orc_lightcone_solve(orc_grid *gr, const double *rho, int orientation, double *gauss)
{
    orc_lightcone_solve_vt(gr, rho, orientation, gauss, NULL);
}

void orc_lightcone_solve_vt(orc_grid *gr, const double *rho, int orientation, double *gauss, double *VT)
{
    const int nc = gr->nc, A = gr->asz, G = gr->gsz, ncomp = nc * nc - 1;
    const int L = gr->dims[0], n1 = gr->dims[1], n2 = gr->dims[2];
    const long T = (long)n1 * n2, n = gr->n;
    const double aL = gr->as[0], aT = gr->as[1];
    /* phi at t = -at/2, sheet by sheet */
    double *phi = (double *)calloc(n * A, sizeof(double));
    double *r2 = (double *)malloc(sizeof(double) * T), *p2 = (double *)malloc(sizeof(double) * T);
    for (int z = 0; z < L; z++)
        for (int c = 0; c < ncomp; c++) {
            for (long i = 0; i < T; i++) r2[i] = orc_alg_get(nc, rho + (z * T + i) * A, c);
            orc_poisson_2d(n1, n2, aT, r2, p2);
            for (long i = 0; i < T; i++) orc_alg_set(nc, phi + (z * T + i) * A, c, p2[i]);
        }
    free(r2);
    free(p2);
    /* Wilson lines V at t = -at/2, path ordered along the beam from the far side */
    double *V = (double *)malloc(sizeof(double) * n * G), *Vc = (double *)malloc(sizeof(double) * n * G);
    for (long c = 0; c < n; c++) {
        double *v = V + c * G;
        memset(v, 0, sizeof(double) * G);
        if (nc == 2) v[0] = 1.0;
        else v[0] = v[4] = v[8] = 1.0;
    }
    const int M = 32;
    for (int k = 1; k < L; k++) {
        const int z = orientation < 0 ? k : L - k - 1;
        for (long i = 0; i < T; i++) {
            const long index = z * T + i;
            const long indexL = orc_shift(gr, index, 0, orientation);
            double W[18], W2[18];
            if (orientation == -1) {
                const long i3 = index, i2 = orc_shift(gr, i3, 0, -1), i1 = orc_shift(gr, i2, 0, -1);
                orc_wilson_segment(gr, z - 1, z - 0.5, M / 2, phi + i1 * A, phi + i2 * A, W);
                orc_wilson_segment(gr, z - 0.5, z, M / 2, phi + i2 * A, phi + i3 * A, W2);
            } else {
                const long i2 = index, i1 = orc_shift(gr, i2, 0, -1), i3 = orc_shift(gr, i2, 0, +1);
                orc_wilson_segment(gr, z + 1, z + 0.5, M / 2, phi + i2 * A, phi + i3 * A, W);
                orc_wilson_segment(gr, z + 0.5, z, M / 2, phi + i1 * A, phi + i2 * A, W2);
            }
            orc_group_mult(nc, W, W2, W);
            orc_group_mult(nc, V + indexL * G, W, V + index * G);
        }
    }
    /* the Wilson lines "far behind" the nucleus (:157-167) */
    if (VT) {
        const int zb = orientation > 0 ? 0 : L - 1;
        memcpy(VT, V + (long)zb * T * G, sizeof(double) * T * G);
    }
    /* links at t = -at/2: U -> V_x U V_{x+i}^dagger; the copy holds V_x V_{x+i}^dagger */
    double *cU = (double *)malloc(sizeof(double) * n * 3 * G), *cUn = (double *)malloc(sizeof(double) * n * 3 * G);
    memcpy(cU, gr->U, sizeof(double) * n * 3 * G);
    memcpy(cUn, gr->Unext, sizeof(double) * n * 3 * G);
    for (long c = 0; c < n; c++)
        for (int d = 1; d < 3; d++) {
            const long is = orc_shift(gr, c, d, 1);
            double a[18], b[18];
            orc_group_adj(nc, V + is * G, b);
            orc_group_mult(nc, V + c * G, GU(gr, gr->U, c, d), a);
            orc_group_mult(nc, a, b, GU(gr, gr->U, c, d));
            orc_group_mult(nc, V + c * G, b, cU + (c * 3 + d) * G);
        }
    /* V at t = +at/2: the extra path of length at along the beam */
    memcpy(Vc, V, sizeof(double) * n * G);
    const int Mfrac = (int)(M * gr->at / aL * 0.5);
    for (long c = 0; c < n; c++) {
        const int zi = (int)(c / T);
        const long i1 = orc_shift(gr, c, 0, -1), i2 = c;
        const double z1 = zi, z2 = orientation == -1 ? z1 + gr->at / aL : z1 - gr->at / aL;
        double W[18];
        orc_wilson_segment(gr, z1, z2, Mfrac, phi + i1 * A, phi + i2 * A, W);
        orc_group_mult(nc, V + c * G, W, V + c * G);
    }
    for (long c = 0; c < n; c++)
        for (int d = 1; d < 3; d++) {
            const long is = orc_shift(gr, c, d, 1);
            double a[18], b[18];
            orc_group_adj(nc, V + is * G, b);
            orc_group_mult(nc, V + c * G, GU(gr, gr->Unext, c, d), a);
            orc_group_mult(nc, a, b, GU(gr, gr->Unext, c, d));
            orc_group_mult(nc, V + c * G, b, cUn + (c * 3 + d) * G);
        }
    /* E at t = 0 from the temporal plaquette: -proj(Unext U^dagger) / at, on the grid and the copy */
    double *cE = (double *)malloc(sizeof(double) * n * 3 * A);
    for (long c = 0; c < n; c++)
        for (int d = 0; d < 3; d++) {
            double a[18], m[18];
            orc_group_adj(nc, GU(gr, gr->U, c, d), a);
            orc_group_mult(nc, GU(gr, gr->Unext, c, d), a, m);
            orc_group_proj(nc, m, GA(gr, gr->E, c, d));
            for (int k = 0; k < A; k++) GA(gr, gr->E, c, d)[k] *= -1.0 / gr->at;
            orc_group_adj(nc, cU + (c * 3 + d) * G, a);
            orc_group_mult(nc, cUn + (c * 3 + d) * G, a, m);
            orc_group_proj(nc, m, cE + (c * 3 + d) * A);
            for (int k = 0; k < A; k++) cE[(c * 3 + d) * A + k] *= -1.0 / gr->at;
        }
    /* Gauss constraint of the copy (links V V^dagger only) */
    orc_grid copy = *gr;
    copy.U = cU;
    copy.Unext = cUn;
    copy.E = cE;
    for (long c = 0; c < n; c++) {
        if (gr->active[c]) orc_gauss_constraint(&copy, c, gauss + c * A);
        else memset(gauss + c * A, 0, sizeof(double) * A);
    }
    free(phi);
    free(V);
    free(Vc);
    free(cU);
    free(cUn);
    free(cE);
}

/* ===================================================================== */
/* super-particle creator                                                 */
/* ===================================================================== */

/* the block of x0 planes that carries charge (LightConeNGPSuperParticleCreator.java:108-143):
 * planes whose largest |G| exceeds 1e-9 of the global largest; returns the first plane and the
 * width (last - first, the reference's count) */
void orc_lightcone_block(const orc_grid *gr, const double *gauss, int *z_start, int *block_width)
{
    const int L = gr->dims[0], A = gr->asz;
    const long T = (long)gr->dims[1] * gr->dims[2];
    double *mx = (double *)calloc(L, sizeof(double));
    double gmax = 0.0;
    for (int z = 0; z < L; z++) {
        double m = 0.0;
        for (long j = 0; j < T; j++) {
            const double q = sqrt(orc_alg_square(gr->nc, gauss + (z * T + j) * A));
            if (m < q) m = q;
        }
        mx[z] = m;
        if (gmax < m) gmax = m;
    }
    const double cut = gmax * 10E-10;
    int zs = 0, ze = L - 1;
    for (int z = 0; z < L; z++)
        if (mx[z] > cut) {
            zs = z;
            break;
        }
    for (int z = ze - 1; z >= 0; z--)
        if (mx[z] > cut) {
            ze = z;
            break;
        }
    *z_start = zs;
    *block_width = ze - zs;
    free(mx);
}

static int pmod(int i, int n) { return (i % n + n) % n; }

/* The refinement sweeps of LightConeNGPSuperParticleCreator.refine2/refine4 (:245-300) and
 * LightConeNGPParticleCreator over one periodic list of nn charges (q[i*A .. i*A+A)): `sweeps`
 * passes of increasing i, each moving dq between entries i and i+1 where (i + skip) % ppc < ppc - 1
 * (skip = ppc/2 for super particles, 0 for particles): order 2 with dq the (-1, 3, -3, 1)/4
 * combination of entries i-1 .. i+2, order 4 with (1, -5, 10, -10, 5, -1)/12 of i-2 .. i+3. */
void orc_refine_list(double *q, int nn, int A, int ppc, int skip, int order, int sweeps)
{
    for (int it = 0; it < sweeps; it++)
        for (int i = 0; i < nn; i++) {
            const int jm = (i + skip) % ppc;
            if (!(jm >= 0 && jm < ppc - 1)) continue;
            if (order == 2) {
                double *Q0 = q + pmod(i - 1, nn) * A, *Q1 = q + pmod(i, nn) * A;
                double *Q2 = q + pmod(i + 1, nn) * A, *Q3 = q + pmod(i + 2, nn) * A;
                for (int c = 0; c < A; c++) {
                    double dq = Q0[c] * -1;
                    dq += Q1[c] * 3;
                    dq += Q2[c] * -3;
                    dq += Q3[c] * 1;
                    dq *= 1.0 / 4.0;
                    Q1[c] += dq * -1.0;
                    Q2[c] += dq * 1.0;
                }
            } else {
                double *Q[6];
                for (int m = 0; m < 6; m++) Q[m] = q + pmod(i - 2 + m, nn) * A;
                for (int c = 0; c < A; c++) {
                    double dq = Q[0][c] * +1;
                    dq += Q[1][c] * -5;
                    dq += Q[2][c] * +10;
                    dq += Q[3][c] * -10;
                    dq += Q[4][c] * +5;
                    dq += Q[5][c] * -1;
                    dq *= 1.0 / 12.0;
                    Q[2][c] += dq * -1.0;
                    Q[3][c] += dq * 1.0;
                }
            }
        }
}

/* 2nd- then 4th-order refinement, 100 sweeps each, of a list held at positions lst[0 .. nn) of
 * `charges` (gathered into tmp and scattered back) */
static void refine_positions(double *charges, const long *lst, int nn, int A, int ppc, int skip, double *tmp)
{
    for (int i = 0; i < nn; i++) memcpy(tmp + (size_t)i * A, charges + lst[i] * A, sizeof(double) * A);
    orc_refine_list(tmp, nn, A, ppc, skip, 2, 100);
    orc_refine_list(tmp, nn, A, ppc, skip, 4, 100);
    for (int i = 0; i < nn; i++) memcpy(charges + lst[i] * A, tmp + (size_t)i * A, sizeof(double) * A);
}

/* LightConeNGPSuperParticleCreator.initializeParticles (:104-245) for a given block: `subdivisions`
 * groups of ppc super particles (descriptor arrays of subdivisions * ppc entries) and their charges
 * (block_width * T * ppc algebra elements in super-particle order), refined to 2nd then 4th order
 * along the beam, 100 sweeps each. */
void orc_lightcone_superparticles(const orc_grid *gr, const double *gauss, int orientation, int z_start,
                                  int block_width, int subdivisions, int ppc, int *sp_orientation, int *sp_num,
                                  int *sp_index_offset, int *sp_per_plane, int *sp_shift, int *sp_ppc, double *charges)
{
    const int nc = gr->nc, A = gr->asz;
    const long T = (long)gr->dims[1] * gr->dims[2];
    const int ws = block_width / subdivisions, wl = block_width / subdivisions + 1;
    const long ps = ws * T, pl = wl * T;
    const int nl = block_width - subdivisions * ws;
    const long nlong = (long)block_width * ppc;
    /* list[t][l] -> position of that charge in `charges` */
    long *list = (long *)malloc(sizeof(long) * T * nlong);
    long qpos = 0;
    long offset = z_start * T;
    for (int j = 0; j < subdivisions; j++) {
        const long num = j < nl ? pl : ps;
        for (int k = 0; k < ppc; k++) {
            const int s = j * ppc + k;
            sp_orientation[s] = orientation;
            sp_num[s] = (int)num;
            sp_index_offset[s] = (int)offset;
            sp_per_plane[s] = (int)T;
            sp_shift[s] = k;
            sp_ppc[s] = ppc;
        }
        for (long i = 0; i < num; i++) {
            const long index = offset + i;
            for (int k = 0; k < ppc; k++) {
                const long ngp = k < ppc / 2 ? index : orc_shift(gr, index, 0, 1);
                /* charge k of particle i of super particle (j, k) */
                const long at = qpos + (long)k * num + i;
                for (int q = 0; q < A; q++) charges[at * A + q] = gauss[ngp * A + q] * (1.0 / ppc);
                const long tr = index % T, sh = index - z_start * T;
                const long lg = (sh / T) * ppc + k;
                list[tr * nlong + lg] = at;
            }
        }
        qpos += num * ppc;
        offset += num;
    }
    /* refinement along each transverse cell's list */
    double *tmp = (double *)malloc(sizeof(double) * nlong * A);
    for (long t = 0; t < T; t++) refine_positions(charges, list + t * nlong, (int)nlong, A, ppc, ppc / 2, tmp);
    free(tmp);
    (void)nc;
    free(list);
}

/* LightConeNGPParticleCreator.initializeParticles (LightConeNGPParticleCreator.java:104-234):
 * ppc particles per cell of the planes z_start..z_end (inclusive), in (z, transverse cell, j)
 * order, at x0 = z a0 + 1e-11 a0 + (j - ppc/2)/ppc a0 (t = 0) and the same plus orientation at
 * (t = at), the transverse coordinates at x a + 1e-11 a, v = orientation along x0; charge
 * G(ngp(pos0)) / ppc, refined along each transverse cell's list (2nd then 4th order, 100
 * sweeps each; the last particle of a cell is not moved). pos0/pos1/vel are [count][3],
 * charges [count][A]; count = (z_end - z_start + 1) * N1 N2 * ppc. */
void orc_lightcone_particles(const orc_grid *gr, const double *gauss, int orientation, int z_start, int z_end, int ppc,
                             double *pos0, double *pos1, double *vel, double *charges)
{
    const int A = gr->asz, n1 = gr->dims[1], n2 = gr->dims[2];
    const long T = (long)n1 * n2;
    const int planes = z_end - z_start + 1;
    const long nlong = (long)planes * ppc;
    long *list = (long *)malloc(sizeof(long) * T * nlong);
    long cnt = 0;
    for (int z = z_start; z <= z_end; z++)
        for (long i = 0; i < T; i++) {
            const int gp[3] = {z, (int)(i / n2), (int)(i % n2)};
            for (int j = 0; j < ppc; j++) {
                const double x = (1.0 * j - ppc / 2) / ppc;
                const double dz = x * gr->as[0];
                double *p0 = pos0 + cnt * 3, *p1 = pos1 + cnt * 3, *v = vel + cnt * 3;
                for (int k = 0; k < 3; k++) {
                    const double fix = 10E-12 * gr->as[k];
                    p0[k] = gp[k] * gr->as[k] + fix;
                    p1[k] = gp[k] * gr->as[k] + fix;
                    v[k] = 0.0;
                    if (k == 0) {
                        p0[k] += 0.0 * orientation + dz;
                        p1[k] += (0.0 + gr->at) * orientation + dz;
                        v[k] = 1.0 * orientation;
                    }
                }
                long ngp[3];
                for (int k = 0; k < 3; k++) {
                    long c = (long)rint(p0[k] / gr->as[k]);
                    c %= gr->dims[k];
                    if (c < 0) c += gr->dims[k];
                    ngp[k] = c;
                }
                const long index = (ngp[0] * n1 + ngp[1]) * n2 + ngp[2];
                for (int q = 0; q < A; q++) charges[cnt * A + q] = gauss[index * A + q] * (1.0 / ppc);
                list[i * nlong + (long)(z - z_start) * ppc + j] = cnt;
                cnt++;
            }
        }
    /* refinement along each transverse cell's list; refine2/refine4 use jmod = i % ppc here */
    double *tmp = (double *)malloc(sizeof(double) * nlong * A);
    for (long t = 0; t < T; t++) refine_positions(charges, list + t * nlong, (int)nlong, A, ppc, 0, tmp);
    free(tmp);
    free(list);
}

/* ===================================================================== */
/* Wilson-line observables                                                */
/* ===================================================================== */

static double real_trace(int nc, const double *g)
{
    return nc == 2 ? 2 * g[0] : g[0] + g[4] + g[8];
}

/* WilsonLineObservables.computeTadpole (WilsonLineObservables.java:36-63): mean real trace */
double orc_wilson_tadpole(int nc, long T, const double *VT)
{
    const int G = orc_group_size(nc);
    double s = 0.0;
    for (long i = 0; i < T; i++) s += real_trace(nc, VT + i * G);
    return s / T;
}

/* WilsonLineObservables.computeDipole (WilsonLineObservables.java:71-131): tr(V_x V_y^dagger)
 * binned by the periodic transverse distance, pairs in (x, y) order */
void orc_wilson_dipole(int nc, int n1, int n2, double a1, double a2, const double *VT, int nbins, double *corr,
                       long *counts)
{
    const int G = orc_group_size(nc);
    const long T = (long)n1 * n2;
    const int nc_[2] = {n1, n2};
    const double as[2] = {a1, a2};
    const double maxd = fmin(n1 * a1, n2 * a2) * 0.5, ds = maxd / ((double)nbins);
    for (int b = 0; b < nbins; b++) {
        corr[b] = 0.0;
        counts[b] = 0;
    }
    for (long i = 0; i < T; i++) {
        const int x[2] = {(int)(i / n2), (int)(i % n2)};
        for (long j = 0; j < T; j++) {
            const int y[2] = {(int)(j / n2), (int)(j % n2)};
            double dist = 0.0;
            for (int k = 0; k < 2; k++) {
                double d = fabs((double)(x[k] - y[k]));
                if (d > nc_[k] / 2) d -= nc_[k];
                dist += d * d * as[k] * as[k];
            }
            dist = sqrt(dist);
            const int bin = (int)(dist / ds);
            if (bin < nbins) {
                double a[18], m[18];
                orc_group_adj(nc, VT + j * G, a);
                orc_group_mult(nc, VT + i * G, a, m);
                corr[bin] += real_trace(nc, m);
                counts[bin]++;
            }
        }
    }
    for (int b = 0; b < nbins; b++) corr[b] /= (double)counts[b];
}
