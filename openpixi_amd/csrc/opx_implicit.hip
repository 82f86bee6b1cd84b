// opx_implicit.hip -- the iteration of ImplicitTYMSolver (physics/fields/ImplicitTYMSolver.java:35-263)
// and the group logarithm / power its damping uses.
//
// One implicit step is: ImplicitBegin (the fused explicit update into the implicit copies of E
// and Unext; the host runs the autotuned fused kernel for it), then per iteration the
// implicit links are swapped and k_implicit_iteration recomputes every active link from
// plaquette sums that mix three time levels (ImplicitStep, :96-121, with the damping of
// ImplicitDamping, :218-231, applied to the same link), then ImplicitEnd copies E and Unext back
// (k_implicit_end). The beam direction is x0, as in the reference.
#include <math.h>

#include <type_traits>

#include "opx_group.h"
#include "opx_kernels.h"
#include "opx_order.h"

namespace opx {

namespace {

constexpr int kIB = 256;

OPX_HD float fasin(float x) { return asinf(x); }
OPX_HD double fasin(double x) { return asin(x); }
OPX_HD float fatan2(float y, float x) { return atan2f(y, x); }
OPX_HD double fatan2(double y, double x) { return atan2(y, x); }
OPX_HD float fpow(float x, float y) { return powf(x, y); }
OPX_HD double fpow(double x, double y) { return pow(x, y); }

// fields in the device layout (fpos); the cell count argument is not needed to address them
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, T>::M ld_link(const T* __restrict__ U, uint32_t, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    typename GR::M m;
    const T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) GR::gref(m, GR::slot(k)) = p[fcomp<T>(dir * GR::S + k, 3 * GR::S)];
    GR::complete(m);
    return m;
}

// the stored reals only (two rows of an FP32 SU(3) link), row 2 left zero
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, T>::M ld_link_r01(const T* __restrict__ U, uint32_t, uint32_t site,
                                                                int dir)
{
    using GR = Group<NC, T>;
    typename GR::M m = GR::zero();
    const T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) GR::gref(m, GR::slot(k)) = p[fcomp<T>(dir * GR::S + k, 3 * GR::S)];
    return m;
}

template <int NC, typename T>
__device__ __forceinline__ void st_link(T* __restrict__ U, uint32_t, uint32_t site, int dir,
                                        const typename Group<NC, T>::M& m)
{
    using GR = Group<NC, T>;
    T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) p[fcomp<T>(dir * GR::S + k, 3 * GR::S)] = GR::gget(m, GR::slot(k));
}

template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, T>::A ld_alg(const T* __restrict__ X, uint32_t, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    typename GR::A a;
    const T* p = fcell(X, site, 3 * GR::AD);
#pragma unroll
    for (int k = 0; k < GR::AD; k++) aref(a, k) = p[fcomp<T>(dir * GR::AD + k, 3 * GR::AD)];
    return a;
}

template <int NC, typename T>
__device__ __forceinline__ void st_alg(T* __restrict__ X, uint32_t, uint32_t site, int dir,
                                       const typename Group<NC, T>::A& a)
{
    using GR = Group<NC, T>;
    T* p = fcell(X, site, 3 * GR::AD);
#pragma unroll
    for (int k = 0; k < GR::AD; k++) p[fcomp<T>(dir * GR::AD + k, 3 * GR::AD)] = aget(a, k);
}

// the same accesses with the non-temporal hint, for data a launch touches once (E, J, iE, iUn):
// the streamed fields then leave the neighbour links of the three link fields in L2
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, T>::A ld_alg_once(const T* __restrict__ X, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    typename GR::A a;
    const T* p = fcell(X, site, 3 * GR::AD);
#pragma unroll
    for (int k = 0; k < GR::AD; k++) aref(a, k) = __builtin_nontemporal_load(p + fcomp<T>(dir * GR::AD + k, 3 * GR::AD));
    return a;
}

template <int NC, typename T>
__device__ __forceinline__ void st_alg_once(T* __restrict__ X, uint32_t site, int dir, const typename Group<NC, T>::A& a)
{
    using GR = Group<NC, T>;
    T* p = fcell(X, site, 3 * GR::AD);
#pragma unroll
    for (int k = 0; k < GR::AD; k++) __builtin_nontemporal_store(aget(a, k), p + fcomp<T>(dir * GR::AD + k, 3 * GR::AD));
}

template <int NC, typename T>
__device__ __forceinline__ void st_link_once(T* __restrict__ U, uint32_t site, int dir, const typename Group<NC, T>::M& m)
{
    using GR = Group<NC, T>;
    T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) __builtin_nontemporal_store(GR::gget(m, GR::slot(k)), p + fcomp<T>(dir * GR::S + k, 3 * GR::S));
}

// ------------------------------------------------------------------ logarithm and power
// SU2GroupElement.getAlgebraElement (math/SU2GroupElement.java:234-251):
// v = 2 asin|e| / |e| e, or 2 e when |e| < 1e-15
template <typename T>
OPX_HD Alg2<T> group_log(const Su2<T>& u)
{
    T norm = u.e1 * u.e1 + u.e2 * u.e2 + u.e3 * u.e3;
    norm = fsqrt(norm);
    if (norm < T(1.E-15)) return Alg2<T>{{2 * u.e1, 2 * u.e2, 2 * u.e3}};
    const T f = T(2) * fasin(norm) / norm;
    return Alg2<T>{{f * u.e1, f * u.e2, f * u.e3}};
}

// log U by the series of log(1 + X), X = U - 1 (math/SU3GroupElement.java:425-449):
// term_m = term_{m-1} (1 - U), term_0 = -1, log U = sum_m term_m / m; H = -i log U
template <typename T>
OPX_HD Su3<T> su3_log_series(const Su3<T>& u, int terms)
{
    Su3<T> term = Group<3, T>::zero(), one_minus_u, sum = Group<3, T>::zero();
    term.r[0] = term.r[4] = term.r[8] = T(-1);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        one_minus_u.r[k] = -u.r[k];
        one_minus_u.i[k] = -u.i[k];
    }
    one_minus_u.r[0] += T(1);
    one_minus_u.r[4] += T(1);
    one_minus_u.r[8] += T(1);
    for (int m = 1; m <= terms; m++) {
        term = mul(term, one_minus_u);
        const T w = T(1) / T(m);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            sum.r[k] += term.r[k] * w;
            sum.i[k] += term.i[k] * w;
        }
    }
    Su3<T> h;   // -i log U
#pragma unroll
    for (int k = 0; k < 9; k++) {
        h.r[k] = sum.i[k];
        h.i[k] = -sum.r[k];
    }
    return h;
}

// SU3GroupElement.getAlgebraElement (math/SU3GroupElement.java:260-453,476-485): H with
// U = exp(i H) as a full matrix (its Hermitian part and real diagonal are kept; the
// series forms leave a trace of the size of their truncation error). Eigenvalues from the depressed characteristic
// cubic (det U = 1), eigenvectors as a non-vanishing column of (U - l_j)(U - l_k),
// phases made traceless by shifting the first; the series form near the identity
// (Re tr U >= 3 - 1e-2, 15 terms) and for nearly degenerate eigenvalues (50 terms).
template <typename T>
OPX_HD Su3<T> su3_log_matrix(const Su3<T>& u)
{
    const T trRe = u.r[0] + u.r[4] + u.r[8], trIm = u.i[0] + u.i[4] + u.i[8];
    if (trRe >= T(3 - 1.E-2)) return su3_log_series(u, 15);
    const Su3<T> u2 = mul(u, u);
    const T sqRe = u2.r[0] + u2.r[4] + u2.r[8], sqIm = u2.i[0] + u2.i[4] + u2.i[8];
    const T pRe = (trRe * trRe - trIm * trIm) / 6 - sqRe / 2;
    const T pIm = trRe * trIm / 3 - sqIm / 2;
    const T qRe = -1 - (trRe * sqRe - trIm * sqIm) / 6 + 5 * trRe * (trRe * trRe - 3 * trIm * trIm) / 54;
    const T qIm = -(trRe * sqIm + trIm * sqRe) / 6 + 5 * trIm * (3 * trRe * trRe - trIm * trIm) / 54;
    T disc = qRe * qRe - qIm * qIm + 4 * pRe * (pRe * pRe - 3 * pIm * pIm) / 27;
    if (disc < 0) disc = 0;
    const T rad = fsqrt(disc);
    const T wRe = (-qRe + rad) / 2, wIm = -qIm / 2;
    const T th = fatan2(wIm, wRe);
    const T r = fpow(wRe * wRe + wIm * wIm, T(1) / T(6));
    T lRe[3], lIm[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T t = (th + 2 * T(M_PI) * i) / 3;
        if (fabs(r) == T(0)) {
            lRe[i] = trRe / 3;
            lIm[i] = trIm / 3;
        } else {
            const T c = fcos(t), s = fsin(t);
            lRe[i] = r * c - (pRe * c + pIm * s) / (3 * r) + trRe / 3;
            lIm[i] = r * s + (pRe * s - pIm * c) / (3 * r) + trIm / 3;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int a = k == 2 ? 1 : 0, b = k == 0 ? 1 : 2;
        if (fabs(1 - lRe[a] / lRe[b]) < T(1.E-4) && fabs(1 - lIm[a] / lIm[b]) < T(1.E-4))
            return su3_log_series(u, 50);
    }
    // H = sum_i ph_i x_i x_i^dagger
    T ph[3];
#pragma unroll
    for (int i = 0; i < 3; i++) ph[i] = fatan2(lIm[i], lRe[i]);
    ph[0] -= ph[0] + ph[1] + ph[2];
    Su3<T> H = Group<3, T>::zero();
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        Su3<T> A = u, B = u;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            A.r[4 * q] -= lRe[j];
            A.i[4 * q] -= lIm[j];
            B.r[4 * q] -= lRe[k];
            B.i[4 * q] -= lIm[k];
        }
        const Su3<T> M = mul(A, B);
        T xr[3], xi[3];
        bool ok = false;
#pragma unroll
        for (int col = 0; col < 3; col++) {
            if (ok) continue;
            T nrm = 0;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                xr[q] = M.r[3 * q + col];
                xi[q] = M.i[3 * q + col];
                nrm += xr[q] * xr[q] + xi[q] * xi[q];
            }
            nrm = fsqrt(nrm);
            if (!(fabs(nrm) < T(1.E-12))) {
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    xr[q] /= nrm;
                    xi[q] /= nrm;
                }
                ok = true;
            }
        }
        if (!ok) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                xr[q] = q == i ? T(1) : T(0);
                xi[q] = 0;
            }
        }
        // H += ph_i x x^dagger
#pragma unroll
        for (int rr = 0; rr < 3; rr++)
#pragma unroll
            for (int ss = 0; ss < 3; ss++) {
                H.r[3 * rr + ss] += ph[i] * (xr[rr] * xr[ss] + xi[rr] * xi[ss]);
                H.i[3 * rr + ss] += ph[i] * (xi[rr] * xr[ss] - xr[rr] * xi[ss]);
            }
    }
    return H;
}

// precision conversions of group elements
template <typename T>
OPX_HD Su2<double> to_double(const Su2<T>& u)
{
    return Su2<double>{(double)u.e0, (double)u.e1, (double)u.e2, (double)u.e3};
}
template <typename T>
OPX_HD Su3<double> to_double(const Su3<T>& u)
{
    Su3<double> o;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        o.r[k] = (double)u.r[k];
        o.i[k] = (double)u.i[k];
    }
    return o;
}
template <typename M>
OPX_HD M from_double(const Su2<double>& u)
{
    using T = decltype(M{}.e0);
    return M{(T)u.e0, (T)u.e1, (T)u.e2, (T)u.e3};
}
template <typename M>
OPX_HD M from_double(const Su3<double>& u)
{
    using T = typename std::remove_reference<decltype(M{}.r[0])>::type;
    M o;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        o.r[k] = (T)u.r[k];
        o.i[k] = (T)u.i[k];
    }
    return o;
}

// GroupElement.pow(x) = getLink(x getAlgebraElement())
template <typename T>
OPX_HD Su2<T> group_pow(const Su2<T>& u, T x)
{
    return exp_link(group_log(u), x);
}

// SU(3): exp(i x H) of the traceless part, times the phase exp(i x tr H / 3) of the trace
// where SU3AlgebraElement.getLink takes its power series, which exponentiates the stored
// matrix with its diagonal (tr (xH)^2 <= 1, math/SU3AlgebraElement.java:389-417)
template <typename T>
OPX_HD Su3<T> group_pow(const Su3<T>& u, T x)
{
    const Su3<T> H = su3_log_matrix(u);
    Su3<T> e = exp_link(herm_coeffs(H), x);
    const T d0 = H.r[0] * x, d4 = H.r[4] * x, d8 = H.r[8] * x;
    const T o1 = (H.r[1] + H.r[3]) / 2 * x, o2 = (H.r[2] + H.r[6]) / 2 * x, o5 = (H.r[5] + H.r[7]) / 2 * x;
    const T o3 = (H.i[3] - H.i[1]) / 2 * x, o6 = (H.i[6] - H.i[2]) / 2 * x, o7 = (H.i[7] - H.i[5]) / 2 * x;
    const T trsq = d0 * d0 + d4 * d4 + d8 * d8 + 2 * (o1 * o1 + o2 * o2 + o3 * o3 + o5 * o5 + o6 * o6 + o7 * o7);
    if (trsq <= T(1)) {
        const T ph = (d0 + d4 + d8) / 3, c = fcos(ph), s = fsin(ph);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const T re = e.r[k] * c - e.i[k] * s, im = e.r[k] * s + e.i[k] * c;
            e.r[k] = re;
            e.i[k] = im;
        }
    }
    return e;
}

// U^damping V^(1 - damping) (ImplicitDamping), evaluated in double in every precision: the
// logarithm's branch decisions (near-identity and degeneracy cut-offs, eigenvector
// normalisation) are tuned for double. Its eigenvectors of nearly degenerate spectra lose
// orthogonality in proportion to the input's distance from the group over the spectral gap,
// so FP32 links (unitary to ~1e-7) are first projected onto the group in double.
template <typename M, typename T>
__device__ M damped_link(const M& u, const M& v, T damping)
{
    auto on_group = [](const M& x) {
        auto xd = to_double(x);
        if (sizeof(T) == 4) xd = reunit(xd);
        return xd;
    };
    const double dd = (double)damping;
    // Damping 0 or 1: one factor is x^0 = getLink(0), the identity exactly, and the other is
    // x^1 = getLink(getAlgebraElement(x)), which is x (projected onto the group) up to the
    // rounding of the logarithm and the exponential, about 1e-16: it is returned as it is.
    if (dd == 0.0) return from_double<M>(on_group(v));
    if (dd == 1.0) return from_double<M>(on_group(u));
    return from_double<M>(mul(group_pow(on_group(u), dd), group_pow(on_group(v), 1.0 - dd)));
}

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ void site_coords(const Lattice& L, uint32_t site, uint32_t x[3])
{
    const uint32_t q = site / L.d2;
    x[2] = site - q * L.d2;
    x[0] = q / L.d1;
    x[1] = q - x[0] * L.d1;
}

__device__ __forceinline__ uint32_t step_site(const Lattice& L, const uint32_t x[3], int d, int o)
{
    uint32_t y[3] = {x[0], x[1], x[2]};
    const uint32_t n = d == 0 ? L.d0 : (d == 1 ? L.d1 : L.d2);
    y[d] = o > 0 ? (y[d] + 1 == n ? 0 : y[d] + 1) : (y[d] == 0 ? n - 1 : y[d] - 1);
    return y[0] * L.s0 + y[1] * L.s1 + y[2];
}

__device__ __forceinline__ bool in_region(const Box& b, const uint32_t x[3])
{
    if (b.all) return true;
    return (int)x[0] >= b.lo[0] && (int)x[0] <= b.hi[0] && (int)x[1] >= b.lo[1] && (int)x[1] <= b.hi[1] &&
           (int)x[2] >= b.lo[2] && (int)x[2] <= b.hi[2];
}

// ImplicitStep + ImplicitDamping for one link. Time levels of the links: -1 = Uold (the links
// before U), 0 = U, +1 = iU (the current implicit estimate). Blocks run in the row-tile order
// of opx_order.h with the direction fastest, so the neighbour links of the three fields that a
// link reads are mostly still in L2 from the blocks just before it.
template <int NC, typename T, bool HAS_J>
__global__ __launch_bounds__(kIB, 2) void k_implicit_iteration(Lattice L, LinkOrder order, Box active,
                                                            const T* __restrict__ U, const T* __restrict__ Uold,
                                                            const T* __restrict__ iU, const T* __restrict__ E,
                                                            const T* __restrict__ J, T* __restrict__ iE,
                                                            T* __restrict__ iUn, T damping)
{
    using GR = Group<NC, T>;
    using M = typename GR::M;
    using A = typename GR::A;
    const uint32_t logical = xcd_logical(blockIdx.x, gridDim.x);
    const int d = (int)(logical % 3);
    uint32_t site;
    if (!tile_site<kIB>(L, order, logical / 3, site)) return;
    uint32_t x[3];
    site_coords(L, site, x);
    if (!in_region(active, x)) return;
    const uint32_t n = L.n;
    auto field = [&](int t) { return t < 0 ? Uold : (t == 0 ? U : iU); };
    auto link = [&](int t, uint32_t s, int dir) -> M { return ld_link<NC, T>(field(t), n, s, dir); };
    const T at = (T)L.at;
    A e = ld_alg_once<NC, T>(E, site, d);
    const uint32_t xpd = step_site(L, x, d, 1);
    uint32_t xp_d[3];
    site_coords(L, xpd, xp_d);
    for (int t = 1; t >= -1; t -= 2) {
        M S = GR::zero();
        // both planes' loads issued together where the registers allow it (FP32)
        constexpr int kPlaneUnroll = sizeof(T) == 4 ? 3 : 1;
#pragma unroll kPlaneUnroll
        for (int i = 0; i < 3; i++) {
            if (i == d) continue;
            // implicit in the transverse plane; in planes with the beam only the beam link
            const int td = (d != 0) && (i != 0) ? t : (d == 0 ? t : 0);
            const int ti = (d != 0) && (i != 0) ? t : (i == 0 ? t : 0);
            const uint32_t xpi = step_site(L, x, i, 1), xmi = step_site(L, x, i, -1), xpdmi = step_site(L, xp_d, i, -1);
            M u1;
            if constexpr (GR::S != GR::G) {
                // two-row storage: each product of group elements is formed in its first two rows
                // and completed by the row-2 identity; left factors are read in the stored rows
                u1 = mul_adj_r01(mul_adj_r01(ld_link_r01<NC, T>(field(ti), n, xpd, i), link(td, xpi, d)),
                                 link(ti, site, i));
                su3_third_row(u1);
                M y = mul_r01(ld_link_r01<NC, T>(field(td), n, xmi, d), link(ti, xpdmi, i));
                su3_third_row(y);
                M u2 = adj_mul_r01(y, link(ti, xmi, i));
                su3_third_row(u2);
                acc(u1, u2);
            } else {
                u1 = mul_adj(mul_adj(link(ti, xpd, i), link(td, xpi, d)), link(ti, site, i));
                const M u2 = adj_mul(mul(link(td, xmi, d), link(ti, xpdmi, i)), link(ti, xmi, i));
                acc(u1, u2);
            }
            acc_scaled(S, mul(link(td, site, d), u1), (T)L.area[i]);
        }
        const A pr = proj(S);
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(e, k) += aget(pr, k) * (at * T(0.5));
    }
    if (HAS_J) {
        const A jj = ld_alg_once<NC, T>(J, site, d);
        const T uf = (T)L.uf[d];
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(e, k) += aget(jj, k) * uf;
    }
    st_alg_once<NC, T>(iE, site, d, e);
    const M V = mul(exp_link(e, -at), ld_link<NC, T>(U, n, site, d));
    // Unext = U^damping Unext^(1 - damping) on the implicit links
    const M out = damped_link(ld_link<NC, T>(iU, n, site, d), V, damping);
    st_link_once<NC, T>(iUn, site, d, out);
}

// ImplicitEnd: E and Unext of the active links from the implicit copies
template <int NC, typename T>
__global__ __launch_bounds__(kIB) void k_implicit_end(Lattice L, Box active, const T* __restrict__ iE,
                                                      const T* __restrict__ iUn, T* __restrict__ E,
                                                      T* __restrict__ Unext)
{
    using GR = Group<NC, T>;
    const uint32_t site = blockIdx.x * kIB + threadIdx.x;
    const int d = blockIdx.y;
    if (site >= L.n) return;
    uint32_t x[3];
    site_coords(L, site, x);
    if (!in_region(active, x)) return;
#pragma unroll
    for (int k = 0; k < GR::AD; k++) {
        const uint64_t i = fposT<T>(d * GR::AD + k, site, 3 * GR::AD);
        E[i] = iE[i];
    }
#pragma unroll
    for (int k = 0; k < GR::S; k++) {
        const uint64_t i = fposT<T>(d * GR::S + k, site, 3 * GR::S);
        Unext[i] = iUn[i];
    }
}

// single-element power for tests of the device arithmetic, in double as the damping evaluates it
template <int NC, typename T>
__global__ void k_group_pow(int count, const T* __restrict__ U, T x, T* __restrict__ out)
{
    using GR = Group<NC, T>;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    typename GR::M u;
#pragma unroll
    for (int q = 0; q < GR::G; q++) GR::gref(u, q) = U[(size_t)k * GR::G + q];
    const typename GR::M o = from_double<typename GR::M>(group_pow(to_double(u), (double)x));
#pragma unroll
    for (int q = 0; q < GR::G; q++) out[(size_t)k * GR::G + q] = GR::gget(o, q);
}

}  // namespace

#define OPX_IDISPATCH(nc, prec, ...)                                                           \
    do {                                                                                       \
        if ((nc) == 2 && (prec) == 0) { constexpr int NC = 2; using T = double; __VA_ARGS__; } \
        else if ((nc) == 2) { constexpr int NC = 2; using T = float; __VA_ARGS__; }           \
        else if ((prec) == 0) { constexpr int NC = 3; using T = double; __VA_ARGS__; }        \
        else { constexpr int NC = 3; using T = float; __VA_ARGS__; }                          \
    } while (0)

hipError_t launch_implicit_iteration(int nc, int prec, const Lattice& L, const Box& active, const void* U,
                                     const void* Uold, const void* iU, const void* E, const void* J, bool has_j,
                                     void* iE, void* iUn, double damping, hipStream_t st)
{
    // 16-row tiles: at 128^3 SU(3) within 1 % of the fastest of 4, 8, 16 and 32 rows in both
    // precisions (the links of three fields around a link come from L2 rather than HBM).
    const LinkOrder order = make_link_order(L, 16, kIB);
    const dim3 grid(3 * link_order_blocks(L, order));
    OPX_IDISPATCH(nc, prec, {
        if (has_j)
            hipLaunchKernelGGL((k_implicit_iteration<NC, T, true>), grid, dim3(kIB), 0, st, L, order, active,
                               (const T*)U, (const T*)Uold, (const T*)iU, (const T*)E, (const T*)J, (T*)iE, (T*)iUn,
                               (T)damping);
        else
            hipLaunchKernelGGL((k_implicit_iteration<NC, T, false>), grid, dim3(kIB), 0, st, L, order, active,
                               (const T*)U, (const T*)Uold, (const T*)iU, (const T*)E, (const T*)E, (T*)iE, (T*)iUn,
                               (T)damping);
    });
    return hipGetLastError();
}

hipError_t launch_implicit_end(int nc, int prec, const Lattice& L, const Box& active, const void* iE, const void* iUn,
                               void* E, void* Unext, hipStream_t st)
{
    const dim3 grid((L.n + kIB - 1) / kIB, 3);
    OPX_IDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_implicit_end<NC, T>), grid, dim3(kIB), 0, st, L, active, (const T*)iE, (const T*)iUn,
                           (T*)E, (T*)Unext);
    });
    return hipGetLastError();
}

hipError_t launch_group_pow(int nc, int prec, int count, const void* U, double x, void* out, hipStream_t st)
{
    OPX_IDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_group_pow<NC, T>), dim3((count + 63) / 64), dim3(64), 0, st, count, (const T*)U, (T)x,
                           (T*)out);
    });
    return hipGetLastError();
}

}  // namespace opx
