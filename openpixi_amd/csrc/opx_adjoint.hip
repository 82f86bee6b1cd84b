// opx_adjoint.hip -- gfx950 kernels of the SU(2) adjoint (reverse-mode derivative)
// of the temporal-gauge leapfrog. The reference has no differentiation; these
// kernels transpose the Jacobians of its forward map (FastTYMSolver.java:40-84).
//
// An SU(2) element and its cotangent are real 4-vectors (e0, e1, e2, e3) with
// the quaternion product of Su2; for r = x y the transposed Jacobians are
// x_bar = r_bar y^+ and y_bar = x^+ r_bar (^+ the quaternion conjugate), which
// holds for any 4-vectors because the product is bilinear.
//
// Force term. The plaquette p = (x, mu < nu) with loop factors a = U_mu(x),
// b = U_nu(x+mu), c = U_mu(x+nu)^+, d = U_nu(x)^+ contributes to the force
// proj(U_l S_l) of its four links with cyclic rotations of the loop:
//   F_mu(x) += w_nu proj(abcd), F_nu(x) -= w_mu proj(abcd),
//   F_nu(x+mu) += w_mu proj(bcda), F_mu(x+nu) -= w_nu proj(dabc),    w_k = 1/a_k^2,
// so sum_l <Ebar_l, F_l> = sum_p <C0, proj abcd> + <C1, proj bcda> + <C3, proj dabc>
// with C0 = w_nu Ebar_a - w_mu Ebar_d, C1 = w_mu Ebar_b, C3 = -w_nu Ebar_c, the
// cotangents on the plaquette's own four links. Each link gathers the derivative
// of that sum from the four plaquettes that contain it (no atomics).
//
// Layout: [dir][component][cell] like the forward fields; the cotangents Ubar
// (4 per link) and Ebar (3 per link) are in the context precision.
#include "opx_group.h"
#include "opx_kernels.h"

namespace opx {

namespace {

constexpr int kAdjBlock = 256;

template <typename T>
using Q = Su2<T>;
template <typename T>
using V3 = Alg2<T>;

struct Site {
    uint32_t x[3];
};

__device__ __forceinline__ Site site_of(const Lattice& L, uint32_t s)
{
    Site c;
    uint32_t q = s / L.d2;
    c.x[2] = s - q * L.d2;
    c.x[0] = q / L.d1;
    c.x[1] = q - c.x[0] * L.d1;
    return c;
}

__device__ __forceinline__ uint32_t index_of(const Lattice& L, const Site& c)
{
    return c.x[0] * L.s0 + c.x[1] * L.s1 + c.x[2];
}

// periodic step of one cell along d (o = +1 or -1)
__device__ __forceinline__ Site step(const Lattice& L, Site c, int d, int o)
{
    const uint32_t n = d == 0 ? L.d0 : (d == 1 ? L.d1 : L.d2);
    uint32_t v = c.x[d];
    c.x[d] = o > 0 ? (v + 1 == n ? 0u : v + 1) : (v == 0 ? n - 1 : v - 1);
    return c;
}

// link fields (U, Ubar: 12 components per cell) and algebra fields (E, Ebar: 9) in the device
// layout (fpos); the cell count argument is not needed to address them
template <typename T>
__device__ __forceinline__ Q<T> ld_q(const T* __restrict__ U, uint32_t, uint32_t s, int d)
{
    const T* p = fcell(U, s, 12);
    return Q<T>{p[fcomp<T>(4 * d, 12)], p[fcomp<T>(4 * d + 1, 12)], p[fcomp<T>(4 * d + 2, 12)], p[fcomp<T>(4 * d + 3, 12)]};
}

template <typename T>
__device__ __forceinline__ void st_q(T* __restrict__ U, uint32_t, uint32_t s, int d, const Q<T>& q)
{
    T* p = fcell(U, s, 12);
    p[fcomp<T>(4 * d, 12)] = q.e0;
    p[fcomp<T>(4 * d + 1, 12)] = q.e1;
    p[fcomp<T>(4 * d + 2, 12)] = q.e2;
    p[fcomp<T>(4 * d + 3, 12)] = q.e3;
}

template <typename T>
__device__ __forceinline__ V3<T> ld_v(const T* __restrict__ X, uint32_t, uint32_t s, int d)
{
    const T* p = fcell(X, s, 9);
    return V3<T>{{p[fcomp<T>(3 * d, 9)], p[fcomp<T>(3 * d + 1, 9)], p[fcomp<T>(3 * d + 2, 9)]}};
}

template <typename T>
__device__ __forceinline__ void st_v(T* __restrict__ X, uint32_t, uint32_t s, int d, const V3<T>& v)
{
    T* p = fcell(X, s, 9);
    p[fcomp<T>(3 * d, 9)] = v.v[0];
    p[fcomp<T>(3 * d + 1, 9)] = v.v[1];
    p[fcomp<T>(3 * d + 2, 9)] = v.v[2];
}

template <typename T>
__device__ __forceinline__ Q<T> conj(const Q<T>& a)
{
    return Q<T>{a.e0, -a.e1, -a.e2, -a.e3};
}

// transpose of proj (proj(q) = 2 (q1, q2, q3))
template <typename T>
__device__ __forceinline__ Q<T> lift(const V3<T>& c)
{
    return Q<T>{T(0), 2 * c.v[0], 2 * c.v[1], 2 * c.v[2]};
}

template <typename T>
__device__ __forceinline__ Q<T> add3(const Q<T>& x, const Q<T>& y, const Q<T>& z)
{
    return Q<T>{x.e0 + y.e0 + z.e0, x.e1 + y.e1 + z.e1, x.e2 + y.e2 + z.e2, x.e3 + y.e3 + z.e3};
}

template <typename T>
struct Plaq {
    Q<T> a, b, c, d;   // loop factors, c and d already conjugated
};

template <typename T>
__device__ __forceinline__ Plaq<T> load_plaq(const Lattice& L, const T* __restrict__ U, const Site& x, int mu, int nu,
                                             uint32_t& ia, uint32_t& ib, uint32_t& ic, uint32_t& id)
{
    const Site xm = step(L, x, mu, 1), xn = step(L, x, nu, 1);
    ia = index_of(L, x);
    ib = index_of(L, xm);
    ic = index_of(L, xn);
    id = ia;
    Plaq<T> p;
    p.a = ld_q(U, L.n, ia, mu);
    p.b = ld_q(U, L.n, ib, nu);
    p.c = conj(ld_q(U, L.n, ic, mu));
    p.d = conj(ld_q(U, L.n, id, nu));
    return p;
}

// cotangent of loop factor ROLE (0 a, 1 b, 2 c, 3 d) of
// <C0, proj abcd> + <C1, proj bcda> + <C3, proj dabc>, lifted c_k = lift(C_k)
template <int ROLE, typename T>
__device__ __forceinline__ Q<T> factor_cotangent(const Plaq<T>& p, const Q<T>& c0, const Q<T>& c1, const Q<T>& c3)
{
    const Q<T> ad = conj(p.a), bd = conj(p.b), cd = conj(p.c), dd = conj(p.d);
    if constexpr (ROLE == 0) {   // c0 (bcd)^+ + (bcd)^+ c1 + d^+ c3 (bc)^+
        const Q<T> bc_d = mul(cd, bd);
        const Q<T> bcd_d = mul(dd, bc_d);
        return add3(mul(c0, bcd_d), mul(bcd_d, c1), mul(mul(dd, c3), bc_d));
    } else if constexpr (ROLE == 1) {   // a^+ c0 (cd)^+ + c1 (cda)^+ + (da)^+ c3 c^+
        const Q<T> da_d = mul(ad, dd);
        return add3(mul(mul(ad, c0), mul(dd, cd)), mul(c1, mul(da_d, cd)), mul(mul(da_d, c3), cd));
    } else if constexpr (ROLE == 2) {   // (ab)^+ c0 d^+ + b^+ c1 (da)^+ + (dab)^+ c3
        const Q<T> ab_d = mul(bd, ad);
        return add3(mul(mul(ab_d, c0), dd), mul(mul(bd, c1), mul(ad, dd)), mul(mul(ab_d, dd), c3));
    } else {   // (abc)^+ c0 + (bc)^+ c1 a^+ + c3 (abc)^+
        const Q<T> bc_d = mul(cd, bd);
        const Q<T> abc_d = mul(bc_d, ad);
        return add3(mul(abc_d, c0), mul(mul(bc_d, c1), ad), mul(c3, abc_d));
    }
}

// C0, C1, C3 of the force term from the cotangents on the plaquette's links
template <typename T>
__device__ __forceinline__ void force_weights(const Lattice& L, const T* __restrict__ Eb, int mu, int nu, uint32_t ia,
                                              uint32_t ib, uint32_t ic, uint32_t id, Q<T>& c0, Q<T>& c1, Q<T>& c3)
{
    const T wmu = (T)L.area[mu], wnu = (T)L.area[nu];
    const V3<T> Ea = ld_v(Eb, L.n, ia, mu), Ebb = ld_v(Eb, L.n, ib, nu);
    const V3<T> Ec = ld_v(Eb, L.n, ic, mu), Ed = ld_v(Eb, L.n, id, nu);
    V3<T> C0, C1, C3;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        C0.v[k] = wnu * Ea.v[k] - wmu * Ed.v[k];
        C1.v[k] = wmu * Ebb.v[k];
        C3.v[k] = -wnu * Ec.v[k];
    }
    c0 = lift(C0);
    c1 = lift(C1);
    c3 = lift(C3);
}

// Magnetic-energy term: C0 = 2 coef proj(abcd), C1 = C3 = 0
template <typename T>
__device__ __forceinline__ Q<T> bsq_c0(const Plaq<T>& p, T coef)
{
    const Q<T> P = mul(mul(p.a, p.b), mul(p.c, p.d));
    const V3<T> pr = proj(P);
    return lift(V3<T>{{2 * coef * pr.v[0], 2 * coef * pr.v[1], 2 * coef * pr.v[2]}});
}

struct BsqCoef {
    double c[3];   // coefficient of |proj P|^2 for the plaquette orthogonal to direction j
};

// Ubar(l) += scale * sum over the 4 plaquettes containing link l of the derivative of
// the force pairing (BSQ = false) or of the magnetic loss term (BSQ = true)
template <typename T, bool BSQ>
__global__ void __launch_bounds__(kAdjBlock) k_adj_gather(Lattice L, const T* __restrict__ U,
                                                           const T* __restrict__ Ebar, T* __restrict__ Ubar, T scale,
                                                           BsqCoef bc)
{
    const uint32_t id = blockIdx.x * kAdjBlock + threadIdx.x;
    if (id >= 3 * L.n) return;
    const int dd = (int)(id / L.n);
    const uint32_t s = id - (uint32_t)dd * L.n;
    const Site y = site_of(L, s);
    Q<T> acc{T(0), T(0), T(0), T(0)};
    const Q<T> zero{T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int e = 0; e < 3; e++) {
        if (e == dd) continue;
        const int mu = dd < e ? dd : e, nu = dd < e ? e : dd;
        const T coef = (T)bc.c[3 - mu - nu];
        uint32_t ia, ib, ic, idx;
        Q<T> c0, c1, c3;
        if (dd < e) {
            // role a in the plaquette at y, role c (conjugated) in the plaquette at y - nu
            Plaq<T> p = load_plaq(L, U, y, mu, nu, ia, ib, ic, idx);
            if constexpr (BSQ) {
                c0 = bsq_c0(p, coef);
                c1 = c3 = zero;
            } else {
                force_weights(L, Ebar, mu, nu, ia, ib, ic, idx, c0, c1, c3);
            }
            const Q<T> ga = factor_cotangent<0>(p, c0, c1, c3);
            p = load_plaq(L, U, step(L, y, nu, -1), mu, nu, ia, ib, ic, idx);
            if constexpr (BSQ) {
                c0 = bsq_c0(p, coef);
            } else {
                force_weights(L, Ebar, mu, nu, ia, ib, ic, idx, c0, c1, c3);
            }
            const Q<T> gc = conj(factor_cotangent<2>(p, c0, c1, c3));
            acc = add3(acc, ga, gc);
        } else {
            // role d (conjugated) in the plaquette at y, role b in the plaquette at y - mu
            Plaq<T> p = load_plaq(L, U, y, mu, nu, ia, ib, ic, idx);
            if constexpr (BSQ) {
                c0 = bsq_c0(p, coef);
                c1 = c3 = zero;
            } else {
                force_weights(L, Ebar, mu, nu, ia, ib, ic, idx, c0, c1, c3);
            }
            const Q<T> gd = conj(factor_cotangent<3>(p, c0, c1, c3));
            p = load_plaq(L, U, step(L, y, mu, -1), mu, nu, ia, ib, ic, idx);
            if constexpr (BSQ) {
                c0 = bsq_c0(p, coef);
            } else {
                force_weights(L, Ebar, mu, nu, ia, ib, ic, idx, c0, c1, c3);
            }
            const Q<T> gb = factor_cotangent<1>(p, c0, c1, c3);
            acc = add3(acc, gd, gb);
        }
    }
    Q<T> u = ld_q(Ubar, L.n, s, dd);
    u.e0 += scale * acc.e0;
    u.e1 += scale * acc.e1;
    u.e2 += scale * acc.e2;
    u.e3 += scale * acc.e3;
    st_q(Ubar, L.n, s, dd, u);
}

// exp(w) = (cos(r/2), w sin(r/2)/r), r = |w|: wbar = J^T Mbar. The series branch
// avoids the cancellation of the closed form of (d/dr)(sin(r/2)/r)/r at small r.
template <typename T>
__device__ __forceinline__ V3<T> exp_vjp(const V3<T>& w, const Q<T>& Mb)
{
    const T r2 = w.v[0] * w.v[0] + w.v[1] * w.v[1] + w.v[2] * w.v[2];
    const T r = fsqrt(r2);
    const T thr = sizeof(T) == 4 ? T(0.5) : T(0.1);
    T f, fpr;
    if (r < thr) {
        f = T(0.5) - r2 / T(48) + r2 * r2 / T(3840) - r2 * r2 * r2 / T(645120);
        fpr = T(-1) / T(24) + r2 / T(960) - r2 * r2 / T(107520) + r2 * r2 * r2 / T(23224320);
    } else {
        const T sh = fsin(T(0.5) * r), ch = fcos(T(0.5) * r);
        f = sh / r;
        fpr = (T(0.5) * r * ch - sh) / (r2 * r);
    }
    const T dot = w.v[0] * Mb.e1 + w.v[1] * Mb.e2 + w.v[2] * Mb.e3;
    const T m0 = Mb.e0;
    return V3<T>{{T(-0.5) * f * w.v[0] * m0 + f * Mb.e1 + fpr * dot * w.v[0],
                  T(-0.5) * f * w.v[1] * m0 + f * Mb.e2 + fpr * dot * w.v[1],
                  T(-0.5) * f * w.v[2] * m0 + f * Mb.e3 + fpr * dot * w.v[2]}};
}

// reverse of the link step U' = exp(-at E) U of every link:
// Ebar += -at J_exp^T (Ubar U^+), Ubar <- exp(-at E)^+ Ubar
template <typename T>
__global__ void __launch_bounds__(kAdjBlock) k_adj_link_step(uint32_t n, const T* __restrict__ U,
                                                              const T* __restrict__ E, T* __restrict__ Ubar,
                                                              T* __restrict__ Ebar, T at)
{
    const uint32_t id = blockIdx.x * kAdjBlock + threadIdx.x;
    if (id >= 3 * n) return;
    const int d = (int)(id / n);
    const uint32_t s = id - (uint32_t)d * n;
    const V3<T> e = ld_v(E, n, s, d);
    const V3<T> w{{e.v[0] * -at, e.v[1] * -at, e.v[2] * -at}};
    const Q<T> M = exp_link(e, -at);
    const Q<T> u = ld_q(U, n, s, d);
    const Q<T> ub = ld_q(Ubar, n, s, d);
    const V3<T> wb = exp_vjp(w, mul(ub, conj(u)));
    V3<T> eb = ld_v(Ebar, n, s, d);
#pragma unroll
    for (int k = 0; k < 3; k++) eb.v[k] += -at * wb.v[k];
    st_v(Ebar, n, s, d, eb);
    st_q(Ubar, n, s, d, mul(conj(M), ub));
}

// Ebar = coefE[dir] * E (dL/dE of the electric term), Ubar = 0
template <typename T>
__global__ void __launch_bounds__(kAdjBlock) k_adj_seed(uint32_t n, const T* __restrict__ E, T* __restrict__ Ebar,
                                                         T* __restrict__ Ubar, BsqCoef ce)
{
    const uint32_t id = blockIdx.x * kAdjBlock + threadIdx.x;
    if (id >= 3 * n) return;
    const int d = (int)(id / n);
    const uint32_t s = id - (uint32_t)d * n;
    const T c = (T)ce.c[d];
    const V3<T> e = ld_v(E, n, s, d);
    st_v(Ebar, n, s, d, V3<T>{{c * e.v[0], c * e.v[1], c * e.v[2]}});
    st_q(Ubar, n, s, d, Q<T>{T(0), T(0), T(0), T(0)});
}

inline dim3 link_grid(uint32_t n) { return dim3((3 * n + kAdjBlock - 1) / kAdjBlock); }

}  // namespace

hipError_t launch_adj_seed(int prec, uint32_t n, const void* E, void* Ebar, void* Ubar, const double coefE[3],
                           hipStream_t st)
{
    BsqCoef c{{coefE[0], coefE[1], coefE[2]}};
    if (prec == 0)
        hipLaunchKernelGGL(k_adj_seed<double>, link_grid(n), dim3(kAdjBlock), 0, st, n, (const double*)E,
                           (double*)Ebar, (double*)Ubar, c);
    else
        hipLaunchKernelGGL(k_adj_seed<float>, link_grid(n), dim3(kAdjBlock), 0, st, n, (const float*)E, (float*)Ebar,
                           (float*)Ubar, c);
    return hipGetLastError();
}

hipError_t launch_adj_link_step(int prec, uint32_t n, const void* U, const void* E, void* Ubar, void* Ebar, double at,
                                hipStream_t st)
{
    if (prec == 0)
        hipLaunchKernelGGL(k_adj_link_step<double>, link_grid(n), dim3(kAdjBlock), 0, st, n, (const double*)U,
                           (const double*)E, (double*)Ubar, (double*)Ebar, at);
    else
        hipLaunchKernelGGL(k_adj_link_step<float>, link_grid(n), dim3(kAdjBlock), 0, st, n, (const float*)U,
                           (const float*)E, (float*)Ubar, (float*)Ebar, (float)at);
    return hipGetLastError();
}

hipError_t launch_adj_force(int prec, const Lattice& L, const void* U, const void* Ebar, void* Ubar, double scale,
                            hipStream_t st)
{
    BsqCoef none{{0, 0, 0}};
    if (prec == 0)
        hipLaunchKernelGGL((k_adj_gather<double, false>), link_grid(L.n), dim3(kAdjBlock), 0, st, L, (const double*)U,
                           (const double*)Ebar, (double*)Ubar, scale, none);
    else
        hipLaunchKernelGGL((k_adj_gather<float, false>), link_grid(L.n), dim3(kAdjBlock), 0, st, L, (const float*)U,
                           (const float*)Ebar, (float*)Ubar, (float)scale, none);
    return hipGetLastError();
}

hipError_t launch_adj_bsq(int prec, const Lattice& L, const void* U, void* Ubar, const double coef[3], hipStream_t st)
{
    BsqCoef c{{coef[0], coef[1], coef[2]}};
    if (prec == 0)
        hipLaunchKernelGGL((k_adj_gather<double, true>), link_grid(L.n), dim3(kAdjBlock), 0, st, L, (const double*)U,
                           (const double*)nullptr, (double*)Ubar, 1.0, c);
    else
        hipLaunchKernelGGL((k_adj_gather<float, true>), link_grid(L.n), dim3(kAdjBlock), 0, st, L, (const float*)U,
                           (const float*)nullptr, (float*)Ubar, 1.0f, c);
    return hipGetLastError();
}

}  // namespace opx
