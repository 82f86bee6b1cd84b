// opx_group.h -- SU(2)/SU(3) group and algebra arithmetic for gfx950 kernels.
//
// Conventions (identical to OpenPixi's math package):
//   SU(2): U = e0 + i e_k sigma_k (4 reals), algebra X = sum_a v_a sigma_a/2 (3 coefficients)
//          (math/SU2GroupElement.java:14, math/SU2AlgebraElement.java:5)
//   SU(3): U = 3x3 complex matrix, algebra X = sum_a c_a lambda_a/2 with the
//          standard Gell-Mann matrices (8 coefficients; the reference stores the
//          9-real Hermitian matrix, math/SU3AlgebraElement.java:3-11, and its
//          get(j) gives these coefficients, :117-129)
//   getLink(X) = exp(+i X)      (SU2AlgebraElement.java:140-152, SU3AlgebraElement.java:407-417)
//   proj(M)    = coefficients of the traceless part of (M - M^dagger)/(2i)
//                (SU2GroupElement.java:254-263, SU3GroupElement.java:494-505)
//   act(X, g)  = g X g^dagger   (AlgebraElement.java:99-110)
//
// All functions are templated on the storage/arithmetic type T (float or double).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OPX_HD __host__ __device__ __forceinline__

namespace opx {

OPX_HD float fsin(float x) { return sinf(x); }
OPX_HD double fsin(double x) { return sin(x); }
OPX_HD float fcos(float x) { return cosf(x); }
OPX_HD double fcos(double x) { return cos(x); }
OPX_HD float fsqrt(float x) { return sqrtf(x); }
OPX_HD double fsqrt(double x) { return sqrt(x); }

// ------------------------------------------------------------------ SU(2)
template <typename T>
struct Su2 {
    T e0, e1, e2, e3;
};

template <typename T>
struct Alg2 {
    T v[3];
};

template <typename T>
OPX_HD Su2<T> mul(const Su2<T>& a, const Su2<T>& b)
{
    Su2<T> r;
    r.e0 = a.e0 * b.e0 - a.e1 * b.e1 - a.e2 * b.e2 - a.e3 * b.e3;
    r.e1 = a.e0 * b.e1 + a.e1 * b.e0 - a.e2 * b.e3 + a.e3 * b.e2;
    r.e2 = a.e0 * b.e2 + a.e2 * b.e0 - a.e3 * b.e1 + a.e1 * b.e3;
    r.e3 = a.e0 * b.e3 + a.e3 * b.e0 - a.e1 * b.e2 + a.e2 * b.e1;
    return r;
}

template <typename T>
OPX_HD Su2<T> adj(const Su2<T>& a)
{
    return Su2<T>{a.e0, -a.e1, -a.e2, -a.e3};
}

// a * b^dagger
template <typename T>
OPX_HD Su2<T> mul_adj(const Su2<T>& a, const Su2<T>& b)
{
    return mul(a, adj(b));
}

// a^dagger * b
template <typename T>
OPX_HD Su2<T> adj_mul(const Su2<T>& a, const Su2<T>& b)
{
    return mul(adj(a), b);
}

template <typename T>
OPX_HD void acc(Su2<T>& s, const Su2<T>& a)
{
    s.e0 += a.e0; s.e1 += a.e1; s.e2 += a.e2; s.e3 += a.e3;
}

// s += a f
template <typename T>
OPX_HD void acc_scaled(Su2<T>& s, const Su2<T>& a, T f)
{
    s.e0 += a.e0 * f; s.e1 += a.e1 * f; s.e2 += a.e2 * f; s.e3 += a.e3 * f;
}

template <typename T>
OPX_HD Su2<T> scaled(const Su2<T>& a, T f)
{
    return Su2<T>{a.e0 * f, a.e1 * f, a.e2 * f, a.e3 * f};
}

template <typename T>
OPX_HD Alg2<T> proj(const Su2<T>& a)
{
    return Alg2<T>{{2 * a.e1, 2 * a.e2, 2 * a.e3}};
}

// exp(i * f * sum_a v_a sigma_a / 2) (SU2AlgebraElement.getLink applied to v*f): with
// m = |w| / 2, e0 = cos m and e_a = w_a sin(m) / (2 m). In float, for m^2 <= kExpPolySu2M2
// both are their Taylor series in m^2 through m^8 (no square root, divide or sine); the
// first omitted terms, m^10 / 10! and m^10 / 11!, are below 3e-10 there.
constexpr double kExpPolySu2M2 = 0.25;
template <typename T>
OPX_HD Su2<T> exp_link(const Alg2<T>& v, T f)
{
    T w0 = v.v[0] * f, w1 = v.v[1] * f, w2 = v.v[2] * f;
    T sum = w0 * w0 + w1 * w1 + w2 * w2;
    const T m2 = sum * T(0.25);
    if (sizeof(T) == 4 && m2 <= T(kExpPolySu2M2)) {
        const T c = T(1) + m2 * (T(-1.0 / 2) + m2 * (T(1.0 / 24) + m2 * (T(-1.0 / 720) + m2 * T(1.0 / 40320))));
        const T h = T(0.5) + m2 * (T(-0.5 / 6) + m2 * (T(0.5 / 120) + m2 * (T(-0.5 / 5040) + m2 * T(0.5 / 362880))));
        const T sinfakt = (m2 < T(1e-40)) ? T(0) : h;
        return Su2<T>{c, w0 * sinfakt, w1 * sinfakt, w2 * sinfakt};
    }
    T mod = fsqrt(sum) / 2;
    T sinfakt = (mod < T(1e-20)) ? T(0) : T(0.5) / mod * fsin(mod);
    return Su2<T>{fcos(mod), w0 * sinfakt, w1 * sinfakt, w2 * sinfakt};
}

// g X g^dagger
template <typename T>
OPX_HD Alg2<T> act(const Alg2<T>& x, const Su2<T>& g)
{
    Su2<T> xm{T(0), x.v[0] / 2, x.v[1] / 2, x.v[2] / 2};
    return proj(mul(g, mul(xm, adj(g))));
}

template <typename T>
OPX_HD T square(const Alg2<T>& x)
{
    return x.v[0] * x.v[0] + x.v[1] * x.v[1] + x.v[2] * x.v[2];
}

// ------------------------------------------------------------------ SU(3)
template <typename T>
struct Su3 {
    T r[9];   // real parts, row-major
    T i[9];   // imaginary parts, row-major
};

template <typename T>
struct Alg3 {
    T c[8];
};

// Complex 3x3 products written as chains of single multiply-adds (one fused
// multiply-add per real term), e.g. re = a*b; re -= c*d; re += e*f; ...

// a * b
template <typename T>
OPX_HD Su3<T> mul(const Su3<T>& a, const Su3<T>& b)
{
    Su3<T> o;
#pragma unroll
    for (int r = 0; r < 3; r++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T re = a.r[3 * r] * b.r[c];
            re -= a.i[3 * r] * b.i[c];
            T im = a.r[3 * r] * b.i[c];
            im += a.i[3 * r] * b.r[c];
#pragma unroll
            for (int k = 1; k < 3; k++) {
                re += a.r[3 * r + k] * b.r[3 * k + c];
                re -= a.i[3 * r + k] * b.i[3 * k + c];
                im += a.r[3 * r + k] * b.i[3 * k + c];
                im += a.i[3 * r + k] * b.r[3 * k + c];
            }
            o.r[3 * r + c] = re;
            o.i[3 * r + c] = im;
        }
    }
    return o;
}

template <typename T>
OPX_HD Su3<T> adj(const Su3<T>& a)
{
    Su3<T> o;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            o.r[3 * r + c] = a.r[3 * c + r];
            o.i[3 * r + c] = -a.i[3 * c + r];
        }
    return o;
}

// entry (r, c) of a * b^dagger: sum_k a[r][k] * conj(b[c][k])
template <typename T>
OPX_HD void mul_adj_entry(const Su3<T>& a, const Su3<T>& b, int r, int c, T& re, T& im)
{
    re = a.r[3 * r] * b.r[3 * c];
    re += a.i[3 * r] * b.i[3 * c];
    im = a.i[3 * r] * b.r[3 * c];
    im -= a.r[3 * r] * b.i[3 * c];
#pragma unroll
    for (int k = 1; k < 3; k++) {
        re += a.r[3 * r + k] * b.r[3 * c + k];
        re += a.i[3 * r + k] * b.i[3 * c + k];
        im += a.i[3 * r + k] * b.r[3 * c + k];
        im -= a.r[3 * r + k] * b.i[3 * c + k];
    }
}

// a * b^dagger
template <typename T>
OPX_HD Su3<T> mul_adj(const Su3<T>& a, const Su3<T>& b)
{
    Su3<T> o;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) mul_adj_entry(a, b, r, c, o.r[3 * r + c], o.i[3 * r + c]);
    return o;
}

// a^dagger * b
template <typename T>
OPX_HD Su3<T> adj_mul(const Su3<T>& a, const Su3<T>& b)
{
    Su3<T> o;
#pragma unroll
    for (int r = 0; r < 3; r++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            // sum_k conj(a[k][r]) * b[k][c]
            T re = a.r[r] * b.r[c];
            re += a.i[r] * b.i[c];
            T im = a.r[r] * b.i[c];
            im -= a.i[r] * b.r[c];
#pragma unroll
            for (int k = 1; k < 3; k++) {
                re += a.r[3 * k + r] * b.r[3 * k + c];
                re += a.i[3 * k + r] * b.i[3 * k + c];
                im += a.r[3 * k + r] * b.i[3 * k + c];
                im -= a.i[3 * k + r] * b.r[3 * k + c];
            }
            o.r[3 * r + c] = re;
            o.i[3 * r + c] = im;
        }
    }
    return o;
}

template <typename T>
OPX_HD void acc(Su3<T>& s, const Su3<T>& a)
{
#pragma unroll
    for (int k = 0; k < 9; k++) {
        s.r[k] += a.r[k];
        s.i[k] += a.i[k];
    }
}

// s += a f
template <typename T>
OPX_HD void acc_scaled(Su3<T>& s, const Su3<T>& a, T f)
{
#pragma unroll
    for (int k = 0; k < 9; k++) {
        s.r[k] += a.r[k] * f;
        s.i[k] += a.i[k] * f;
    }
}

template <typename T>
OPX_HD Su3<T> scaled(const Su3<T>& a, T f)
{
    Su3<T> o;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        o.r[k] = a.r[k] * f;
        o.i[k] = a.i[k] * f;
    }
    return o;
}

template <typename T>
OPX_HD T inv_sqrt3()
{
    return T(0.57735026918962576450914878050195746);
}

// Gell-Mann coefficients of the traceless part of (M - M^dagger)/(2i).
template <typename T>
OPX_HD Alg3<T> proj(const Su3<T>& m)
{
    Alg3<T> a;
    a.c[0] = m.i[1] + m.i[3];
    a.c[1] = m.r[1] - m.r[3];
    a.c[2] = m.i[0] - m.i[4];
    a.c[3] = m.i[2] + m.i[6];
    a.c[4] = m.r[2] - m.r[6];
    a.c[5] = m.i[5] + m.i[7];
    a.c[6] = m.r[5] - m.r[7];
    a.c[7] = (m.i[0] + m.i[4] - 2 * m.i[8]) * inv_sqrt3<T>();
    return a;
}

// Hermitian matrix H = sum_a c_a lambda_a / 2 (scaled by f)
template <typename T>
OPX_HD Su3<T> herm(const Alg3<T>& x, T f)
{
    const T h = T(0.5) * f;
    const T d8 = x.c[7] * inv_sqrt3<T>() * h;
    Su3<T> m;
    m.r[0] = x.c[2] * h + d8;  m.i[0] = 0;
    m.r[4] = -x.c[2] * h + d8; m.i[4] = 0;
    m.r[8] = -2 * d8;          m.i[8] = 0;
    m.r[1] = x.c[0] * h; m.i[1] = -x.c[1] * h;
    m.r[3] = x.c[0] * h; m.i[3] = x.c[1] * h;
    m.r[2] = x.c[3] * h; m.i[2] = -x.c[4] * h;
    m.r[6] = x.c[3] * h; m.i[6] = x.c[4] * h;
    m.r[5] = x.c[5] * h; m.i[5] = -x.c[6] * h;
    m.r[7] = x.c[5] * h; m.i[7] = x.c[6] * h;
    return m;
}

// Coefficients of a Hermitian matrix (inverse of herm with f = 1); uses the
// Hermitian part of the off-diagonal entries.
template <typename T>
OPX_HD Alg3<T> herm_coeffs(const Su3<T>& r)
{
    Alg3<T> a;
    a.c[0] = r.r[1] + r.r[3];
    a.c[1] = r.i[3] - r.i[1];
    a.c[2] = r.r[0] - r.r[4];
    a.c[3] = r.r[2] + r.r[6];
    a.c[4] = r.i[6] - r.i[2];
    a.c[5] = r.r[5] + r.r[7];
    a.c[6] = r.i[7] - r.i[5];
    a.c[7] = (r.r[0] + r.r[4] - 2 * r.r[8]) * inv_sqrt3<T>();
    return a;
}

// g X g^dagger
template <typename T>
OPX_HD Alg3<T> act(const Alg3<T>& x, const Su3<T>& g)
{
    Su3<T> h = herm(x, T(1));
    return herm_coeffs(mul(g, mul_adj(h, g)));
}

template <typename T>
OPX_HD T square(const Alg3<T>& x)
{
    T s = x.c[0] * x.c[0];
#pragma unroll
    for (int k = 1; k < 8; k++) s += x.c[k] * x.c[k];
    return s;
}

template <typename T> struct ExpTerms;
template <> struct ExpTerms<double> { static constexpr int n = 18; static constexpr double eps = 1e-18; };
template <> struct ExpTerms<float> { static constexpr int n = 10; static constexpr double eps = 1e-9; };

// Q^2 for Hermitian Q: real diagonal, upper triangle computed, lower = conj(upper)
template <typename T>
OPX_HD Su3<T> herm_square(const Su3<T>& q)
{
    Su3<T> o;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        T d = q.r[3 * r] * q.r[3 * r];
        d += q.i[3 * r] * q.i[3 * r];
        d += q.r[3 * r + 1] * q.r[3 * r + 1];
        d += q.i[3 * r + 1] * q.i[3 * r + 1];
        d += q.r[3 * r + 2] * q.r[3 * r + 2];
        d += q.i[3 * r + 2] * q.i[3 * r + 2];
        o.r[4 * r] = d;
        o.i[4 * r] = 0;
    }
    const int rr[3] = {0, 0, 1}, cc[3] = {1, 2, 2};
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const int r = rr[t], c = cc[t];
        T re = q.r[3 * r] * q.r[c];
        re -= q.i[3 * r] * q.i[c];
        T im = q.r[3 * r] * q.i[c];
        im += q.i[3 * r] * q.r[c];
#pragma unroll
        for (int k = 1; k < 3; k++) {
            re += q.r[3 * r + k] * q.r[3 * k + c];
            re -= q.i[3 * r + k] * q.i[3 * k + c];
            im += q.r[3 * r + k] * q.i[3 * k + c];
            im += q.i[3 * r + k] * q.r[3 * k + c];
        }
        o.r[3 * r + c] = re;
        o.i[3 * r + c] = im;
        o.r[3 * c + r] = re;
        o.i[3 * c + r] = -im;
    }
    return o;
}

// exp(i Q) for Q = H(x) * f traceless Hermitian, via the Cayley-Hamilton reduced
// power series: (iQ)^n / n! = a_n + b_n Q + c_n Q^2 with
//   Q^3 = k1 Q + k0,  k1 = tr(Q^2)/2, k0 = det Q = tr(Q^3)/3,
//   (a,b,c)_{n+1} = i (c_n k0, a_n + c_n k1, b_n) / (n+1).
// The series stops once a term is below the working precision (at most
// ExpTerms<T>::n terms). When tr(Q^2) > 1 the argument is halved s times and the
// result squared s times, so the series always runs with spectral radius <= 1.
// In float, for tr(Q^2) <= kExpPolyTrQ2 the terms n <= 8 are summed in closed form:
//   f0 = 1 - k0^2/720 + k0^2 k1/20160 + i (-k0/6 + k0 k1/120 - k0 k1^2/5040)
//   f1 = k0/24 - k0 k1/360 + k0 k1^2/13440 + i (1 - k1/6 + k1^2/120 - (k0^2 + k1^3)/5040)
//   f2 = -1/2 + k1/24 - k1^2/720 + (k0^2 + k1^3)/40320 + i (k0/120 - k0 k1/2520)
// (the recursion above unrolled); the spectral radius is then at most sqrt(2/3 tr Q^2) = 0.41,
// so the first omitted term is below 1e-9, under float rounding.
constexpr double kExpPolyTrQ2 = 0.25;
template <typename T>
OPX_HD Su3<T> exp_link(const Alg3<T>& x, T f)
{
    Su3<T> q = herm(x, f);
    Su3<T> q2 = herm_square(q);
    T trq2 = q2.r[0] + q2.r[4] + q2.r[8];
    const bool poly = sizeof(T) == 4 && trq2 <= T(kExpPolyTrQ2);
    int s = 0;
    T sc = T(1);
    while (!poly && trq2 * sc * sc > T(1) && s < 30) {
        sc *= T(0.5);
        s++;
    }
    if (s > 0) {
        q = scaled(q, sc);
        q2 = scaled(q2, sc * sc);
        trq2 = trq2 * sc * sc;
    }
    const T k1 = trq2 / 2;
    // tr(Q^3) = sum_ij (Q^2)_ij Q_ji, real for Hermitian Q
    T trq3 = q2.r[0] * q.r[0];
    trq3 += q2.r[4] * q.r[4];
    trq3 += q2.r[8] * q.r[8];
    {
        const int rr[3] = {0, 0, 1}, cc[3] = {1, 2, 2};
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const int r = rr[t], c = cc[t];
            // (Q^2)_rc Q_cr + (Q^2)_cr Q_rc = 2 Re((Q^2)_rc Q_cr)
            T v = q2.r[3 * r + c] * q.r[3 * c + r];
            v -= q2.i[3 * r + c] * q.i[3 * c + r];
            trq3 += 2 * v;
        }
    }
    const T k0 = trq3 / 3;
    T f0r = 1, f0i = 0, f1r = 0, f1i = 0, f2r = 0, f2i = 0;
    if (poly) {
        const T k00 = k0 * k0, k11 = k1 * k1;
        const T k0k1 = k0 * k1;
        const T h = k00 + k11 * k1;
        f0r = T(1) - k00 * (T(1.0 / 720) - k1 * T(1.0 / 20160));
        f0i = k0 * (T(-1.0 / 6) + k1 * T(1.0 / 120) - k11 * T(1.0 / 5040));
        f1r = k0 * (T(1.0 / 24) - k1 * T(1.0 / 360) + k11 * T(1.0 / 13440));
        f1i = T(1) - k1 * T(1.0 / 6) + k11 * T(1.0 / 120) - h * T(1.0 / 5040);
        f2r = T(-0.5) + k1 * T(1.0 / 24) - k11 * T(1.0 / 720) + h * T(1.0 / 40320);
        f2i = k0 * T(1.0 / 120) - k0k1 * T(1.0 / 2520);
    } else {
    // bound on the size of the terms: |a| + |b| rho + |c| rho^2 with rho^2 <= tr Q^2
    const T rho = fsqrt(trq2);
    T ar = 1, ai = 0, br = 0, bi = 0, cr = 0, ci = 0;
#pragma unroll
    for (int n = 0; n < ExpTerms<T>::n; n++) {
        const T inv = T(1) / T(n + 1);
        // new a = i c k0; new b = i (a + c k1); new c = i b
        const T nar = -ci * k0, nai = cr * k0;
        T tbr = ar;
        tbr += cr * k1;
        T tbi = ai;
        tbi += ci * k1;
        const T ncr = -bi, nci = br;
        ar = nar * inv;
        ai = nai * inv;
        br = -tbi * inv;
        bi = tbr * inv;
        cr = ncr * inv;
        ci = nci * inv;
        f0r += ar; f0i += ai;
        f1r += br; f1i += bi;
        f2r += cr; f2i += ci;
        const T mag = (fabs(ar) + fabs(ai)) + (fabs(br) + fabs(bi)) * rho + (fabs(cr) + fabs(ci)) * trq2;
        if (mag < T(ExpTerms<T>::eps)) break;
    }
    }
    // e = f0 + f1 Q + f2 Q^2 (Q and Q^2 Hermitian: real diagonals)
    Su3<T> e;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        T re = f1r * q.r[k];
        re -= f1i * q.i[k];
        re += f2r * q2.r[k];
        re -= f2i * q2.i[k];
        T im = f1r * q.i[k];
        im += f1i * q.r[k];
        im += f2r * q2.i[k];
        im += f2i * q2.r[k];
        e.r[k] = re;
        e.i[k] = im;
    }
    e.r[0] += f0r; e.i[0] += f0i;
    e.r[4] += f0r; e.i[4] += f0i;
    e.r[8] += f0r; e.i[8] += f0i;
    for (int k = 0; k < s; k++) e = mul(e, e);
    return e;
}

// Third row of an SU(3) matrix from its first two: u_2j = conj(u_0a u_1b - u_0b u_1a),
// (j, a, b) cyclic. Exact for det = 1 unitary matrices.
// Each real and imaginary part is one multiply and three fused multiply-adds.
template <typename T>
OPX_HD void su3_third_row(Su3<T>& m)
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int a = (j + 1) % 3, b = (j + 2) % 3;
        // Re(u_0a u_1b - u_0b u_1a)
        T cr = m.r[a] * m.r[3 + b];
        cr -= m.i[a] * m.i[3 + b];
        cr -= m.r[b] * m.r[3 + a];
        cr += m.i[b] * m.i[3 + a];
        // -Im(u_0a u_1b - u_0b u_1a)
        T ci = m.r[b] * m.i[3 + a];
        ci += m.i[b] * m.r[3 + a];
        ci -= m.r[a] * m.i[3 + b];
        ci -= m.i[a] * m.r[3 + b];
        m.r[6 + j] = cr;
        m.i[6 + j] = ci;
    }
}

// Products of SU(3) elements through their first two rows: rows 0 and 1 are computed and row 2,
// where wanted, is rebuilt as conj(row0 x row1), exact for group elements. The left factor of a
// * b and a * b^dagger is read in its first two rows only.
// rows 0 and 1 of a * b^dagger (row 2 of the result left unset)
template <typename T>
OPX_HD Su3<T> mul_adj_r01(const Su3<T>& a, const Su3<T>& b)
{
    Su3<T> o;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) mul_adj_entry(a, b, r, c, o.r[3 * r + c], o.i[3 * r + c]);
#pragma unroll
    for (int c = 6; c < 9; c++) o.r[c] = o.i[c] = 0;
    return o;
}

// rows 0 and 1 of a * b (row 2 of the result left unset)
template <typename T>
OPX_HD Su3<T> mul_r01(const Su3<T>& a, const Su3<T>& b)
{
    Su3<T> o;
#pragma unroll
    for (int r = 0; r < 2; r++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T re = a.r[3 * r] * b.r[c];
            re -= a.i[3 * r] * b.i[c];
            T im = a.r[3 * r] * b.i[c];
            im += a.i[3 * r] * b.r[c];
#pragma unroll
            for (int k = 1; k < 3; k++) {
                re += a.r[3 * r + k] * b.r[3 * k + c];
                re -= a.i[3 * r + k] * b.i[3 * k + c];
                im += a.r[3 * r + k] * b.i[3 * k + c];
                im += a.i[3 * r + k] * b.r[3 * k + c];
            }
            o.r[3 * r + c] = re;
            o.i[3 * r + c] = im;
        }
    }
#pragma unroll
    for (int c = 6; c < 9; c++) o.r[c] = o.i[c] = 0;
    return o;
}

// rows 0 and 1 of a^dagger * b (reads columns 0 and 1 of a, all of b; row 2 left unset)
template <typename T>
OPX_HD Su3<T> adj_mul_r01(const Su3<T>& a, const Su3<T>& b)
{
    Su3<T> o;
#pragma unroll
    for (int r = 0; r < 2; r++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T re = a.r[r] * b.r[c];
            re += a.i[r] * b.i[c];
            T im = a.r[r] * b.i[c];
            im -= a.i[r] * b.r[c];
#pragma unroll
            for (int k = 1; k < 3; k++) {
                re += a.r[3 * k + r] * b.r[3 * k + c];
                re += a.i[3 * k + r] * b.i[3 * k + c];
                im += a.r[3 * k + r] * b.i[3 * k + c];
                im -= a.i[3 * k + r] * b.r[3 * k + c];
            }
            o.r[3 * r + c] = re;
            o.i[3 * r + c] = im;
        }
    }
#pragma unroll
    for (int c = 6; c < 9; c++) o.r[c] = o.i[c] = 0;
    return o;
}

// Reprojection onto the group in double precision: SU(2) normalised; SU(3) rows 0 and 1
// Gram-Schmidt orthonormalised, row 2 = conj(row0 x row1).
OPX_HD Su2<double> reunit(Su2<double> u)
{
    double nrm = sqrt(u.e0 * u.e0 + u.e1 * u.e1 + u.e2 * u.e2 + u.e3 * u.e3);
    double f = 1.0 / nrm;
    return Su2<double>{u.e0 * f, u.e1 * f, u.e2 * f, u.e3 * f};
}

OPX_HD Su3<double> reunit(Su3<double> u)
{
    // row 0
    double n0 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) n0 += u.r[k] * u.r[k] + u.i[k] * u.i[k];
    double f0 = 1.0 / sqrt(n0);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        u.r[k] *= f0;
        u.i[k] *= f0;
    }
    // row 1 -= <row0, row1> row0
    double pr = 0, pi = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        // conj(r0_k) * r1_k
        pr += u.r[k] * u.r[3 + k] + u.i[k] * u.i[3 + k];
        pi += u.r[k] * u.i[3 + k] - u.i[k] * u.r[3 + k];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        u.r[3 + k] -= pr * u.r[k] - pi * u.i[k];
        u.i[3 + k] -= pr * u.i[k] + pi * u.r[k];
    }
    double n1 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) n1 += u.r[3 + k] * u.r[3 + k] + u.i[3 + k] * u.i[3 + k];
    double f1 = 1.0 / sqrt(n1);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        u.r[3 + k] *= f1;
        u.i[3 + k] *= f1;
    }
    // row 2 = conj(row0 x row1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        double cr = (u.r[a] * u.r[3 + b] - u.i[a] * u.i[3 + b]) - (u.r[b] * u.r[3 + a] - u.i[b] * u.i[3 + a]);
        double ci = (u.r[a] * u.i[3 + b] + u.i[a] * u.r[3 + b]) - (u.r[b] * u.i[3 + a] + u.i[b] * u.r[3 + a]);
        u.r[6 + k] = cr;
        u.i[6 + k] = -ci;
    }
    return u;
}

// --------------------------------------------------- group-generic traits
// G: reals of a group element; S: reals stored per link in device fields, slot(k) the
// element index of stored value k and complete() what rebuilds the rest after a load.
// FP32 SU(3) links store their first two rows (12 reals, S = 12); every other form
// stores the whole element.
template <int NC, typename T> struct Group;

template <typename T>
struct Group<2, T> {
    using M = Su2<T>;
    using A = Alg2<T>;
    static constexpr int G = 4;   // reals per group element
    static constexpr int S = 4;   // reals stored per link
    static constexpr int AD = 3;  // algebra coefficients
    OPX_HD static constexpr int slot(int k) { return k; }
    OPX_HD static void complete(M&) {}
    OPX_HD static M zero() { return M{0, 0, 0, 0}; }
    OPX_HD static M identity() { return M{1, 0, 0, 0}; }
    OPX_HD static A azero() { return A{{0, 0, 0}}; }
    OPX_HD static T& gref(M& m, int k) { return k == 0 ? m.e0 : (k == 1 ? m.e1 : (k == 2 ? m.e2 : m.e3)); }
    OPX_HD static T gget(const M& m, int k) { return k == 0 ? m.e0 : (k == 1 ? m.e1 : (k == 2 ? m.e2 : m.e3)); }
};

template <typename T>
struct Group<3, T> {
    using M = Su3<T>;
    using A = Alg3<T>;
    static constexpr int G = 18;
    static constexpr int S = sizeof(T) == 4 ? 12 : 18;
    static constexpr int AD = 8;
    OPX_HD static constexpr int slot(int k) { return S == G ? k : (k < 6 ? k : k + 3); }
    OPX_HD static void complete(M& m)
    {
        if constexpr (S != G) su3_third_row(m);
    }
    OPX_HD static M zero()
    {
        M m;
#pragma unroll
        for (int k = 0; k < 9; k++) m.r[k] = m.i[k] = 0;
        return m;
    }
    OPX_HD static M identity()
    {
        M m = zero();
        m.r[0] = m.r[4] = m.r[8] = 1;
        return m;
    }
    OPX_HD static A azero()
    {
        A a;
#pragma unroll
        for (int k = 0; k < 8; k++) a.c[k] = 0;
        return a;
    }
    OPX_HD static T& gref(M& m, int k) { return k < 9 ? m.r[k] : m.i[k - 9]; }
    OPX_HD static T gget(const M& m, int k) { return k < 9 ? m.r[k] : m.i[k - 9]; }
};

// ---------------------------------------------------------------- reduced products
// proj(a b^dagger): only the entries the projection reads.
template <typename T>
OPX_HD Alg2<T> proj_mul_adj(const Su2<T>& a, const Su2<T>& b)
{
    return proj(mul(a, adj(b)));
}

template <typename T>
OPX_HD Alg3<T> proj_mul_adj(const Su3<T>& a, const Su3<T>& b)
{
    // projection of m = a b^dagger: imaginary diagonal and the off-diagonal pairs
    T d[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        T im = a.i[3 * r] * b.r[3 * r];
        im -= a.r[3 * r] * b.i[3 * r];
#pragma unroll
        for (int k = 1; k < 3; k++) {
            im += a.i[3 * r + k] * b.r[3 * r + k];
            im -= a.r[3 * r + k] * b.i[3 * r + k];
        }
        d[r] = im;
    }
    T m01r, m01i, m10r, m10i, m02r, m02i, m20r, m20i, m12r, m12i, m21r, m21i;
    mul_adj_entry(a, b, 0, 1, m01r, m01i);
    mul_adj_entry(a, b, 1, 0, m10r, m10i);
    mul_adj_entry(a, b, 0, 2, m02r, m02i);
    mul_adj_entry(a, b, 2, 0, m20r, m20i);
    mul_adj_entry(a, b, 1, 2, m12r, m12i);
    mul_adj_entry(a, b, 2, 1, m21r, m21i);
    Alg3<T> o;
    o.c[0] = m01i + m10i;
    o.c[1] = m01r - m10r;
    o.c[2] = d[0] - d[1];
    o.c[3] = m02i + m20i;
    o.c[4] = m02r - m20r;
    o.c[5] = m12i + m21i;
    o.c[6] = m12r - m21r;
    o.c[7] = (d[0] + d[1] - 2 * d[2]) * inv_sqrt3<T>();
    return o;
}

// act(x, v^dagger) = v^dagger X v, X = sum_a x_a t_a
template <typename T>
OPX_HD Alg2<T> act_adj(const Alg2<T>& x, const Su2<T>& v)
{
    return act(x, adj(v));
}

template <typename T>
OPX_HD Alg3<T> act_adj(const Alg3<T>& x, const Su3<T>& v)
{
    // W = H v, then R = v^dagger W (Hermitian: diagonal real parts and upper triangle)
    const Su3<T> h = herm(x, T(1));
    T wr[9], wi[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            // h_rr real
            T re = h.r[4 * r] * v.r[3 * r + c];
            T im = h.r[4 * r] * v.i[3 * r + c];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if (k == r) continue;
                re += h.r[3 * r + k] * v.r[3 * k + c];
                re -= h.i[3 * r + k] * v.i[3 * k + c];
                im += h.r[3 * r + k] * v.i[3 * k + c];
                im += h.i[3 * r + k] * v.r[3 * k + c];
            }
            wr[3 * r + c] = re;
            wi[3 * r + c] = im;
        }
    // R_rc = sum_k conj(v_kr) w_kc
    T d[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        T s = v.r[r] * wr[r];
        s += v.i[r] * wi[r];
        s += v.r[3 + r] * wr[3 + r];
        s += v.i[3 + r] * wi[3 + r];
        s += v.r[6 + r] * wr[6 + r];
        s += v.i[6 + r] * wi[6 + r];
        d[r] = s;
    }
    T orr[3], oii[3];  // R_01, R_02, R_12
    const int rr[3] = {0, 0, 1}, cc[3] = {1, 2, 2};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int r = rr[q], c = cc[q];
        T re = v.r[r] * wr[c];
        re += v.i[r] * wi[c];
        T im = v.r[r] * wi[c];
        im -= v.i[r] * wr[c];
#pragma unroll
        for (int k = 1; k < 3; k++) {
            re += v.r[3 * k + r] * wr[3 * k + c];
            re += v.i[3 * k + r] * wi[3 * k + c];
            im += v.r[3 * k + r] * wi[3 * k + c];
            im -= v.i[3 * k + r] * wr[3 * k + c];
        }
        orr[q] = re;
        oii[q] = im;
    }
    // coefficients of the Hermitian R (R_10 = conj R_01, ...)
    Alg3<T> o;
    o.c[0] = 2 * orr[0];
    o.c[1] = -2 * oii[0];
    o.c[2] = d[0] - d[1];
    o.c[3] = 2 * orr[1];
    o.c[4] = -2 * oii[1];
    o.c[5] = 2 * orr[2];
    o.c[6] = -2 * oii[2];
    o.c[7] = (d[0] + d[1] - 2 * d[2]) * inv_sqrt3<T>();
    return o;
}

template <typename T>
OPX_HD T& aref(Alg2<T>& a, int k) { return a.v[k]; }
template <typename T>
OPX_HD T& aref(Alg3<T>& a, int k) { return a.c[k]; }
template <typename T>
OPX_HD T aget(const Alg2<T>& a, int k) { return a.v[k]; }
template <typename T>
OPX_HD T aget(const Alg3<T>& a, int k) { return a.c[k]; }

}  // namespace opx
