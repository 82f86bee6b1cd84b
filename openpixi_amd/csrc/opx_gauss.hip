// opx_gauss.hip -- Gauss-law projection of the electric field (double precision).
//
// The Gauss constraint of Grid.getGaussConstraint (physics/grid/Grid.java:944-968) is
//   G(x) = sum_i w_i [E_i(x) - U_i(x-i)^+ E_i(x-i) U_i(x-i)] - rho(x),  w_i = 1/(g a_i^2),
// i.e. G = D^T E - rho with the covariant difference
//   (D phi)_i(x) = w_i [phi(x) - U_i(x) phi(x+i) U_i(x)^+]
// (the adjoint action is orthogonal on generator coefficients, so D^T is the
// transpose of D). The projected field E' = E - D lambda with D^T D lambda = G
// satisfies G' = 0 and is the closest such field to E. D^T D is the covariant
// lattice Laplacian; it is solved by conjugate gradients on the host driving the
// kernels below. Every reduction is a fixed-order two-level sum. The context's fields (U, Unext,
// E, rho) are in the device block layout (fpos); the double work arrays of the solver (phi, F,
// G, R, the CG vectors) are [component][cell] with n cells each.
#include "opx_group.h"
#include "opx_kernels.h"

namespace opx {

namespace {

constexpr int kB = 256;
constexpr int kDotBlocks = 1024;

__device__ __forceinline__ uint32_t nb_fwd(const Lattice& L, uint32_t site, uint32_t x, int d)
{
    const uint32_t n = d == 0 ? L.d0 : (d == 1 ? L.d1 : L.d2);
    const uint32_t s = d == 0 ? L.s0 : (d == 1 ? L.s1 : 1u);
    return (x == n - 1) ? site - (n - 1) * s : site + s;
}
__device__ __forceinline__ uint32_t nb_bwd(const Lattice& L, uint32_t site, uint32_t x, int d)
{
    const uint32_t n = d == 0 ? L.d0 : (d == 1 ? L.d1 : L.d2);
    const uint32_t s = d == 0 ? L.s0 : (d == 1 ? L.s1 : 1u);
    return (x == 0) ? site + (n - 1) * s : site - s;
}
__device__ __forceinline__ void coords3(const Lattice& L, uint32_t site, uint32_t* x)
{
    const uint32_t q = site / L.d2;
    x[2] = site - q * L.d2;
    x[0] = q / L.d1;
    x[1] = q - x[0] * L.d1;
}

template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, double>::M link_d(const T* __restrict__ U, uint32_t site, int d)
{
    using GD = Group<NC, double>;
    using GR = Group<NC, T>;
    typename GD::M m;
    const T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) GD::gref(m, GR::slot(k)) = (double)p[fcomp<T>(d * GR::S + k, 3 * GR::S)];
    if constexpr (GR::S != GR::G) su3_third_row(m);   // rebuilt in double
    return m;
}

template <int NC>
__device__ __forceinline__ typename Group<NC, double>::A alg_at(const double* __restrict__ X, uint32_t n, uint32_t site,
                                                                int d)
{
    using GD = Group<NC, double>;
    typename GD::A a;
#pragma unroll
    for (int k = 0; k < GD::AD; k++) aref(a, k) = X[((size_t)d * GD::AD + k) * n + site];
    return a;
}

// F = D phi   (F: [dir][coef][cell], phi: [coef][cell])
template <int NC, typename T>
__global__ __launch_bounds__(kB) void k_apply_D(Lattice L, const T* __restrict__ U, const double* __restrict__ phi,
                                                double* __restrict__ F)
{
    using GD = Group<NC, double>;
    const uint32_t site = blockIdx.x * kB + threadIdx.x;
    const int d = blockIdx.y;
    if (site >= L.n) return;
    uint32_t x[3];
    coords3(L, site, x);
    const double w = 1.0 / (L.g * L.as[d] * L.as[d]);
    const uint32_t xp = nb_fwd(L, site, x[d], d);
    const auto u = link_d<NC, T>(U, site, d);
    const auto t = act(alg_at<NC>(phi, L.n, xp, 0), u);
    const auto p = alg_at<NC>(phi, L.n, site, 0);
#pragma unroll
    for (int k = 0; k < GD::AD; k++) F[((size_t)d * GD::AD + k) * L.n + site] = w * (aget(p, k) - aget(t, k));
}

// out = D^T F - rho  (rho may be null), or D^T F when sub_rho == 0
template <int NC, typename T>
__global__ __launch_bounds__(kB) void k_apply_Dt(Lattice L, const T* __restrict__ U, const double* __restrict__ F,
                                                 const T* __restrict__ rho, double* __restrict__ out)
{
    using GD = Group<NC, double>;
    const uint32_t site = blockIdx.x * kB + threadIdx.x;
    if (site >= L.n) return;
    uint32_t x[3];
    coords3(L, site, x);
    typename GD::A g = GD::azero();
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double w = 1.0 / (L.g * L.as[d] * L.as[d]);
        const uint32_t xm = nb_bwd(L, site, x[d], d);
        const auto u = link_d<NC, T>(U, xm, d);
        const auto t = act(alg_at<NC>(F, L.n, xm, d), adj(u));
        const auto f = alg_at<NC>(F, L.n, site, d);
#pragma unroll
        for (int k = 0; k < GD::AD; k++) aref(g, k) += w * (aget(f, k) - aget(t, k));
    }
    if (rho) {
#pragma unroll
        for (int k = 0; k < GD::AD; k++) aref(g, k) -= (double)rho[fposT<T>(k, site, GD::AD)];
    }
#pragma unroll
    for (int k = 0; k < GD::AD; k++) out[(size_t)k * L.n + site] = aget(g, k);
}

// E (precision T) -> double copy, and back with E -= F
// GaussConstraintRestoration.computeGaussViolation (diagnostics/methods/GaussConstraintRestoration.java:
// 117-135): the violation C = sum_j [Unext_j(x) U_j(x)^+ - U_j(x-j)^+ Unext_j(x-j)] / (a_t a) enters
// only through its projection (the projection is linear, so it is the sum of the terms'
// projections): G = proj C, and R = G - rho, whose mean square is the total violation.
template <int NC, typename T>
__global__ __launch_bounds__(kB) void k_restore_violation(Lattice L, const T* __restrict__ U,
                                                          const T* __restrict__ Unext, const T* __restrict__ rho,
                                                          double factor, double* __restrict__ G,
                                                          double* __restrict__ R)
{
    using GD = Group<NC, double>;
    const uint32_t site = blockIdx.x * kB + threadIdx.x;
    if (site >= L.n) return;
    uint32_t x[3];
    coords3(L, site, x);
    typename GD::A g = GD::azero();
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const uint32_t xm = nb_bwd(L, site, x[j], j);
        const auto c1 = proj(mul_adj(link_d<NC, T>(Unext, site, j), link_d<NC, T>(U, site, j)));
        const auto c2 = proj(adj_mul(link_d<NC, T>(U, xm, j), link_d<NC, T>(Unext, xm, j)));
#pragma unroll
        for (int k = 0; k < GD::AD; k++) aref(g, k) += aget(c1, k) - aget(c2, k);
    }
#pragma unroll
    for (int k = 0; k < GD::AD; k++) {
        const double v = aget(g, k) * factor;
        G[(size_t)k * L.n + site] = v;
        R[(size_t)k * L.n + site] = v - (rho ? (double)rho[fposT<T>(k, site, GD::AD)] : 0.0);
    }
}

// GaussConstraintRestoration.applyCorrection (:142-158), every cell and direction:
//   E_j(x) += f2 proj(U_j(x) C(x+j) U_j(x)^+ - C(x)) = f2 [act(G(x+j), U_j(x)) - G(x)]
template <int NC, typename T>
__global__ __launch_bounds__(kB) void k_restore_correct(Lattice L, const T* __restrict__ U,
                                                        const double* __restrict__ G, double f2, T* __restrict__ E)
{
    using GD = Group<NC, double>;
    const uint32_t site = blockIdx.x * kB + threadIdx.x;
    const int d = blockIdx.y;
    if (site >= L.n) return;
    uint32_t x[3];
    coords3(L, site, x);
    const uint32_t xp = nb_fwd(L, site, x[d], d);
    const auto t = act(alg_at<NC>(G, L.n, xp, 0), link_d<NC, T>(U, site, d));
    const auto gi = alg_at<NC>(G, L.n, site, 0);
    T* e = fcell(E, site, 3 * GD::AD);
#pragma unroll
    for (int k = 0; k < GD::AD; k++) {
        const double v = (double)e[fcomp<T>(d * GD::AD + k, 3 * GD::AD)] + (aget(t, k) - aget(gi, k)) * f2;
        e[fcomp<T>(d * GD::AD + k, 3 * GD::AD)] = (T)v;
    }
}

// UnitarityTester.calculate (diagnostics/methods/UnitarityTester.java:31-46): flag (cell, dir) when
// |Re tr U| or |Re tr Unext| is not below N + tol (NaN included)
template <int NC, typename T>
__global__ __launch_bounds__(kB) void k_trace_flags(uint32_t n, const T* __restrict__ U, const T* __restrict__ Unext,
                                                    double tol, unsigned char* __restrict__ flags,
                                                    unsigned long long* __restrict__ count)
{
    const uint32_t site = blockIdx.x * kB + threadIdx.x;
    const int d = blockIdx.y;
    if (site >= n) return;
    bool bad = false;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const auto u = link_d<NC, T>(t ? Unext : U, site, d);
        double tr;
        if constexpr (NC == 2)
            tr = 2.0 * u.e0;
        else
            tr = u.r[0] + u.r[4] + u.r[8];
        if (!(fabs(tr) < NC + tol)) bad = true;
    }
    flags[(size_t)site * 3 + d] = bad ? 1 : 0;
    if (bad) atomicAdd(count, 1ull);
}

// field a (block layout, m / n components per cell) -> [component][cell] double copy b
template <typename T>
__global__ __launch_bounds__(kB) void k_to_double(uint64_t m, uint32_t n, const T* __restrict__ a, double* __restrict__ b)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const uint32_t c = (uint32_t)(i / n), s = (uint32_t)(i - (uint64_t)c * n);
    b[i] = (double)a[fposT<T>(c, s, (uint32_t)(m / n))];
}

// E (block layout) = Ed - F ([component][cell])
template <typename T>
__global__ __launch_bounds__(kB) void k_sub_into(uint64_t m, uint32_t n, const double* __restrict__ Ed,
                                                 const double* __restrict__ F, T* __restrict__ E)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const uint32_t c = (uint32_t)(i / n), s = (uint32_t)(i - (uint64_t)c * n);
    E[fposT<T>(c, s, (uint32_t)(m / n))] = (T)(Ed[i] - F[i]);
}

__device__ __forceinline__ void two_sum_merge(double& s1, double& c1, double s2, double c2)
{
    const double s = s1 + s2;
    const double bp = s - s1;
    const double err = (s1 - (s - bp)) + (s2 - bp);
    s1 = s;
    c1 = c1 + c2 - err;
}

// per-block partial sums of a.b (compensated), then a fixed-order final sum
__global__ __launch_bounds__(kB) void k_dot_partial(uint64_t m, const double* __restrict__ a, const double* __restrict__ b,
                                                    double* __restrict__ partial)
{
    double s = 0.0, c = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kB) {
        const double y = a[i] * b[i] - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    double cc = -c;
    for (int off = 32; off > 0; off >>= 1) {
        const double s2 = __shfl_down(s, off, 64);
        const double c2 = __shfl_down(cc, off, 64);
        two_sum_merge(s, cc, s2, c2);
    }
    __shared__ double sh[kB / 64][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sh[wave][0] = s;
        sh[wave][1] = cc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ts = sh[0][0], tc = sh[0][1];
        for (int w = 1; w < kB / 64; w++) two_sum_merge(ts, tc, sh[w][0], sh[w][1]);
        partial[2 * blockIdx.x] = ts;
        partial[2 * blockIdx.x + 1] = tc;
    }
}

// one wave: lane l merges partials l, l + 64, ... in order, then a fixed shuffle tree
__global__ __launch_bounds__(64) void k_dot_final(int nb, const double* __restrict__ partial, double* __restrict__ out)
{
    const int lane = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int b = lane; b < nb; b += 64) two_sum_merge(s, c, partial[2 * b], partial[2 * b + 1]);
    for (int off = 32; off > 0; off >>= 1) {
        const double s2 = __shfl_down(s, off, 64);
        const double c2 = __shfl_down(c, off, 64);
        two_sum_merge(s, c, s2, c2);
    }
    if (lane == 0) *out = s + c;
}

// y += alpha x ; or p = r + beta p
__global__ __launch_bounds__(kB) void k_axpy(uint64_t m, double alpha, const double* __restrict__ x, double* __restrict__ y)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i < m) y[i] += alpha * x[i];
}
__global__ __launch_bounds__(kB) void k_xpby(uint64_t m, const double* __restrict__ x, double beta, double* __restrict__ y)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i < m) y[i] = x[i] + beta * y[i];
}

dim3 g1(uint64_t m) { return dim3((unsigned)((m + kB - 1) / kB)); }

}  // namespace

#define OPX_GDISPATCH(nc, prec, ...)                                                              \
    do {                                                                                          \
        if ((nc) == 2 && (prec) == 0) { constexpr int NC = 2; using T = double; __VA_ARGS__; }    \
        else if ((nc) == 2) { constexpr int NC = 2; using T = float; __VA_ARGS__; }              \
        else if ((prec) == 0) { constexpr int NC = 3; using T = double; __VA_ARGS__; }           \
        else { constexpr int NC = 3; using T = float; __VA_ARGS__; }                             \
    } while (0)

hipError_t gauss_apply_D(int nc, int prec, const Lattice& L, const void* U, const double* phi, double* F, hipStream_t st)
{
    OPX_GDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_apply_D<NC, T>), dim3((L.n + kB - 1) / kB, 3), dim3(kB), 0, st, L, (const T*)U, phi, F);
    });
    return hipGetLastError();
}

hipError_t gauss_apply_Dt(int nc, int prec, const Lattice& L, const void* U, const double* F, const void* rho,
                          double* out, hipStream_t st)
{
    OPX_GDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_apply_Dt<NC, T>), dim3((L.n + kB - 1) / kB), dim3(kB), 0, st, L, (const T*)U, F,
                           (const T*)rho, out);
    });
    return hipGetLastError();
}

hipError_t gauss_to_double(int prec, uint64_t m, uint32_t n, const void* a, double* b, hipStream_t st)
{
    if (prec == 0)
        hipLaunchKernelGGL(k_to_double<double>, g1(m), dim3(kB), 0, st, m, n, (const double*)a, b);
    else
        hipLaunchKernelGGL(k_to_double<float>, g1(m), dim3(kB), 0, st, m, n, (const float*)a, b);
    return hipGetLastError();
}

hipError_t gauss_sub_into(int prec, uint64_t m, uint32_t n, const double* Ed, const double* F, void* E, hipStream_t st)
{
    if (prec == 0)
        hipLaunchKernelGGL(k_sub_into<double>, g1(m), dim3(kB), 0, st, m, n, Ed, F, (double*)E);
    else
        hipLaunchKernelGGL(k_sub_into<float>, g1(m), dim3(kB), 0, st, m, n, Ed, F, (float*)E);
    return hipGetLastError();
}

size_t gauss_dot_scratch() { return 2 * kDotBlocks + 8; }

hipError_t gauss_dot(uint64_t m, const double* a, const double* b, double* scratch, double* out, hipStream_t st)
{
    int nb = (int)((m + kB - 1) / kB);
    if (nb > kDotBlocks) nb = kDotBlocks;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(kB), 0, st, m, a, b, scratch);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(64), 0, st, nb, scratch, out);
    return hipGetLastError();
}

hipError_t gauss_axpy(uint64_t m, double alpha, const double* x, double* y, hipStream_t st)
{
    hipLaunchKernelGGL(k_axpy, g1(m), dim3(kB), 0, st, m, alpha, x, y);
    return hipGetLastError();
}

hipError_t gauss_xpby(uint64_t m, const double* x, double beta, double* y, hipStream_t st)
{
    hipLaunchKernelGGL(k_xpby, g1(m), dim3(kB), 0, st, m, x, beta, y);
    return hipGetLastError();
}

hipError_t restore_violation(int nc, int prec, const Lattice& L, const void* U, const void* Unext, const void* rho,
                             double factor, double* G, double* R, hipStream_t st)
{
    const dim3 g((L.n + kB - 1) / kB);
    OPX_GDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_restore_violation<NC, T>), g, dim3(kB), 0, st, L, (const T*)U, (const T*)Unext,
                           (const T*)rho, factor, G, R);
    });
    return hipGetLastError();
}

hipError_t restore_correct(int nc, int prec, const Lattice& L, const void* U, const double* G, double f2, void* E,
                           hipStream_t st)
{
    const dim3 g((L.n + kB - 1) / kB, 3);
    OPX_GDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_restore_correct<NC, T>), g, dim3(kB), 0, st, L, (const T*)U, G, f2, (T*)E);
    });
    return hipGetLastError();
}

hipError_t unitarity_flags(int nc, int prec, uint32_t n, const void* U, const void* Unext, double tol,
                           unsigned char* flags, unsigned long long* count, hipStream_t st)
{
    const dim3 g((n + kB - 1) / kB, 3);
    OPX_GDISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_trace_flags<NC, T>), g, dim3(kB), 0, st, n, (const T*)U, (const T*)Unext, tol, flags,
                           count);
    });
    return hipGetLastError();
}

}  // namespace opx
