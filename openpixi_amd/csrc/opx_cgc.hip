// opx_cgc.hip -- gfx950 kernels of the CGC initial condition with the beam along x0:
//   - MV charges regulated in Fourier space (physics/initial/CGC/MVModel.java:101-157,
//     FourierFunctions.java:86-128),
//   - the light-cone Poisson solve with path-ordered Wilson lines
//     (LightConePoissonSolverImprovedFull.java:40-291),
//   - the super-particle charges with their 2nd/4th-order refinement
//     (LightConeNGPSuperParticleCreator.java:104-307).
//
// Arithmetic is double throughout; the context's links, E and rho are read and written in its
// own precision and the device block layout (fpos). Work arrays are component-major
// [component][cell] (complex ones as double2),
// group-element arrays [real][cell]. The discrete Fourier transforms run one x-axis at a time:
// a workgroup stages a few lines in LDS and runs radix-2 butterflies for power-of-two lengths,
// or the direct sum against a twiddle table for other lengths.
#include <hip/hip_runtime.h>
#include <math.h>

#include <utility>

#include "opx_group.h"
#include "opx_kernels.h"

namespace opx {
namespace {

constexpr int kB = 256;
constexpr int kFftThreads = 256;
constexpr int kWilsonM = 32;   // midpoints of one lattice spacing (LightConePoissonSolverImprovedFull.java:72)

// ------------------------------------------------------------------ lattice helpers
__device__ __forceinline__ void site_xyz(const Lattice& L, uint32_t site, uint32_t x[3])
{
    x[0] = site / L.s0;
    const uint32_t r = site - x[0] * L.s0;
    x[1] = r / L.s1;
    x[2] = r - x[1] * L.s1;
}

// periodic neighbour of `site` along d (Grid.shift, physics/grid/Grid.java:794-823)
__device__ __forceinline__ uint32_t nb(const Lattice& L, uint32_t site, int d, int o)
{
    uint32_t x[3];
    site_xyz(L, site, x);
    const uint32_t len = d == 0 ? L.d0 : (d == 1 ? L.d1 : L.d2);
    const uint32_t st = d == 0 ? L.s0 : (d == 1 ? L.s1 : 1u);
    const uint32_t c = x[d];
    if (o > 0) return c + 1 == len ? site - c * st : site + st;
    return c == 0 ? site + (len - 1) * st : site - st;
}

__device__ __forceinline__ bool in_active(const Box& b, const uint32_t x[3])
{
    if (b.all) return true;
    for (int d = 0; d < 3; d++)
        if ((int)x[d] < b.lo[d] || (int)x[d] > b.hi[d]) return false;
    return true;
}

// a context link (precision T, S stored reals, device block layout) in double
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, double>::M ld_link_d(const T* __restrict__ U, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    using GD = Group<NC, double>;
    typename GD::M m = GD::zero();
    const T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) GD::gref(m, GR::slot(k)) = (double)p[fcomp<T>(dir * GR::S + k, 3 * GR::S)];
    if constexpr (GR::S != GR::G) su3_third_row(m);
    return m;
}

template <int NC, typename T>
__device__ __forceinline__ void st_link_d(T* __restrict__ U, uint32_t site, int dir,
                                          const typename Group<NC, double>::M& m)
{
    using GR = Group<NC, T>;
    using GD = Group<NC, double>;
    T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) p[fcomp<T>(dir * GR::S + k, 3 * GR::S)] = (T)GD::gget(m, GR::slot(k));
}

// algebra coefficients of direction dir of a context field in the device block layout (E: comps
// 3 AD; the charge density: AD)
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, double>::A ld_field_alg(const T* __restrict__ X, uint32_t site, int dir,
                                                                     uint32_t comps)
{
    using GD = Group<NC, double>;
    typename GD::A a;
    const T* p = fcell(X, site, comps);
#pragma unroll
    for (int k = 0; k < GD::AD; k++) aref(a, k) = (double)p[fcomp<T>(dir * GD::AD + k, comps)];
    return a;
}

template <int NC, typename T>
__device__ __forceinline__ void st_field_alg(T* __restrict__ X, uint32_t site, int dir, uint32_t comps,
                                             const typename Group<NC, double>::A& a)
{
    using GD = Group<NC, double>;
    T* p = fcell(X, site, comps);
#pragma unroll
    for (int k = 0; k < GD::AD; k++) p[fcomp<T>(dir * GD::AD + k, comps)] = (T)aget(a, k);
}

// group element of a double [real][cell] array
template <int NC>
__device__ __forceinline__ typename Group<NC, double>::M ld_g(const double* __restrict__ V, uint32_t n, uint32_t site)
{
    using GD = Group<NC, double>;
    typename GD::M m;
#pragma unroll
    for (int k = 0; k < GD::G; k++) GD::gref(m, k) = V[(size_t)k * n + site];
    return m;
}

template <int NC>
__device__ __forceinline__ void st_g(double* __restrict__ V, uint32_t n, uint32_t site,
                                     const typename Group<NC, double>::M& m)
{
    using GD = Group<NC, double>;
#pragma unroll
    for (int k = 0; k < GD::G; k++) V[(size_t)k * n + site] = GD::gget(m, k);
}

// algebra coefficients of a [component][cell] work array
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, double>::A ld_coef(const T* __restrict__ X, uint32_t n, uint32_t site,
                                                                int dir = 0)
{
    using GD = Group<NC, double>;
    typename GD::A a;
    const T* p = X + (size_t)dir * GD::AD * n + site;
#pragma unroll
    for (int k = 0; k < GD::AD; k++) aref(a, k) = (double)p[(size_t)k * n];
    return a;
}


// ------------------------------------------------------------------ discrete Fourier transforms
// tw[t] = (cos 2 pi t / n, sin 2 pi t / n), t < n
__global__ __launch_bounds__(kB) void k_twiddles(uint32_t n, double2* __restrict__ tw)
{
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= n) return;
    double s, c;
    sincospi(2.0 * (double)t / (double)n, &s, &c);
    tw[t] = make_double2(c, s);
}

// X_k = sum_m x_m exp(sign 2 pi i k m / n) along every line of a strided complex array: line l starts
// at (l / stride) * n * stride + l % stride, its elements are `stride` apart. A workgroup handles B
// consecutive lines, staged in LDS with row pitch `pitch`.
template <bool POW2>
__global__ __launch_bounds__(kFftThreads) void k_fft(double2* __restrict__ x, uint64_t lines, uint32_t n,
                                                     uint32_t stride, uint32_t B, uint32_t pitch, uint32_t logn,
                                                     const double2* __restrict__ tw, double sign)
{
    extern __shared__ double2 sm[];
    const uint64_t l0 = (uint64_t)blockIdx.x * B;
    const uint32_t tot = B * n;
    auto addr = [&](uint32_t b, uint32_t k) -> uint64_t {
        const uint64_t l = l0 + b;
        return (l / stride) * (uint64_t)n * stride + (l % stride) + (uint64_t)k * stride;
    };
    // element f of the workgroup: consecutive threads take consecutive addresses
    auto split = [&](uint32_t f, uint32_t& b, uint32_t& k) {
        if (stride == 1) {
            b = f / n;
            k = f - b * n;
        } else {
            k = f / B;
            b = f - k * B;
        }
    };
    for (uint32_t f = threadIdx.x; f < tot; f += kFftThreads) {
        uint32_t b, k;
        split(f, b, k);
        double2 v = make_double2(0.0, 0.0);
        if (l0 + b < lines) v = x[addr(b, k)];
        const uint32_t pos = POW2 ? (logn == 0 ? 0u : (__brev(k) >> (32 - logn))) : k;
        sm[b * pitch + pos] = v;
    }
    __syncthreads();
    if constexpr (POW2) {
        const uint32_t hn = n / 2;
        for (uint32_t half = 1; half < n; half <<= 1) {
            const uint32_t step = n / (2 * half);
            for (uint32_t u = threadIdx.x; u < B * hn; u += kFftThreads) {
                const uint32_t b = u / hn, j = u - b * hn;
                const uint32_t grp = j / half, pos = j - grp * half;
                const uint32_t i0 = b * pitch + grp * 2 * half + pos, i1 = i0 + half;
                const double2 w = tw[pos * step];
                const double wr = w.x, wi = sign * w.y;
                const double2 a = sm[i0], c = sm[i1];
                const double vr = c.x * wr - c.y * wi, vi = c.x * wi + c.y * wr;
                sm[i0] = make_double2(a.x + vr, a.y + vi);
                sm[i1] = make_double2(a.x - vr, a.y - vi);
            }
            __syncthreads();
        }
        for (uint32_t f = threadIdx.x; f < tot; f += kFftThreads) {
            uint32_t b, k;
            split(f, b, k);
            if (l0 + b < lines) x[addr(b, k)] = sm[b * pitch + k];
        }
    } else {
        for (uint32_t f = threadIdx.x; f < tot; f += kFftThreads) {
            uint32_t b, k;
            split(f, b, k);
            if (l0 + b >= lines) continue;
            const double2* line = sm + b * pitch;
            double sr = 0.0, si = 0.0;
            uint32_t t = 0;
            for (uint32_t m = 0; m < n; m++) {
                const double2 w = tw[t];
                const double wr = w.x, wi = sign * w.y;
                const double2 v = line[m];
                sr += v.x * wr - v.y * wi;
                si += v.x * wi + v.y * wr;
                t += k;
                if (t >= n) t -= n;
            }
            x[addr(b, k)] = make_double2(sr, si);
        }
    }
}

// ------------------------------------------------------------------ momentum-space factors
// FourierFunctions.computeEffectiveTransverseMomentumSquared / computeTransverseMomentumSquared /
// computeLatticeMomentum1D (FourierFunctions.java:182-235) on the transverse plane (x1, x2)
__device__ __forceinline__ double kt_eff2(uint32_t p1, uint32_t p2, uint32_t n1, uint32_t n2, double aT)
{
    double m = 4.0;
    m -= 2.0 * cos((2.0 * M_PI * p1) / n1);
    m -= 2.0 * cos((2.0 * M_PI * p2) / n2);
    return m / (aT * aT);
}

__device__ __forceinline__ double kt2(uint32_t p1, uint32_t p2, uint32_t n1, uint32_t n2, double aT)
{
    const double twopi = 2.0 * M_PI;
    const double c1 = p1 < n1 / 2 ? twopi * p1 / (aT * n1) : twopi * (double)(n1 - p1) / (aT * n1);
    const double c2 = p2 < n2 / 2 ? twopi * p2 / (aT * n2) : twopi * (double)(n2 - p2) / (aT * n2);
    return c1 * c1 + c2 * c2;
}

__device__ __forceinline__ double k_long(uint32_t pos, uint32_t n, double aL)
{
    const double delta = pos / (double)n;
    if (delta < 0.5) return 2.0 * delta * M_PI / aL;
    return 2.0 * (delta - 1.0) * M_PI / aL;
}

// MVModel.initialize (MVModel.java:119-128): noise * mu g / sqrt(aL aT^2) * sqrt(Gaussian(location, width)(x0 aL))
__global__ __launch_bounds__(kB) void k_mv_fill(Lattice L, uint64_t m, const double* __restrict__ noise, double rcw,
                                                double location, double norm, double i2s2, double2* __restrict__ x)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const uint32_t c = (uint32_t)(i % L.n);
    const double xm = (double)(c / L.s0) * L.as[0] - location;
    const double profile = sqrt(norm * exp(-xm * xm * i2s2));
    x[i] = make_double2(noise[i] * rcw * profile, 0.0);
}

// FourierFunctions.regulateChargeDensityGaussian (FourierFunctions.java:104-123); inv_n folds in the
// normalization of the inverse transform
__global__ __launch_bounds__(kB) void k_mv_regulate(Lattice L, uint64_t m, double uvt2, double long_width, double ir2,
                                                    double inv_n, double2* __restrict__ x)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    uint32_t p[3];
    site_xyz(L, (uint32_t)(i % L.n), p);
    const double aL = L.as[0], aT = L.as[1];
    const double kte = kt_eff2(p[1], p[2], L.d1, L.d2, aT), kt = kt2(p[1], p[2], L.d1, L.d2, aT);
    const double kl = fabs(k_long(p[0], L.d0, aL));
    double f = 0.0;
    if (kt <= uvt2 && kt > 0) f = kte / (kte + ir2) * exp(-0.25 * kl * kl * long_width * long_width) * inv_n;
    const double2 v = x[i];
    x[i] = make_double2(v.x * f, v.y * f);
}

// the box exclusion of MVModel.initialize (MVModel.java:138-149)
__global__ __launch_bounds__(kB) void k_mv_store(Lattice L, Box active, uint64_t m, int lmin, int lmax,
                                                 const double2* __restrict__ x, double* __restrict__ rho)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    uint32_t p[3];
    site_xyz(L, (uint32_t)(i % L.n), p);
    const int z = (int)p[0];
    rho[i] = (lmin < z && z < lmax && in_active(active, p)) ? x[i].x : 0.0;
}

__global__ __launch_bounds__(kB) void k_to_complex(uint64_t m, const double* __restrict__ a, double2* __restrict__ x)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i < m) x[i] = make_double2(a[i], 0.0);
}

__global__ __launch_bounds__(kB) void k_real_part(uint64_t m, const double2* __restrict__ x, double* __restrict__ a)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i < m) a[i] = x[i].x;
}

// FourierFunctions.solvePoisson2D (FourierFunctions.java:146-154) on every x0 sheet: 1 / kT_eff^2,
// the zero mode removed; inv_n folds in the normalization of the inverse transform
__global__ __launch_bounds__(kB) void k_poisson_scale(Lattice L, uint64_t m, double inv_n, double2* __restrict__ x)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    uint32_t p[3];
    site_xyz(L, (uint32_t)(i % L.n), p);
    double f = 0.0;
    if (p[1] != 0 || p[2] != 0) f = 1.0 / kt_eff2(p[1], p[2], L.d1, L.d2, L.as[1]) * inv_n;
    const double2 v = x[i];
    x[i] = make_double2(v.x * f, v.y * f);
}

// ------------------------------------------------------------------ Wilson lines
// W(z1, z2, M, P1, P2) (LightConePoissonSolverImprovedFull.java:271-291): the ordered product of the
// links of the linear interpolation between P1 and P2 at M midpoints, centred on round((z1 + z2) / 2)
template <int NC>
__device__ typename Group<NC, double>::M wilson_segment(double z1, double z2, int M,
                                                        const typename Group<NC, double>::A& P1,
                                                        const typename Group<NC, double>::A& P2, double g, double aL)
{
    using GD = Group<NC, double>;
    typename GD::M W = GD::identity();
    if (M <= 0) return W;
    const double dz = fabs(z2 - z1) / (double)M * aL;
    const double gf = -g * dz;
    const double ui = floor((z1 + z2) * 0.5 + 0.5);
    const double u1 = z1 - ui, u2 = z2 - ui;
    typename GD::A PM, PD, P;
#pragma unroll
    for (int k = 0; k < GD::AD; k++) {
        aref(PM, k) = (aget(P1, k) + aget(P2, k)) * 0.5;
        aref(PD, k) = aget(P2, k) - aget(P1, k);
    }
#pragma unroll 1
    for (int m = 0; m < M; m++) {
        const double u = u1 + (u2 - u1) * (m + 0.5) / (double)M;
#pragma unroll
        for (int k = 0; k < GD::AD; k++) aref(P, k) = (aget(PM, k) + aget(PD, k) * u) * gf;
        W = mul(W, exp_link(P, 1.0));
    }
    return W;
}

// the step of the Wilson line into plane x0 from its neighbour against the orientation
// (LightConePoissonSolverImprovedFull.java:88-118); identity on the plane the line starts from
template <int NC>
__global__ __launch_bounds__(kB) void k_wilson_steps(Lattice L, int orientation, const double* __restrict__ phi,
                                                     double* __restrict__ Wz)
{
    using GD = Group<NC, double>;
    const uint32_t c = blockIdx.x * kB + threadIdx.x;
    if (c >= L.n) return;
    const uint32_t z = c / L.s0, i = c - z * L.s0;
    const int Lz = (int)L.d0;
    auto plane = [&](int zz) -> uint32_t { return (uint32_t)(((zz % Lz) + Lz) % Lz) * L.s0 + i; };
    const int h = kWilsonM / 2;
    const double aL = L.as[0];
    typename GD::M W = GD::identity();
    const double zd = (double)z;
    if (orientation < 0) {
        if (z != 0) {
            const uint32_t i3 = c, i2 = plane((int)z - 1), i1 = plane((int)z - 2);
            const auto p1 = ld_coef<NC, double>(phi, L.n, i1), p2 = ld_coef<NC, double>(phi, L.n, i2),
                       p3 = ld_coef<NC, double>(phi, L.n, i3);
            W = mul(wilson_segment<NC>(zd - 1, zd - 0.5, h, p1, p2, L.g, aL),
                    wilson_segment<NC>(zd - 0.5, zd, h, p2, p3, L.g, aL));
        }
    } else {
        if ((int)z != Lz - 1) {
            const uint32_t i2 = c, i1 = plane((int)z - 1), i3 = plane((int)z + 1);
            const auto p1 = ld_coef<NC, double>(phi, L.n, i1), p2 = ld_coef<NC, double>(phi, L.n, i2),
                       p3 = ld_coef<NC, double>(phi, L.n, i3);
            W = mul(wilson_segment<NC>(zd + 1, zd + 0.5, h, p2, p3, L.g, aL),
                    wilson_segment<NC>(zd + 0.5, zd, h, p1, p2, L.g, aL));
        }
    }
    st_g<NC>(Wz, L.n, c, W);
}

// V(x0) = V(x0 - orientation) W(x0), from the identity on the starting plane: one thread per transverse cell
template <int NC>
__global__ __launch_bounds__(kB) void k_wilson_scan(Lattice L, int orientation, const double* __restrict__ Wz,
                                                    double* __restrict__ V)
{
    using GD = Group<NC, double>;
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= L.s0) return;
    const uint32_t Lz = L.d0;
    typename GD::M v = GD::identity();
    st_g<NC>(V, L.n, (orientation < 0 ? 0u : Lz - 1) * L.s0 + i, v);
#pragma unroll 1
    for (uint32_t k = 1; k < Lz; k++) {
        const uint32_t z = orientation < 0 ? k : Lz - k - 1;
        const uint32_t c = z * L.s0 + i;
        v = mul(v, ld_g<NC>(Wz, L.n, c));
        st_g<NC>(V, L.n, c, v);
    }
}

// V at t = +at/2: the extra path of length at along the beam (LightConePoissonSolverImprovedFull.java:167-190)
template <int NC>
__global__ __launch_bounds__(kB) void k_wilson_extra(Lattice L, int orientation, int mfrac, const double* __restrict__ phi,
                                                     const double* __restrict__ Vm, double* __restrict__ Vp)
{
    const uint32_t c = blockIdx.x * kB + threadIdx.x;
    if (c >= L.n) return;
    const uint32_t i1 = nb(L, c, 0, -1);
    const double z1 = (double)(c / L.s0);
    const double z2 = orientation == -1 ? z1 + L.at / L.as[0] : z1 - L.at / L.as[0];
    const auto W = wilson_segment<NC>(z1, z2, mfrac, ld_coef<NC, double>(phi, L.n, i1),
                                      ld_coef<NC, double>(phi, L.n, c), L.g, L.as[0]);
    st_g<NC>(Vp, L.n, c, mul(ld_g<NC>(Vm, L.n, c), W));
}

// transverse links U -> V_x U V_{x+i}^+ at both time levels and E = -proj(Unext U^+) / at in every
// direction (LightConePoissonSolverImprovedFull.java:142-165, 192-230)
template <int NC, typename T>
__global__ __launch_bounds__(kB) void k_cgc_links(Lattice L, const double* __restrict__ Vm, const double* __restrict__ Vp,
                                                  T* __restrict__ U, T* __restrict__ Unext, T* __restrict__ E)
{
    using GD = Group<NC, double>;
    const uint32_t c = blockIdx.x * kB + threadIdx.x;
    if (c >= L.n) return;
    const double f = -1.0 / L.at;
#pragma unroll 1
    for (int d = 0; d < 3; d++) {
        typename GD::M u = ld_link_d<NC, T>(U, c, d), un = ld_link_d<NC, T>(Unext, c, d);
        if (d > 0) {
            const uint32_t xd = nb(L, c, d, 1);
            u = mul_adj(mul(ld_g<NC>(Vm, L.n, c), u), ld_g<NC>(Vm, L.n, xd));
            un = mul_adj(mul(ld_g<NC>(Vp, L.n, c), un), ld_g<NC>(Vp, L.n, xd));
            st_link_d<NC, T>(U, c, d, u);
            st_link_d<NC, T>(Unext, c, d, un);
        }
        typename GD::A e = proj(mul_adj(un, u));
#pragma unroll
        for (int k = 0; k < GD::AD; k++) aref(e, k) *= f;
        st_field_alg<NC, T>(E, c, d, 3 * GD::AD, e);
    }
}

// Gauss constraint of the Wilson-line copy of the fields (transverse links V_x V_{x+i}^+, the grid's
// own longitudinal links, E from its temporal plaquettes) minus the grid's rho
// (LightConePoissonSolverImprovedFull.java:232-262, Grid.getGaussConstraint)
template <int NC, typename T>
__global__ __launch_bounds__(kB) void k_cgc_gauss(Lattice L, Box active, const double* __restrict__ Vm,
                                                  const double* __restrict__ Vp, const T* __restrict__ U,
                                                  const T* __restrict__ Unext, const T* __restrict__ rho,
                                                  double* __restrict__ gauss)
{
    using GD = Group<NC, double>;
    using M = typename GD::M;
    using A = typename GD::A;
    const uint32_t c = blockIdx.x * kB + threadIdx.x;
    if (c >= L.n) return;
    uint32_t x[3];
    site_xyz(L, c, x);
    A g = GD::azero();
    if (in_active(active, x)) {
        auto copy_u = [&](const double* V, const T* grid, uint32_t s, int d) -> M {
            if (d == 0) return ld_link_d<NC, T>(grid, s, 0);
            return mul_adj(ld_g<NC>(V, L.n, s), ld_g<NC>(V, L.n, nb(L, s, d, 1)));
        };
        auto copy_e = [&](uint32_t s, int d, const M& u) -> A {
            A e = proj(mul_adj(copy_u(Vp, Unext, s, d), u));
#pragma unroll
            for (int k = 0; k < GD::AD; k++) aref(e, k) *= -1.0 / L.at;
            return e;
        };
#pragma unroll 1
        for (int d = 0; d < 3; d++) {
            const uint32_t xm = nb(L, c, d, -1);
            const M um = copy_u(Vm, U, xm, d);
            const A t = act(copy_e(xm, d, um), adj(um));
            const A ex = copy_e(c, d, copy_u(Vm, U, c, d));
            const double f = 1.0 / (L.as[d] * (L.g * L.as[d]));
#pragma unroll
            for (int k = 0; k < GD::AD; k++) aref(g, k) += (aget(ex, k) - aget(t, k)) * f;
        }
        if (rho) {
            const A r = ld_field_alg<NC, T>(rho, c, 0, GD::AD);
#pragma unroll
            for (int k = 0; k < GD::AD; k++) aref(g, k) -= aget(r, k);
        }
    }
#pragma unroll
    for (int k = 0; k < GD::AD; k++) gauss[(size_t)k * L.n + c] = aget(g, k);
}

// ------------------------------------------------------------------ super particles
// largest |G| of each x0 plane (LightConeNGPSuperParticleCreator.java:110-127)
__global__ __launch_bounds__(kB) void k_plane_max(uint32_t nT, uint32_t n, int ad, const double* __restrict__ gauss,
                                                  double* __restrict__ mx)
{
    __shared__ double red[kB];
    const uint32_t z = blockIdx.x;
    double m = 0.0;
    for (uint32_t j = threadIdx.x; j < nT; j += kB) {
        const size_t c = (size_t)z * nT + j;
        double s = 0.0;
        for (int k = 0; k < ad; k++) {
            const double v = gauss[(size_t)k * n + c];
            s += v * v;
        }
        m = fmax(m, sqrt(s));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) mx[z] = red[0];
}

__device__ __forceinline__ int wrap(int i, int n)
{
    while (i < 0) i += n;
    while (i >= n) i -= n;
    return i;
}

// One sweep of the 2nd- or 4th-order charge refinement over a periodic list of nn charges
// (refine2/refine4, LightConeNGPSuperParticleCreator.java:248-307 with shift = ppc / 2,
// LightConeNGPParticleCreator.java:237-297 with shift = 0), element l at col[l * st]; the
// charge at position i is not moved when (i + shift) % ppc = ppc - 1. Written through memory as the reference does,
// so that short lists whose stencil wraps onto itself behave the same.
template <int ORDER>
__device__ __forceinline__ void refine_sweep(double* col, size_t st, int nn, int ppc, int shift)
{
#pragma unroll 1
    for (int i = 0; i < nn; i++) {
        const int jm = (i + shift) % ppc;
        if (jm >= ppc - 1) continue;
        if constexpr (ORDER == 2) {
            double* Q0 = col + (size_t)wrap(i - 1, nn) * st;
            double* Q1 = col + (size_t)wrap(i, nn) * st;
            double* Q2 = col + (size_t)wrap(i + 1, nn) * st;
            double* Q3 = col + (size_t)wrap(i + 2, nn) * st;
            double dq = *Q0 * -1;
            dq += *Q1 * 3;
            dq += *Q2 * -3;
            dq += *Q3 * 1;
            dq *= 1.0 / 4.0;
            *Q1 += dq * -1.0;
            *Q2 += dq * 1.0;
        } else {
            double* Q[6];
#pragma unroll
            for (int m = 0; m < 6; m++) Q[m] = col + (size_t)wrap(i - 2 + m, nn) * st;
            double dq = *Q[0] * +1;
            dq += *Q[1] * -5;
            dq += *Q[2] * +10;
            dq += *Q[3] * -10;
            dq += *Q[4] * +5;
            dq += *Q[5] * -1;
            dq *= 1.0 / 12.0;
            *Q[2] += dq * -1.0;
            *Q[3] += dq * 1.0;
        }
    }
}

// The initial charges of a block (LightConeNGPSuperParticleCreator.java:197-215): one thread per
// (component, transverse cell) g fills its list of bw * ppc charges, work[l * ad * nT + g] =
// G(ngp) / ppc, with ngp the cell itself for the first ppc / 2 charges of a cell and its +x0
// neighbour for the rest.
__global__ __launch_bounds__(kB) void k_sp_init(uint32_t nT, uint32_t n, uint32_t Lz, int ad, SpBlock b,
                                                const double* __restrict__ gauss, double* __restrict__ work)
{
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= (uint32_t)ad * nT) return;
    const uint32_t comp = g / nT, t = g - comp * nT;
    const int ppc = b.ppc;
    const size_t st = (size_t)ad * nT;
    const double inv = 1.0 / ppc;
    for (int p = 0; p < b.bw; p++) {
        const uint32_t z = (uint32_t)(b.zs + p), zn = (z + 1) % Lz;
        const double qa = gauss[(size_t)comp * n + (size_t)z * nT + t] * inv;
        const double qb = gauss[(size_t)comp * n + (size_t)zn * nT + t] * inv;
        for (int k = 0; k < ppc; k++) work[(size_t)(p * ppc + k) * st + g] = k < ppc / 2 ? qa : qb;
    }
}

// 100 sweeps of each order on every list, one thread per list (long lists)
__global__ __launch_bounds__(kB) void k_sp_refine(uint32_t cols, int nn, int ppc, int shift, double* __restrict__ work)
{
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= cols) return;
#pragma unroll 1
    for (int it = 0; it < 100; it++) refine_sweep<2>(work + g, cols, nn, ppc, shift);
#pragma unroll 1
    for (int it = 0; it < 100; it++) refine_sweep<4>(work + g, cols, nn, ppc, shift);
}

// The matrix of one sweep, S[l * nn + c] = (sweep e_c)[l]: thread c runs the sweep on the unit vector c.
template <int ORDER>
__global__ __launch_bounds__(kB) void k_sweep_matrix(int nn, int ppc, int shift, double* __restrict__ S)
{
    const int c = blockIdx.x * kB + threadIdx.x;
    if (c >= nn) return;
    for (int l = 0; l < nn; l++) S[(size_t)l * nn + c] = l == c ? 1.0 : 0.0;
    refine_sweep<ORDER>(S + c, (size_t)nn, nn, ppc, shift);
}

__global__ __launch_bounds__(kB) void k_identity(int nn, double* __restrict__ R)
{
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i < (uint64_t)nn * nn) R[i] = (i / nn == i % nn) ? 1.0 : 0.0;
}

// C[M x N] = A[M x K] B[K x N], row-major doubles, fixed summation order over K. 64 x 64 tiles,
// 256 threads of 4 x 4 outputs (rows ty + 16 i, columns tx + 16 j), K in LDS slabs of 16.
constexpr int kGemmTile = 64, kGemmK = 16;
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const double* __restrict__ A,
                                              const double* __restrict__ B, double* __restrict__ C)
{
    __shared__ double As[kGemmK][kGemmTile + 1];
    __shared__ double Bs[kGemmK][kGemmTile];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int row0 = blockIdx.y * kGemmTile, col0 = blockIdx.x * kGemmTile;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < K; k0 += kGemmK) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int f = tid + 256 * e;
            // A slab: 64 rows x 16 k, k fastest in memory
            const int ar = f / kGemmK, ak = f % kGemmK;
            const int gr = row0 + ar, gk = k0 + ak;
            As[ak][ar] = (gr < M && gk < K) ? A[(size_t)gr * K + gk] : 0.0;
            // B slab: 16 k x 64 columns, columns fastest
            const int bk = f / kGemmTile, bc = f % kGemmTile;
            const int hk = k0 + bk, hc = col0 + bc;
            Bs[bk][bc] = (hk < K && hc < N) ? B[(size_t)hk * N + hc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kGemmK; k++) {
            double a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[k][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; j++) bb[j] = Bs[k][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * bb[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = row0 + ty + 16 * i, c = col0 + tx + 16 * j;
            if (r < M && c < N) C[(size_t)r * N + c] = acc[i][j];
        }
}

// Each refined charge to its super particle (LightConeNGPSuperParticleCreator.java:177-215):
// subdivision j of the block covers planes [P_j, P_j + w_j), particle k of it holds w_j * nT
// charges starting at q_base + ppc nT P_j + k w_j nT.
template <typename T>
__global__ __launch_bounds__(kB) void k_sp_scatter(uint32_t nT, int ad, SpBlock b, const double* __restrict__ work,
                                                   T* __restrict__ q, int64_t q_total, int64_t q_base)
{
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= (uint32_t)ad * nT) return;
    const uint32_t comp = g / nT, t = g - comp * nT;
    const int ppc = b.ppc, wl = b.ws + 1;
    const size_t st = (size_t)ad * nT;
    for (int p = 0; p < b.bw; p++) {
        int P, w;
        if (p < b.nl * wl) {
            P = (p / wl) * wl;
            w = wl;
        } else {
            P = b.nl * wl + ((p - b.nl * wl) / b.ws) * b.ws;
            w = b.ws;
        }
        for (int k = 0; k < ppc; k++) {
            const int64_t at = q_base + (int64_t)ppc * nT * P + (int64_t)k * w * nT + (int64_t)(p - P) * nT + t;
            q[(size_t)comp * q_total + at] = (T)work[(size_t)(p * ppc + k) * st + g];
        }
    }
}

// The particles of LightConeNGPParticleCreator.initializeParticles (LightConeNGPParticleCreator.java:155-205)
// for planes zs .. zs + bw - 1: particle p of the new range is (plane, transverse cell i, j) in
// list order, at x0 = z a0 + 1e-11 a0 + (j - ppc/2)/ppc a0 (t = 0) and that plus orientation
// at (t = at), transverse coordinates x a + 1e-11 a, velocity orientation along x0. Its charge
// G(ngp(pos0)) / ppc goes to the refinement lists: work[(plane ppc + j) cols + comp nT + i].
// Arrays [3][count] (positions, velocities) are written at `base + p`.
__global__ __launch_bounds__(kB) void k_pp_init(Lattice L, int ad, SpBlock b, int orientation, const double* __restrict__ gauss,
                                                double* __restrict__ work, int* __restrict__ dir, double* __restrict__ pos0,
                                                double* __restrict__ pos1, double* __restrict__ vel, uint64_t count,
                                                uint64_t base)
{
#pragma clang fp contract(off)
    const uint32_t nT = L.s0, ppc = (uint32_t)b.ppc;
    const uint64_t p = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= (uint64_t)b.bw * nT * ppc) return;
    const uint32_t plane = (uint32_t)(p / ((uint64_t)nT * ppc));
    const uint32_t r = (uint32_t)(p - (uint64_t)plane * nT * ppc);
    const uint32_t i = r / ppc, j = r - i * ppc;
    const int gp[3] = {b.zs + (int)plane, (int)(i / L.d2), (int)(i % L.d2)};
    const double x = (1.0 * (int)j - (int)(ppc / 2)) / (int)ppc;
    const double dz = x * L.as[0];
    const uint64_t q = base + p;
    uint32_t ngp[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double fix = 10E-12 * L.as[k];
        double a = gp[k] * L.as[k] + fix, c = gp[k] * L.as[k] + fix, v = 0.0;
        if (k == 0) {
            a += 0.0 * orientation + dz;
            c += (0.0 + L.at) * orientation + dz;
            v = 1.0 * orientation;
        }
        pos0[(size_t)k * count + q] = a;
        pos1[(size_t)k * count + q] = c;
        vel[(size_t)k * count + q] = v;
        const uint32_t len = k == 0 ? L.d0 : (k == 1 ? L.d1 : L.d2);
        long long g = (long long)rint(a / L.as[k]) % (long long)len;
        if (g < 0) g += len;
        ngp[k] = (uint32_t)g;
    }
    dir[q] = 0;
    const uint32_t cell = ngp[0] * L.s0 + ngp[1] * L.s1 + ngp[2];
    const size_t cols = (size_t)ad * nT;
    const double inv = 1.0 / (int)ppc;
    for (int comp = 0; comp < ad; comp++)
        work[((size_t)plane * ppc + j) * cols + (size_t)comp * nT + i] = gauss[(size_t)comp * L.n + cell] * inv;
}

// the refined charges into Q0 and Q1 of the new particles ([coefficient][count], precision T)
template <typename T>
__global__ __launch_bounds__(kB) void k_pp_scatter(uint32_t nT, int ad, SpBlock b, const double* __restrict__ work,
                                                   T* __restrict__ Q0, T* __restrict__ Q1, uint64_t count, uint64_t base)
{
    const uint32_t ppc = (uint32_t)b.ppc;
    const uint64_t p = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= (uint64_t)b.bw * nT * ppc) return;
    const uint32_t plane = (uint32_t)(p / ((uint64_t)nT * ppc));
    const uint32_t r = (uint32_t)(p - (uint64_t)plane * nT * ppc);
    const uint32_t i = r / ppc, j = r - i * ppc;
    const size_t cols = (size_t)ad * nT;
    for (int comp = 0; comp < ad; comp++) {
        const T v = (T)work[((size_t)plane * ppc + j) * cols + (size_t)comp * nT + i];
        Q0[(size_t)comp * count + base + p] = v;
        Q1[(size_t)comp * count + base + p] = v;
    }
}

// ------------------------------------------------------------------ java.util.Random
// The generator java.util.Random specifies: seed' = seed * 0x5DEECE66D + 0xB (mod 2^48),
// next(bits) = the top `bits` of the 48, nextDouble = (next(26) 2^27 + next(27)) 2^-53, and
// nextGaussian by the polar method, which consumes two doubles per trial and returns both
// values of an accepted pair in turn. Trial p consumes LCG steps 4p+1 .. 4p+4, reached by a
// jump (the affine step composed with itself); accepted trials are then compacted in order.
constexpr uint64_t kJMult = 0x5DEECE66DULL, kJAdd = 0xBULL, kJMask = (1ULL << 48) - 1;

__device__ __forceinline__ uint64_t lcg_jump(uint64_t s, uint64_t k)
{
    uint64_t ra = 1, rc = 0, ba = kJMult, bc = kJAdd;   // arithmetic mod 2^64; 2^48 divides it
    while (k) {
        if (k & 1) {
            ra = ra * ba;
            rc = rc * ba + bc;
        }
        bc = bc * ba + bc;
        ba = ba * ba;
        k >>= 1;
    }
    return (ra * s + rc) & kJMask;
}

// StrictMath.log for positive finite x: fdlibm's e_log.c (Sun Microsystems, freely
// distributable), x = 2^k (1 + f), log(1 + f) = f - s (f - R(s^2)) with s = f / (2 + f), k ln 2
// added in two parts; evaluated without contractions, as Java does.
__device__ double strict_log(double x)
{
#pragma clang fp contract(off)
    const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01, Lg3 = 2.857142874366239149e-01,
                 Lg4 = 2.222219843214978396e-01, Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                 Lg7 = 1.479819860511658591e-01;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    int k = 0;
    long long bits = __double_as_longlong(x);
    int hx = (int)(bits >> 32);
    if (hx < 0x00100000) {   // subnormal
        k -= 54;
        x *= 18014398509481984.0;
        bits = __double_as_longlong(x);
        hx = (int)(bits >> 32);
    }
    k += (hx >> 20) - 1023;
    hx &= 0x000fffff;
    const int i = (hx + 0x95f64) & 0x100000;
    x = __longlong_as_double(((long long)(hx | (i ^ 0x3ff00000)) << 32) | (bits & 0xffffffffLL));
    k += i >> 20;
    const double f = x - 1.0;
    const double dk = (double)k;
    if ((0x000fffff & (2 + hx)) < 3) {   // |f| < 2^-20
        if (f == 0.0) return k == 0 ? 0.0 : dk * ln2_hi + dk * ln2_lo;
        const double R = f * f * (0.5 - 0.33333333333333333 * f);
        return k == 0 ? f - R : dk * ln2_hi - ((R - dk * ln2_lo) - f);
    }
    const double s = f / (2.0 + f);
    const double z = s * s;
    const double w = z * z;
    const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    const double R = t2 + t1;
    if (((hx - 0x6147a) | (0x6b851 - hx)) > 0) {
        const double hfsq = 0.5 * f * f;
        return k == 0 ? f - (hfsq - s * (hfsq + R)) : dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
    }
    return k == 0 ? f - s * (f - R) : dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

__device__ __forceinline__ double java_double(uint64_t s26, uint64_t s27)
{
    return (double)(((s26 >> 22) << 27) + (s27 >> 21)) * 0x1.0p-53;
}

// trial p: its two values when accepted (flag 1)
__global__ __launch_bounds__(kB) void k_jr_trials(uint64_t seed, uint64_t trials, double2* __restrict__ vals,
                                                  uint32_t* __restrict__ flag)
{
#pragma clang fp contract(off)
    const uint64_t p = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= trials) return;
    const uint64_t s1 = lcg_jump(seed, 4 * p + 1);
    const uint64_t s2 = (s1 * kJMult + kJAdd) & kJMask;
    const uint64_t s3 = (s2 * kJMult + kJAdd) & kJMask;
    const uint64_t s4 = (s3 * kJMult + kJAdd) & kJMask;
    const double v1 = 2 * java_double(s1, s2) - 1;
    const double v2 = 2 * java_double(s3, s4) - 1;
    const double sq = v1 * v1 + v2 * v2;
    const bool ok = sq < 1 && sq != 0;
    flag[p] = ok ? 1u : 0u;
    if (ok) {
        const double m = sqrt(-2 * strict_log(sq) / sq);
        vals[p] = make_double2(v1 * m, v2 * m);
    }
}

// exclusive prefix sums of `flag` within blocks of kScan, block totals to `sums`
constexpr int kScan = 1024;
__global__ __launch_bounds__(kB) void k_scan_blocks(uint64_t m, const uint32_t* __restrict__ flag,
                                                    uint32_t* __restrict__ pos, uint32_t* __restrict__ sums)
{
    __shared__ uint32_t sh[kScan];
    const uint64_t base = (uint64_t)blockIdx.x * kScan;
    for (int e = threadIdx.x; e < kScan; e += kB) sh[e] = base + e < m ? flag[base + e] : 0u;
    __syncthreads();
    // Hillis-Steele inclusive scan over kScan values, four per thread
    for (int off = 1; off < kScan; off <<= 1) {
        uint32_t v[kScan / kB];
#pragma unroll
        for (int r = 0; r < kScan / kB; r++) {
            const int e = threadIdx.x + r * kB;
            v[r] = sh[e] + (e >= off ? sh[e - off] : 0u);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kScan / kB; r++) sh[threadIdx.x + r * kB] = v[r];
        __syncthreads();
    }
    for (int e = threadIdx.x; e < kScan; e += kB)
        if (base + e < m) pos[base + e] = sh[e] - flag[base + e];
    if (threadIdx.x == 0) sums[blockIdx.x] = sh[kScan - 1];
}

// block offsets: one workgroup scans the block totals chunk by chunk (exclusive, carried over the
// chunks), the grand total last
__global__ __launch_bounds__(kScan) void k_scan_sums(uint32_t nb, uint32_t* __restrict__ sums,
                                                     uint64_t* __restrict__ total)
{
    __shared__ uint64_t sh[kScan];
    uint64_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += kScan) {
        const uint32_t e = c0 + threadIdx.x;
        const uint64_t v = e < nb ? sums[e] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kScan; off <<= 1) {
            const uint64_t a = sh[threadIdx.x] + (threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0);
            __syncthreads();
            sh[threadIdx.x] = a;
            __syncthreads();
        }
        if (e < nb) sums[e] = (uint32_t)(carry + sh[threadIdx.x] - v);
        carry += sh[kScan - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// out[2 r], out[2 r + 1] of the r-th accepted trial, r < count / 2 rounded up
__global__ __launch_bounds__(kB) void k_jr_scatter(uint64_t trials, uint64_t count, const double2* __restrict__ vals,
                                                   const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                                   const uint32_t* __restrict__ sums, double* __restrict__ out)
{
    const uint64_t p = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= trials || !flag[p]) return;
    const uint64_t r = (uint64_t)pos[p] + sums[p / kScan];
    const double2 v = vals[p];
    if (2 * r < count) out[2 * r] = v.x;
    if (2 * r + 1 < count) out[2 * r + 1] = v.y;
}

// ------------------------------------------------------------------ Wilson-line observables
// Re tr(V_i V_j^+) of two elements of a [real][T] array (GroupElement.mult(adj()).getRealTrace())
template <int NC>
__device__ __forceinline__ double re_tr_adj(const double* __restrict__ V, uint32_t T, uint32_t i, uint32_t j)
{
    using GD = Group<NC, double>;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < GD::G; k++) s += V[(size_t)k * T + i] * V[(size_t)k * T + j];
    return NC == 2 ? 2.0 * s : s;
}

// sum over the plane of Re tr V (GroupElement.getRealTrace): one workgroup, fixed order
template <int NC>
__global__ __launch_bounds__(kB) void k_tadpole(uint32_t T, const double* __restrict__ V, double* __restrict__ out)
{
    __shared__ double red[kB];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < T; i += kB)
        s += NC == 2 ? 2.0 * V[i] : V[i] + V[(size_t)4 * T + i] + V[(size_t)8 * T + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// binned tr(V_x V_y^+) over all pairs (WilsonLineObservables.computeDipole,
// WilsonLineObservables.java:86-115): thread x walks every y in order and adds into its own
// column of the [bin][T] partial sums and counts
template <int NC>
__global__ __launch_bounds__(kB) void k_dipole_pairs(uint32_t n1, uint32_t n2, double a1, double a2, double ds,
                                                     int nbins, const double* __restrict__ V, double* __restrict__ psum,
                                                     unsigned long long* __restrict__ pcnt)
{
    const uint32_t T = n1 * n2;
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= T) return;
    for (int b = 0; b < nbins; b++) {
        psum[(size_t)b * T + i] = 0.0;
        pcnt[(size_t)b * T + i] = 0ull;
    }
    const int x0 = (int)(i / n2), x1 = (int)(i % n2);
    const int h1 = (int)n1 / 2, h2 = (int)n2 / 2;
#pragma unroll 1
    for (uint32_t j = 0; j < T; j++) {
        const int y0 = (int)(j / n2), y1 = (int)(j % n2);
        double d0 = fabs((double)(x0 - y0)), d1 = fabs((double)(x1 - y1));
        if (d0 > h1) d0 -= n1;
        if (d1 > h2) d1 -= n2;
        double dist = 0.0;
        dist += d0 * d0 * a1 * a1;
        dist += d1 * d1 * a2 * a2;
        dist = sqrt(dist);
        const int bin = (int)(dist / ds);
        if (bin < nbins) {
            psum[(size_t)bin * T + i] += re_tr_adj<NC>(V, T, i, j);
            pcnt[(size_t)bin * T + i] += 1ull;
        }
    }
}

// per bin: the mean over all pairs (0/0 = NaN for an empty bin, as the reference divides)
__global__ __launch_bounds__(kB) void k_dipole_bins(uint32_t T, const double* __restrict__ psum,
                                                    const unsigned long long* __restrict__ pcnt,
                                                    double* __restrict__ corr, unsigned long long* __restrict__ cnt)
{
    __shared__ double rs[kB];
    __shared__ unsigned long long rc[kB];
    const uint32_t b = blockIdx.x;
    double s = 0.0;
    unsigned long long c = 0;
    for (uint32_t i = threadIdx.x; i < T; i += kB) {
        s += psum[(size_t)b * T + i];
        c += pcnt[(size_t)b * T + i];
    }
    rs[threadIdx.x] = s;
    rc[threadIdx.x] = c;
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            rs[threadIdx.x] += rs[threadIdx.x + w];
            rc[threadIdx.x] += rc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        corr[b] = rs[0] / (double)rc[0];
        cnt[b] = rc[0];
    }
}

inline unsigned blocks(uint64_t m) { return (unsigned)((m + kB - 1) / kB); }

}  // namespace

// ------------------------------------------------------------------ launchers
hipError_t cgc_twiddles(uint32_t n, double2* tw, hipStream_t st)
{
    k_twiddles<<<blocks(n), kB, 0, st>>>(n, tw);
    return hipGetLastError();
}

bool cgc_fft_supported(uint32_t n) { return n >= 1 && n <= kCgcMaxFft; }

hipError_t cgc_fft_axis(double2* x, uint64_t total, uint32_t n, uint32_t stride, const double2* tw, int sign,
                        hipStream_t st)
{
    if (n == 1) return hipSuccess;
    if (!cgc_fft_supported(n)) return hipErrorInvalidValue;
    const bool pow2 = (n & (n - 1)) == 0;
    uint32_t logn = 0;
    while ((1u << logn) < n) logn++;
    // lines per workgroup: a power of two with B (n + 1) complex values within 64 KiB of LDS; for
    // strided axes at most the largest power of two dividing the stride
    uint32_t B = 1;
    while (B < 16 && (2 * B) * (n + 1) <= 4096u) B *= 2;
    if (stride > 1)
        while (B > 1 && stride % B != 0) B /= 2;
    const uint32_t pitch = B == 1 ? n : n + 1;
    const uint64_t lines = total / n;
    const unsigned grid = (unsigned)((lines + B - 1) / B);
    const size_t lds = (size_t)B * pitch * sizeof(double2);
    if (pow2)
        k_fft<true><<<grid, kFftThreads, lds, st>>>(x, lines, n, stride, B, pitch, logn, tw, (double)sign);
    else
        k_fft<false><<<grid, kFftThreads, lds, st>>>(x, lines, n, stride, B, pitch, logn, tw, (double)sign);
    return hipGetLastError();
}

hipError_t cgc_mv_fill(const Lattice& L, uint64_t m, const double* noise, double rcw, double location, double norm,
                       double i2s2, double2* x, hipStream_t st)
{
    k_mv_fill<<<blocks(m), kB, 0, st>>>(L, m, noise, rcw, location, norm, i2s2, x);
    return hipGetLastError();
}

hipError_t cgc_mv_regulate(const Lattice& L, uint64_t m, double uvt2, double long_width, double ir2, double inv_n,
                           double2* x, hipStream_t st)
{
    k_mv_regulate<<<blocks(m), kB, 0, st>>>(L, m, uvt2, long_width, ir2, inv_n, x);
    return hipGetLastError();
}

hipError_t cgc_mv_store(const Lattice& L, const Box& active, uint64_t m, int lmin, int lmax, const double2* x,
                        double* rho, hipStream_t st)
{
    k_mv_store<<<blocks(m), kB, 0, st>>>(L, active, m, lmin, lmax, x, rho);
    return hipGetLastError();
}

hipError_t cgc_to_complex(uint64_t m, const double* a, double2* x, hipStream_t st)
{
    k_to_complex<<<blocks(m), kB, 0, st>>>(m, a, x);
    return hipGetLastError();
}

hipError_t cgc_real_part(uint64_t m, const double2* x, double* a, hipStream_t st)
{
    k_real_part<<<blocks(m), kB, 0, st>>>(m, x, a);
    return hipGetLastError();
}

hipError_t cgc_poisson_scale(const Lattice& L, uint64_t m, double inv_n, double2* x, hipStream_t st)
{
    k_poisson_scale<<<blocks(m), kB, 0, st>>>(L, m, inv_n, x);
    return hipGetLastError();
}

hipError_t cgc_wilson_lines(int nc, const Lattice& L, int orientation, int mfrac, const double* phi, double* Wz,
                            double* Vm, double* Vp, hipStream_t st)
{
    if (nc == 2) {
        k_wilson_steps<2><<<blocks(L.n), kB, 0, st>>>(L, orientation, phi, Wz);
        k_wilson_scan<2><<<blocks(L.s0), kB, 0, st>>>(L, orientation, Wz, Vm);
        k_wilson_extra<2><<<blocks(L.n), kB, 0, st>>>(L, orientation, mfrac, phi, Vm, Vp);
    } else {
        k_wilson_steps<3><<<blocks(L.n), kB, 0, st>>>(L, orientation, phi, Wz);
        k_wilson_scan<3><<<blocks(L.s0), kB, 0, st>>>(L, orientation, Wz, Vm);
        k_wilson_extra<3><<<blocks(L.n), kB, 0, st>>>(L, orientation, mfrac, phi, Vm, Vp);
    }
    return hipGetLastError();
}

hipError_t cgc_links(int nc, int prec, const Lattice& L, const double* Vm, const double* Vp, void* U, void* Unext,
                     void* E, hipStream_t st)
{
#define OPX_CGC_LINKS(NC, T) \
    k_cgc_links<NC, T><<<blocks(L.n), kB, 0, st>>>(L, Vm, Vp, (T*)U, (T*)Unext, (T*)E)
    if (nc == 2 && prec == 0) OPX_CGC_LINKS(2, double);
    else if (nc == 2) OPX_CGC_LINKS(2, float);
    else if (prec == 0) OPX_CGC_LINKS(3, double);
    else OPX_CGC_LINKS(3, float);
#undef OPX_CGC_LINKS
    return hipGetLastError();
}

hipError_t cgc_gauss(int nc, int prec, const Lattice& L, const Box& active, const double* Vm, const double* Vp,
                     const void* U, const void* Unext, const void* rho, double* gauss, hipStream_t st)
{
#define OPX_CGC_GAUSS(NC, T)                                                                                  \
    k_cgc_gauss<NC, T><<<blocks(L.n), kB, 0, st>>>(L, active, Vm, Vp, (const T*)U, (const T*)Unext, (const T*)rho, \
                                                   gauss)
    if (nc == 2 && prec == 0) OPX_CGC_GAUSS(2, double);
    else if (nc == 2) OPX_CGC_GAUSS(2, float);
    else if (prec == 0) OPX_CGC_GAUSS(3, double);
    else OPX_CGC_GAUSS(3, float);
#undef OPX_CGC_GAUSS
    return hipGetLastError();
}

hipError_t cgc_plane_max(int nc, const Lattice& L, const double* gauss, double* mx, hipStream_t st)
{
    k_plane_max<<<L.d0, kB, 0, st>>>(L.s0, L.n, nc == 2 ? 3 : 8, gauss, mx);
    return hipGetLastError();
}

size_t cgc_sp_work(const SpBlock& b, uint32_t nT, int ad)
{
    const size_t nn = (size_t)b.bw * b.ppc, cols = (size_t)ad * nT;
    return nn * cols * (nn <= (size_t)kRefineMatrixMax ? 2 : 1) + (nn <= (size_t)kRefineMatrixMax ? 5 * nn * nn : 0);
}

namespace {

hipError_t gemm(int M, int N, int K, const double* A, const double* B, double* C, hipStream_t st)
{
    dim3 grid((unsigned)((N + kGemmTile - 1) / kGemmTile), (unsigned)((M + kGemmTile - 1) / kGemmTile));
    k_gemm<<<grid, 256, 0, st>>>(M, N, K, A, B, C);
    return hipGetLastError();
}

// R = S^e by binary powering; P and tmp are scratch matrices; returns the buffer holding R
double* matrix_power(int nn, double* S, int e, double* R, double* P, double* tmp, hipStream_t st, hipError_t* err)
{
    const size_t m = (size_t)nn * nn;
    k_identity<<<blocks(m), kB, 0, st>>>(nn, R);
    hipError_t r = hipMemcpyAsync(P, S, m * sizeof(double), hipMemcpyDeviceToDevice, st);
    while (e && r == hipSuccess) {
        if (e & 1) {
            r = gemm(nn, nn, nn, R, P, tmp, st);
            std::swap(R, tmp);
        }
        e >>= 1;
        if (e && r == hipSuccess) {
            r = gemm(nn, nn, nn, P, P, tmp, st);
            std::swap(P, tmp);
        }
    }
    *err = r == hipSuccess ? hipGetLastError() : r;
    return R;
}

}  // namespace

namespace {

// the 2nd- then 4th-order refinement of `cols` lists of nn = bw * ppc charges in work
// ([position][list]); returns the buffer holding the result (work or the one after it)
hipError_t refine(int nn, int ppc, int shift, uint32_t cols, double* work, const double** refined, hipStream_t st)
{
    *refined = work;
    if (nn > kRefineMatrixMax) {
        k_sp_refine<<<blocks(cols), kB, 0, st>>>(cols, nn, ppc, shift, work);
        return hipGetLastError();
    }
    // every list goes through the same linear map: one sweep's matrix per order, raised to the
    // 100th power, and (R4 R2) applied to all lists at once
    double* out = work + (size_t)nn * cols;
    double* m = out + (size_t)nn * cols;
    const size_t mm = (size_t)nn * nn;
    double *S = m, *X = m + mm, *Y = m + 2 * mm, *Z = m + 3 * mm, *W = m + 4 * mm;
    hipError_t e = hipSuccess;
    const unsigned sg = blocks((uint64_t)nn);
    k_sweep_matrix<2><<<sg, kB, 0, st>>>(nn, ppc, shift, S);
    double* R2 = matrix_power(nn, S, 100, X, Y, Z, st, &e);
    if (e != hipSuccess) return e;
    // keep R2 in W while the 4th-order power runs in the other buffers
    if ((e = hipMemcpyAsync(W, R2, mm * sizeof(double), hipMemcpyDeviceToDevice, st)) != hipSuccess) return e;
    k_sweep_matrix<4><<<sg, kB, 0, st>>>(nn, ppc, shift, S);
    double* R4 = matrix_power(nn, S, 100, X, Y, Z, st, &e);
    if (e != hipSuccess) return e;
    double* Rt = R4 == X ? Y : X;
    if ((e = gemm(nn, nn, nn, R4, W, Rt, st)) != hipSuccess) return e;
    if ((e = gemm(nn, (int)cols, nn, Rt, work, out, st)) != hipSuccess) return e;
    *refined = out;
    return hipSuccess;
}

}  // namespace

hipError_t cgc_sp_charges(int nc, int prec, const Lattice& L, const SpBlock& b, const double* gauss, double* work,
                          void* q, int64_t q_total, int64_t q_base, hipStream_t st)
{
    const int ad = nc == 2 ? 3 : 8;
    const uint32_t cols = (uint32_t)ad * L.s0;
    const unsigned grid = blocks(cols);
    k_sp_init<<<grid, kB, 0, st>>>(L.s0, L.n, L.d0, ad, b, gauss, work);
    const double* refined = nullptr;
    hipError_t e = refine(b.bw * b.ppc, b.ppc, b.ppc / 2, cols, work, &refined, st);
    if (e != hipSuccess) return e;
    if (prec == 0)
        k_sp_scatter<double><<<grid, kB, 0, st>>>(L.s0, ad, b, refined, (double*)q, q_total, q_base);
    else
        k_sp_scatter<float><<<grid, kB, 0, st>>>(L.s0, ad, b, refined, (float*)q, q_total, q_base);
    return hipGetLastError();
}

hipError_t cgc_particles(int nc, int prec, const Lattice& L, const SpBlock& b, int orientation, const double* gauss,
                         double* work, int* dir, double* pos0, double* pos1, double* vel, void* Q0, void* Q1,
                         uint64_t count, uint64_t base, hipStream_t st)
{
    const int ad = nc == 2 ? 3 : 8;
    const uint32_t cols = (uint32_t)ad * L.s0;
    const uint64_t np = (uint64_t)b.bw * L.s0 * b.ppc;
    k_pp_init<<<blocks(np), kB, 0, st>>>(L, ad, b, orientation, gauss, work, dir, pos0, pos1, vel, count, base);
    const double* refined = nullptr;
    hipError_t e = refine(b.bw * b.ppc, b.ppc, 0, cols, work, &refined, st);
    if (e != hipSuccess) return e;
    if (prec == 0)
        k_pp_scatter<double><<<blocks(np), kB, 0, st>>>(L.s0, ad, b, refined, (double*)Q0, (double*)Q1, count, base);
    else
        k_pp_scatter<float><<<blocks(np), kB, 0, st>>>(L.s0, ad, b, refined, (float*)Q0, (float*)Q1, count, base);
    return hipGetLastError();
}

size_t java_gaussian_work(uint64_t trials)
{
    const uint64_t nb = (trials + kScan - 1) / kScan;
    return trials * (sizeof(double2) + 2 * sizeof(uint32_t)) + nb * sizeof(uint32_t) + sizeof(uint64_t) + 64;
}

hipError_t java_gaussians(uint64_t scrambled_seed, uint64_t count, uint64_t trials, void* work, double* out,
                          uint64_t* accepted, hipStream_t st)
{
    const uint64_t nb = (trials + kScan - 1) / kScan;
    char* w = (char*)work;
    double2* vals = (double2*)w;
    w += trials * sizeof(double2);
    uint32_t* flag = (uint32_t*)w;
    w += trials * sizeof(uint32_t);
    uint32_t* pos = (uint32_t*)w;
    w += trials * sizeof(uint32_t);
    uint32_t* sums = (uint32_t*)w;
    w += nb * sizeof(uint32_t);
    uint64_t* total = (uint64_t*)(((uintptr_t)w + 7) & ~(uintptr_t)7);
    k_jr_trials<<<blocks(trials), kB, 0, st>>>(scrambled_seed, trials, vals, flag);
    k_scan_blocks<<<(unsigned)nb, kB, 0, st>>>(trials, flag, pos, sums);
    k_scan_sums<<<1, kScan, 0, st>>>((uint32_t)nb, sums, total);
    k_jr_scatter<<<blocks(trials), kB, 0, st>>>(trials, count, vals, flag, pos, sums, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(accepted, total, sizeof(uint64_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

hipError_t cgc_tadpole(int nc, uint32_t T, const double* V, double* out, hipStream_t st)
{
    if (nc == 2)
        k_tadpole<2><<<1, kB, 0, st>>>(T, V, out);
    else
        k_tadpole<3><<<1, kB, 0, st>>>(T, V, out);
    return hipGetLastError();
}

hipError_t cgc_dipole(int nc, uint32_t n1, uint32_t n2, double a1, double a2, double ds, int nbins, const double* V,
                      double* psum, unsigned long long* pcnt, double* corr, unsigned long long* cnt, hipStream_t st)
{
    const uint32_t T = n1 * n2;
    if (nc == 2)
        k_dipole_pairs<2><<<blocks(T), kB, 0, st>>>(n1, n2, a1, a2, ds, nbins, V, psum, pcnt);
    else
        k_dipole_pairs<3><<<blocks(T), kB, 0, st>>>(n1, n2, a1, a2, ds, nbins, V, psum, pcnt);
    if (nbins > 0) k_dipole_bins<<<(unsigned)nbins, kB, 0, st>>>(T, psum, pcnt, corr, cnt);
    return hipGetLastError();
}

}  // namespace opx
