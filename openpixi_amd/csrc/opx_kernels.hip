// opx_kernels.hip -- gfx950 kernels of the temporal-gauge Yang-Mills leapfrog.
//
// Device layout: 64-cell blocks, double components in pairs per cell (fpos in opx_kernels.h). A
// link field has 3 S components per cell (component k of link (cell, dir) is dir*S + k), an
// algebra field 3 AD (coefficient k of direction dir is dir*AD + k), the charge density AD.
// Consecutive lanes of a wavefront touch consecutive cells of one block, so every access of a
// component (or pair) is one contiguous wavefront access: 256 B (float), 1 KiB (double pair).
#include <mutex>
#include <type_traits>

#include "opx_buf.h"
#include "opx_group.h"
#include "opx_kernels.h"
#include "opx_order.h"

namespace opx {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ bool in_box(const Box& b, uint32_t x0, uint32_t x1, uint32_t x2)
{
    if (b.all) return true;
    return (int)x0 >= b.lo[0] && (int)x0 <= b.hi[0] && (int)x1 >= b.lo[1] && (int)x1 <= b.hi[1] &&
           (int)x2 >= b.lo[2] && (int)x2 <= b.hi[2];
}

struct Coord {
    uint32_t x0, x1, x2;
};

// whether plane x0 lies in one of the lattice's current slabs
__device__ __forceinline__ bool in_current(const Lattice& L, uint32_t x0)
{
    bool r = false;
#pragma unroll
    for (int i = 0; i < kMaxCurrentSlabs; i++) r |= (i < (int)L.j_slabs) & (x0 >= L.j_lo[i]) & (x0 < L.j_hi[i]);
    return r;
}

__device__ __forceinline__ Coord coords(const Lattice& L, uint32_t site)
{
    Coord c;
    uint32_t q = site / L.d2;
    c.x2 = site - q * L.d2;
    c.x0 = q / L.d1;
    c.x1 = q - c.x0 * L.d1;
    return c;
}

__device__ __forceinline__ uint32_t dim_of(const Lattice& L, int d) { return d == 0 ? L.d0 : (d == 1 ? L.d1 : L.d2); }
__device__ __forceinline__ uint32_t stride_of(const Lattice& L, int d) { return d == 0 ? L.s0 : (d == 1 ? L.s1 : 1u); }
__device__ __forceinline__ uint32_t coord_of(const Coord& c, int d) { return d == 0 ? c.x0 : (d == 1 ? c.x1 : c.x2); }

// periodic neighbour of `site` along d (Grid.shift, physics/grid/Grid.java:802-823)
__device__ __forceinline__ uint32_t fwd(const Lattice& L, uint32_t site, uint32_t c, int d)
{
    uint32_t n = dim_of(L, d), s = stride_of(L, d);
    return (c == n - 1) ? site - (n - 1) * s : site + s;
}
__device__ __forceinline__ uint32_t bwd(const Lattice& L, uint32_t site, uint32_t c, int d)
{
    uint32_t n = dim_of(L, d), s = stride_of(L, d);
    return (c == 0) ? site + (n - 1) * s : site - s;
}

// link (site, dir) of a link field / coefficients of direction dir of an algebra field (3 AD
// components per cell; dir 0 with comps = AD for the charge density)
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, T>::M load_link(const T* __restrict__ U, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    typename GR::M m;
    const T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) GR::gref(m, GR::slot(k)) = p[fcomp<T>(dir * GR::S + k, 3 * GR::S)];
    GR::complete(m);
    return m;
}

template <int NC, typename T>
__device__ __forceinline__ void store_link(T* __restrict__ U, uint32_t site, int dir, const typename Group<NC, T>::M& m)
{
    using GR = Group<NC, T>;
    T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) p[fcomp<T>(dir * GR::S + k, 3 * GR::S)] = GR::gget(m, GR::slot(k));
}

template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, T>::A load_alg(const T* __restrict__ X, uint32_t site, int dir,
                                                             uint32_t comps = 3 * Group<NC, T>::AD)
{
    using GR = Group<NC, T>;
    typename GR::A a;
    const T* p = fcell(X, site, comps);
#pragma unroll
    for (int k = 0; k < GR::AD; k++) aref(a, k) = p[fcomp<T>(dir * GR::AD + k, comps)];
    return a;
}

template <int NC, typename T>
__device__ __forceinline__ void store_alg(T* __restrict__ X, uint32_t site, int dir, typename Group<NC, T>::A a,
                                          uint32_t comps = 3 * Group<NC, T>::AD)
{
    using GR = Group<NC, T>;
    T* p = fcell(X, site, comps);
#pragma unroll
    for (int k = 0; k < GR::AD; k++) p[fcomp<T>(dir * GR::AD + k, comps)] = aref(a, k);
}

// Non-temporal cache policy on the per-link kernel's once-touched data (E read and
// written, Unext written): 1 to enable.
#ifndef OPX_STREAMING
#define OPX_STREAMING 1
#endif

// Field accessors of the fused kernels: component k of cell s of a field of n cells with
// `comps` components per cell. FieldB uses buffer instructions (the component's offset in an
// SGPR, the cell's in a VGPR); FieldG plain global pointers (fields of 4 GiB or more). FieldB's G
// is the field's group width (fvec; kGroup<NC, T> for the link and algebra fields of the fused
// kernels): with vw() > 1, ldv/stv move the G components of group k0 / G of a cell at once, as
// one 16-byte load and two 8-byte stores. Stores are kept to 8 bytes: 16-byte
// stores whose last data register was rewritten by the next instruction stored a wrong value in
// some lanes on MI355X.
template <typename T, uint32_t G = 1>
struct FieldB {
    Buf<T> b;
    uint32_t bstride;   // bytes per 64-cell block
    __device__ __forceinline__ FieldB(const T* p, uint32_t n_, uint32_t comps_)
        : b(p, comps_ * (uint32_t)padded_cells(n_) * (uint32_t)sizeof(T)), bstride(comps_ * 64u * (uint32_t)sizeof(T)) {}
    __device__ __forceinline__ uint32_t vw() const { return G; }
    __device__ __forceinline__ uint32_t vo(uint32_t s) const { return (s >> 6) * bstride + (s & 63u) * G * (uint32_t)sizeof(T); }
    __device__ __forceinline__ uint32_t so(uint32_t k) const { return ((k / G) * G * 64u + k % G) * (uint32_t)sizeof(T); }
    __device__ __forceinline__ T ld(uint32_t k, uint32_t s) const { return b.ld(vo(s), so(k)); }
    __device__ __forceinline__ void st(uint32_t k, uint32_t s, T x) const { b.st(x, vo(s), so(k)); }
    __device__ __forceinline__ T ld_once(uint32_t k, uint32_t s) const { return b.ld(vo(s), so(k), OPX_STREAMING); }
    __device__ __forceinline__ void st_once(uint32_t k, uint32_t s, T x) const { b.st(x, vo(s), so(k), OPX_STREAMING); }
    __device__ __forceinline__ void ldv(uint32_t k0, uint32_t s, T* out, int once) const { b.ld16(out, vo(s), so(k0), once); }
    __device__ __forceinline__ void stv(uint32_t k0, uint32_t s, const T* in, int once) const { b.st16(in, vo(s), so(k0), once); }
};
template <typename T>
struct FieldG {
    T* p;
    uint32_t comps;
    __device__ __forceinline__ FieldG(const T* p_, uint32_t, uint32_t comps_) : p(const_cast<T*>(p_)), comps(comps_) {}
    __device__ __forceinline__ uint32_t vw() const { return 1; }
    __device__ __forceinline__ T ld(uint32_t k, uint32_t s) const { return fcell(p, s, comps)[fcomp<T>(k, comps)]; }
    __device__ __forceinline__ void st(uint32_t k, uint32_t s, T v) const { fcell(p, s, comps)[fcomp<T>(k, comps)] = v; }
    __device__ __forceinline__ T ld_once(uint32_t k, uint32_t s) const { return ld(k, s); }
    __device__ __forceinline__ void st_once(uint32_t k, uint32_t s, T v) const { st(k, s, v); }
    __device__ __forceinline__ void ldv(uint32_t, uint32_t, T*, int) const {}
    __device__ __forceinline__ void stv(uint32_t, uint32_t, const T*, int) const {}
};

// N consecutive components from `base` of cell `site` (base and N multiples of the group width)
template <int N, typename T, typename F>
__device__ __forceinline__ void fld_n(const F& f, uint32_t base, uint32_t site, T* out, int once)
{
    const uint32_t w = f.vw();
    if (w > 1) {
#pragma unroll
        for (int k = 0; k < N; k += (int)(16 / sizeof(T))) f.ldv(base + k, site, out + k, once);
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) out[k] = once ? f.ld_once(base + k, site) : f.ld(base + k, site);
    }
}
template <int N, typename T, typename F>
__device__ __forceinline__ void fst_n(const F& f, uint32_t base, uint32_t site, const T* in, int once)
{
    const uint32_t w = f.vw();
    if (w > 1) {
#pragma unroll
        for (int k = 0; k < N; k += (int)(16 / sizeof(T))) f.stv(base + k, site, in + k, once);
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) {
            if (once)
                f.st_once(base + k, site, in[k]);
            else
                f.st(base + k, site, in[k]);
        }
    }
}

// Each accessor moves the components one by one when the field's group width is 1 and group by
// group (ldv / stv) otherwise.
template <int NC, typename T, typename F>
__device__ __forceinline__ typename Group<NC, T>::M fload_link(const F& f, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    typename GR::M m;
    if (f.vw() > 1) {
        T v[GR::S];
        fld_n<GR::S, T>(f, (uint32_t)(dir * GR::S), site, v, 0);
#pragma unroll
        for (int k = 0; k < GR::S; k++) GR::gref(m, GR::slot(k)) = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < GR::S; k++) GR::gref(m, GR::slot(k)) = f.ld((uint32_t)(dir * GR::S + k), site);
    }
    GR::complete(m);
    return m;
}
// the stored reals only (two rows of an FP32 SU(3) link), without rebuilding row 2
template <int NC, typename T, typename F>
__device__ __forceinline__ typename Group<NC, T>::M fload_link_r01(const F& f, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    typename GR::M m = GR::zero();
    if (f.vw() > 1) {
        T v[GR::S];
        fld_n<GR::S, T>(f, (uint32_t)(dir * GR::S), site, v, 0);
#pragma unroll
        for (int k = 0; k < GR::S; k++) GR::gref(m, GR::slot(k)) = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < GR::S; k++) GR::gref(m, GR::slot(k)) = f.ld((uint32_t)(dir * GR::S + k), site);
    }
    return m;
}
template <int NC, typename T, typename F>
__device__ __forceinline__ void fstore_link(const F& f, uint32_t site, int dir, const typename Group<NC, T>::M& m)
{
    using GR = Group<NC, T>;
    if (f.vw() > 1) {
        T v[GR::S];
#pragma unroll
        for (int k = 0; k < GR::S; k++) v[k] = GR::gget(m, GR::slot(k));
        fst_n<GR::S, T>(f, (uint32_t)(dir * GR::S), site, v, 0);
    } else {
#pragma unroll
        for (int k = 0; k < GR::S; k++) f.st((uint32_t)(dir * GR::S + k), site, GR::gget(m, GR::slot(k)));
    }
}
template <int NC, typename T, typename F>
__device__ __forceinline__ typename Group<NC, T>::A fload_alg(const F& f, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    typename GR::A a;
    if (f.vw() > 1) {
        T v[GR::AD];
        fld_n<GR::AD, T>(f, (uint32_t)(dir * GR::AD), site, v, 0);
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(a, k) = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(a, k) = f.ld((uint32_t)(dir * GR::AD + k), site);
    }
    return a;
}
template <int NC, typename T, typename F>
__device__ __forceinline__ typename Group<NC, T>::A fload_alg_once(const F& f, uint32_t site, int dir)
{
    using GR = Group<NC, T>;
    typename GR::A a;
    if (f.vw() > 1) {
        T v[GR::AD];
        fld_n<GR::AD, T>(f, (uint32_t)(dir * GR::AD), site, v, 1);
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(a, k) = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(a, k) = f.ld_once((uint32_t)(dir * GR::AD + k), site);
    }
    return a;
}
template <int NC, typename T, typename F>
__device__ __forceinline__ void fstore_alg_once(const F& f, uint32_t site, int dir, typename Group<NC, T>::A a)
{
    using GR = Group<NC, T>;
    if (f.vw() > 1) {
        T v[GR::AD];
#pragma unroll
        for (int k = 0; k < GR::AD; k++) v[k] = aref(a, k);
        fst_n<GR::AD, T>(f, (uint32_t)(dir * GR::AD), site, v, 1);
    } else {
#pragma unroll
        for (int k = 0; k < GR::AD; k++) f.st_once((uint32_t)(dir * GR::AD + k), site, aref(a, k));
    }
}
template <int NC, typename T, typename F>
__device__ __forceinline__ void fstore_link_once(const F& f, uint32_t site, int dir, const typename Group<NC, T>::M& m)
{
    using GR = Group<NC, T>;
    if (f.vw() > 1) {
        T v[GR::S];
#pragma unroll
        for (int k = 0; k < GR::S; k++) v[k] = GR::gget(m, GR::slot(k));
        fst_n<GR::S, T>(f, (uint32_t)(dir * GR::S), site, v, 1);
    } else {
#pragma unroll
        for (int k = 0; k < GR::S; k++) f.st_once((uint32_t)(dir * GR::S + k), site, GR::gget(m, GR::slot(k)));
    }
}
template <int NC, typename T, typename F>
__device__ __forceinline__ void fstore_alg(const F& f, uint32_t site, int dir, typename Group<NC, T>::A a)
{
    using GR = Group<NC, T>;
    if (f.vw() > 1) {
        T v[GR::AD];
#pragma unroll
        for (int k = 0; k < GR::AD; k++) v[k] = aref(a, k);
        fst_n<GR::AD, T>(f, (uint32_t)(dir * GR::AD), site, v, 0);
    } else {
#pragma unroll
        for (int k = 0; k < GR::AD; k++) f.st((uint32_t)(dir * GR::AD + k), site, aref(a, k));
    }
}

// ---------------------------------------------------------------- kernel 1
// FastTYMSolver.CombinedUpdate.execute (physics/fields/FastTYMSolver.java:40-62)
// for one link (cell, dir = blockIdx.y):
//   S   = sum_{j != dir} a_j^-2 [U_j(x+d) U_d(x+j)^+ U_j(x)^+ + (U_d(x-j) U_j(x+d-j))^+ U_j(x-j)]
//   E  += at proj(U_d S) ;  E += uf_d J   (when a current is present)
//   Unext = exp(-i at E) U_d
// waves per SIMD requested from the register allocator (OPX_LINK_WAVES, default 4 for fp32)
#ifndef OPX_LINK_WAVES
#define OPX_LINK_WAVES 4
#endif
#ifndef OPX_LINK_WAVES_FP64
#define OPX_LINK_WAVES_FP64 2
#endif
// The update of one link (cell `site`, direction d) on lattice L.
template <int NC, typename T, bool HAS_J, typename Fld>
__device__ __forceinline__ void link_update(const Lattice& L, const Fld& U, const Fld& Unext, const Fld& E,
                                            const Fld& J, uint32_t site, int d)
{
    using GR = Group<NC, T>;
    using M = typename GR::M;
    using A = typename GR::A;
    const Coord c = coords(L, site);
    const uint32_t xpd = fwd(L, site, coord_of(c, d), d);
    M S = GR::zero();
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (j == d) continue;
        const uint32_t cj = coord_of(c, j);
        const uint32_t xpj = fwd(L, site, cj, j);
        const uint32_t xmj = bwd(L, site, cj, j);
        const uint32_t xpd_mj = bwd(L, xpd, cj, j);
        M u1, u2;
        if constexpr (GR::S != GR::G) {
            // two-row storage: every staple factor is a group element, so each product is
            // formed in its first two rows and completed by the row-2 identity; the left
            // factors are read in the two rows stored
            u1 = mul_adj_r01(mul_adj_r01(fload_link_r01<NC, T>(U, xpd, j), fload_link<NC, T>(U, xpj, d)),
                             fload_link<NC, T>(U, site, j));
            su3_third_row(u1);
            M y = mul_r01(fload_link_r01<NC, T>(U, xmj, d), fload_link<NC, T>(U, xpd_mj, j));
            su3_third_row(y);
            u2 = adj_mul_r01(y, fload_link<NC, T>(U, xmj, j));
            su3_third_row(u2);
        } else {
            u1 = mul_adj(mul_adj(fload_link<NC, T>(U, xpd, j), fload_link<NC, T>(U, xpj, d)),
                         fload_link<NC, T>(U, site, j));
            u2 = adj_mul(mul(fload_link<NC, T>(U, xmj, d), fload_link<NC, T>(U, xpd_mj, j)),
                         fload_link<NC, T>(U, xmj, j));
        }
        acc(u1, u2);
        acc_scaled(S, u1, (T)L.area[j]);
    }
    const M Ud = fload_link<NC, T>(U, site, d);
    const A pr = proj(mul(Ud, S));
    A e = fload_alg_once<NC, T>(E, site, d);
    const T at = (T)L.at;
#pragma unroll
    for (int k = 0; k < GR::AD; k++) aref(e, k) += aget(pr, k) * at;
    if (HAS_J && in_current(L, c.x0)) {
        const A jj = fload_alg<NC, T>(J, site, d);
        const T uf = (T)L.uf[d];
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(e, k) += aget(jj, k) * uf;
    }
    fstore_alg_once<NC, T>(E, site, d, e);
    if constexpr (GR::S != GR::G)
        fstore_link_once<NC, T>(Unext, site, d, mul_r01(exp_link(e, -at), Ud));   // the two stored rows
    else
        fstore_link_once<NC, T>(Unext, site, d, mul(exp_link(e, -at), Ud));
}

template <int NC, typename T, bool HAS_J, bool BUF>
__global__ __launch_bounds__(kBlock, (sizeof(T) == 4 ? OPX_LINK_WAVES : OPX_LINK_WAVES_FP64)) void k_combined_update(Lattice L, LinkOrder order, Box active, const T* __restrict__ Up,
                                                            T* __restrict__ Unextp, T* __restrict__ Ep,
                                                            const T* __restrict__ Jp)
{
    using Fld = typename std::conditional<BUF, FieldB<T, kGroup<NC, T>>, FieldG<T>>::type;
    const Fld U(Up, L.n, 3 * Group<NC, T>::S), Unext(Unextp, L.n, 3 * Group<NC, T>::S);
    const Fld E(Ep, L.n, 3 * Group<NC, T>::AD), J(Jp, L.n, 3 * Group<NC, T>::AD);
    // Block order (opx_order.h) with the direction fastest: the three links of a cell
    // are updated by consecutive logical blocks. XCD k (= blockIdx % 8) sweeps its
    // tiles plane by plane, so the x0 +- 1 neighbours of a block were (or are about to
    // be) read by blocks a few dozen positions away on the same XCD, i.e. from its L2.
    const uint32_t logical = xcd_logical(blockIdx.x, gridDim.x);
    const int d = (int)(logical % 3);
    uint32_t site;
    if (!tile_site<kBlock>(L, order, logical / 3, site)) return;
    const Coord c = coords(L, site);
    if (!in_box(active, c.x0, c.x1, c.x2)) return;
    link_update<NC, T, HAS_J>(L, U, Unext, E, J, site, d);
}

// ---------------------------------------------------------------- cell form
// The three link updates of one cell in one thread. In each plane (a, b), a < b, the two upper
// staples share one product: with P = U_b(x+a) U_a(x+b)^+
//   link a: U_b(x+a) U_a(x+b)^+ U_b(x)^+ = P U_b(x)^+,   link b: U_a(x+b) U_b(x+a)^+ U_a(x)^+ = (U_a(x) P)^+,
// and the lower staples are those of the per-link form. Planes run (0,1), (0,2), (1,2); a link
// is finished as soon as its staple sum is complete, which keeps at most two sums live. The
// staple sums accumulate in the per-link form's order. A cell reads 33 links instead of the 39
// of three per-link threads, and forms 21 staple products instead of 24.
#ifndef OPX_CELL_WAVES
#define OPX_CELL_WAVES 2
#endif
// FP64 SU(3): one wave per SIMD, its 512 registers hold a cell without spilling (at two waves
// it spills; one is 7 % faster than the per-link form at 128^3, measured)
#ifndef OPX_CELL_WAVES_SU3_FP64
#define OPX_CELL_WAVES_SU3_FP64 1
#endif
// The new link reprojected onto the group in double precision (reunit), as k_reunitarize does
// to a stored link: only the rows the storage keeps enter.
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, T>::M reunit_link(const typename Group<NC, T>::M& u)
{
    using GR = Group<NC, T>;
    using GD = Group<NC, double>;
    typename GD::M ud;
#pragma unroll
    for (int k = 0; k < GR::G; k++) GD::gref(ud, k) = (double)GR::gget(u, k);
    ud = reunit(ud);
    typename GR::M o;
#pragma unroll
    for (int k = 0; k < GR::G; k++) GR::gref(o, k) = (T)GD::gget(ud, k);
    return o;
}

template <int NC, typename T, bool HAS_J, bool REUNIT, typename Fld>
__device__ __forceinline__ void cell_update(const Lattice& L, const Fld& U, const Fld& Unext, const Fld& E,
                                            const Fld& J, uint32_t site)
{
    using GR = Group<NC, T>;
    using M = typename GR::M;
    using A = typename GR::A;
    constexpr bool TWO_ROWS = GR::S != GR::G;
    const Coord c = coords(L, site);
    uint32_t xp[3], xm[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        xp[d] = fwd(L, site, coord_of(c, d), d);
        xm[d] = bwd(L, site, coord_of(c, d), d);
    }
    // P = U_b(x+a) U_a(x+b)^+
    auto upper = [&](int a, int b) -> M {
        if constexpr (TWO_ROWS) {
            M P = mul_adj_r01(fload_link_r01<NC, T>(U, xp[a], b), fload_link<NC, T>(U, xp[b], a));
            su3_third_row(P);
            return P;
        } else {
            return mul_adj(fload_link<NC, T>(U, xp[a], b), fload_link<NC, T>(U, xp[b], a));
        }
    };
    // staple of link a in plane (a, b): P U_b(x)^+ + (U_a(x-b) U_b(x+a-b))^+ U_b(x-b); ub() gives U_b(x)
    auto staple_a = [&](int a, int b, const M& P, auto ub) -> M {
        const uint32_t xpa_mb = bwd(L, xp[a], coord_of(c, b), b);
        M u, l;
        if constexpr (TWO_ROWS) {
            u = mul_adj_r01(P, ub());
            su3_third_row(u);
            M y = mul_r01(fload_link_r01<NC, T>(U, xm[b], a), fload_link<NC, T>(U, xpa_mb, b));
            su3_third_row(y);
            l = adj_mul_r01(y, fload_link<NC, T>(U, xm[b], b));
            su3_third_row(l);
        } else {
            u = mul_adj(P, ub());
            l = adj_mul(mul(fload_link<NC, T>(U, xm[b], a), fload_link<NC, T>(U, xpa_mb, b)),
                        fload_link<NC, T>(U, xm[b], b));
        }
        acc(u, l);
        return u;
    };
    // staple of link b in plane (a, b): (U_a(x) P)^+ + (U_b(x-a) U_a(x+b-a))^+ U_a(x-a); ua() gives
    // U_a(x), of which the two stored rows are read in the two-row form
    auto staple_b = [&](int a, int b, const M& P, auto ua) -> M {
        const uint32_t xpb_ma = bwd(L, xp[b], coord_of(c, a), a);
        M u, l;
        if constexpr (TWO_ROWS) {
            M t = mul_r01(ua(), P);
            su3_third_row(t);
            u = adj(t);
            M y = mul_r01(fload_link_r01<NC, T>(U, xm[a], b), fload_link<NC, T>(U, xpb_ma, a));
            su3_third_row(y);
            l = adj_mul_r01(y, fload_link<NC, T>(U, xm[a], a));
            su3_third_row(l);
        } else {
            u = adj(mul(ua(), P));
            l = adj_mul(mul(fload_link<NC, T>(U, xm[a], b), fload_link<NC, T>(U, xpb_ma, a)),
                        fload_link<NC, T>(U, xm[a], a));
        }
        acc(u, l);
        return u;
    };
    const T at = (T)L.at;
    // E(x) of direction d plus the current term, read before any store of the cell
    const bool in_j = HAS_J && in_current(L, c.x0);
    auto e_start = [&](int d) -> A {
        A e = fload_alg_once<NC, T>(E, site, d);
        if (in_j) {
            const A jj = fload_alg<NC, T>(J, site, d);
            const T uf = (T)L.uf[d];
#pragma unroll
            for (int k = 0; k < GR::AD; k++) aref(e, k) += aget(jj, k) * uf;
        }
        return e;
    };
    // E(x) += fs proj(U_d(x) S) and the new link; ud() gives U_d(x)
    auto finish = [&](int d, const M& S, A e, auto ud, T fs) {
        const M Ud = ud();
        const A pr = proj(mul(Ud, S));
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(e, k) += aget(pr, k) * fs;
        fstore_alg_once<NC, T>(E, site, d, e);
        M nu;
        if constexpr (TWO_ROWS)
            nu = mul_r01(exp_link(e, -at), Ud);
        else
            nu = mul(exp_link(e, -at), Ud);
        if constexpr (REUNIT) nu = reunit_link<NC, T>(nu);
        fstore_link_once<NC, T>(Unext, site, d, nu);
    };
    // In float the staple sum of direction d over its planes (d, j1), (d, j2), j1 < j2, is kept as
    // S_d = s_j1 + (a_j2^-2 / a_j1^-2) s_j2 and E gets at a_j1^-2 proj(U_d S_d) (Lattice::cell_ratio
    // and cell_scale): one scaled accumulation per direction instead of two. In float each own link
    // is also read once per use group: U_0 serves finish(0) and then staple_b(0, 2), U_1 finish(1)
    // and then staple_b(1, 2), U_2 staple_a(1, 2) and finish(2) (a read after a store of the cell
    // is not merged with an earlier one). The double form sums a_j^-2 (upper + lower) plane by plane
    // and reads an own link at each use: with the float arrangement its launch takes 2 % longer.
    // E of all three directions (and the current term) is read before the first store to E: a load
    // of E cannot move above a store to it, and read late its latency is exposed at every update.
    constexpr bool FLT = sizeof(T) == 4;
    const A e0 = e_start(0), e1 = e_start(1), e2 = e_start(2);
    auto own = [&](int d) { return [&, d]() -> M { return fload_link<NC, T>(U, site, d); }; };
    auto held = [](const M& m) { return [&m]() -> M { return m; }; };
    auto first = [&](M& S, const M& s, int j) {
        if constexpr (FLT) {
            S = s;
        } else {
            S = GR::zero();
            acc_scaled(S, s, (T)L.area[j]);
        }
    };
    auto second = [&](M& S, const M& s, int d, int j) { acc_scaled(S, s, FLT ? (T)L.cell_ratio[d] : (T)L.area[j]); };
    auto scale = [&](int d) -> T { return FLT ? (T)L.cell_scale[d] : at; };
    M S0, S1, S2;
    {
        const M P = upper(0, 1);
        first(S0, staple_a(0, 1, P, own(1)), 1);
        if constexpr (TWO_ROWS)
            first(S1, staple_b(0, 1, P, [&]() -> M { return fload_link_r01<NC, T>(U, site, 0); }), 0);
        else
            first(S1, staple_b(0, 1, P, own(0)), 0);
    }
    if constexpr (FLT) {
        {
            const M P = upper(0, 2);
            second(S0, staple_a(0, 2, P, own(2)), 0, 2);
            const M U0 = own(0)();
            finish(0, S0, e0, held(U0), scale(0));
            first(S2, staple_b(0, 2, P, held(U0)), 0);
        }
        {
            const M P = upper(1, 2);
            const M U2 = own(2)();
            second(S1, staple_a(1, 2, P, held(U2)), 1, 2);
            const M U1 = own(1)();
            finish(1, S1, e1, held(U1), scale(1));
            second(S2, staple_b(1, 2, P, held(U1)), 2, 1);
            finish(2, S2, e2, held(U2), scale(2));
        }
    } else {
        {
            const M P = upper(0, 2);
            second(S0, staple_a(0, 2, P, own(2)), 0, 2);
            finish(0, S0, e0, own(0), scale(0));
            first(S2, staple_b(0, 2, P, own(0)), 0);
        }
        {
            const M P = upper(1, 2);
            second(S1, staple_a(1, 2, P, own(2)), 1, 2);
            finish(1, S1, e1, own(1), scale(1));
            second(S2, staple_b(1, 2, P, own(1)), 2, 1);
            finish(2, S2, e2, own(2), scale(2));
        }
    }
}

// ---------------------------------------------------------------- plaquette form
// The force of the fused update from plaquettes computed once per cell. With
// P_dj(x) = U_d(x) U_j(x+d) U_d(x+j)^+ U_j(x)^+ the staple term is
//   proj(U_d S_d)(x) = sum_{j != d} a_j^-2 [proj P_dj(x) - act(proj P_dj(x-j), U_j(x-j)^+)],
// the second term being the backward staple U_d(x) U_j(x+d-j)^+ U_d(x-j)^+ U_j(x-j)
// = U_j(x-j)^+ P_dj(x-j)^+ U_j(x-j). Kernel A stores proj P for the three planes
// (0,1), (0,2), (1,2) of every cell; P_jd = P_dj^+ flips the sign. Kernel B does one
// link: the force from four stored plaquettes and two adjoint actions, then the
// E and link updates of FastTYMSolver.CombinedUpdate.

__device__ __forceinline__ int plane_of(int a, int b) { return a + b - 1; }   // (0,1)->0, (0,2)->1, (1,2)->2

// one thread per cell, blocks in the per-link kernel's tile order without the direction;
// REUNIT: the new links are reprojected onto the group in double before they are stored
template <int NC, typename T, bool HAS_J, bool BUF, bool REUNIT>
__global__ __launch_bounds__(kBlock, (NC == 3 && sizeof(T) == 8) ? OPX_CELL_WAVES_SU3_FP64 : OPX_CELL_WAVES) void k_cell_update(Lattice L, LinkOrder order, Box active,
                                                                        const T* __restrict__ Up, T* __restrict__ Unextp,
                                                                        T* __restrict__ Ep, const T* __restrict__ Jp)
{
    using Fld = typename std::conditional<BUF, FieldB<T, kGroup<NC, T>>, FieldG<T>>::type;
    const Fld U(Up, L.n, 3 * Group<NC, T>::S), Unext(Unextp, L.n, 3 * Group<NC, T>::S);
    const Fld E(Ep, L.n, 3 * Group<NC, T>::AD), J(Jp, L.n, 3 * Group<NC, T>::AD);
    uint32_t site;
    if (!tile_site(L, order, xcd_logical(blockIdx.x, gridDim.x), site)) return;
    const Coord c = coords(L, site);
    if (!in_box(active, c.x0, c.x1, c.x2)) return;
    cell_update<NC, T, HAS_J, REUNIT>(L, U, Unext, E, J, site);
}

#ifndef OPX_PLAQ_WAVES
#define OPX_PLAQ_WAVES 4
#endif
#ifndef OPX_PLAQ_UPDATE_WAVES
#define OPX_PLAQ_UPDATE_WAVES 6
#endif
template <int NC, typename T, bool BUF>
__global__ __launch_bounds__(kBlock, OPX_PLAQ_WAVES) void k_plaquettes(Lattice L, LinkOrder order, const T* __restrict__ Up,
                                                       T* __restrict__ Pp)
{
    using GR = Group<NC, T>;
    using Fld = typename std::conditional<BUF, FieldB<T, kGroup<NC, T>>, FieldG<T>>::type;
    const Fld U(Up, L.n, 3 * GR::S), P(Pp, L.n, 3 * GR::AD);
    uint32_t site;
    if (!tile_site(L, order, xcd_logical(blockIdx.x, gridDim.x), site)) return;
    const Coord c = coords(L, site);
    const uint32_t x0p = fwd(L, site, c.x0, 0), x1p = fwd(L, site, c.x1, 1), x2p = fwd(L, site, c.x2, 2);
    const typename GR::M u0 = fload_link<NC, T>(U, site, 0);
    const typename GR::M u1 = fload_link<NC, T>(U, site, 1);
    // P_01 = U_0(x) U_1(x+0) [U_1(x) U_0(x+1)]^+
    fstore_alg<NC, T>(P, site, 0,
                      proj_mul_adj(mul(u0, fload_link<NC, T>(U, x0p, 1)), mul(u1, fload_link<NC, T>(U, x1p, 0))));
    const typename GR::M u2 = fload_link<NC, T>(U, site, 2);
    // P_02 = U_0(x) U_2(x+0) [U_2(x) U_0(x+2)]^+
    fstore_alg<NC, T>(P, site, 1,
                      proj_mul_adj(mul(u0, fload_link<NC, T>(U, x0p, 2)), mul(u2, fload_link<NC, T>(U, x2p, 0))));
    // P_12 = U_1(x) U_2(x+1) [U_2(x) U_1(x+2)]^+
    fstore_alg<NC, T>(P, site, 2,
                      proj_mul_adj(mul(u1, fload_link<NC, T>(U, x1p, 2)), mul(u2, fload_link<NC, T>(U, x2p, 1))));
}

template <int NC, typename T, bool HAS_J, bool BUF>
__global__ __launch_bounds__(kBlock, (sizeof(T) == 4 ? OPX_PLAQ_UPDATE_WAVES : OPX_LINK_WAVES_FP64)) void k_update_plaq(
    Lattice L, LinkOrder order, Box active, const T* __restrict__ Up, const T* __restrict__ Pp,
    T* __restrict__ Unextp, T* __restrict__ Ep, const T* __restrict__ Jp)
{
    using GR = Group<NC, T>;
    using A = typename GR::A;
    using Fld = typename std::conditional<BUF, FieldB<T, kGroup<NC, T>>, FieldG<T>>::type;
    const Fld U(Up, L.n, 3 * GR::S), Unext(Unextp, L.n, 3 * GR::S);
    const Fld P(Pp, L.n, 3 * GR::AD), E(Ep, L.n, 3 * GR::AD), J(Jp, L.n, 3 * GR::AD);
    const uint32_t logical = xcd_logical(blockIdx.x, gridDim.x);
    const int d = (int)(logical % 3);
    uint32_t site;
    if (!tile_site(L, order, logical / 3, site)) return;
    const Coord c = coords(L, site);
    if (!in_box(active, c.x0, c.x1, c.x2)) return;
    A f = GR::azero();
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (j == d) continue;
        const int pl = d < j ? plane_of(d, j) : plane_of(j, d);
        const uint32_t xmj = bwd(L, site, coord_of(c, j), j);
        const A back = act_adj(fload_alg<NC, T>(P, xmj, pl), fload_link<NC, T>(U, xmj, j));
        const A here = fload_alg<NC, T>(P, site, pl);
        const T w = d < j ? (T)L.area[j] : -(T)L.area[j];
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(f, k) += w * (aget(here, k) - aget(back, k));
    }
    A e = fload_alg<NC, T>(E, site, d);
    const T at = (T)L.at;
#pragma unroll
    for (int k = 0; k < GR::AD; k++) aref(e, k) += aget(f, k) * at;
    if (HAS_J) {
        const A jj = fload_alg<NC, T>(J, site, d);
        const T uf = (T)L.uf[d];
#pragma unroll
        for (int k = 0; k < GR::AD; k++) aref(e, k) += aget(jj, k) * uf;
    }
    fstore_alg<NC, T>(E, site, d, e);
    const typename GR::M Ud = fload_link<NC, T>(U, site, d);
    fstore_link<NC, T>(Unext, site, d, mul(exp_link(e, -at), Ud));
}

// ---------------------------------------------------------------- batched events
// Persistent form: a fixed grid loops over the work items of all events. A work
// item is kBlock consecutive cells of one direction of one event; items are
// numbered event by event (cells-block major, direction minor). XCD k
// (= blockIdx % 8) owns the contiguous item range [k T/8, (k+1) T/8) and its
// workgroups stride through it, so every workgroup exits after a bounded number
// of items and neighbouring cells are handled by the same XCD.
template <int NC, typename T>
__global__ __launch_bounds__(kBlock, (sizeof(T) == 4 ? OPX_LINK_WAVES : OPX_LINK_WAVES_FP64)) void k_batch_update(
    const BatchEvent* __restrict__ ev, int nev, uint32_t total_items, const T* __restrict__ Ubase,
    T* __restrict__ Unbase, T* __restrict__ Ebase)
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t xcd = b & 7, q8 = nb >> 3, r8 = nb & 7;
    const uint32_t wg_on_xcd = q8 + (xcd < r8 ? 1u : 0u);
    const uint32_t w = b >> 3;
    const uint32_t i0 = (uint32_t)(((uint64_t)total_items * xcd) / 8), i1 = (uint32_t)(((uint64_t)total_items * (xcd + 1)) / 8);
    for (uint32_t item = i0 + w; item < i1; item += wg_on_xcd) {
        // event of this item: last e with ev[e].item_start <= item
        int lo = 0, hi = nev - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ev[mid].item_start <= item) lo = mid;
            else hi = mid - 1;
        }
        const BatchEvent& E = ev[lo];
        const uint32_t local = item - E.item_start;
        const int d = (int)(local % 3);
        const uint32_t site = (local / 3) * kBlock + threadIdx.x;
        if (site >= E.L.n) continue;
        const FieldG<T> U(Ubase + E.link_offset, E.L.n, 3 * Group<NC, T>::S), Un(Unbase + E.link_offset, E.L.n, 3 * Group<NC, T>::S);
        const FieldG<T> Ef(Ebase + E.alg_offset, E.L.n, 3 * Group<NC, T>::AD);
        link_update<NC, T, false>(E.L, U, Un, Ef, Ef, site, d);
    }
}


// Persistent cell form over a batch: work item i of event e covers cells
// [256 (i - item_start/3), +256) of that event, all three links of each (the per-link items of
// an event number 3 ceil(n/256), so item_start/3 numbers the cell items); same XCD-aware
// static schedule as k_batch_update.
template <int NC, typename T, bool BUF, bool REUNIT>
__global__ __launch_bounds__(kBlock, OPX_CELL_WAVES) void k_batch_cell(const BatchEvent* __restrict__ ev, int nev,
                                                                       uint32_t total_items, const T* __restrict__ Ubase,
                                                                       T* __restrict__ Unbase, T* __restrict__ Ebase)
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t xcd = b & 7, q8 = nb >> 3, r8 = nb & 7;
    const uint32_t wg_on_xcd = q8 + (xcd < r8 ? 1u : 0u);
    const uint32_t w = b >> 3;
    const uint32_t i0 = (uint32_t)(((uint64_t)total_items * xcd) / 8), i1 = (uint32_t)(((uint64_t)total_items * (xcd + 1)) / 8);
    for (uint32_t item = i0 + w; item < i1; item += wg_on_xcd) {
        int lo = 0, hi = nev - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ev[mid].item_start / 3 <= item) lo = mid;
            else hi = mid - 1;
        }
        const BatchEvent& E = ev[lo];
        const uint32_t site = (item - E.item_start / 3) * kBlock + threadIdx.x;
        if (site >= E.L.n) continue;
        // the event's fields, by buffer instructions when every event is below 4 GiB
        using GR = Group<NC, T>;
        using Fld = typename std::conditional<BUF, FieldB<T, kGroup<NC, T>>, FieldG<T>>::type;
        const Fld U(Ubase + E.link_offset, E.L.n, 3 * GR::S), Un(Unbase + E.link_offset, E.L.n, 3 * GR::S);
        const Fld Ef(Ebase + E.alg_offset, E.L.n, 3 * GR::AD);
        cell_update<NC, T, false, REUNIT>(E.L, U, Un, Ef, Ef, site);
    }
}

// ---------------------------------------------------------------- kernel 2
// FastTYMSolver.UpdateLinks.execute (physics/fields/FastTYMSolver.java:65-84)
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_step_links(Lattice L, Box active, const T* __restrict__ U,
                                                       T* __restrict__ Unext, const T* __restrict__ E)
{
    const uint32_t site = blockIdx.x * kBlock + threadIdx.x;
    const int d = blockIdx.y;
    if (site >= L.n) return;
    const Coord c = coords(L, site);
    if (!in_box(active, c.x0, c.x1, c.x2)) return;
    const auto e = load_alg<NC, T>(E, site, d);
    const auto u = load_link<NC, T>(U, site, d);
    store_link<NC, T>(Unext, site, d, mul(exp_link(e, -(T)L.at), u));
}

// ---------------------------------------------------------------- kernel 3

template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_reunitarize(uint32_t n, T* __restrict__ U)
{
    using GR = Group<NC, T>;
    using GD = Group<NC, double>;
    const uint32_t site = blockIdx.x * kBlock + threadIdx.x;
    const int d = blockIdx.y;
    if (site >= n) return;
    typename GR::M u = load_link<NC, T>(U, site, d);
    typename GD::M ud;
#pragma unroll
    for (int k = 0; k < GR::G; k++) GD::gref(ud, k) = (double)GR::gget(u, k);
    ud = reunit(ud);
#pragma unroll
    for (int k = 0; k < GR::G; k++) GR::gref(u, k) = (T)GD::gget(ud, k);
    store_link<NC, T>(U, site, d, u);
}

// reprojection of every link of a batch (same static schedule as k_batch_update)
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_batch_reunitarize(const BatchEvent* __restrict__ ev, int nev,
                                                              uint32_t total_items, T* __restrict__ Ubase)
{
    using GR = Group<NC, T>;
    using GD = Group<NC, double>;
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t xcd = b & 7, q8 = nb >> 3, r8 = nb & 7;
    const uint32_t wg_on_xcd = q8 + (xcd < r8 ? 1u : 0u);
    const uint32_t w = b >> 3;
    const uint32_t i0 = (uint32_t)(((uint64_t)total_items * xcd) / 8), i1 = (uint32_t)(((uint64_t)total_items * (xcd + 1)) / 8);
    for (uint32_t item = i0 + w; item < i1; item += wg_on_xcd) {
        int lo = 0, hi = nev - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ev[mid].item_start <= item) lo = mid;
            else hi = mid - 1;
        }
        const BatchEvent& E = ev[lo];
        const uint32_t local = item - E.item_start;
        const int d = (int)(local % 3);
        const uint32_t site = (local / 3) * kBlock + threadIdx.x;
        if (site >= E.L.n) continue;
        T* U = Ubase + E.link_offset;
        typename GR::M u = load_link<NC, T>(U, site, d);
        typename GD::M ud;
#pragma unroll
        for (int k = 0; k < GR::G; k++) GD::gref(ud, k) = (double)GR::gget(u, k);
        ud = reunit(ud);
#pragma unroll
        for (int k = 0; k < GR::G; k++) GR::gref(u, k) = (T)GD::gget(ud, k);
        store_link<NC, T>(U, site, d, u);
    }
}

// ---------------------------------------------------------------- init helper
// U = exp(i sum_a c_a t_a) from double coefficients in [dir][coef][cell] layout.
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_links_from_coeffs(uint32_t n, const double* __restrict__ C,
                                                              T* __restrict__ U)
{
    using GD = Group<NC, double>;
    using GR = Group<NC, T>;
    const uint32_t site = blockIdx.x * kBlock + threadIdx.x;
    const int d = blockIdx.y;
    if (site >= n) return;
    typename GD::A a;
#pragma unroll
    for (int k = 0; k < GD::AD; k++) aref(a, k) = C[((size_t)d * GD::AD + k) * n + site];
    typename GD::M u = exp_link(a, 1.0);
    typename GR::M o;
#pragma unroll
    for (int k = 0; k < GR::G; k++) GR::gref(o, k) = (T)GD::gget(u, k);
    store_link<NC, T>(U, site, d, o);
}

// ---------------------------------------------------------------- measurements
// Compensated accumulation in double.
struct Kahan {
    double s, c;
    __device__ void add(double x)
    {
        double y = x - c;
        double t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// (s1,c1) + (s2,c2) with an error-free transformation of the head sums
__device__ __forceinline__ void merge(double& s1, double& c1, double s2, double c2)
{
    double s = s1 + s2;
    double bp = s - s1;
    double err = (s1 - (s - bp)) + (s2 - bp);
    s1 = s;
    c1 = c1 + c2 - err;
}

constexpr int kMeasureBlocks = 4096;
constexpr int kNQ = 8;

template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, double>::M load_link_d(const T* __restrict__ U, uint32_t site, int dir)
{
    using GD = Group<NC, double>;
    using GR = Group<NC, T>;
    typename GD::M m;
    const T* p = fcell(U, site, 3 * GR::S);
#pragma unroll
    for (int k = 0; k < GR::S; k++) GD::gref(m, GR::slot(k)) = (double)p[fcomp<T>(dir * GR::S + k, 3 * GR::S)];
    if constexpr (GR::S != GR::G) su3_third_row(m);   // rebuilt in double
    return m;
}

// plaquette P_{jk}(x) = U_j(x) U_k(x+j) U_j(x+k)^+ U_k(x)^+ (Grid.getPlaquette with o1 = o2 = +1)
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, double>::M plaquette_d(const Lattice& L, const T* __restrict__ U,
                                                                     uint32_t site, const Coord& c, int j, int k)
{
    uint32_t xpj = fwd(L, site, coord_of(c, j), j);
    uint32_t xpk = fwd(L, site, coord_of(c, k), k);
    return mul_adj(mul_adj(mul(load_link_d<NC, T>(U, site, j), load_link_d<NC, T>(U, xpj, k)),
                           load_link_d<NC, T>(U, xpk, j)),
                   load_link_d<NC, T>(U, site, k));
}

// comps: 3 AD for E (the default), AD for the charge density
template <int NC, typename T>
__device__ __forceinline__ typename Group<NC, double>::A load_alg_d(const T* __restrict__ X, uint32_t site, int dir,
                                                                    uint32_t comps = 3 * Group<NC, T>::AD)
{
    using GD = Group<NC, double>;
    typename GD::A a;
    const T* p = fcell(X, site, comps);
#pragma unroll
    for (int k = 0; k < GD::AD; k++) aref(a, k) = (double)p[fcomp<T>(dir * GD::AD + k, comps)];
    return a;
}


// B_i^2 at one time level: proj(P_jk) a_i/(a_j a_k), (j,k) cyclic (Grid.getB, Grid.java:877-900)
template <int NC, typename T>
__device__ __forceinline__ double b2(const Lattice& L, const T* __restrict__ U, uint32_t site, const Coord& c, int i)
{
    int j = (i + 1) % 3, k = (i + 2) % 3;
    auto p = proj(plaquette_d<NC, T>(L, U, site, c, j, k));
    double f = L.as[i] / (L.as[j] * L.as[k]);
    return square(p) * f * f;
}

// Gauss constraint G(x) (Grid.getGaussConstraint, Grid.java:944-968)
template <int NC, typename T>
__device__ __forceinline__ double gauss2(const Lattice& L, const T* __restrict__ U, const T* __restrict__ E,
                                         const T* __restrict__ rho, uint32_t site, const Coord& c)
{
    using GD = Group<NC, double>;
    typename GD::A g = GD::azero();
#pragma unroll 1
    for (int i = 0; i < 3; i++) {
        uint32_t xm = bwd(L, site, coord_of(c, i), i);
        auto em = load_alg_d<NC, T>(E, xm, i);
        auto u = load_link_d<NC, T>(U, xm, i);
        auto t = act(em, adj(u));
        auto ex = load_alg_d<NC, T>(E, site, i);
        double f = 1.0 / (L.as[i] * (L.g * L.as[i]));
#pragma unroll
        for (int k = 0; k < GD::AD; k++) aref(g, k) += (aref(ex, k) - aref(t, k)) * f;
    }
    if (rho) {
        auto r = load_alg_d<NC, T>(rho, site, 0, Group<NC, T>::AD);
#pragma unroll
        for (int k = 0; k < GD::AD; k++) aref(g, k) -= aref(r, k);
    }
    return square(g);
}

template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_measure(Lattice L, Box eval, const T* __restrict__ U,
                                                    const T* __restrict__ Unext, const T* __restrict__ E,
                                                    const T* __restrict__ rho, double* __restrict__ partials)
{
    Kahan acc[kNQ];
#pragma unroll
    for (int q = 0; q < kNQ; q++) acc[q] = Kahan{0.0, 0.0};
    double uf[3];
    for (int i = 0; i < 3; i++) uf[i] = 1.0 / ((L.g * L.as[i]) * (L.g * L.as[i]));
    for (uint32_t site = blockIdx.x * kBlock + threadIdx.x; site < L.n; site += gridDim.x * kBlock) {
        const Coord c = coords(L, site);
        if (!in_box(eval, c.x0, c.x1, c.x2)) continue;
#pragma unroll
        for (int i = 0; i < 3; i++) acc[i].add(square(load_alg_d<NC, T>(E, site, i)) * uf[i]);
        // one direction at a time: the plaquettes of both time levels are the largest live state
#pragma unroll 1
        for (int i = 0; i < 3; i++) {
            const double v = 0.5 * (b2<NC, T>(L, U, site, c, i) + b2<NC, T>(L, Unext, site, c, i)) * uf[i];
            if (i == 0)
                acc[3].add(v);
            else if (i == 1)
                acc[4].add(v);
            else
                acc[5].add(v);
        }
        acc[6].add(gauss2<NC, T>(L, U, E, rho, site, c));
        if (rho) acc[7].add(square(load_alg_d<NC, T>(rho, site, 0, Group<NC, T>::AD)));
    }
    // wavefront reduction (64 lanes), then across the 4 waves of the block
    __shared__ double sh[kBlock / 64][kNQ][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kNQ; q++) {
        double s = acc[q].s, cc = -acc[q].c;
        for (int off = 32; off > 0; off >>= 1) {
            double s2 = __shfl_down(s, off, 64);
            double c2 = __shfl_down(cc, off, 64);
            merge(s, cc, s2, c2);
        }
        if (lane == 0) {
            sh[wave][q][0] = s;
            sh[wave][q][1] = cc;
        }
    }
    __syncthreads();
    if (threadIdx.x < kNQ) {
        int q = threadIdx.x;
        double s = sh[0][q][0], cc = sh[0][q][1];
        for (int w = 1; w < kBlock / 64; w++) merge(s, cc, sh[w][q][0], sh[w][q][1]);
        partials[((size_t)blockIdx.x * kNQ + q) * 2 + 0] = s;
        partials[((size_t)blockIdx.x * kNQ + q) * 2 + 1] = cc;
    }
}

// Fixed-order reduction of the per-block partials; writes normalised results. Wave q
// reduces quantity q: lane l merges blocks l, l + 64, ... in order, then a fixed
// shuffle tree.
__global__ __launch_bounds__(64 * kNQ) void k_measure_final(int nblocks, double inv_n,
                                                            const double* __restrict__ partials,
                                                            double* __restrict__ out)
{
    const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s = 0.0, cc = 0.0;
    for (int b = lane; b < nblocks; b += 64)
        merge(s, cc, partials[((size_t)b * kNQ + q) * 2], partials[((size_t)b * kNQ + q) * 2 + 1]);
    for (int off = 32; off > 0; off >>= 1) {
        const double s2 = __shfl_down(s, off, 64);
        const double c2 = __shfl_down(cc, off, 64);
        merge(s, cc, s2, c2);
    }
    if (lane != 0) return;
    double tot = s + cc;
    if (q == 7)
        out[q] = sqrt(tot) * inv_n;
    else
        out[q] = tot * inv_n;
}

// ProjectedEnergyDensity: one block per slice along `dir`, fixed-order reduction.
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_projected_energy(Lattice L, Box eval, int dir, const T* __restrict__ U,
                                                             const T* __restrict__ Unext, const T* __restrict__ E,
                                                             double* __restrict__ out)
{
    const uint32_t slice = blockIdx.x;
    const uint32_t nd = dim_of(L, dir);
    const uint32_t per = L.n / nd;
    double area = (double)L.n / (double)nd;
    double uf[3];
    for (int i = 0; i < 3; i++) uf[i] = 1.0 / ((L.g * L.as[i]) * (L.g * L.as[i])) / area;
    Kahan a[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    // the other two dimensions enumerate the slice
    const int da = dir == 0 ? 1 : 0, db = dir == 2 ? 1 : 2;
    const uint32_t nb = dim_of(L, db);
    for (uint32_t m = threadIdx.x; m < per; m += kBlock) {
        uint32_t xa = m / nb, xb = m - (m / nb) * nb;
        uint32_t x[3];
        x[dir] = slice;
        x[da] = xa;
        x[db] = xb;
        uint32_t site = x[0] * L.s0 + x[1] * L.s1 + x[2];
        Coord c{x[0], x[1], x[2]};
        if (!in_box(eval, c.x0, c.x1, c.x2)) continue;
        for (int j = 0; j < 3; j++) {
            double el = 0.5 * square(load_alg_d<NC, T>(E, site, j)) * uf[j];
            double mag = 0.25 * (b2<NC, T>(L, U, site, c, j) + b2<NC, T>(L, Unext, site, c, j)) * uf[j];
            if (j == dir) {
                a[2].add(el);
                a[3].add(mag);
            } else {
                a[0].add(el);
                a[1].add(mag);
            }
        }
    }
    __shared__ double sh[kBlock / 64][4][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < 4; q++) {
        double s = a[q].s, cc = -a[q].c;
        for (int off = 32; off > 0; off >>= 1) {
            double s2 = __shfl_down(s, off, 64);
            double c2 = __shfl_down(cc, off, 64);
            merge(s, cc, s2, c2);
        }
        if (lane == 0) {
            sh[wave][q][0] = s;
            sh[wave][q][1] = cc;
        }
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        int q = threadIdx.x;
        double s = sh[0][q][0], cc = sh[0][q][1];
        for (int w = 1; w < kBlock / 64; w++) merge(s, cc, sh[w][q][0], sh[w][q][1]);
        out[(size_t)q * nd + slice] = s + cc;
    }
}

// ---------------------------------------------------------------- Poynting vector
// ProjectedEnergyDensity.PoyntingComputation (diagnostics/methods/ProjectedEnergyDensity.java:178-262):
// the longitudinal Poynting vector per slice along `dir`, spatially and time averaged
// (out[0, nd)) and only time averaged (out[nd, 2 nd)). Every cell counts: the reference's
// computation has no evaluation-region test. SU(2) only, as in the reference, whose SU(3)
// algebra dot product throws.
__device__ __forceinline__ Coord shifted(const Lattice& L, Coord c, int d, int o)
{
    const uint32_t n = dim_of(L, d);
    uint32_t& v = d == 0 ? c.x0 : (d == 1 ? c.x1 : c.x2);
    v = o > 0 ? (v + 1 == n ? 0 : v + 1) : (v == 0 ? n - 1 : v - 1);
    return c;
}
__device__ __forceinline__ uint32_t site_of(const Lattice& L, const Coord& c) { return c.x0 * L.s0 + c.x1 * L.s1 + c.x2; }

// Grid.getLink(index, d, o, t) (Grid.java:712-727): U_d(x), or U_d(x - d)^+ for o < 0
template <typename T>
__device__ __forceinline__ Su2<double> link_or(const Lattice& L, const T* __restrict__ U, const Coord& c, int d, int o)
{
    if (o > 0) return load_link_d<2, T>(U, site_of(L, c), d);
    return adj(load_link_d<2, T>(U, site_of(L, shifted(L, c, d, -1)), d));
}

// Grid.getPlaquette(index, d1, d2, o1, o2, t) (Grid.java:609-642)
template <typename T>
__device__ __forceinline__ Su2<double> plaq_or(const Lattice& L, const T* __restrict__ U, const Coord& c1, int d1,
                                               int d2, int o1, int o2)
{
    const Coord c2 = shifted(L, c1, d1, o1), c3 = shifted(L, c2, d2, o2), c4 = shifted(L, c3, d1, -o1);
    return mul(mul(mul(link_or<T>(L, U, c1, d1, o1), link_or<T>(L, U, c2, d2, o2)), link_or<T>(L, U, c3, d1, -o1)),
               link_or<T>(L, U, c4, d2, -o2));
}

__device__ __forceinline__ double dot(const Alg2<double>& a, const Alg2<double>& b)
{
    return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_projected_poynting(Lattice L, int dir, const T* __restrict__ U,
                                                               const T* __restrict__ Unext, const T* __restrict__ E,
                                                               double* __restrict__ out)
{
    const uint32_t slice = blockIdx.x;
    const uint32_t nd = dim_of(L, dir);
    const uint32_t per = L.n / nd;
    const double root_area = sqrt((double)L.n / (double)nd);
    double ufF[3], ufEB[3];
    for (int i = 0; i < 3; i++) {
        ufF[i] = 1.0 / ((L.as[i] * L.as[dir]) * L.g * root_area);
        ufEB[i] = 1.0 / ((L.g * L.as[i]) * root_area);
    }
    const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
    Kahan a[2] = {{0, 0}, {0, 0}};
    const int da = dir == 0 ? 1 : 0, db = dir == 2 ? 1 : 2;
    const uint32_t nb = dim_of(L, db);
    for (uint32_t m = threadIdx.x; m < per; m += kBlock) {
        uint32_t x[3];
        x[dir] = slice;
        x[da] = m / nb;
        x[db] = m - (m / nb) * nb;
        const Coord c{x[0], x[1], x[2]};
        const uint32_t site = site_of(L, c);
        double s_avg = 0.0;
#pragma unroll 1
        for (int i = 0; i < 3; i++) {
            if (i == dir) continue;
            Su2<double> fg{0, 0, 0, 0};
            for (int t = 0; t < 2; t++) {
                const T* Ut = t == 0 ? U : Unext;
                acc(fg, plaq_or<T>(L, Ut, c, dir, i, 1, 1));
                acc(fg, plaq_or<T>(L, Ut, c, i, dir, -1, 1));
                acc(fg, plaq_or<T>(L, Ut, c, i, dir, 1, -1));
                acc(fg, plaq_or<T>(L, Ut, c, dir, i, -1, -1));
            }
            Alg2<double> F = proj(fg);
            const double ff = ufF[i] / 8.0;
            for (int k = 0; k < 3; k++) F.v[k] *= ff;
            const Coord cm = shifted(L, c, i, -1);
            const Alg2<double> e2 = act(load_alg_d<2, T>(E, site_of(L, cm), i), link_or<T>(L, U, c, i, -1));
            const Alg2<double> e1 = load_alg_d<2, T>(E, site, i);
            Alg2<double> e;
            for (int k = 0; k < 3; k++) e.v[k] = (e1.v[k] + e2.v[k]) * (0.5 * ufEB[i]);
            s_avg += dot(e, F);
        }
        Alg2<double> E1 = load_alg_d<2, T>(E, site, d1), E2 = load_alg_d<2, T>(E, site, d2);
        for (int k = 0; k < 3; k++) {
            E1.v[k] *= ufEB[d1];
            E2.v[k] *= ufEB[d2];
        }
        // Grid.getB(index, k, t) (Grid.java:877-900)
        auto bfield = [&](const T* Ut, int k) {
            const int j = (k + 1) % 3, l = (k + 2) % 3;
            Alg2<double> b = proj(plaq_or<T>(L, Ut, c, j, l, 1, 1));
            const double f = L.as[k] / (L.as[j] * L.as[l]);
            for (int q = 0; q < 3; q++) b.v[q] *= f;
            return b;
        };
        const Alg2<double> b10 = bfield(U, d1), b11 = bfield(Unext, d1);
        const Alg2<double> b20 = bfield(U, d2), b21 = bfield(Unext, d2);
        Alg2<double> B1, B2;
        for (int k = 0; k < 3; k++) {
            B1.v[k] = (b10.v[k] + b11.v[k]) * (ufEB[d1] * 0.5);
            B2.v[k] = (b20.v[k] + b21.v[k]) * (ufEB[d2] * 0.5);
        }
        a[0].add(s_avg);
        a[1].add(dot(E1, B2) - dot(E2, B1));
    }
    __shared__ double sh[kBlock / 64][2][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < 2; q++) {
        double s = a[q].s, cc = -a[q].c;
        for (int off = 32; off > 0; off >>= 1) {
            double s2 = __shfl_down(s, off, 64);
            double c2 = __shfl_down(cc, off, 64);
            merge(s, cc, s2, c2);
        }
        if (lane == 0) {
            sh[wave][q][0] = s;
            sh[wave][q][1] = cc;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int q = threadIdx.x;
        double s = sh[0][q][0], cc = sh[0][q][1];
        for (int w = 1; w < kBlock / 64; w++) merge(s, cc, sh[w][q][0], sh[w][q][1]);
        out[(size_t)q * nd + slice] = s + cc;
    }
}

// ---------------------------------------------------------------- health monitor
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_health(uint32_t n, const T* __restrict__ U, double tol,
                                                   unsigned long long* __restrict__ counters)
{
    using GD = Group<NC, double>;
    const uint32_t site = blockIdx.x * kBlock + threadIdx.x;
    const int d = blockIdx.y;
    if (site >= n) return;
    auto u = load_link_d<NC, T>(U, site, d);
    bool finite = true;
#pragma unroll
    for (int k = 0; k < GD::G; k++) finite = finite && isfinite(GD::gget(u, k));
    auto p = mul_adj(u, u);
    auto id = GD::identity();
    double dev = 0;
#pragma unroll
    for (int k = 0; k < GD::G; k++) {
        double x = GD::gget(p, k) - GD::gget(id, k);
        dev = fmax(dev, fabs(x));
    }
    // SU(2) parametrisation: U U^+ = (sum e^2) * 1 with e1..e3 components zero
    if (!finite) atomicAdd(&counters[0], 1ull);
    if (!(dev <= tol)) atomicAdd(&counters[1], 1ull);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_health_alg(uint32_t nv, const T* __restrict__ X,
                                                       unsigned long long* __restrict__ counters)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nv) return;
    if (!isfinite((double)X[i])) atomicAdd(&counters[0], 1ull);
}

// ---------------------------------------------------------------- CPIC: super particles
// CGCSuperParticle.getCurrentOffset / needsUpdate
// (physics/particles/CGCSuperParticle.java:71-77, 100-105); getCurrentNGPOffset is
// evaluated on the host, where the charge-density segments are built.
__device__ __forceinline__ int floor_div(int a, int b)
{
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}
__device__ __forceinline__ int sp_current_offset(int o, int off, int pp, int k, int ppc, int t)
{
    if (o > 0) return off + ((t + k) / ppc) * pp;
    return off + floor_div(-t + k, ppc) * pp;
}
__device__ __forceinline__ bool sp_needs_update(int o, int k, int ppc, int t)
{
    if (o > 0) return 0 == (t + k + ppc / 2 + 1) % ppc;
    return 0 == (t - k + ppc / 2) % ppc;
}

// interpolateToGrid (physics/grid/CGCSuperParticleInterpolationNGP.java:19-73):
// one block row per super particle (blockIdx.y), one thread per charge.
// Charges are [coef][total] in T. Within one beam direction the active super
// particles of a step cover disjoint cells; opposite beams may meet in one cell,
// where the two additions onto the reset current commute exactly.
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_sp_to_grid(Lattice L, SuperParticles sp, int t, T* __restrict__ q,
                                                       const T* __restrict__ Unext, T* __restrict__ J)
{
    using GR = Group<NC, T>;
    const int p = blockIdx.y;
    const int o = sp.orientation[p], k = sp.shift[p], ppc = sp.ppc[p];
    if (!sp_needs_update(o, k, ppc, t)) return;
    const int off = sp_current_offset(o, sp.index_offset[p], sp.per_plane[p], k, ppc, t);
    const int quickfix = (ppc == 1) ? sp.per_plane[p] : 0;
    const int imin = (off < 0) ? -off : 0;
    long lim = (long)off + sp.num[p];
    if (lim > (long)L.n - quickfix) lim = (long)L.n - quickfix;
    long imax = lim - off;
    if (imax < 0) imax = 0;
    const long ireg = imax - imin;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= ireg) return;
    const long j = i + imin;
    const uint32_t index = (uint32_t)(off + j);
    const size_t qi = (size_t)sp.q_offset[p] + j;
    typename GR::A Q;
#pragma unroll
    for (int m = 0; m < GR::AD; m++) aref(Q, m) = q[(size_t)m * sp.total + qi];
    const auto U = load_link<NC, T>(Unext, index, 0);
    const T f = (T)(L.as[0] / L.at);
    if (o > 0) {
#pragma unroll
        for (int m = 0; m < GR::AD; m++) atomicAdd(J + fposT<T>(m, index + quickfix, 3 * GR::AD), aref(Q, m) * f);
        Q = act(Q, adj(U));
    } else {
        Q = act(Q, U);
        const T fm = (T)(-L.as[0] / L.at);
#pragma unroll
        for (int m = 0; m < GR::AD; m++) atomicAdd(J + fposT<T>(m, index, 3 * GR::AD), aref(Q, m) * fm);
    }
#pragma unroll
    for (int m = 0; m < GR::AD; m++) q[(size_t)m * sp.total + qi] = aref(Q, m);
}

// interpolateChargedensity (physics/grid/CGCSuperParticleInterpolationNGP.java:75-93)
// as a gather: each cell sums the charges that sit on it in particle-list order,
// the order of the sequential reference loop, with no atomics. Threads cover the
// segments of `cs` back to back; a thread finds its segment by bisection and adds
// only the particles that cover it (O(cells x particles per cell)).
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_sp_chargedensity(Lattice L, int64_t q_total, const T* __restrict__ q,
                                                             T* __restrict__ rho, CellSegments cs)
{
    using GR = Group<NC, T>;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= cs.cells) return;
    uint32_t lo = 0, hi = cs.count;   // cells_prefix[lo] <= i < cells_prefix[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cs.cells_prefix[mid] <= i) lo = mid;
        else hi = mid;
    }
    const uint32_t cell = cs.first[lo] + (i - cs.cells_prefix[lo]);
    typename GR::A a;
#pragma unroll
    for (int m = 0; m < GR::AD; m++) aref(a, m) = rho[fposT<T>(m, cell, GR::AD)];
    const uint32_t e1 = cs.list_prefix[lo + 1];
    for (uint32_t e = cs.list_prefix[lo]; e < e1; e++) {
        const size_t qi = (size_t)(cs.qbase[e] + (int64_t)cell);
#pragma unroll
        for (int m = 0; m < GR::AD; m++) aref(a, m) += q[(size_t)m * q_total + qi];
    }
#pragma unroll
    for (int m = 0; m < GR::AD; m++) rho[fposT<T>(m, cell, GR::AD)] = aref(a, m);
}

dim3 grid_links(uint32_t n) { return dim3((n + kBlock - 1) / kBlock, 3, 1); }

// component blockIdx.y (of gridDim.y per cell) of the cells of the ranges set to zero; thread i takes
// the i-th cell of the ranges laid end to end
template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_cells(T* __restrict__ f, CellRanges r)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= r.start[r.count]) return;
    int k = 0;
#pragma unroll
    for (int q = 1; q < kMaxCurrentSlabs; q++) k += (q < r.count) & (i >= r.start[q]);
    f[fposT<T>(blockIdx.y, r.first[k] + (i - r.start[k]), gridDim.y)] = T(0);
}

}  // namespace

// ---------------------------------------------------------------- dispatch
#define OPX_DISPATCH(nc, prec, ...)                                  \
    do {                                                             \
        if ((nc) == 2 && (prec) == 0) { constexpr int NC = 2; using T = double; __VA_ARGS__; } \
        else if ((nc) == 2) { constexpr int NC = 2; using T = float; __VA_ARGS__; }           \
        else if ((prec) == 0) { constexpr int NC = 3; using T = double; __VA_ARGS__; }        \
        else { constexpr int NC = 3; using T = float; __VA_ARGS__; }                          \
    } while (0)

hipError_t launch_combined_update(int nc, int prec, const Lattice& L, const Box& active, const void* U, void* Unext,
                                  void* E, const void* J, bool has_j, hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        LinkOrder order;
        order.rows = L.d1 < 16 ? L.d1 : 16;
        order.blocks_per_tile_plane = (order.rows * L.d2 + kBlock - 1) / kBlock;
        const uint32_t tiles = (L.d1 + order.rows - 1) / order.rows;
        const dim3 g1(3 * order.blocks_per_tile_plane * L.d0 * tiles);
        const bool buf = !L.force_global && padded_cells(L.n) * 3 * Group<NC, T>::S * sizeof(T) < (1ull << 32);
        if (buf && has_j)
            hipLaunchKernelGGL((k_combined_update<NC, T, true, true>), g1, dim3(kBlock), 0, st, L, order, active,
                               (const T*)U, (T*)Unext, (T*)E, (const T*)J);
        else if (buf)
            hipLaunchKernelGGL((k_combined_update<NC, T, false, true>), g1, dim3(kBlock), 0, st, L, order, active,
                               (const T*)U, (T*)Unext, (T*)E, (const T*)E);
        else if (has_j)
            hipLaunchKernelGGL((k_combined_update<NC, T, true, false>), g1, dim3(kBlock), 0, st, L, order, active,
                               (const T*)U, (T*)Unext, (T*)E, (const T*)J);
        else
            hipLaunchKernelGGL((k_combined_update<NC, T, false, false>), g1, dim3(kBlock), 0, st, L, order, active,
                               (const T*)U, (T*)Unext, (T*)E, (const T*)E);
    });
    return hipGetLastError();
}

template <int NC, typename T, bool REUNIT>
static void launch_cell_variant(const Lattice& L, const Box& active, const LinkOrder& order, dim3 g1, bool buf,
                                const void* U, void* Unext, void* E, const void* J, bool has_j, hipStream_t st)
{
    if (buf && has_j)
        hipLaunchKernelGGL((k_cell_update<NC, T, true, true, REUNIT>), g1, dim3(kBlock), 0, st, L, order, active,
                           (const T*)U, (T*)Unext, (T*)E, (const T*)J);
    else if (buf)
        hipLaunchKernelGGL((k_cell_update<NC, T, false, true, REUNIT>), g1, dim3(kBlock), 0, st, L, order, active,
                           (const T*)U, (T*)Unext, (T*)E, (const T*)E);
    else if (has_j)
        hipLaunchKernelGGL((k_cell_update<NC, T, true, false, REUNIT>), g1, dim3(kBlock), 0, st, L, order, active,
                           (const T*)U, (T*)Unext, (T*)E, (const T*)J);
    else
        hipLaunchKernelGGL((k_cell_update<NC, T, false, false, REUNIT>), g1, dim3(kBlock), 0, st, L, order, active,
                           (const T*)U, (T*)Unext, (T*)E, (const T*)E);
}

hipError_t launch_cell_update(int nc, int prec, const Lattice& L, const Box& active, int tile, const void* U,
                              void* Unext, void* E, const void* J, bool has_j, bool reunit, hipStream_t st)
{
    const uint32_t rows = kCellRows[tile >= 0 && tile < kCellTiles ? tile : 0];
    OPX_DISPATCH(nc, prec, {
        LinkOrder order;
        order.rows = L.d1 < rows ? L.d1 : rows;
        order.blocks_per_tile_plane = (order.rows * L.d2 + kBlock - 1) / kBlock;
        const uint32_t tiles = (L.d1 + order.rows - 1) / order.rows;
        const dim3 g1(order.blocks_per_tile_plane * L.d0 * tiles);
        const bool buf = !L.force_global && padded_cells(L.n) * 3 * Group<NC, T>::S * sizeof(T) < (1ull << 32);
        if (reunit)
            launch_cell_variant<NC, T, true>(L, active, order, g1, buf, U, Unext, E, J, has_j, st);
        else
            launch_cell_variant<NC, T, false>(L, active, order, g1, buf, U, Unext, E, J, has_j, st);
    });
    return hipGetLastError();
}

hipError_t launch_plaq_update(int nc, int prec, const Lattice& L, const Box& active, const void* U, void* P,
                              void* Unext, void* E, const void* J, bool has_j, hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        LinkOrder order;
        order.rows = L.d1 < 16 ? L.d1 : 16;
        order.blocks_per_tile_plane = (order.rows * L.d2 + kBlock - 1) / kBlock;
        const uint32_t tiles = (L.d1 + order.rows - 1) / order.rows;
        const uint32_t cell_blocks = order.blocks_per_tile_plane * L.d0 * tiles;
        const bool buf = !L.force_global && padded_cells(L.n) * 3 * Group<NC, T>::S * sizeof(T) < (1ull << 32);
        const T* Jt = has_j ? (const T*)J : (const T*)E;
        if (buf) {
            hipLaunchKernelGGL((k_plaquettes<NC, T, true>), dim3(cell_blocks), dim3(kBlock), 0, st, L, order,
                               (const T*)U, (T*)P);
            if (has_j)
                hipLaunchKernelGGL((k_update_plaq<NC, T, true, true>), dim3(3 * cell_blocks), dim3(kBlock), 0, st, L,
                                   order, active, (const T*)U, (const T*)P, (T*)Unext, (T*)E, Jt);
            else
                hipLaunchKernelGGL((k_update_plaq<NC, T, false, true>), dim3(3 * cell_blocks), dim3(kBlock), 0, st, L,
                                   order, active, (const T*)U, (const T*)P, (T*)Unext, (T*)E, Jt);
        } else {
            hipLaunchKernelGGL((k_plaquettes<NC, T, false>), dim3(cell_blocks), dim3(kBlock), 0, st, L, order,
                               (const T*)U, (T*)P);
            if (has_j)
                hipLaunchKernelGGL((k_update_plaq<NC, T, true, false>), dim3(3 * cell_blocks), dim3(kBlock), 0, st, L,
                                   order, active, (const T*)U, (const T*)P, (T*)Unext, (T*)E, Jt);
            else
                hipLaunchKernelGGL((k_update_plaq<NC, T, false, false>), dim3(3 * cell_blocks), dim3(kBlock), 0, st, L,
                                   order, active, (const T*)U, (const T*)P, (T*)Unext, (T*)E, Jt);
        }
    });
    return hipGetLastError();
}

hipError_t launch_step_links(int nc, int prec, const Lattice& L, const Box& active, const void* U, void* Unext,
                             const void* E, hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_step_links<NC, T>), grid_links(L.n), dim3(kBlock), 0, st, L, active, (const T*)U,
                           (T*)Unext, (const T*)E);
    });
    return hipGetLastError();
}

hipError_t launch_reunitarize(int nc, int prec, uint32_t n, void* U, hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_reunitarize<NC, T>), grid_links(n), dim3(kBlock), 0, st, n, (T*)U);
    });
    return hipGetLastError();
}

hipError_t launch_links_from_coeffs(int nc, int prec, uint32_t n, const double* C, void* U, hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_links_from_coeffs<NC, T>), grid_links(n), dim3(kBlock), 0, st, n, C, (T*)U);
    });
    return hipGetLastError();
}

size_t measure_partials_size(uint32_t) { return (size_t)kMeasureBlocks * kNQ * 2; }

hipError_t launch_measure(int nc, int prec, const Lattice& L, const Box& eval, const void* U, const void* Unext,
                          const void* E, const void* rho, double* partials, double* out, hipStream_t st)
{
    int nb = (int)((L.n + kBlock - 1) / kBlock);
    if (nb > kMeasureBlocks) nb = kMeasureBlocks;
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_measure<NC, T>), dim3(nb), dim3(kBlock), 0, st, L, eval, (const T*)U,
                           (const T*)Unext, (const T*)E, (const T*)rho, partials);
    });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_measure_final, dim3(1), dim3(64 * kNQ), 0, st, nb, 1.0 / (double)L.n, partials, out);
    return hipGetLastError();
}

hipError_t launch_projected_poynting(int nc, int prec, const Lattice& L, int dir, const void* U, const void* Unext,
                                     const void* E, double* out2, hipStream_t st)
{
    if (nc != 2) return hipErrorInvalidValue;
    const uint32_t nd = dir == 0 ? L.d0 : (dir == 1 ? L.d1 : L.d2);
    if (prec == 0)
        hipLaunchKernelGGL(k_projected_poynting<double>, dim3(nd), dim3(kBlock), 0, st, L, dir, (const double*)U,
                           (const double*)Unext, (const double*)E, out2);
    else
        hipLaunchKernelGGL(k_projected_poynting<float>, dim3(nd), dim3(kBlock), 0, st, L, dir, (const float*)U,
                           (const float*)Unext, (const float*)E, out2);
    return hipGetLastError();
}

// The per-slice sums of opx_projected_energy_density_noalloc: a static device array, one copy per
// GPU, addressed through the pointer hipGetSymbolAddress returns on that GPU (its name is only a
// host-side handle). Every block of k_projected_energy writes its four sums, so nothing in it
// needs clearing; calls that overlap on one GPU race on it.
static __device__ __align__(256) double scratch[4 * kNoAllocMaxSlices];

hipError_t noalloc_scratch(int device, double** out)
{
    static std::mutex lock;
    static double* ptr[kNoAllocMaxDevices] = {};
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(lock);
    if (!ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;   // the address is looked up on the current GPU
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptr[device] = static_cast<double*>(p);
    }
    *out = ptr[device];
    return hipSuccess;
}

hipError_t launch_projected_energy(int nc, int prec, const Lattice& L, const Box& eval, int dir, const void* U,
                                   const void* Unext, const void* E, double* out4, hipStream_t st)
{
    uint32_t nd = dir == 0 ? L.d0 : (dir == 1 ? L.d1 : L.d2);
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_projected_energy<NC, T>), dim3(nd), dim3(kBlock), 0, st, L, eval, dir, (const T*)U,
                           (const T*)Unext, (const T*)E, out4);
    });
    return hipGetLastError();
}

hipError_t launch_health(int nc, int prec, uint32_t n, const void* U, double tol, unsigned long long* counters,
                         hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_health<NC, T>), grid_links(n), dim3(kBlock), 0, st, n, (const T*)U, tol, counters);
    });
    return hipGetLastError();
}

hipError_t launch_health_alg(int nc, int prec, uint32_t nv, const void* X, unsigned long long* counters,
                             hipStream_t st)
{
    dim3 g((nv + kBlock - 1) / kBlock);
    if (prec == 0)
        hipLaunchKernelGGL(k_health_alg<double>, g, dim3(kBlock), 0, st, nv, (const double*)X, counters);
    else
        hipLaunchKernelGGL(k_health_alg<float>, g, dim3(kBlock), 0, st, nv, (const float*)X, counters);
    (void)nc;
    return hipGetLastError();
}

hipError_t launch_sp_to_grid(int nc, int prec, const Lattice& L, const SuperParticles& sp, int t, void* q,
                             const void* Unext, void* J, hipStream_t st)
{
    if (sp.count == 0 || sp.max_num == 0) return hipSuccess;
    dim3 g((sp.max_num + kBlock - 1) / kBlock, sp.count, 1);
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_sp_to_grid<NC, T>), g, dim3(kBlock), 0, st, L, sp, t, (T*)q, (const T*)Unext,
                           (T*)J);
    });
    return hipGetLastError();
}

hipError_t launch_zero_cell_ranges(int prec, void* f, uint32_t comps, const CellRanges& r, hipStream_t st)
{
    if (r.count <= 0 || comps == 0) return hipSuccess;
    const uint32_t total = r.start[r.count];
    if (total == 0) return hipSuccess;
    const dim3 g((total + kBlock - 1) / kBlock, comps);
    if (prec == 0)
        hipLaunchKernelGGL((k_zero_cells<double>), g, dim3(kBlock), 0, st, (double*)f, r);
    else
        hipLaunchKernelGGL((k_zero_cells<float>), g, dim3(kBlock), 0, st, (float*)f, r);
    return hipGetLastError();
}

hipError_t launch_sp_chargedensity(int nc, int prec, const Lattice& L, int64_t q_total, const void* q, void* rho,
                                   const CellSegments& cs, hipStream_t st)
{
    if (cs.count == 0 || cs.cells == 0) return hipSuccess;
    dim3 g((cs.cells + kBlock - 1) / kBlock);
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_sp_chargedensity<NC, T>), g, dim3(kBlock), 0, st, L, q_total, (const T*)q, (T*)rho,
                           cs);
    });
    return hipGetLastError();
}

hipError_t launch_batch_update(int nc, int prec, const BatchEvent* ev, int nev, uint32_t total_items, uint32_t grid,
                               const void* U, void* Unext, void* E, hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_batch_update<NC, T>), dim3(grid), dim3(kBlock), 0, st, ev, nev, total_items, (const T*)U,
                           (T*)Unext, (T*)E);
    });
    return hipGetLastError();
}

hipError_t launch_batch_cell(int nc, int prec, const BatchEvent* ev, int nev, uint32_t total_items, uint32_t grid,
                             bool buf, const void* U, void* Unext, void* E, bool reunit, hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        if (buf && reunit)
            hipLaunchKernelGGL((k_batch_cell<NC, T, true, true>), dim3(grid), dim3(kBlock), 0, st, ev, nev,
                               total_items / 3, (const T*)U, (T*)Unext, (T*)E);
        else if (buf)
            hipLaunchKernelGGL((k_batch_cell<NC, T, true, false>), dim3(grid), dim3(kBlock), 0, st, ev, nev,
                               total_items / 3, (const T*)U, (T*)Unext, (T*)E);
        else if (reunit)
            hipLaunchKernelGGL((k_batch_cell<NC, T, false, true>), dim3(grid), dim3(kBlock), 0, st, ev, nev,
                               total_items / 3, (const T*)U, (T*)Unext, (T*)E);
        else
            hipLaunchKernelGGL((k_batch_cell<NC, T, false, false>), dim3(grid), dim3(kBlock), 0, st, ev, nev,
                               total_items / 3, (const T*)U, (T*)Unext, (T*)E);
    });
    return hipGetLastError();
}

hipError_t launch_batch_reunitarize(int nc, int prec, const BatchEvent* ev, int nev, uint32_t total_items,
                                    uint32_t grid, void* U, hipStream_t st)
{
    OPX_DISPATCH(nc, prec, {
        hipLaunchKernelGGL((k_batch_reunitarize<NC, T>), dim3(grid), dim3(kBlock), 0, st, ev, nev, total_items,
                           (T*)U);
    });
    return hipGetLastError();
}

}  // namespace opx
