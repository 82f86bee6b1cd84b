// opx_plane.hip -- plane-tile form of the fused update.
//
// FastTYMSolver.CombinedUpdate (physics/fields/FastTYMSolver.java:40-62) for the
// three links of every cell, computed as the per-link kernel computes them, with
// the neighbour links of the cell's own x0 plane taken from LDS: a block owns a
// TY x TX tile of (x1, x2) cells of one plane, stages the links of the tile and a
// one-cell halo around it (every direction) in LDS, and its 3 TY TX threads (one
// per link, the direction uniform over a wave) read the in-plane staple links
// from there. Only the links on the planes x0 +- 1 (four per
// link) come from global memory; blocks are ordered plane-fastest per tile, so
// those planes were just staged by the neighbouring blocks and come from L2.
// Against the per-link kernel this cuts the vector-L1 traffic per link from 13
// links to about 5.3 (the LDS fill plus the four x0 neighbours).
#include <type_traits>

#include "opx_buf.h"
#include "opx_group.h"
#include "opx_kernels.h"

namespace opx {

namespace {

__device__ __forceinline__ uint32_t pwrap(int v, uint32_t n) { return (uint32_t)(v + (int)n) % n; }

__device__ __forceinline__ bool in_box_plane(const Box& b, uint32_t x0, uint32_t x1, uint32_t x2)
{
    if (b.all) return true;
    return (int)x0 >= b.lo[0] && (int)x0 <= b.hi[0] && (int)x1 >= b.lo[1] && (int)x1 <= b.hi[1] &&
           (int)x2 >= b.lo[2] && (int)x2 <= b.hi[2];
}

// One link of direction D at (p, x1, x2); `c` is the cell's position in the LDS tile
// (rows of PW cells, NS cells per component).
// Every offset is a compile-time constant, so the loads of all neighbour links can
// be scheduled ahead of the arithmetic.
template <int NC, typename T, int NS, int PW, bool HAS_J, int D>
__device__ __forceinline__ void plane_link(const Lattice& L, const T (*sl)[NS], const Buf<T>& bu, const Buf<T>& bun,
                                           const Buf<T>& be, const T* __restrict__ Jp, uint32_t p, uint32_t x1,
                                           uint32_t x2, int c)
{
    using GR = Group<NC, T>;
    using M = typename GR::M;
    using A = typename GR::A;
    constexpr int S = GR::S, AD = GR::AD;
    const uint32_t sz = sizeof(T);
    const uint32_t pp = p + 1 == L.d0 ? 0 : p + 1, pm = p == 0 ? L.d0 - 1 : p - 1;
    const uint32_t x1p = x1 + 1 == L.d1 ? 0 : x1 + 1, x1m = x1 == 0 ? L.d1 - 1 : x1 - 1;
    const uint32_t x2p = x2 + 1 == L.d2 ? 0 : x2 + 1, x2m = x2 == 0 ? L.d2 - 1 : x2 - 1;
    // the link of direction DIR at offset (O0, O1, O2): LDS in this plane, global off it
    auto link_at = [&](auto o0, auto o1, auto o2, auto dir) -> M {
        constexpr int O0 = decltype(o0)::value, O1 = decltype(o1)::value, O2 = decltype(o2)::value;
        constexpr int DIR = decltype(dir)::value;
        M m;
        if constexpr (O0 == 0) {
            const int pos = c + O1 * PW + O2;
#pragma unroll
            for (int k = 0; k < S; k++) GR::gref(m, GR::slot(k)) = sl[DIR * S + k][pos];
        } else {
            const uint32_t y1 = O1 > 0 ? x1p : (O1 < 0 ? x1m : x1);
            const uint32_t y2 = O2 > 0 ? x2p : (O2 < 0 ? x2m : x2);
            const uint32_t site = (O0 > 0 ? pp : pm) * L.s0 + y1 * L.s1 + y2;
#pragma unroll
            for (int k = 0; k < S; k++) GR::gref(m, GR::slot(k)) = bu.ld(fvoff(site, 3 * S, sz), fsoff(DIR * S + k, 3 * S, sz));
        }
        GR::complete(m);
        return m;
    };
    using I = std::integral_constant<int, 0>;
    using dD = std::integral_constant<int, D>;
    constexpr int ED0 = D == 0, ED1 = D == 1, ED2 = D == 2;
    M St = GR::zero();
    // U_j(x+d) U_d(x+j)^+ U_j(x)^+ + (U_d(x-j) U_j(x+d-j))^+ U_j(x-j), j != d ascending
    auto staple = [&](auto jj) {
        constexpr int J = decltype(jj)::value;
        constexpr int EJ0 = J == 0, EJ1 = J == 1, EJ2 = J == 2;
        using C = std::integral_constant<int, 0>;
        M u1 = mul_adj(mul_adj(link_at(std::integral_constant<int, ED0>(), std::integral_constant<int, ED1>(),
                                       std::integral_constant<int, ED2>(), jj),
                               link_at(std::integral_constant<int, EJ0>(), std::integral_constant<int, EJ1>(),
                                       std::integral_constant<int, EJ2>(), dD())),
                       link_at(C(), C(), C(), jj));
        M u2 = adj_mul(mul(link_at(std::integral_constant<int, -EJ0>(), std::integral_constant<int, -EJ1>(),
                                   std::integral_constant<int, -EJ2>(), dD()),
                           link_at(std::integral_constant<int, ED0 - EJ0>(), std::integral_constant<int, ED1 - EJ1>(),
                                   std::integral_constant<int, ED2 - EJ2>(), jj)),
                       link_at(std::integral_constant<int, -EJ0>(), std::integral_constant<int, -EJ1>(),
                               std::integral_constant<int, -EJ2>(), jj));
        acc(u1, u2);
        acc_scaled(St, u1, (T)L.area[J]);
    };
    if constexpr (D != 0) staple(std::integral_constant<int, 0>());
    if constexpr (D != 1) staple(std::integral_constant<int, 1>());
    if constexpr (D != 2) staple(std::integral_constant<int, 2>());
    const M Ud = link_at(I(), I(), I(), dD());
    const A pr = proj(mul(Ud, St));
    const uint32_t site = p * L.s0 + x1 * L.s1 + x2;
    const T at = (T)L.at;
    A e;
#pragma unroll
    for (int k = 0; k < AD; k++) aref(e, k) = be.ld(fvoff(site, 3 * AD, sz), fsoff(D * AD + k, 3 * AD, sz), 1);
#pragma unroll
    for (int k = 0; k < AD; k++) aref(e, k) += aget(pr, k) * at;
    if (HAS_J) {
        const T uf = (T)L.uf[D];
#pragma unroll
        for (int k = 0; k < AD; k++) aref(e, k) += Jp[fposT<T>(D * AD + k, site, 3 * AD)] * uf;
    }
#pragma unroll
    for (int k = 0; k < AD; k++) be.st(aget(e, k), fvoff(site, 3 * AD, sz), fsoff(D * AD + k, 3 * AD, sz), 1);
    const M un = mul(exp_link(e, -at), Ud);
#pragma unroll
    for (int k = 0; k < S; k++) bun.st(GR::gget(un, GR::slot(k)), fvoff(site, 3 * S, sz), fsoff(D * S + k, 3 * S, sz), 1);
}

template <int NC, typename T, int TY, int TX, bool HAS_J>
__global__ __launch_bounds__(3 * TY * TX) void k_plane_update(Lattice L, Box active, PlaneGeom pg,
                                                              const T* __restrict__ Up, T* __restrict__ Unp,
                                                              T* __restrict__ Ep, const T* __restrict__ Jp)
{
    using GR = Group<NC, T>;
    constexpr int S = GR::S, AD = GR::AD;
    constexpr int PW = TX + 2, NS = (TY + 2) * (TX + 2), CELLS = TY * TX, NTHREADS = 3 * CELLS;
    static_assert(CELLS % 64 == 0, "a wave updates links of one direction");
    __shared__ T sl[3 * S][NS];   // [dir * S + stored value][position]
    const uint32_t np = (uint32_t)padded_cells(L.n), sz = sizeof(T);
    const Buf<T> bu(Up, 3u * S * np * sz), bun(Unp, 3u * S * np * sz), be(Ep, 3u * AD * np * sz);
    // XCD k takes a contiguous range of (tile, plane) work, planes fastest
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
    const uint32_t logical = xcd * q8 + (xcd < r8 ? xcd : r8) + (b >> 3);
    const uint32_t p = logical % L.d0;
    const uint32_t rest = logical / L.d0;
    const uint32_t t2 = rest % pg.tiles2, t1 = rest / pg.tiles2;
    const int row0 = (int)(t1 * TY), col0 = (int)(t2 * TX);
    const int tid = threadIdx.x;
    // stage plane p: tile plus a one-cell halo, all three directions
    for (int i = tid; i < NS; i += NTHREADS) {
        const int hy = i / PW, hx = i - hy * PW;
        const uint32_t site = p * L.s0 + pwrap(row0 - 1 + hy, L.d1) * L.s1 + pwrap(col0 - 1 + hx, L.d2);
#pragma unroll
        for (int k = 0; k < 3 * S; k++) sl[k][i] = bu.ld(fvoff(site, 3 * S, sz), fsoff(k, 3 * S, sz));
    }
    __syncthreads();
    // direction of this wave's links (uniform: CELLS is a multiple of the wave size)
    const int d = __builtin_amdgcn_readfirstlane(tid / CELLS);
    const int cell = tid - d * CELLS;
    const int ty = cell / TX, tx = cell - ty * TX;
    const uint32_t x1 = (uint32_t)(row0 + ty), x2 = (uint32_t)(col0 + tx);
    if (x1 >= L.d1 || x2 >= L.d2 || !in_box_plane(active, p, x1, x2)) return;
    const int c = (ty + 1) * PW + (tx + 1);
    const T(*slp)[NS] = sl;
    if (d == 0)
        plane_link<NC, T, NS, PW, HAS_J, 0>(L, slp, bu, bun, be, Jp, p, x1, x2, c);
    else if (d == 1)
        plane_link<NC, T, NS, PW, HAS_J, 1>(L, slp, bu, bun, be, Jp, p, x1, x2, c);
    else
        plane_link<NC, T, NS, PW, HAS_J, 2>(L, slp, bu, bun, be, Jp, p, x1, x2, c);
}

template <int NC, typename T, int TY, int TX>
hipError_t plane_launch(const Lattice& L, const Box& active, const void* U, void* Unext, void* E, const void* J,
                        bool has_j, hipStream_t st)
{
    PlaneGeom pg;
    pg.tiles1 = (L.d1 + TY - 1) / TY;
    pg.tiles2 = (L.d2 + TX - 1) / TX;
    const dim3 grid(L.d0 * pg.tiles1 * pg.tiles2), block(3 * TY * TX);
    if (has_j)
        hipLaunchKernelGGL((k_plane_update<NC, T, TY, TX, true>), grid, block, 0, st, L, active, pg, (const T*)U,
                           (T*)Unext, (T*)E, (const T*)J);
    else
        hipLaunchKernelGGL((k_plane_update<NC, T, TY, TX, false>), grid, block, 0, st, L, active, pg, (const T*)U,
                           (T*)Unext, (T*)E, (const T*)E);
    return hipGetLastError();
}

}  // namespace

// tile shapes (TY x TX cells, 3 TY TX threads) of the form; LDS = 3 S (TY + 2)(TX + 2) reals
//   shape 0: 8 x 16, 1: 8 x 8, 2: 4 x 16 (SU(3): FP32 26 / 14 / 16 KB, FP64 78 / 43 / 47 KB)
bool plane_supported(int nc, int prec, const Lattice& L, int shape)
{
    if (shape < 0 || shape >= kPlaneShapes) return false;
    // buffer offsets are 32-bit
    const uint64_t bytes = padded_cells(L.n) * 3 * (nc == 2 ? 4 : (prec == 0 ? 18 : 12)) * (prec == 0 ? 8 : 4);
    return bytes < (1ull << 32);
}

hipError_t launch_plane_update(int nc, int prec, const Lattice& L, const Box& active, int shape, const void* U,
                               void* Unext, void* E, const void* J, bool has_j, hipStream_t st)
{
    if (!plane_supported(nc, prec, L, shape)) return hipErrorInvalidConfiguration;
#define OPX_PLANE(NC_, T_, TY_, TX_) return plane_launch<NC_, T_, TY_, TX_>(L, active, U, Unext, E, J, has_j, st)
#define OPX_PLANE_SHAPES(NC_, T_)              \
    if (shape == 0) OPX_PLANE(NC_, T_, 8, 16); \
    if (shape == 1) OPX_PLANE(NC_, T_, 8, 8);  \
    OPX_PLANE(NC_, T_, 4, 16)
    if (prec == 0) {
        if (nc == 3) {
            OPX_PLANE_SHAPES(3, double);
        }
        OPX_PLANE_SHAPES(2, double);
    }
    if (nc == 3) {
        OPX_PLANE_SHAPES(3, float);
    }
    OPX_PLANE_SHAPES(2, float);
#undef OPX_PLANE_SHAPES
#undef OPX_PLANE
}

}  // namespace opx
