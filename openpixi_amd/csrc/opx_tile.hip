// opx_tile.hip -- SU(2) form of the fused update with links staged in LDS.
//
// FastTYMSolver.CombinedUpdate (physics/fields/FastTYMSolver.java:40-62) for every
// link, as the per-link kernel computes it, but every neighbour link comes from
// LDS instead of the vector L1: a block owns a TY x TX tile of (x1, x2) cells and
// marches along x0 over a chunk of planes, keeping the links of planes p-1, p and
// p+1 (tile plus a one-cell halo on every side) in a rolling three-slot buffer.
// Each link is read from HBM/L2 once per block (halo cells by the neighbouring
// blocks too); the thread of a cell updates its three links. SU(2) links are 16 B
// (FP32), so the buffer of a 16 x 16 tile is 47 KB; SU(3) links would not fit.
#include "opx_buf.h"
#include "opx_group.h"
#include "opx_kernels.h"

namespace opx {

namespace {

template <typename T>
using Q2 = Su2<T>;

// E and Unext are written once per step: non-temporal stores (NT) keep them out of
// the caches, ordinary ones let a lattice that fits the last-level cache be re-read
// from it by the next step; the autotuner picks per lattice.
template <bool NT, typename T>
__device__ __forceinline__ void st_once(T v, T* p)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// v mod n for v >= -1 (tiles may be wider than the lattice, so v can exceed n several times)
__device__ __forceinline__ uint32_t wrap(int v, uint32_t n)
{
    return (uint32_t)(v + (int)n) % n;
}

__device__ __forceinline__ bool in_box_tile(const Box& b, uint32_t x0, uint32_t x1, uint32_t x2)
{
    if (b.all) return true;
    return (int)x0 >= b.lo[0] && (int)x0 <= b.hi[0] && (int)x1 >= b.lo[1] && (int)x1 <= b.hi[1] &&
           (int)x2 >= b.lo[2] && (int)x2 <= b.hi[2];
}

// SU(2) link field (12 components per cell) in the device layout (fpos)
template <typename T>
__device__ __forceinline__ Q2<T> ld_link2(const T* __restrict__ U, uint32_t site, int d)
{
    const T* p = fcell(U, site, 12);
    return Q2<T>{p[fcomp<T>(4 * d, 12)], p[fcomp<T>(4 * d + 1, 12)], p[fcomp<T>(4 * d + 2, 12)], p[fcomp<T>(4 * d + 3, 12)]};
}

// One block's march over planes [pb, pe) of the tile at (row0, col0) of lattice L.
template <typename T, int TY, int TX, bool HAS_J, bool NT>
__device__ __forceinline__ void tile_march(const Lattice& L, const Box& active, const T* __restrict__ U,
                                           T* __restrict__ Un, T* __restrict__ E, const T* __restrict__ J,
                                           uint32_t row0, uint32_t col0, uint32_t pb, uint32_t pe,
                                           Q2<T> (*sl)[3][(TY + 2) * (TX + 2)])
{
    using A = Alg2<T>;
    constexpr int PW = TX + 2, NS = (TY + 2) * (TX + 2), NTHREADS = TY * TX;
    const int tid = threadIdx.x;
    const int ty = tid / TX, tx = tid - ty * TX;
    auto load_plane = [&](uint32_t plane, int slot) {
        for (int i = tid; i < NS; i += NTHREADS) {
            const int hy = i / PW, hx = i - hy * PW;
            const uint32_t x1 = wrap((int)row0 - 1 + hy, L.d1), x2 = wrap((int)col0 - 1 + hx, L.d2);
            const uint32_t site = plane * L.s0 + x1 * L.s1 + x2;
#pragma unroll
            for (int d = 0; d < 3; d++) sl[slot][d][i] = ld_link2(U, site, d);
        }
    };
    const uint32_t x1 = row0 + (uint32_t)ty, x2 = col0 + (uint32_t)tx;
    const bool inside = x1 < L.d1 && x2 < L.d2;
    const int c = (ty + 1) * PW + (tx + 1);
    int sp = 0, sc = 1, sn = 2;   // slots of planes p-1, p, p+1
    load_plane(pb == 0 ? L.d0 - 1 : pb - 1, sp);
    load_plane(pb, sc);
    const T at = (T)L.at;
    for (uint32_t p = pb; p < pe; p++) {
        load_plane(p + 1 == L.d0 ? 0 : p + 1, sn);
        __syncthreads();
        if (inside && in_box_tile(active, p, x1, x2)) {
            const uint32_t site = p * L.s0 + x1 * L.s1 + x2;
            // the link of direction `dir` at offset (o0, o1, o2) from this cell
            auto at_off = [&](int o0, int o1, int o2, int dir) -> Q2<T> {
                const int slot = o0 < 0 ? sp : (o0 > 0 ? sn : sc);
                return sl[slot][dir][c + o1 * PW + o2];
            };
#pragma unroll
            for (int d = 0; d < 3; d++) {
                Q2<T> S{T(0), T(0), T(0), T(0)};
                const int ed0 = d == 0, ed1 = d == 1, ed2 = d == 2;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    if (j == d) continue;
                    const int ej0 = j == 0, ej1 = j == 1, ej2 = j == 2;
                    // U_j(x+d) U_d(x+j)^+ U_j(x)^+ + (U_d(x-j) U_j(x+d-j))^+ U_j(x-j)
                    Q2<T> u1 = mul_adj(mul_adj(at_off(ed0, ed1, ed2, j), at_off(ej0, ej1, ej2, d)),
                                       at_off(0, 0, 0, j));
                    Q2<T> u2 = adj_mul(mul(at_off(-ej0, -ej1, -ej2, d), at_off(ed0 - ej0, ed1 - ej1, ed2 - ej2, j)),
                                       at_off(-ej0, -ej1, -ej2, j));
                    acc(u1, u2);
                    acc(S, scaled(u1, (T)L.area[j]));
                }
                const Q2<T> Ud = at_off(0, 0, 0, d);
                const A pr = proj(mul(Ud, S));
                T* pe_ = fcell(E, site, 9);
                A e{{pe_[fcomp<T>(3 * d, 9)], pe_[fcomp<T>(3 * d + 1, 9)], pe_[fcomp<T>(3 * d + 2, 9)]}};
#pragma unroll
                for (int k = 0; k < 3; k++) e.v[k] += pr.v[k] * at;
                if (HAS_J) {
                    const T uf = (T)L.uf[d];
#pragma unroll
                    for (int k = 0; k < 3; k++) e.v[k] += fcell(J, site, 9)[fcomp<T>(3 * d + k, 9)] * uf;
                }
#pragma unroll
                for (int k = 0; k < 3; k++) st_once<NT>(e.v[k], pe_ + fcomp<T>(3 * d + k, 9));
                const Q2<T> un = mul(exp_link(e, -at), Ud);
                T* pu = fcell(Un, site, 12);
                st_once<NT>(un.e0, pu + fcomp<T>(4 * d, 12));
                st_once<NT>(un.e1, pu + fcomp<T>(4 * d + 1, 12));
                st_once<NT>(un.e2, pu + fcomp<T>(4 * d + 2, 12));
                st_once<NT>(un.e3, pu + fcomp<T>(4 * d + 3, 12));
            }
        }
        __syncthreads();   // plane p-1 is read before it is overwritten
        const int t = sp;
        sp = sc;
        sc = sn;
        sn = t;
    }
}

template <typename T, int TY, int TX, bool HAS_J, bool NT>
__global__ __launch_bounds__(TY* TX) void k_tile_update(Lattice L, Box active, TileGeom tg, const T* __restrict__ U,
                                                        T* __restrict__ Un, T* __restrict__ E,
                                                        const T* __restrict__ J)
{
    __shared__ Q2<T> sl[3][3][(TY + 2) * (TX + 2)];
    // XCD k (= blockIdx % 8) takes a contiguous range of (chunk, tile) work, tiles fastest
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
    const uint32_t logical = xcd * q8 + (xcd < r8 ? xcd : r8) + (b >> 3);
    const uint32_t t2 = logical % tg.tiles2;
    const uint32_t rest = logical / tg.tiles2;
    const uint32_t t1 = rest % tg.tiles1;
    const uint32_t chunk = rest / tg.tiles1;
    const uint32_t pb = (uint32_t)(((uint64_t)chunk * L.d0) / tg.nchunks);
    const uint32_t pe = (uint32_t)(((uint64_t)(chunk + 1) * L.d0) / tg.nchunks);
    tile_march<T, TY, TX, HAS_J, NT>(L, active, U, Un, E, J, t1 * TY, t2 * TX, pb, pe, sl);
}

// Persistent form over the (event, tile, chunk) items of a batch: a fixed grid
// strides through the item table, XCD k owning a contiguous eighth of it.
template <typename T, int TY, int TX, bool NT>
__global__ __launch_bounds__(TY* TX) void k_batch_tile(const BatchEvent* __restrict__ ev, const TileItem* __restrict__ items,
                                                       uint32_t nitems, const T* __restrict__ Ubase,
                                                       T* __restrict__ Unbase, T* __restrict__ Ebase)
{
    __shared__ Q2<T> sl[3][3][(TY + 2) * (TX + 2)];
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t xcd = b & 7, q8 = nb >> 3, r8 = nb & 7;
    const uint32_t wg_on_xcd = q8 + (xcd < r8 ? 1u : 0u);
    const uint32_t i0 = (uint32_t)(((uint64_t)nitems * xcd) / 8), i1 = (uint32_t)(((uint64_t)nitems * (xcd + 1)) / 8);
    Box all{{0, 0, 0}, {0, 0, 0}, 1};
    for (uint32_t it = i0 + (b >> 3); it < i1; it += wg_on_xcd) {
        const TileItem w = items[it];
        const BatchEvent& e = ev[w.event];
        tile_march<T, TY, TX, false, NT>(e.L, all, Ubase + e.link_offset, Unbase + e.link_offset, Ebase + e.alg_offset,
                                     nullptr, w.row0, w.col0, w.pb, w.pe, sl);
        __syncthreads();
    }
}

}  // namespace

bool tile_supported(int nc) { return nc == 2; }

TileGeom tile_geom(int prec, const Lattice& L, uint32_t nchunks)
{
    const int ty = prec == 0 ? kTileTY64 : kTileTY32, tx = kTileTX;
    TileGeom g;
    g.tiles1 = (L.d1 + ty - 1) / ty;
    g.tiles2 = (L.d2 + tx - 1) / tx;
    g.nchunks = nchunks < 1 ? 1 : (nchunks > L.d0 ? L.d0 : nchunks);
    return g;
}

namespace {

template <typename T, int TY, bool NT>
void tile_launch(const Lattice& L, const Box& active, const TileGeom& tg, const void* U, void* Unext, void* E,
                 const void* J, bool has_j, hipStream_t st)
{
    const dim3 grid(tg.tiles1 * tg.tiles2 * tg.nchunks), block(TY * kTileTX);
    if (has_j)
        hipLaunchKernelGGL((k_tile_update<T, TY, kTileTX, true, NT>), grid, block, 0, st, L, active, tg, (const T*)U,
                           (T*)Unext, (T*)E, (const T*)J);
    else
        hipLaunchKernelGGL((k_tile_update<T, TY, kTileTX, false, NT>), grid, block, 0, st, L, active, tg, (const T*)U,
                           (T*)Unext, (T*)E, (const T*)E);
}

template <typename T, int TY, bool NT>
void batch_tile_launch(const BatchEvent* ev, const TileItem* items, uint32_t nitems, uint32_t grid, const void* U,
                       void* Unext, void* E, hipStream_t st)
{
    hipLaunchKernelGGL((k_batch_tile<T, TY, kTileTX, NT>), dim3(grid), dim3(TY * kTileTX), 0, st, ev, items, nitems,
                       (const T*)U, (T*)Unext, (T*)E);
}

}  // namespace

hipError_t launch_tile_update(int nc, int prec, const Lattice& L, const Box& active, uint32_t nchunks, bool nt,
                              const void* U, void* Unext, void* E, const void* J, bool has_j, hipStream_t st)
{
    if (nc != 2) return hipErrorInvalidConfiguration;
    const TileGeom tg = tile_geom(prec, L, nchunks);
    if (prec == 0) {
        if (nt) tile_launch<double, kTileTY64, true>(L, active, tg, U, Unext, E, J, has_j, st);
        else tile_launch<double, kTileTY64, false>(L, active, tg, U, Unext, E, J, has_j, st);
    } else {
        if (nt) tile_launch<float, kTileTY32, true>(L, active, tg, U, Unext, E, J, has_j, st);
        else tile_launch<float, kTileTY32, false>(L, active, tg, U, Unext, E, J, has_j, st);
    }
    return hipGetLastError();
}

hipError_t launch_batch_tile(int prec, bool nt, const BatchEvent* ev, const TileItem* items, uint32_t nitems,
                             uint32_t grid, const void* U, void* Unext, void* E, hipStream_t st)
{
    if (nitems == 0) return hipSuccess;
    if (prec == 0) {
        if (nt) batch_tile_launch<double, kTileTY64, true>(ev, items, nitems, grid, U, Unext, E, st);
        else batch_tile_launch<double, kTileTY64, false>(ev, items, nitems, grid, U, Unext, E, st);
    } else {
        if (nt) batch_tile_launch<float, kTileTY32, true>(ev, items, nitems, grid, U, Unext, E, st);
        else batch_tile_launch<float, kTileTY32, false>(ev, items, nitems, grid, U, Unext, E, st);
    }
    return hipGetLastError();
}

}  // namespace opx
