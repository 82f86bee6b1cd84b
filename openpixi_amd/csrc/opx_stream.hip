// opx_stream.hip -- 2.5D streaming form of the fused E-field / link update.
//
// Same result as FastTYMSolver.CombinedUpdate (physics/fields/FastTYMSolver.java:40-62)
// on every active cell, reorganised so that each plaquette is computed once:
//
//   proj(U_d(x) S_d(x)) = sum_{j != d} a_j^-2 [ proj P_dj(x) + proj B_dj(x) ]
//
// with P_dj(x) = U_d(x) U_j(x+d) U_d(x+j)^+ U_j(x)^+ the plaquette at x and the
// backward staple term B_dj(x) = U_j(y)^+ P_dj(y)^+ U_j(y) evaluated at y = x - j,
// so proj B_dj(x) = -act(proj P_dj(y), U_j(y)^+) (proj commutes with the adjoint
// action). Every site y computes its three plaquettes P_01, P_02, P_12 from links
// at y, y+e0, y+e1, y+e2 only, keeps the two "forward" projections for its own
// links, and sends the transported projections to y+e0 (own thread, next plane,
// in registers), y+e1 and y+e2 (neighbour threads, through LDS).
//
// A block owns a TY x TX tile of (x1, x2) columns (plus a one-column halo on the
// low side, whose plaquettes feed the tile) and marches along x0 over a chunk of
// planes, starting one plane early to fill the carried x0 contributions. Each
// link of a plane is read from HBM once per block; links of the +1 halo row and
// column are read again by the neighbouring block (mostly an L2 hit: blocks are
// remapped so that each XCD gets a contiguous range of tiles).
// Loads and stores use buffer instructions with the component offset in an SGPR.
#include "opx_buf.h"
#include "opx_group.h"
#include "opx_kernels.h"

namespace opx {

namespace {


__device__ __forceinline__ bool in_box_dev(const Box& b, uint32_t x0, uint32_t x1, uint32_t x2)
{
    if (b.all) return true;
    return (int)x0 >= b.lo[0] && (int)x0 <= b.hi[0] && (int)x1 >= b.lo[1] && (int)x1 <= b.hi[1] &&
           (int)x2 >= b.lo[2] && (int)x2 <= b.hi[2];
}

template <typename A, typename T>
__device__ __forceinline__ void axpy(A& y, const A& x, T f)
{
#pragma unroll
    for (int k = 0; k < (int)(sizeof(A) / sizeof(T)); k++) aref(y, k) += aget(x, k) * f;
}

}  // namespace

// Launch geometry of the streaming kernel.
struct StreamGeom {
    uint32_t tiles1, tiles2, nchunks;
};

namespace {

template <int NC, typename T, int TY, int TX, bool HAS_J>
__global__ __launch_bounds__(TY* TX, (sizeof(T) == 4 ? 2 : 1)) void k_stream_update(Lattice L, Box active, StreamGeom sg,
                                                          const T* __restrict__ Up, T* __restrict__ Unp,
                                                          T* __restrict__ Ep, const T* __restrict__ Jp)
{
    using GR = Group<NC, T>;
    using M = typename GR::M;
    using A = typename GR::A;
    constexpr int G = GR::G;
    constexpr int AD = GR::AD;
    constexpr int NS = (TY + 1) * (TX + 1);  // link slots (tile + high halo row/column)
    constexpr int NT = TY * TX;
    constexpr int LINK_WORDS = 3 * G * NS;
    constexpr int CONTRIB_WORDS = 4 * AD * NT;
    constexpr int SMEM = LINK_WORDS > CONTRIB_WORDS ? LINK_WORDS : CONTRIB_WORDS;
    __shared__ T smem[SMEM];

    const uint32_t np = (uint32_t)padded_cells(L.n);
    const uint32_t sz = sizeof(T);
    const Buf<T> bu(Up, 3u * GR::S * np * sz);
    const Buf<T> bun(Unp, 3u * GR::S * np * sz);
    const Buf<T> be(Ep, 3u * AD * np * sz);

    // XCD-aware block order: XCD k (= blockIdx % 8) gets a contiguous range of tiles
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q8 = nb >> 3, r8 = nb & 7, xcd = b & 7, idx = b >> 3;
    const uint32_t logical = xcd * q8 + (xcd < r8 ? xcd : r8) + idx;
    const uint32_t t2 = logical % sg.tiles2;
    const uint32_t rest = logical / sg.tiles2;
    const uint32_t t1 = rest % sg.tiles1;
    const uint32_t chunk = rest / sg.tiles1;
    const uint32_t pb = (uint32_t)(((uint64_t)chunk * L.d0) / sg.nchunks);
    const uint32_t pe = (uint32_t)(((uint64_t)(chunk + 1) * L.d0) / sg.nchunks);

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int raw1 = (int)(t1 * (TY - 1)) - 1 + ty;
    const int raw2 = (int)(t2 * (TX - 1)) - 1 + tx;
    const uint32_t x1 = (uint32_t)((raw1 + (int)L.d1) % (int)L.d1);
    const uint32_t x2 = (uint32_t)((raw2 + (int)L.d2) % (int)L.d2);
    const bool interior = ty >= 1 && tx >= 1 && raw1 < (int)L.d1 && raw2 < (int)L.d2;
    const uint32_t col = x1 * L.s1 + x2;  // cell index within a plane
    // halo slot served by this thread (row TY: tid < TX; column TX: TX <= tid < TX+TY)
    int hslot = -1;
    uint32_t hcol = 0;
    if (tid < TX) {
        const int h1 = (int)(t1 * (TY - 1)) - 1 + TY, h2 = (int)(t2 * (TX - 1)) - 1 + tid;
        hcol = (uint32_t)((h1 + (int)L.d1) % (int)L.d1) * L.s1 + (uint32_t)((h2 + (int)L.d2) % (int)L.d2);
        hslot = TY * (TX + 1) + tid;
    } else if (tid < TX + TY) {
        const int hy = tid - TX;
        const int h1 = (int)(t1 * (TY - 1)) - 1 + hy, h2 = (int)(t2 * (TX - 1)) - 1 + TX;
        hcol = (uint32_t)((h1 + (int)L.d1) % (int)L.d1) * L.s1 + (uint32_t)((h2 + (int)L.d2) % (int)L.d2);
        hslot = hy * (TX + 1) + TX;
    }
    const int slot = ty * (TX + 1) + tx;

    const T w0 = (T)L.area[0], w1 = (T)L.area[1], w2 = (T)L.area[2];
    const T at = (T)L.at;

    auto load_links = [&](uint32_t site, M* u) {
        const uint32_t vo = fvoff(site, 3 * GR::S, sz);
#pragma unroll
        for (int l = 0; l < 3; l++) {
#pragma unroll
            for (int k = 0; k < GR::S; k++) GR::gref(u[l], GR::slot(k)) = bu.ld(vo, fsoff(l * GR::S + k, 3 * GR::S, sz));
            GR::complete(u[l]);
        }
    };
    auto lds_link = [&](int s, int l) {
        M m;
#pragma unroll
        for (int k = 0; k < G; k++) GR::gref(m, k) = smem[(l * G + k) * NS + s];
        return m;
    };

    uint32_t p = (pb + L.d0 - 1) % L.d0;  // the halo plane below the chunk
    M ucur[3], unx[3];
    load_links(p * L.s0 + col, ucur);
    A carry1 = GR::azero(), carry2 = GR::azero();  // E_1, E_2 contributions from plane p-1

    const uint32_t iters = pe - pb + 1;
    for (uint32_t it = 0; it < iters; it++) {
        const uint32_t pn = (p + 1 == L.d0) ? 0 : p + 1;
        const uint32_t site = p * L.s0 + col;
        const bool upd = it > 0 && interior;
        // next plane (own column) first: its links 1 and 2 enter this plane's plaquettes
        load_links(pn * L.s0 + col, unx);
        // links of this plane into LDS (own column + high halo)
#pragma unroll
        for (int l = 0; l < 3; l++)
#pragma unroll
            for (int k = 0; k < G; k++) smem[(l * G + k) * NS + slot] = GR::gget(ucur[l], k);
        if (hslot >= 0) {
            M hu[3];
            load_links(p * L.s0 + hcol, hu);
#pragma unroll
            for (int l = 0; l < 3; l++)
#pragma unroll
                for (int k = 0; k < G; k++) smem[(l * G + k) * NS + hslot] = GR::gget(hu[l], k);
        }
        __syncthreads();
        // plaquettes at y: P_jk = U_j(y) U_k(y+j) U_j(y+k)^+ U_k(y)^+
        const A a12 = proj_mul_adj(mul(ucur[1], lds_link(slot + (TX + 1), 2)), mul(ucur[2], lds_link(slot + 1, 1)));
        const A a01 = proj_mul_adj(mul(ucur[0], unx[1]), mul(ucur[1], lds_link(slot + (TX + 1), 0)));
        const A a02 = proj_mul_adj(mul(ucur[0], unx[2]), mul(ucur[2], lds_link(slot + 1, 0)));
        __syncthreads();  // link slots are reused for the exchanged contributions
        // this plane's E (and J): issued here so that the transports below cover the latency
        A e[3], jv[3];
        if (upd) {
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int k = 0; k < AD; k++) aref(e[d], k) = be.ld(fvoff(site, 3 * AD, sz), fsoff(d * AD + k, 3 * AD, sz));
            if (HAS_J) {
                const Buf<T> bj(Jp, 3u * AD * np * sz);
#pragma unroll
                for (int d = 0; d < 3; d++)
#pragma unroll
                    for (int k = 0; k < AD; k++) aref(jv[d], k) = bj.ld(fvoff(site, 3 * AD, sz), fsoff(d * AD + k, 3 * AD, sz));
            }
        }
        // transported contributions
        //   E_1(y+e0) += act(A01, U0^+)/a0^2   E_2(y+e0) += act(A02, U0^+)/a0^2
        //   E_0(y+e1) -= act(A01, U1^+)/a1^2   E_2(y+e1) += act(A12, U1^+)/a1^2
        //   E_0(y+e2) -= act(A02, U2^+)/a2^2   E_1(y+e2) -= act(A12, U2^+)/a2^2
        A c[4];
        c[0] = act_adj(a01, ucur[1]);
        c[1] = act_adj(a12, ucur[1]);
        c[2] = act_adj(a02, ucur[2]);
        c[3] = act_adj(a12, ucur[2]);
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int k = 0; k < AD; k++) smem[(q * AD + k) * NT + tid] = aget(c[q], k);
        __syncthreads();
        if (upd && in_box_dev(active, p, x1, x2)) {
            const int tm1 = tid - TX, tm2 = tid - 1;  // threads owning x-e1 and x-e2
            A f[3];
            // forward terms at x
            f[0] = GR::azero();
            axpy(f[0], a01, w1);
            axpy(f[0], a02, w2);
            f[1] = GR::azero();
            axpy(f[1], a01, -w0);
            axpy(f[1], a12, w2);
            f[2] = GR::azero();
            axpy(f[2], a02, -w0);
            axpy(f[2], a12, -w1);
            // from x-e0 (carried), x-e1 and x-e2 (LDS)
            axpy(f[1], carry1, w0);
            axpy(f[2], carry2, w0);
#pragma unroll
            for (int k = 0; k < AD; k++) {
                aref(f[0], k) -= smem[(0 * AD + k) * NT + tm1] * w1;
                aref(f[2], k) += smem[(1 * AD + k) * NT + tm1] * w1;
                aref(f[0], k) -= smem[(2 * AD + k) * NT + tm2] * w2;
                aref(f[1], k) -= smem[(3 * AD + k) * NT + tm2] * w2;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                axpy(e[d], f[d], at);
                if (HAS_J) axpy(e[d], jv[d], (T)L.uf[d]);
#pragma unroll
                for (int k = 0; k < AD; k++) be.st(aget(e[d], k), fvoff(site, 3 * AD, sz), fsoff(d * AD + k, 3 * AD, sz));
                const M un = mul(exp_link(e[d], -at), ucur[d]);
#pragma unroll
                for (int k = 0; k < GR::S; k++)
                    bun.st(GR::gget(un, GR::slot(k)), fvoff(site, 3 * GR::S, sz), fsoff(d * GR::S + k, 3 * GR::S, sz));
            }
        }
        carry1 = act_adj(a01, ucur[0]);
        carry2 = act_adj(a02, ucur[0]);
#pragma unroll
        for (int l = 0; l < 3; l++) ucur[l] = unx[l];
        p = pn;
        __syncthreads();  // contributions read before the next plane's links overwrite them
    }
}

}  // namespace

// ---------------------------------------------------------------- launcher
namespace {

// LDS bytes of one block of the streaming kernel
constexpr size_t stream_lds_bytes(int nc, size_t tsize, int ty, int tx)
{
    const size_t g = nc == 2 ? 4 : 18, ad = nc == 2 ? 3 : 8;
    const size_t links = 3 * g * (size_t)(ty + 1) * (tx + 1), contrib = 4 * ad * (size_t)ty * tx;
    return (links > contrib ? links : contrib) * tsize;
}
constexpr size_t kMaxLds = 160 * 1024;

template <int TY, int TX>
hipError_t launch_stream_tile(int nc, int prec, const Lattice& L, const Box& active, uint32_t nchunks, const void* U,
                              void* Unext, void* E, const void* J, bool has_j, hipStream_t st)
{
    StreamGeom sg;
    sg.tiles1 = (L.d1 + (TY - 1) - 1) / (TY - 1);
    sg.tiles2 = (L.d2 + (TX - 1) - 1) / (TX - 1);
    sg.nchunks = nchunks;
    const dim3 grid(sg.tiles1 * sg.tiles2 * nchunks), block(TY * TX);
#define OPX_STREAM_LAUNCH(NC_, T_)                                                                              \
    do {                                                                                                        \
        if constexpr (stream_lds_bytes(NC_, sizeof(T_), TY, TX) > kMaxLds) return hipErrorInvalidConfiguration; \
        else if (has_j)                                                                                         \
            hipLaunchKernelGGL((k_stream_update<NC_, T_, TY, TX, true>), grid, block, 0, st, L, active, sg,    \
                               (const T_*)U, (T_*)Unext, (T_*)E, (const T_*)J);                                 \
        else                                                                                                    \
            hipLaunchKernelGGL((k_stream_update<NC_, T_, TY, TX, false>), grid, block, 0, st, L, active, sg,   \
                               (const T_*)U, (T_*)Unext, (T_*)E, (const T_*)nullptr);                           \
    } while (0)
    if (nc == 2 && prec == 0) OPX_STREAM_LAUNCH(2, double);
    else if (nc == 2) OPX_STREAM_LAUNCH(2, float);
    else if (prec == 0) OPX_STREAM_LAUNCH(3, double);
    else OPX_STREAM_LAUNCH(3, float);
#undef OPX_STREAM_LAUNCH
    return hipGetLastError();
}

}  // namespace

bool stream_supported(int nc, int prec, const Lattice& L, int tile)
{
    const uint64_t G = nc == 2 ? 4 : (prec == 0 ? 18 : 12);   // reals stored per link
    const uint64_t sz = prec == 0 ? 8 : 4;
    if (padded_cells(L.n) * 3 * G * sz >= (1ull << 32)) return false;  // one buffer descriptor per field
    if (tile < 0 || tile >= kStreamTiles) return false;
    int ty, tx;
    stream_tile_shape(tile, &ty, &tx);
    if (stream_lds_bytes(nc, sz, ty, tx) > kMaxLds) return false;
    return L.d0 >= 1 && L.d1 >= 1 && L.d2 >= 1;
}

void stream_tile_shape(int tile, int* ty, int* tx)
{
    static const int shapes[kStreamTiles][2] = {{16, 16}, {8, 32}, {16, 32}, {8, 64}, {8, 16}};
    *ty = shapes[tile][0];
    *tx = shapes[tile][1];
}

hipError_t launch_stream_update(int nc, int prec, const Lattice& L, const Box& active, int tile, uint32_t nchunks,
                                const void* U, void* Unext, void* E, const void* J, bool has_j, hipStream_t st)
{
    if (nchunks < 1) nchunks = 1;
    if (nchunks > L.d0) nchunks = L.d0;
    switch (tile) {
    case 0: return launch_stream_tile<16, 16>(nc, prec, L, active, nchunks, U, Unext, E, J, has_j, st);
    case 1: return launch_stream_tile<8, 32>(nc, prec, L, active, nchunks, U, Unext, E, J, has_j, st);
    case 2: return launch_stream_tile<16, 32>(nc, prec, L, active, nchunks, U, Unext, E, J, has_j, st);
    case 3: return launch_stream_tile<8, 64>(nc, prec, L, active, nchunks, U, Unext, E, J, has_j, st);
    case 4: return launch_stream_tile<8, 16>(nc, prec, L, active, nchunks, U, Unext, E, J, has_j, st);
    default: return hipErrorInvalidValue;
    }
}

uint32_t stream_blocks_per_chunk(const Lattice& L, int tile)
{
    int ty, tx;
    stream_tile_shape(tile, &ty, &tx);
    return ((L.d1 + (ty - 1) - 1) / (ty - 1)) * ((L.d2 + (tx - 1) - 1) / (tx - 1));
}

}  // namespace opx
