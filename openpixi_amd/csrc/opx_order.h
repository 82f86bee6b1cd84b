// opx_order.h -- block order of the per-link and per-cell kernels.
//
// The lattice is cut into tiles of `rows` consecutive x1 rows (rows * d2 cells); a tile is
// swept along x0, and the blocks of one x0 plane of a tile are consecutive. Consecutive
// blocks therefore read the links of neighbouring planes soon after one another, while they
// still sit in L2. xcd_logical hands each XCD (the hardware deals blockIdx round robin over
// 8 XCDs, each with its own L2) one contiguous range of that order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "opx_kernels.h"

namespace opx {

struct LinkOrder {
    uint32_t rows;                   // x1 rows per tile
    uint32_t blocks_per_tile_plane;  // blocks of B threads covering rows * d2 cells
};

// host side: the order for tiles of at most `rows` rows and blocks of `block` threads, and
// the number of logical blocks (times `per_cell` when a cell has several blocks' work)
inline LinkOrder make_link_order(const Lattice& L, uint32_t rows, uint32_t block)
{
    LinkOrder o;
    o.rows = L.d1 < rows ? L.d1 : rows;
    o.blocks_per_tile_plane = (o.rows * L.d2 + block - 1) / block;
    return o;
}
inline uint32_t link_order_blocks(const Lattice& L, const LinkOrder& o)
{
    return o.blocks_per_tile_plane * L.d0 * ((L.d1 + o.rows - 1) / o.rows);
}

// the cell of thread threadIdx.x in logical block `logical`; false past a partial tile's end
template <uint32_t B = 256>
__device__ __forceinline__ bool tile_site(const Lattice& L, const LinkOrder& order, uint32_t logical, uint32_t& site)
{
    const uint32_t sb = logical % order.blocks_per_tile_plane;
    const uint32_t rest2 = logical / order.blocks_per_tile_plane;
    const uint32_t x0b = rest2 % L.d0;
    const uint32_t tile = rest2 / L.d0;
    const uint32_t row0 = tile * order.rows;
    const uint32_t tile_rows = (L.d1 - row0 < order.rows) ? L.d1 - row0 : order.rows;
    const uint32_t in_tile = sb * B + threadIdx.x;
    if (in_tile >= tile_rows * L.d2) return false;
    site = x0b * L.s0 + row0 * L.d2 + in_tile;
    return true;
}

// logical block of hardware block b of nb: XCD b % 8 takes the b-th eighth of the order
__device__ __forceinline__ uint32_t xcd_logical(uint32_t b, uint32_t nb)
{
    const uint32_t q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
    return xcd * q8 + (xcd < r8 ? xcd : r8) + (b >> 3);
}

}  // namespace opx
