// opx_buf.h -- raw buffer loads/stores with the component offset in an SGPR.
//
// A field is one buffer resource in the 64-cell block layout; element (component k, cell s) is
// read with voffset = the byte offset of cell s's block plus (s % 64) * sizeof(T) (per lane) and
// soffset = k * 64 * sizeof(T) (uniform, held in a scalar register), so the components of a cell
// share one vector address. Usable while the field is smaller than 4 GiB.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace opx {

constexpr int kBufFlags = 0x00020000;  // gfx9 raw buffer descriptor word 3

template <typename T>
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Buf(const T* p, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), (short)0, (int)bytes, kBufFlags))
    {
    }
    __device__ __forceinline__ T ld(uint32_t voff, uint32_t soff, int aux = 0) const;
    __device__ __forceinline__ void st(T v, uint32_t voff, uint32_t soff, int aux = 0) const;
};

// cache-policy operand of the buffer instructions: non-temporal (data touched once)
constexpr int kBufNonTemporal = 2;

template <>
__device__ __forceinline__ float Buf<float>::ld(uint32_t voff, uint32_t soff, int aux) const
{
    return aux ? __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, kBufNonTemporal))
               : __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double Buf<double>::ld(uint32_t voff, uint32_t soff, int aux) const
{
    return aux ? __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kBufNonTemporal))
               : __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
typedef int v2i __attribute__((ext_vector_type(2)));
template <>
__device__ __forceinline__ void Buf<float>::st(float v, uint32_t voff, uint32_t soff, int aux) const
{
    if (aux)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, voff, soff, kBufNonTemporal);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, voff, soff, 0);
}
template <>
__device__ __forceinline__ void Buf<double>::st(double v, uint32_t voff, uint32_t soff, int aux) const
{
    if (aux)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, voff, soff, kBufNonTemporal);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, voff, soff, 0);
}

}  // namespace opx
