// opx_buf.h -- raw buffer loads/stores with the component offset in an SGPR.
//
// A field is one buffer resource in the 64-cell block layout; element (component k, cell s) is
// read with voffset = the byte offset of cell s's block plus the cell's offset in its component
// group (per lane) and soffset = the component's offset (uniform, held in a scalar register), so
// the components of a cell share one vector address. Usable while the field is smaller than 4 GiB.
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
    // 16 bytes (16 / sizeof(T) values) at voff + soff: one load; two 8-byte stores
    __device__ __forceinline__ void ld16(T* out, uint32_t voff, uint32_t soff, int aux) const;
    __device__ __forceinline__ void st16(const T* in, uint32_t voff, uint32_t soff, int aux) const;
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

typedef int v4i __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ void Buf<T>::ld16(T* out, uint32_t voff, uint32_t soff, int aux) const
{
    const v4i raw = aux ? __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, kBufNonTemporal)
                        : __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    T t[16 / sizeof(T)];
    __builtin_memcpy(t, &raw, 16);
#pragma unroll
    for (unsigned q = 0; q < 16 / sizeof(T); q++) out[q] = t[q];
}
template <typename T>
__device__ __forceinline__ void Buf<T>::st16(const T* in, uint32_t voff, uint32_t soff, int aux) const
{
    T t[16 / sizeof(T)];
#pragma unroll
    for (unsigned q = 0; q < 16 / sizeof(T); q++) t[q] = in[q];
    v2i lo, hi;
    __builtin_memcpy(&lo, t, 8);
    __builtin_memcpy(&hi, reinterpret_cast<const char*>(t) + 8, 8);
    // the second half's 8 bytes go into the scalar offset: two stores with different scalar
    // offsets are not merged into one 16-byte store
    if (aux) {
        __builtin_amdgcn_raw_buffer_store_b64(lo, r, voff, soff, kBufNonTemporal);
        __builtin_amdgcn_raw_buffer_store_b64(hi, r, voff, soff + 8, kBufNonTemporal);
    } else {
        __builtin_amdgcn_raw_buffer_store_b64(lo, r, voff, soff, 0);
        __builtin_amdgcn_raw_buffer_store_b64(hi, r, voff, soff + 8, 0);
    }
}

}  // namespace opx
