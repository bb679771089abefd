// pw_buffer.h -- buffer-resource loads and stores shared by the 2D field kernels.
//
// Fields are read and written through buffer resources: a scalar base and byte size, a 32-bit
// per-lane offset and a scalar (per-column) offset.  A lane whose cell lies outside its field gets
// the out-of-range offset kOOB, which the hardware answers with 0 (the value a zero boundary
// gives) without a memory access; stores at kOOB are dropped.  Every size stays below 2 GiB.
#pragma once

#include <hip/hip_runtime.h>

namespace pwbuf {

constexpr unsigned kOOB = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t win(const void* base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T bld(__amdgpu_buffer_rsrc_t r, unsigned off, bool ok);
template <>
__device__ __forceinline__ float bld<float>(__amdgpu_buffer_rsrc_t r, unsigned off, bool ok)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, ok ? off : kOOB, 0, 0));
}
template <>
__device__ __forceinline__ double bld<double>(__amdgpu_buffer_rsrc_t r, unsigned off, bool ok)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, ok ? off : kOOB, 0, 0));
}

// the same with a scalar offset added to the lane offset
template <typename T>
__device__ __forceinline__ T bldo(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float bldo<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double bldo<double>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// non-temporal (streamed once: not kept in the caches for reuse) load and store of an f32 or f64
template <typename T>
__device__ __forceinline__ T bldo_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float bldo_nt<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2));
}
template <>
__device__ __forceinline__ double bldo_nt<double>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}
__device__ __forceinline__ void bsto_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 2);
}
__device__ __forceinline__ void bsto_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double v)
{
    using u32x2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 2);
}

__device__ __forceinline__ void bsto(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
__device__ __forceinline__ void bsto(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double v)
{
    using u32x2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

}  // namespace pwbuf
