// Storage codecs of the 4-channel fields (float4, binary16 x4, bfloat16 x4): load as
// float4, store with round-to-nearest-even.  Shared by the step and diagnostics kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "fs_internal.h"

// ---- storage codecs -----------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (unsigned short)(x >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short u) { return __uint_as_float((unsigned)u << 16); }

// Word 3 of a raw (untyped) buffer resource on gfx9-class chips: 32-bit data format, no swizzle.
constexpr int FS_BUF_WORD3 = 0x00020000;
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

template <class S> struct Codec;
template <> struct Codec<StF32> {
    static __device__ __forceinline__ float4 ld(const float4* p, int i) { return p[i]; }
    static __device__ __forceinline__ float4 ldb(__amdgpu_buffer_rsrc_t r, int off) {
        const u32x4_t u = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
    }
    static __device__ __forceinline__ void st(float4* p, int i, float4 v) { p[i] = v; }
    static __device__ __forceinline__ float rt(float v) { return v; }          // the stored value of v
};
template <> struct Codec<StF16> {
    static __device__ __forceinline__ float4 ld(const half4_t* p, int i) {
        half4_t h = p[i];
        return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
    }
    static __device__ __forceinline__ float4 dec(u32x2_t u) {
        const half4_t h = __builtin_bit_cast(half4_t, u);
        return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
    }
    static __device__ __forceinline__ float4 ldb(__amdgpu_buffer_rsrc_t r, int off) {
        return dec(__builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    static __device__ __forceinline__ u32x2_t enc(float4 v) {   // the texel st writes, as two words
        half4_t h;
        h.x = (_Float16)v.x; h.y = (_Float16)v.y; h.z = (_Float16)v.z; h.w = (_Float16)v.w;
        return __builtin_bit_cast(u32x2_t, h);
    }
    static __device__ __forceinline__ void st(half4_t* p, int i, float4 v) {
        half4_t h;
        h.x = (_Float16)v.x; h.y = (_Float16)v.y; h.z = (_Float16)v.z; h.w = (_Float16)v.w;
        p[i] = h;
    }
    static __device__ __forceinline__ float rt(float v) { return (float)(_Float16)v; }
};
template <> struct Codec<StBF16> {
    static __device__ __forceinline__ float4 ld(const ushort4_t* p, int i) {
        ushort4_t u = p[i];
        return make_float4(bf2f(u.x), bf2f(u.y), bf2f(u.z), bf2f(u.w));
    }
    static __device__ __forceinline__ float4 dec(u32x2_t w) {
        const ushort4_t u = __builtin_bit_cast(ushort4_t, w);
        return make_float4(bf2f(u.x), bf2f(u.y), bf2f(u.z), bf2f(u.w));
    }
    static __device__ __forceinline__ float4 ldb(__amdgpu_buffer_rsrc_t r, int off) {
        return dec(__builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    static __device__ __forceinline__ u32x2_t enc(float4 v) {   // the texel st writes, as two words
        ushort4_t u;
        u.x = f2bf(v.x); u.y = f2bf(v.y); u.z = f2bf(v.z); u.w = f2bf(v.w);
        return __builtin_bit_cast(u32x2_t, u);
    }
    static __device__ __forceinline__ void st(ushort4_t* p, int i, float4 v) {
        ushort4_t u;
        u.x = f2bf(v.x); u.y = f2bf(v.y); u.z = f2bf(v.z); u.w = f2bf(v.w);
        p[i] = u;
    }
    static __device__ __forceinline__ float rt(float v) { return bf2f(f2bf(v)); }
};
