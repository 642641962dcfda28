// a3d_store.hpp -- device-side storage helpers shared by the 3D kernels:
// element access by 32-bit offsets, FP32/BF16/FP16 storage with per-field
// power-of-two scales (single values and z-adjacent pairs), optional
// stochastic rounding, and the velocity update of one cell.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hip/hip_bf16.h"
#include "hip/hip_fp16.h"

namespace sm {

template <typename S>
struct Store;

// Element access by a 32-bit element index turned into a 32-bit byte
// offset from a uniform base, so loads use the SGPR-base + VGPR-offset form
// (every device array here is well below 4 GiB).
template <typename T>
__device__ __forceinline__ T ld_at(const T *base, unsigned idx)
{
    return *(const T *)((const char *)base + idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st_at(T *base, unsigned idx, T v)
{
    *(T *)((char *)base + idx * (unsigned)sizeof(T)) = v;
}

// Stochastic rounding of 16-bit stores (optional): the discarded low bits
// of the FP32 value are compared with a pseudo-random threshold, so a value
// rounds up in magnitude with probability equal to its fraction of an ulp.
// Increments smaller than half an ulp then survive on average instead of
// being rounded away every step.  The random bits are a hash of the
// element index and a per-step, per-buffer salt, so results are
// reproducible and independent of the launch shape.
constexpr unsigned SALT_MV = 0x68E31DA4u;   // velocity-side memory variables
constexpr unsigned SALT_MP = 0xB5297A4Du;   // pressure-side memory variables

__device__ __forceinline__ unsigned sr_bits(unsigned i, unsigned salt)
{
    unsigned h = (i ^ salt) * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    return h;
}

// Store<S>: ld/st of element i with the field's scale; enc/dec convert
// between a physical FP32 value and the stored bits (Bits), which buffer
// loads and stores move.  st is enc followed by the store, so every kernel
// rounds the same way.
template <>
struct Store<float> {
    using Bits = unsigned;
    __device__ static __forceinline__ float dec(Bits b, float) { return __uint_as_float(b); }
    __device__ static __forceinline__ Bits enc(float v, float, unsigned, unsigned, int)
    {
        return __float_as_uint(v);
    }
    __device__ static __forceinline__ float ld(const float *p, unsigned i, float) { return ld_at(p, i); }
    __device__ static __forceinline__ void st(float *p, unsigned i, float v, float, unsigned, int)
    {
        st_at(p, i, v);
    }
};

// BF16 keeps FP32's exponent range, so its fields are never scaled (the
// host leaves every BF16 scale at 1): the scale arguments are ignored, which
// spares a multiply per access and keeps the scales out of the registers.
template <>
struct Store<__hip_bfloat16> {
    using Bits = unsigned short;
    __device__ static __forceinline__ float dec(Bits b, float) { return __uint_as_float((unsigned)b << 16); }
    __device__ static __forceinline__ Bits enc(float v, float, unsigned i, unsigned salt, int sr)
    {
        const float x = v;
        if (sr) {
            // add 16 random bits below the BF16 mantissa, then truncate
            const unsigned b = __float_as_uint(x) + (sr_bits(i, salt) & 0xFFFFu);
            return (Bits)(b >> 16);
        }
        return __builtin_bit_cast(Bits, __float2bfloat16(x));
    }
    __device__ static __forceinline__ float ld(const __hip_bfloat16 *p, unsigned i, float)
    {
        return __bfloat162float(ld_at(p, i));
    }
    __device__ static __forceinline__ void st(__hip_bfloat16 *p, unsigned i, float v, float inv,
                                     unsigned salt, int sr)
    {
        st_at((Bits *)p, i, enc(v, inv, i, salt, sr));
    }
};

template <>
struct Store<__half> {
    using Bits = unsigned short;
    __device__ static __forceinline__ float dec(Bits b, float scale)
    {
        return __half2float(__builtin_bit_cast(__half, b)) * scale;
    }
    __device__ static __forceinline__ Bits enc(float v, float inv, unsigned i, unsigned salt, int sr)
    {
        float x = v * inv;
        if (sr) {
            // normal FP16 range [2^-14, 65504): 13 random bits below the FP16
            // mantissa, then truncate to a value FP16 holds exactly
            unsigned b = __float_as_uint(x);
            const unsigned mag = b & 0x7FFFFFFFu;
            if (mag >= 0x38800000u && mag < 0x477FE000u) {
                b = (b + (sr_bits(i, salt) & 0x1FFFu)) & ~0x1FFFu;
                x = __uint_as_float(b);
            }
        }
        return __builtin_bit_cast(Bits, __float2half(x));
    }
    __device__ static __forceinline__ float ld(const __half *p, unsigned i, float scale)
    {
        return __half2float(ld_at(p, i)) * scale;
    }
    __device__ static __forceinline__ void st(__half *p, unsigned i, float v, float inv, unsigned salt,
                                     int sr)
    {
        st_at((Bits *)p, i, enc(v, inv, i, salt, sr));
    }
};

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// one velocity component at one cell: v - 2dt/(rho0+rho1) * dP, with the
// CPML memory variable m (updated in place) where slab is true
__device__ __forceinline__ float vel_update(float v_old, float p0, float p1, float r0, float r1,
                                            float ih, float two_dt, bool slab, float bh, float ah,
                                            float ikh, float &m)
{
    float dP = (p1 - p0) * ih;
    if (slab) {
        m = bh * m + ah * dP;
        dP = dP * ikh + m;
    }
    return v_old - two_dt * fast_rcp(r0 + r1) * dP;
}

// Two z-adjacent elements (i even, i and i+1) with one memory access:
// dwordx2 for FP32, one dword holding two 16-bit values otherwise.  The
// element index must be even (the access is naturally aligned).
template <typename S>
struct Pair;

// Raw: the pair as stored (kept in registers between the load and its use;
// one VGPR for two 16-bit values); unpack() gives the two physical values.
template <>
struct Pair<float> {
    using Raw = float2;
    __device__ static __forceinline__ Raw raw(const float *p, unsigned i)
    {
        return ld_at((const float2 *)p, i >> 1);
    }
    __device__ static __forceinline__ float2 unpack(Raw r, float) { return r; }
    __device__ static __forceinline__ float2 ld(const float *p, unsigned i, float)
    {
        return ld_at((const float2 *)p, i >> 1);
    }
    __device__ static __forceinline__ void st(float *p, unsigned i, float2 v, float, unsigned, int)
    {
        st_at((float2 *)p, i >> 1, v);
    }
};

template <>
struct Pair<__hip_bfloat16> {
    using Raw = unsigned;
    __device__ static __forceinline__ Raw raw(const __hip_bfloat16 *p, unsigned i)
    {
        return ld_at((const unsigned *)p, i >> 1);
    }
    // unscaled, as Store<__hip_bfloat16>
    __device__ static __forceinline__ float2 unpack(Raw w, float)
    {
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xFFFF0000u));
    }
    __device__ static __forceinline__ float2 ld(const __hip_bfloat16 *p, unsigned i, float scale)
    {
        return unpack(raw(p, i), scale);
    }
    __device__ static __forceinline__ void st(__hip_bfloat16 *p, unsigned i, float2 v, float,
                                     unsigned salt, int sr)
    {
        unsigned lo, hi;
        const float x0 = v.x, x1 = v.y;
        if (sr) {
            lo = (__float_as_uint(x0) + (sr_bits(i, salt) & 0xFFFFu)) >> 16;
            hi = (__float_as_uint(x1) + (sr_bits(i + 1, salt) & 0xFFFFu)) >> 16;
        } else {
            lo = __builtin_bit_cast(unsigned short, __float2bfloat16(x0));
            hi = __builtin_bit_cast(unsigned short, __float2bfloat16(x1));
        }
        st_at((unsigned *)p, i >> 1, lo | (hi << 16));
    }
};

template <>
struct Pair<__half> {
    using Raw = unsigned;
    __device__ static __forceinline__ Raw raw(const __half *p, unsigned i)
    {
        return ld_at((const unsigned *)p, i >> 1);
    }
    __device__ static __forceinline__ float2 unpack(Raw w, float scale)
    {
        const float2 f = __half22float2(__builtin_bit_cast(__half2, w));
        return make_float2(f.x * scale, f.y * scale);
    }
    __device__ static __forceinline__ float2 ld(const __half *p, unsigned i, float scale)
    {
        return unpack(raw(p, i), scale);
    }
    __device__ static __forceinline__ void st(__half *p, unsigned i, float2 v, float inv, unsigned salt,
                                     int sr)
    {
        if (sr) {
            // element by element, as the single-value store rounds
            Store<__half>::st(p, i, v.x, inv, salt, sr);
            Store<__half>::st(p, i + 1, v.y, inv, salt, sr);
            return;
        }
        st_at((__half2 *)p, i >> 1, __floats2half2_rn(v.x * inv, v.y * inv));
    }
};

}  // namespace sm
