// pop_store.h — population storage forms of the fused step.
//
//   WAAM_POP_F32   float, the populations themselves (the plain method);
//   WAAM_POP_F16   IEEE binary16 holding (f - w_q) * 2^15 (shifted and scaled), rounded to
//                  nearest even, saturating at +-65504 (|f - w_q| < 2);
//   WAAM_POP_BF16  bfloat16 holding f - w_q, rounded to nearest even.
//
// Collision arithmetic is fp32 in every form: a shifted population is widened to fp32 and w_q
// added back on load, and w_q subtracted in fp32 before the narrowing store. Shifting around
// the rest equilibrium keeps the stored magnitudes small (|f - w_q| << w_q in a melt pool), so
// the 11- or 8-bit significand resolves the departure from equilibrium rather than w_q itself.
// The binary16 scale 2^15 (a power of two: exact in fp32 both ways) moves binary16's normal
// range [2^-14, 65504] onto departures in [2^-29, 2): every non-zero departure the fp32 collision
// can produce is a multiple of w_q's fp32 spacing (2^-25 for w_0 = 1/3, 2^-29 for 1/36), so it
// is stored with the full 11-bit significand instead of falling into binary16's subnormals
// (below 6.1e-5 unscaled, where a melt pool's departures of 1e-6..1e-4 would keep only a few
// bits). bfloat16 has fp32's exponent range and needs no scale.
// Streaming moves stored values unchanged (bounce-back pairs q with opp(q), and w_opp(q) = w_q),
// so the pull copies 16-bit words without converting them.
#pragma once
#include "waam_common.h"

namespace waam {

__host__ __device__ __forceinline__ float bf16_to_float(uint16_t b)
{
    return __uint_as_float((uint32_t)b << 16);
}

/* round to nearest even; NaN stays a (quiet) NaN */
__host__ __device__ __forceinline__ uint16_t float_to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float f16_to_float(uint16_t h)
{
    return (float)__builtin_bit_cast(_Float16, h);
}

__device__ __forceinline__ uint16_t float_to_f16(float x)
{
    return __builtin_bit_cast(uint16_t, (_Float16)x);  // v_cvt_f16_f32: round to nearest even
}

template <int PD> struct PopStore;

template <> struct PopStore<WAAM_POP_F32> {
    using E = float;
    static __device__ __forceinline__ float load(const E *p, size_t i, int) { return p[i]; }
    static __device__ __forceinline__ void store(E *p, size_t i, int, float v) { p[i] = v; }
};

/* binary16 storage scale of f - w_q (a power of two) and the saturation bound */
constexpr float kF16Scale = 32768.0f;
constexpr float kF16Unscale = 1.0f / 32768.0f;
constexpr float kF16Max = 65504.0f;

template <> struct PopStore<WAAM_POP_F16> {
    using E = uint16_t;
    static __device__ __forceinline__ float load(const E *p, size_t i, int q)
    {
        return f16_to_float(p[i]) * kF16Unscale + wq(q);
    }
    static __device__ __forceinline__ void store(E *p, size_t i, int q, float v)
    {
        float x = (v - wq(q)) * kF16Scale;
        x = x > kF16Max ? kF16Max : (x < -kF16Max ? -kF16Max : x);  // NaN passes through
        p[i] = float_to_f16(x);
    }
};

template <> struct PopStore<WAAM_POP_BF16> {
    using E = uint16_t;
    static __device__ __forceinline__ float load(const E *p, size_t i, int q)
    {
        return bf16_to_float(p[i]) + wq(q);
    }
    static __device__ __forceinline__ void store(E *p, size_t i, int q, float v)
    {
        p[i] = float_to_bf16(v - wq(q));
    }
};

/* wavefront shuffle of one stored element */
__device__ __forceinline__ float shfl_elem(float v, int src) { return __shfl(v, src); }
__device__ __forceinline__ uint16_t shfl_elem(uint16_t v, int src)
{
    return (uint16_t)__shfl((int)v, src);
}

}  // namespace waam
