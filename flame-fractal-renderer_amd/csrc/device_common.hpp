// Device helpers shared by the walker kernels: the per-walker counter-based
// random stream and precision-overloaded math.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#else
// compiled at run time (a flame's own kernel): the runtime header is
// implicit, the <math.h> constants are not
#ifndef M_PI
#define M_PI 3.14159265358979323846
#define M_PI_2 1.57079632679489661923
#define M_PI_4 0.78539816339744830962
#define M_1_PI 0.31830988618379067154
#define M_2_PI 0.63661977236758134308
#endif
#ifndef INFINITY
#define INFINITY __builtin_huge_valf()
#endif
// the run-time compiler declares the fixed-width integers in its own
// namespace
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

namespace ffr {

// ---------------------------------------------------------------------------
// Walker stream.  Word n of walker w under seed s is half (n & 1) of the
// Philox4x32-10 block (Salmon et al., SC'11) with key (lo32 s, hi32 s) and
// counter (lo32 (n>>1), hi32 (n>>1), lo32 w, hi32 w); half 0 = out0|out1<<32,
// half 1 = out2|out3<<32.  This stream replaces the reference's clock-seeded
// thread-local ISAAC (src/rng/flame_rng.hpp:29-36): every walker (reference
// "batch", buffer_renderer.hpp:155) owns one, so renders are reproducible.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                              uint32_t c3, uint32_t k0, uint32_t k1,
                                              uint32_t &o0, uint32_t &o1, uint32_t &o2,
                                              uint32_t &o3)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        // one 32x32->64 product gives both halves (v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32);
        uint32_t lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32);
        uint32_t lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    o0 = c0;
    o1 = c1;
    o2 = c2;
    o3 = c3;
}

struct Stream {
    uint32_t k0, k1, w0, w1;
    uint64_t n;       // index of the next word
    uint64_t cache;   // word cache_n, when loaded with its even partner
    uint64_t cache_n;

    __device__ __forceinline__ void init(uint64_t seed, uint64_t walker)
    {
        // the key is the same for every lane: scalar registers, scalar adds
        k0 = __builtin_amdgcn_readfirstlane((uint32_t)seed);
        k1 = __builtin_amdgcn_readfirstlane((uint32_t)(seed >> 32));
        w0 = (uint32_t)walker;
        w1 = (uint32_t)(walker >> 32);
        n = 0;
        cache = 0;
        cache_n = ~0ull;
    }

    __device__ __forceinline__ uint64_t next()
    {
        uint64_t w;
        if (n == cache_n) {
            w = cache;
        } else {
            uint64_t blk = n >> 1;
            uint32_t o0, o1, o2, o3;
            philox4x32_10((uint32_t)blk, (uint32_t)(blk >> 32), w0, w1, k0, k1, o0, o1,
                          o2, o3);
            uint64_t lo = (uint64_t)o0 | ((uint64_t)o1 << 32);
            uint64_t hi = (uint64_t)o2 | ((uint64_t)o3 << 32);
            if (n & 1) {
                w = hi;
            } else {
                w = lo;
                cache = hi;
                cache_n = n + 1;
            }
        }
        ++n;
        return w;
    }
};

// randNum (flame_rng.hpp:67-87): top 53 bits over 2^53 in double, top 24
// bits over 2^24 in float; both exact
__device__ __forceinline__ double rand_from_word(uint64_t w, double)
{
    return (double)(w >> 11) * 0x1p-53;
}
__device__ __forceinline__ float rand_from_word(uint64_t w, float)
{
    return (float)(w >> 40) * 0x1p-24f;
}
__device__ __forceinline__ double rand_num(Stream &s, double d) { return rand_from_word(s.next(), d); }
__device__ __forceinline__ float rand_num(Stream &s, float f) { return rand_from_word(s.next(), f); }

// ---------------------------------------------------------------------------
// math overloads: the argument type picks the precision (C++ <cmath> rules)
// ---------------------------------------------------------------------------
#define FFR_M1(name, dfn, ffn)                                                   \
    __device__ __forceinline__ double m_##name(double x) { return dfn(x); }      \
    __device__ __forceinline__ float m_##name(float x) { return ffn(x); }
FFR_M1(sin, sin, sinf)
FFR_M1(cos, cos, cosf)
FFR_M1(tan, tan, tanf)
FFR_M1(sqrt, sqrt, sqrtf)
FFR_M1(exp, exp, expf)
FFR_M1(log, log, logf)
FFR_M1(log10, log10, log10f)
FFR_M1(sinh, sinh, sinhf)
FFR_M1(cosh, cosh, coshf)
FFR_M1(acos, acos, acosf)
FFR_M1(floor, floor, floorf)
FFR_M1(trunc, trunc, truncf)
FFR_M1(rint, rint, rintf)
FFR_M1(fabs, fabs, fabsf)
#undef FFR_M1

__device__ __forceinline__ double m_atan2(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ float m_atan2(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ double m_pow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float m_pow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double m_fmod(double a, double b) { return fmod(a, b); }
__device__ __forceinline__ float m_fmod(float a, float b) { return fmodf(a, b); }
__device__ __forceinline__ double m_hypot(double a, double b) { return hypot(a, b); }
__device__ __forceinline__ float m_hypot(float a, float b) { return hypotf(a, b); }
__device__ __forceinline__ double m_copysign(double a, double b) { return copysign(a, b); }
__device__ __forceinline__ void m_sincos(double a, double *s, double *c) { sincos(a, s, c); }
__device__ __forceinline__ void m_sincos(float a, float *s, float *c) { sincosf(a, s, c); }

// FP32 trajectories with the default build's 8-byte buffer words (no
// reference counterpart): the six basic variations and the colour blend
// compute in float throughout.  In both reference builds a double literal
// stays double and promotes float operands, as the reference's expressions
// do (src/types/types.hpp:29 sets num_t; its literals are double).
template <typename R, typename CT> __host__ __device__ constexpr bool native_f32()
{
    return sizeof(R) == 4 && sizeof(CT) == 8;
}
// the type a reference double literal takes in an expression
template <typename R, bool NATIVE> struct LitT { using type = double; };
template <typename R> struct LitT<R, true> { using type = R; };
template <typename R, bool NATIVE> using lit_t = typename LitT<R, NATIVE>::type;

// the colour blend (1 - s) c + s col of RenderIterator::iterate
// (render_iterator.hpp:112-138), in the reference's operand types
template <typename R, bool NATIVE> __device__ __forceinline__ R colour_blend(R c, R s, R col)
{
    using L = lit_t<R, NATIVE>;
    return (R)(((L)1.0 - s) * c + s * col);
}

template <typename R> struct Consts;
template <> struct Consts<double> {
    static constexpr double eps = 1e-20;           // constants.hpp:17-22
    static constexpr double bad = 1e20;            // constants.hpp:50-56
};
template <> struct Consts<float> {
    static constexpr float eps = 1e-10f;
    static constexpr float bad = 1e10f;
};

template <typename R> __device__ __forceinline__ bool bad_value(R v)
{
    // utils/flame.hpp:26-29
    return m_fabs(v) > Consts<R>::bad || v != v;
}

} // namespace ffr
