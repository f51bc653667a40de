// dev_math.h -- device-side vector / colour / RNG helpers for the gfx950 kernels.
//
// Every float expression keeps the operand order of the reference's helpers
// (src/utility/helper_math.h op semantics: dot = x*x'+y*y'+z*z' left to
// right, normalize = v*(1/sqrt(dot(v,v))), host-style min/max a<b?a:b) and the
// whole library is compiled with -ffp-contract=off, so the kernels produce the
// same bits as a host build of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define GPR_HD __host__ __device__ __forceinline__
#define GPR_D __device__ __forceinline__

namespace gpr {

// Load through a global-address-space pointer.  A pointer read from device
// memory (the multi-job scene table) is generic to the compiler, which then
// emits flat loads, counted against the LDS / scalar-memory wait counter as
// well; every buffer the kernels read lives in global memory.
template <class T>
GPR_D T ldg(const T* p) {
    if constexpr (std::is_scalar<T>::value) {
        return *(const __attribute__((address_space(1))) T*)p;
    } else {
        // records: word by word (the compiler merges neighbouring words),
        // assembled with memcpy
        static_assert(sizeof(T) % 4 == 0 && alignof(T) >= 4, "word-wise global load");
        static_assert(std::is_trivially_copyable<T>::value, "record load by bytes");
        uint32_t w[sizeof(T) / 4];
        const __attribute__((address_space(1))) uint32_t* src = (const __attribute__((address_space(1))) uint32_t*)p;
#pragma unroll
        for (int k = 0; k < (int)(sizeof(T) / 4); ++k) w[k] = src[k];
        T r;
        __builtin_memcpy(&r, w, sizeof(T));
        return r;
    }
}
// float4 / int4 are structs: load them as native 4-vectors, so the load
// stays one global 16-byte load instead of a struct copy
typedef float gpr_f4v __attribute__((ext_vector_type(4)));
typedef int gpr_i4v __attribute__((ext_vector_type(4)));
template <>
GPR_D float4 ldg<float4>(const float4* p) {
    const gpr_f4v v = *(const __attribute__((address_space(1))) gpr_f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
template <>
GPR_D int4 ldg<int4>(const int4* p) {
    const gpr_i4v v = *(const __attribute__((address_space(1))) gpr_i4v*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

constexpr float kPi = 3.14159265358979323846f;    // src/renderer/fwd.h:19
constexpr float kInvPi = 0.31830988618379067154f; // src/renderer/fwd.h:20

struct v3 {
    float x, y, z;
};
GPR_HD v3 V(float x, float y, float z) { return v3{x, y, z}; }
GPR_HD v3 operator+(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
GPR_HD v3 operator-(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
GPR_HD v3 operator-(v3 a) { return V(-a.x, -a.y, -a.z); }
GPR_HD v3 operator*(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
GPR_HD v3 operator*(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
GPR_HD v3 operator/(v3 a, v3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
GPR_HD v3 operator/(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
GPR_HD float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GPR_HD v3 cross3(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
GPR_HD float fmin_h(float a, float b) { return a < b ? a : b; }
GPR_HD float fmax_h(float a, float b) { return a > b ? a : b; }
GPR_HD v3 vmin3(v3 a, v3 b) { return V(fmin_h(a.x, b.x), fmin_h(a.y, b.y), fmin_h(a.z, b.z)); }
GPR_HD v3 vmax3(v3 a, v3 b) { return V(fmax_h(a.x, b.x), fmax_h(a.y, b.y), fmax_h(a.z, b.z)); }
GPR_HD v3 norm3(v3 v) {
    float inv = 1.0f / sqrtf(dot3(v, v));
    return v * inv;
}
GPR_HD float len3(v3 v) { return sqrtf(dot3(v, v)); }
GPR_HD v3 load3(const float* p) { return V(p[0], p[1], p[2]); }

// RGB spectrum (src/renderer/spectrum.h:4-54)
struct c3 {
    float r, g, b;
};
GPR_HD c3 C(float c) { return c3{c, c, c}; }
GPR_HD c3 C(float r, float g, float b) { return c3{r, g, b}; }
GPR_HD c3 operator+(c3 a, c3 b) { return C(a.r + b.r, a.g + b.g, a.b + b.b); }
GPR_HD c3 operator-(c3 a, c3 b) { return C(a.r - b.r, a.g - b.g, a.b - b.b); }
GPR_HD c3 operator*(c3 a, c3 b) { return C(a.r * b.r, a.g * b.g, a.b * b.b); }
GPR_HD c3 operator*(c3 a, float s) { return C(a.r * s, a.g * s, a.b * s); }
GPR_HD c3 operator/(c3 a, c3 b) { return C(a.r / b.r, a.g / b.g, a.b / b.b); }
GPR_HD c3 operator/(c3 a, float s) { return C(a.r / s, a.g / s, a.b / s); }
GPR_HD c3 csqrt(c3 a) { return C(sqrtf(a.r), sqrtf(a.g), sqrtf(a.b)); }
GPR_HD float cmax(c3 a) { return fmax_h(a.r, fmax_h(a.g, a.b)); }
GPR_HD bool black(c3 a) { return a.r == 0 && a.g == 0 && a.b == 0; }
GPR_HD c3 loadc(const float* p) { return C(p[0], p[1], p[2]); }

// sin / cos of a float: quadrant reduction and Taylor polynomials evaluated
// in double, rounded once.  Bit-identical between host and device because
// only IEEE double add/mul/floor are used (no contraction).
GPR_HD void sincos_d(float x, double* s, double* c) {
    const double xd = (double)x;
    const double q = floor(xd * 0.6366197723675814 + 0.5);
    double r = xd - q * 1.5707963267948966;
    r = r - q * 6.123233995736766e-17;
    const double r2 = r * r;
    // sin r = r + r^3 P(r^2), cos r = 1 + r^2 Q(r^2): Taylor coefficients to r^15 / r^14
    double ps = -1.0 / 1307674368000.0;
    ps = 1.0 / 6227020800.0 + r2 * ps;
    ps = -1.0 / 39916800.0 + r2 * ps;
    ps = 1.0 / 362880.0 + r2 * ps;
    ps = -1.0 / 5040.0 + r2 * ps;
    ps = 1.0 / 120.0 + r2 * ps;
    ps = -1.0 / 6.0 + r2 * ps;
    const double sr = r + r * r2 * ps;
    double pc = -1.0 / 87178291200.0;
    pc = 1.0 / 479001600.0 + r2 * pc;
    pc = -1.0 / 3628800.0 + r2 * pc;
    pc = 1.0 / 40320.0 + r2 * pc;
    pc = -1.0 / 720.0 + r2 * pc;
    pc = 1.0 / 24.0 + r2 * pc;
    pc = -0.5 + r2 * pc;
    const double cr = 1.0 + r2 * pc;
    const int qq = (int)((long long)q & 3);
    if (qq == 0) { *s = sr; *c = cr; }
    else if (qq == 1) { *s = cr; *c = -sr; }
    else if (qq == 2) { *s = -sr; *c = -cr; }
    else { *s = -cr; *c = sr; }
}
GPR_HD float sin_f(float x) { double s, c; sincos_d(x, &s, &c); return (float)s; }
GPR_HD float cos_f(float x) { double s, c; sincos_d(x, &s, &c); return (float)c; }

// GammaCorrect (spectrum.h:63-72) and the byte conversion of FrameBuffer::output
// and filter_kernel (framebuffer.cpp:18-22, pathtracer_kernel.h:805-810):
// (uint8)clamp(255 * gamma + 0.5, 0, 255), clamp mapping NaN to 255.  powf is
// taken as a double pow rounded once (within one ulp of powf), so the host
// writer and the device resolve give the same bytes.
GPR_HD float gamma_correct(float v) {
    if (v <= 0.0031308f) return 12.92f * v;
    const float p = (float)pow((double)v, (double)(1.f / 2.4f));
    return 1.055f * p - 0.055f;
}
GPR_HD uint32_t srgb_byte(float v) {
    const float x = 255.f * gamma_correct(v) + 0.5f;
    return (uint32_t)fmaxf(0.f, fminf(x, 255.f));
}

// ---------------------------------------------------------------- RNG ---
// InitRandom (src/renderer/sampling.h:31-46): 16 TEA rounds, fixed key.
GPR_HD uint32_t tea16(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
// NextRandom (src/renderer/sampling.h:48-54)
GPR_HD float lcg_next(uint32_t& seed) {
    seed = 1664525u * seed + 1013904223u;
    return float(seed & 0x00FFFFFFu) / float(0x01000000);
}
// PCG32 XSH-RR; the stream (increment) of a camera path is derived from
// (sample, pixel) and is recomputed wherever the state is advanced.
GPR_HD uint32_t pcg_step(uint64_t& st, uint64_t inc) {
    const uint64_t old = st;
    st = old * 6364136223846793005ull + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
GPR_HD uint64_t pcg_inc(uint32_t sample, uint32_t pixel) {
    return ((((uint64_t)sample << 32) | (uint64_t)pixel) << 1u) | 1u;
}
GPR_HD uint64_t pcg_init(uint64_t seed, uint64_t inc) {
    uint64_t st = 0;
    pcg_step(st, inc);
    st += seed;
    pcg_step(st, inc);
    return st;
}
GPR_HD float pcg_unit(uint64_t& st, uint64_t inc) {
    return float(pcg_step(st, inc) >> 8) * (1.0f / 16777216.0f);
}

}  // namespace gpr
