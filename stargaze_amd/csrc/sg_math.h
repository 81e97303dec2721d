// Small vector algebra shared by the HIP kernels (device) and the host
// library (host). Plain structs keep the generated code free of any
// vector-type overload ambiguity.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Free functions have internal linkage: a kernel translation unit may be
// compiled with or without the landscape (SG_LANDSCAPE), and each keeps its
// own definitions. Member functions use SG_MEMBER.
#define SG_FN static __host__ __device__ __forceinline__
#define SG_MEMBER __host__ __device__ __forceinline__

// 0 compiles the landscape (ground scenery, terrain, sky cache) out of a
// kernel translation unit: its kernels serve frames without a ground scene.
#ifndef SG_LANDSCAPE
#define SG_LANDSCAPE 1
#endif
// 0 compiles the solid-body surface styles (2-6) out: their code raises the
// register pressure of the shading kernels even for frames that never use
// them, so the kernels without the landscape are built both ways.
#ifndef SG_STYLES
#define SG_STYLES 1
#endif

// Event hooks of the work model: host builds of the test harness define
// SG_COUNT_EVENTS and sg_count_event to count traversal and shading events;
// everywhere else SG_COUNT compiles to nothing.
enum SgEvent : uint32_t {
    SG_EV_SPHERE, SG_EV_RING, SG_EV_LANDSCAPE, SG_EV_TERRAIN_BOX, SG_EV_TRIANGLE, SG_EV_BVH_BOX, SG_EV_PRIM,
    SG_EV_SHADOW, SG_EV_DENSITY_STEP, SG_EV_VIEW_CELL, SG_EV_LIGHT_SAMPLE, SG_EV_COUNT
};
#ifdef SG_COUNT_EVENTS
void sg_count_event(uint32_t event);
#define SG_COUNT(e) sg_count_event(e)
#else
#define SG_COUNT(e) ((void)0)
#endif

template <class T> struct V3 { T x, y, z; };
typedef V3<float> f3;
typedef V3<double> d3;

template <class T> SG_FN V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
SG_FN f3 f3s(float s) { return f3{s, s, s}; }
template <class T> SG_FN V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> SG_FN V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> SG_FN V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> SG_FN V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> SG_FN V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> SG_FN V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> SG_FN V3<T> operator*(T s, V3<T> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> SG_FN V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> SG_FN V3<T> &operator+=(V3<T> &a, V3<T> b) { a = a + b; return a; }
template <class T> SG_FN V3<T> &operator*=(V3<T> &a, V3<T> b) { a = a * b; return a; }
template <class T> SG_FN V3<T> &operator*=(V3<T> &a, T s) { a = a * s; return a; }
template <class T> SG_FN V3<T> &operator/=(V3<T> &a, T s) { a = a / s; return a; }
template <class T> SG_FN T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> SG_FN V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
SG_FN uint32_t sg_f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
SG_FN float sg_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
SG_FN uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
SG_FN uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
SG_FN int imin(int a, int b) { return a < b ? a : b; }
SG_FN int imax(int a, int b) { return a > b ? a : b; }
SG_FN float sqrtT(float x) { return sqrtf(x); }
SG_FN double sqrtT(double x) { return sqrt(x); }
SG_FN float length(f3 a) { return sqrtf(dot(a, a)); }
SG_FN double length(d3 a) { return sqrt(dot(a, a)); }
SG_FN f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
SG_FN d3 normalize(d3 a) { return a * (1.0 / sqrt(dot(a, a))); }
SG_FN f3 tof(d3 a) { return {(float)a.x, (float)a.y, (float)a.z}; }
SG_FN d3 tod(f3 a) { return {(double)a.x, (double)a.y, (double)a.z}; }
SG_FN f3 xyz(float4 a) { return {a.x, a.y, a.z}; }
SG_FN f3 fabs3(f3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
SG_FN f3 fmin3(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
SG_FN f3 fmax3(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
// e^x to the accuracy WGSL's exp promises (3 + 2|x| ULP), which the
// reference's shaders are written against: on the GPU one v_exp_f32 of
// x log2(e) (about 1 + 1.5|x| ULP), instead of the library's extended-
// precision reduction. Host builds use expf.
SG_FN float sg_exp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(x);
#else
    return expf(x);
#endif
}
SG_FN f3 exp3(f3 a) { return {sg_exp(a.x), sg_exp(a.y), sg_exp(a.z)}; }
// a / b correctly rounded whatever the FP32 division accuracy of the build:
// the FP64 quotient rounded to FP32 (double rounding is innocuous for
// division, 53 >= 2 * 24 + 2 bits). For quotients that must equal the
// host's bit for bit.
SG_FN float div_rn(float a, float b) { return (float)((double)a / (double)b); }
SG_FN float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
SG_FN float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
SG_FN f3 clamp3(f3 a, float lo, float hi) { return {clampf(a.x, lo, hi), clampf(a.y, lo, hi), clampf(a.z, lo, hi)}; }
// WGSL mix/smoothstep/fract semantics.
// Exact IEEE evaluation, no implicit multiply-add contraction, for the noise
// hashes: they take fract() of sums of products near 50, so one contraction
// moves a hash by about 1e-4, and seed offsets scale hashes by 512. Placed
// first in a function body; explicit fmaf calls stay fused.
#define SG_EXACT _Pragma("clang fp contract(off)")
SG_FN float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
SG_FN f3 mix3(f3 a, f3 b, float t) { return a * (1.0f - t) + b * t; }
SG_FN float smoothstepf(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
SG_FN float fractf(float x) { return x - floorf(x); }
SG_FN float luminance(f3 c) { return dot(c, f3{0.2126f, 0.7152f, 0.0722f}); }
SG_FN float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
