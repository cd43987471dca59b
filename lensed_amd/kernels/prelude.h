// Plugin prelude: the OpenCL-C vocabulary that object plugins (objects/*.cl)
// are written in, restated as C++ inside namespace `lcl`.
//
// The plugin dialect is the one Lensed defines for its objects
// (kernel/constants.cl, kernel/object.cl, docs/create.md:12-153): vector
// types float2/float4, the 2x2 matrix type mat22 with mv22(), the object and
// parameter-type enums, `struct param` and the bound macros.  The names and
// meanings are kept because plugins are written against them.
//
// The same header compiles in three passes:
//   * hipRTC (device program, __HIPCC_RTC__),
//   * hipcc (__HIP__),
//   * a plain host C++ compiler (used to test translated plugins on the CPU).
// For the first two, every function is made __host__ __device__ by the clang
// pragma, so plugin code needs no annotations.
#pragma once

#if defined(__HIPCC_RTC__) || defined(__HIP__)
#define LCL_GPU 1
#pragma clang force_cuda_host_device begin
#else
#include <cmath>
#include <cfloat>
#endif

#ifndef FLT_MAX
#define FLT_MAX 3.402823466e+38F
#endif
#ifndef FLT_MIN
#define FLT_MIN 1.175494351e-38F
#endif
#ifndef FLT_EPSILON
#define FLT_EPSILON 1.192092896e-07F
#endif
#ifndef HUGE_VALF
#define HUGE_VALF __builtin_huge_valf()
#endif
#ifndef INFINITY
#define INFINITY __builtin_huge_valf()
#endif
#ifndef MAXFLOAT
#define MAXFLOAT FLT_MAX
#endif

namespace lcl {

typedef unsigned int uint;
typedef unsigned long ulong;
typedef unsigned short ushort;
typedef unsigned char uchar;

// ---------------------------------------------------------------- vectors
// OpenCL alignment: float2 on 8 bytes, float4 on 16 bytes, so that object
// data structs have the same layout as an OpenCL compiler gives them.
struct alignas(8) float2 {
    union {
        struct { float x, y; };
        struct { float s0, s1; };
        float s[2];
    };
    float2() = default;
    inline float2(float v) { x = v; y = v; }
    inline float2(float a, float b) { x = a; y = b; }
};

struct alignas(16) float4 {
    union {
        struct { float x, y, z, w; };
        struct { float s0, s1, s2, s3; };
        float s[4];
    };
    float4() = default;
    inline float2 lo() const { return float2(x, y); }
    inline float2 hi() const { return float2(z, w); }
    inline float4(float v) { x = v; y = v; z = v; w = v; }
    inline float4(float a, float b, float c, float d) { x = a; y = b; z = c; w = d; }
    inline float4(float2 a, float2 b) { x = a.x; y = a.y; z = b.x; w = b.y; }
    inline float4(float2 a, float c, float d) { x = a.x; y = a.y; z = c; w = d; }
    inline float4(float a, float b, float2 c) { x = a; y = b; z = c.x; w = c.y; }
};

#define LCL_VEC2_OP(OP)                                                          \
    inline float2 operator OP(float2 a, float2 b) { return float2(a.x OP b.x, a.y OP b.y); } \
    inline float2 operator OP(float2 a, float b) { return float2(a.x OP b, a.y OP b); }      \
    inline float2 operator OP(float a, float2 b) { return float2(a OP b.x, a OP b.y); }
#define LCL_VEC4_OP(OP)                                                          \
    inline float4 operator OP(float4 a, float4 b) { return float4(a.x OP b.x, a.y OP b.y, a.z OP b.z, a.w OP b.w); } \
    inline float4 operator OP(float4 a, float b) { return float4(a.x OP b, a.y OP b, a.z OP b, a.w OP b); }          \
    inline float4 operator OP(float a, float4 b) { return float4(a OP b.x, a OP b.y, a OP b.z, a OP b.w); }
LCL_VEC2_OP(+) LCL_VEC2_OP(-) LCL_VEC2_OP(*) LCL_VEC2_OP(/)
LCL_VEC4_OP(+) LCL_VEC4_OP(-) LCL_VEC4_OP(*) LCL_VEC4_OP(/)
#undef LCL_VEC2_OP
#undef LCL_VEC4_OP

inline float2 operator-(float2 a) { return float2(-a.x, -a.y); }
inline float4 operator-(float4 a) { return float4(-a.x, -a.y, -a.z, -a.w); }
inline float2 operator+(float2 a) { return a; }
inline float4 operator+(float4 a) { return a; }

#define LCL_COMPOUND(T, OP)                                              \
    inline T& operator OP##=(T& a, T b) { a = a OP b; return a; }        \
    inline T& operator OP##=(T& a, float b) { a = a OP b; return a; }
LCL_COMPOUND(float2, +) LCL_COMPOUND(float2, -) LCL_COMPOUND(float2, *) LCL_COMPOUND(float2, /)
LCL_COMPOUND(float4, +) LCL_COMPOUND(float4, -) LCL_COMPOUND(float4, *) LCL_COMPOUND(float4, /)
#undef LCL_COMPOUND

// ---------------------------------------------------------------- scalar maths
// OpenCL built-in names.  The device passes compile with fast-math style
// flags (approximate functions, flush-to-zero), the counterpart of the
// -cl-fast-relaxed-math -cl-denorms-are-zero build of the OpenCL program.
#define LCL_F1(NAME, IMPL) inline float NAME(float a) { return IMPL(a); }
LCL_F1(sqrt, ::sqrtf)
LCL_F1(exp, ::expf)
LCL_F1(exp2, ::exp2f)
LCL_F1(log, ::logf)
LCL_F1(log2, ::log2f)
LCL_F1(log10, ::log10f)
LCL_F1(sin, ::sinf)
LCL_F1(cos, ::cosf)
LCL_F1(tan, ::tanf)
LCL_F1(asin, ::asinf)
LCL_F1(acos, ::acosf)
LCL_F1(atan, ::atanf)
LCL_F1(sinh, ::sinhf)
LCL_F1(cosh, ::coshf)
LCL_F1(tanh, ::tanhf)
LCL_F1(asinh, ::asinhf)
LCL_F1(acosh, ::acoshf)
LCL_F1(atanh, ::atanhf)
LCL_F1(tgamma, ::tgammaf)
LCL_F1(lgamma, ::lgammaf)
LCL_F1(erf, ::erff)
LCL_F1(erfc, ::erfcf)
LCL_F1(fabs, ::fabsf)
LCL_F1(floor, ::floorf)
LCL_F1(ceil, ::ceilf)
LCL_F1(round, ::roundf)
LCL_F1(trunc, ::truncf)
LCL_F1(cbrt, ::cbrtf)
LCL_F1(expm1, ::expm1f)
LCL_F1(log1p, ::log1pf)
#undef LCL_F1
inline float exp10(float a) { return ::exp2f(a * 3.3219280948873623479f); }
inline float rsqrt(float a) { return 1.0f / ::sqrtf(a); }
inline float recip(float a) { return 1.0f / a; }
inline float atan2(float y, float x) { return ::atan2f(y, x); }
inline float pow(float a, float b) { return ::powf(a, b); }
inline float powr(float a, float b) { return ::exp2f(b * ::log2f(a)); }
inline float pown(float a, int n) { return ::powf(a, (float)n); }
inline float rootn(float a, int n) { return ::powf(a, 1.0f / (float)n); }
inline float hypot(float a, float b) { return ::sqrtf(a * a + b * b); }
inline float fmod(float a, float b) { return ::fmodf(a, b); }
inline float copysign(float a, float b) { return ::copysignf(a, b); }
inline float fma(float a, float b, float c) { return ::fmaf(a, b, c); }
inline float mad(float a, float b, float c) { return a * b + c; }
inline float fmin(float a, float b) { return ::fminf(a, b); }
inline float fmax(float a, float b) { return ::fmaxf(a, b); }
inline float min(float a, float b) { return ::fminf(a, b); }
inline float max(float a, float b) { return ::fmaxf(a, b); }
template <class T> inline T min(T a, T b) { return b < a ? b : a; }
template <class T> inline T max(T a, T b) { return a < b ? b : a; }
inline float clamp(float v, float lo, float hi) { return ::fminf(::fmaxf(v, lo), hi); }
inline float mix(float a, float b, float t) { return a + (b - a) * t; }
inline float step(float edge, float v) { return v < edge ? 0.0f : 1.0f; }
inline float sign(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
inline float degrees(float r) { return r * 57.295779513082320876798f; }
inline float radians(float d) { return d * 0.017453292519943295769237f; }
// OpenCL sincos returns the sine and stores the cosine.
inline float sincos(float a, float* c) { *c = ::cosf(a); return ::sinf(a); }

// native_ and half_ variants carry no accuracy promise in OpenCL
#define native_sqrt sqrt
#define native_rsqrt rsqrt
#define native_exp exp
#define native_exp2 exp2
#define native_exp10 exp10
#define native_log log
#define native_log2 log2
#define native_log10 log10
#define native_sin sin
#define native_cos cos
#define native_tan tan
#define native_powr powr
#define native_recip recip
#define native_divide(a, b) ((a) / (b))
#define half_sqrt sqrt
#define half_rsqrt rsqrt
#define half_exp exp
#define half_log log
#define half_sin sin
#define half_cos cos
#define half_powr powr
#define half_recip recip

// component-wise maths on float2
#define LCL_V1(NAME) inline float2 NAME(float2 a) { return float2(NAME(a.x), NAME(a.y)); }
LCL_V1(sqrt) LCL_V1(exp) LCL_V1(log) LCL_V1(sin) LCL_V1(cos) LCL_V1(fabs)
LCL_V1(floor) LCL_V1(ceil) LCL_V1(atan) LCL_V1(atanh) LCL_V1(rsqrt) LCL_V1(exp2) LCL_V1(log2)
#undef LCL_V1
inline float2 min(float2 a, float2 b) { return float2(min(a.x, b.x), min(a.y, b.y)); }
inline float2 max(float2 a, float2 b) { return float2(max(a.x, b.x), max(a.y, b.y)); }
inline float2 min(float2 a, float b) { return float2(min(a.x, b), min(a.y, b)); }
inline float2 max(float2 a, float b) { return float2(max(a.x, b), max(a.y, b)); }

// ---------------------------------------------------------------- geometry
inline float dot(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }
inline float dot(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
inline float dot(float a, float b) { return a * b; }

// The ray-shooting guard of the generated compute(): a deflection is used
// when it is finite and small enough, else the ray goes to (1e10, 1e10).
// The test |a_x| + |a_y| < 2^64 rejects exactly the non-finite deflections
// that dot(a, a) < HUGE_VALF rejects (any inf or NaN component); the two
// differ only for finite |a| between 1.3e19 and 1.8e19 pixels, where either
// ray ends far outside every source.
inline bool lensed_deflection_ok(float2 a) { return fabs(a.x) + fabs(a.y) < 1.8446744e19f; }
inline float length(float2 a) { return ::sqrtf(a.x * a.x + a.y * a.y); }
inline float length(float4 a) { return ::sqrtf(dot(a, a)); }
inline float fast_length(float2 a) { return length(a); }
inline float distance(float2 a, float2 b) { return length(a - b); }
inline float2 normalize(float2 a) { return a * (1.0f / ::sqrtf(a.x * a.x + a.y * a.y)); }
inline float2 fast_normalize(float2 a) { return normalize(a); }

// ---------------------------------------------------------------- constants
#define PI      3.14159265358979323846264338327950288f
#define PI_HALF 1.57079632679489661923132169163975144f
#define DEG2RAD 0.01745329251994329576923690768488612713f
#define LOG_10  2.30258509299404568401799145468436421f
#define LOG_PI  1.14472988584940017414342735135305871f
#define LOG_2PI 1.83787706640934548356065947281123528f
#ifndef M_PI_F
#define M_PI_F PI
#endif
#if defined(IMAGE_WIDTH) && defined(IMAGE_HEIGHT)
#define IMAGE_CENTER (float2(0.5f * (IMAGE_WIDTH + 1), 0.5f * (IMAGE_HEIGHT + 1)))
#endif

// ---------------------------------------------------------------- matrices
typedef float4 mat22;

// (m.x m.y; m.z m.w) times the column vector v
inline float2 mv22(mat22 m, float2 v) { return float2(m.x * v.x + m.y * v.y, m.z * v.x + m.w * v.y); }

// ---------------------------------------------------------------- objects
enum { LENS = 'L', SOURCE = 'S', FOREGROUND = 'F' };

enum { PARAMETER = 0, POSITION_X, POSITION_Y, RADIUS, MAGNITUDE, AXIS_RATIO, POS_ANGLE };

#define UNBOUNDED {0, 0}
#define POS_BOUND {0, +FLT_MAX}
#define NEG_BOUND {-FLT_MAX, 0}

struct param {
    char name[16];
    int type;
    float bounds[2];
    float defval;
};

}  // namespace lcl

#ifdef LCL_GPU
#pragma clang force_cuda_host_device end
#endif
