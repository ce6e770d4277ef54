// Device-side vector arithmetic with the reference's exact semantics.
//
// Component-wise operators, left-to-right dot, and cross follow
// src/vector_arithmetic.h:36-415.  length() narrows the FP64 dot product to
// FP32 and takes a correctly rounded FP32 square root (the reference passes a
// double to sqrtf, src/vector_arithmetic.h:317-324).  Everything here must
// be compiled with -ffp-contract=off so that no multiply-add is fused.
#ifndef SPHERE_AMD_DEM_MATH_HPP_
#define SPHERE_AMD_DEM_MATH_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sphx {

struct d3 {
    double x, y, z;
};

__device__ __forceinline__ d3 mk3(double x, double y, double z) {
    d3 r;
    r.x = x; r.y = y; r.z = z;
    return r;
}
__device__ __forceinline__ d3 xyz(const double4& a) { return mk3(a.x, a.y, a.z); }

__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
// Scalar broadcast add (src/vector_arithmetic.h:61-64, the P12 wall velocity).
__device__ __forceinline__ d3 operator+(d3 a, double b) { return mk3(a.x + b, a.y + b, a.z + b); }
__device__ __forceinline__ d3 operator*(d3 a, double b) { return mk3(a.x * b, a.y * b, a.z * b); }
__device__ __forceinline__ d3 operator*(double b, d3 a) { return mk3(b * a.x, b * a.y, b * a.z); }
__device__ __forceinline__ d3 operator/(d3 a, double b) { return mk3(a.x / b, a.y / b, a.z / b); }
__device__ __forceinline__ d3 operator-(d3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ void operator+=(d3& a, d3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
__device__ __forceinline__ void operator*=(d3& a, double b) { a.x *= b; a.y *= b; a.z *= b; }

__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Correctly rounded FP32 square root.  A candidate y is checked against the
// exact midpoints to its FP32 neighbours and moved by one ulp when the root
// lies beyond one: for y with a 24-bit significand the midpoints carry at
// most 26 bits, so their squares (<= 52 bits) are exact in FP64, and a
// square root never ties.  This gives the correctly rounded root whenever
// the candidate is within one FP32 ulp of it.
__device__ __forceinline__ float sqrt_rn_fix(float x, float y) {
    const uint32_t b = __float_as_uint(y);
    const float yu = __uint_as_float(b + 1u);
    const float yd = __uint_as_float(b - 1u);
    const double xd = (double)x;
    const double mu = 0.5 * ((double)y + (double)yu);
    const double md = 0.5 * ((double)y + (double)yd);
    if (xd > mu * mu) return yu;
    if (xd < md * md) return yd;
    return y;
}

// The candidate: the hardware FP64 reciprocal square root r0 (relative
// error below 2^-22) gives y0 = x r0, and one Newton step on the residual,
// y1 = y0 + (x - y0^2) r0 / 2, has a relative error below 2^-43, so y1
// rounded to FP32 is within one ulp of the correctly rounded root.
// Zero, negative, infinite and NaN inputs follow IEEE sqrt.  Checked on
// the device against the correctly rounded FP64 root rounded to FP32 (which
// is correctly rounded: 53 >= 2 * 24 + 2) for every positive FP32 value by
// sph_selftest_sqrt32.
__device__ __forceinline__ float sqrt_rn_f32(float x) {
    if (!(x > 0.0f && x < __builtin_huge_valf()))
        return x == 0.0f ? x : (x < 0.0f ? __builtin_nanf("") : x);
    const double xd = (double)x;
    const double r0 = __builtin_amdgcn_rsq(xd);
    const double y0 = xd * r0;
    const double e = __builtin_fma(-y0, y0, xd);
    const double y1 = __builtin_fma(e, 0.5 * r0, y0);
    return sqrt_rn_fix(x, (float)y1);
}

// The reference for that check: correctly rounded FP64 root (llvm.sqrt.f64
// without fast-math flags), rounded to FP32.
__device__ __forceinline__ float sqrt_rn_f32_ref(float x) {
    return (float)__builtin_sqrt((double)x);
}

__device__ __forceinline__ double sqrt32(double v) { return (double)sqrt_rn_f32((float)v); }
__device__ __forceinline__ double length(d3 v) { return sqrt32(dot(v, v)); }

// Correctly rounded FP64 square root (LLVM lowers llvm.sqrt.f64 for AMDGPU
// to a correctly rounded sequence when no fast-math flags are present).
__device__ __forceinline__ double sqrt_rn_f64(double v) { return __builtin_sqrt(v); }

}  // namespace sphx

#endif
