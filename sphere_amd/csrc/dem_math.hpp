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

// Correctly rounded FP32 square root.  The candidate comes from the FP64
// square root rounded to FP32; it is then checked against the exact
// midpoints to its FP32 neighbours.  For y with a 24-bit significand the
// midpoints carry at most 26 bits, so their squares (<= 52 bits) are exact
// in FP64, and a square root never ties.  The result does not depend on the
// accuracy of the hardware FP64 square root beyond one FP32 ulp.
__device__ __forceinline__ float sqrt_rn_f32(float x) {
    float y = (float)__builtin_sqrt((double)x);
    if (x > 0.0f && x < __builtin_huge_valf()) {
        const uint32_t b = __float_as_uint(y);
        const float yu = __uint_as_float(b + 1u);
        const float yd = __uint_as_float(b - 1u);
        const double xd = (double)x;
        const double mu = 0.5 * ((double)y + (double)yu);
        const double md = 0.5 * ((double)y + (double)yd);
        if (xd > mu * mu) {
            y = yu;
        } else if (xd < md * md) {
            y = yd;
        }
    }
    return y;
}

// The reference's length(): FP32 root of the FP64 dot, widened back.
__device__ __forceinline__ double sqrt32(double v) { return (double)sqrt_rn_f32((float)v); }
__device__ __forceinline__ double length(d3 v) { return sqrt32(dot(v, v)); }

// Correctly rounded FP64 square root (LLVM lowers llvm.sqrt.f64 for AMDGPU
// to a correctly rounded sequence when no fast-math flags are present).
__device__ __forceinline__ double sqrt_rn_f64(double v) { return __builtin_sqrt(v); }

}  // namespace sphx

#endif
