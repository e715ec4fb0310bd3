#pragma once

// Small float3 helpers for gfx950 device code. Everything is written out
// component-wise so the arithmetic order is explicit; the library is built
// with -ffp-contract=off and uses fmaf only where the algorithm calls for a
// fused multiply-add.

#include <hip/hip_runtime.h>

#include "numi/cloth_bag_gpu.h"

namespace numi_dev {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 xyz(const float4& v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ V3 xyz(const mr_float4& v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ bool finite3(V3 a) {
    return isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
}
__device__ __forceinline__ bool finite4(const mr_float4& a) {
    return isfinite(a.x) && isfinite(a.y) && isfinite(a.z) && isfinite(a.w);
}
__device__ __forceinline__ V3 fma3(V3 a, float s, V3 b) {
    return {fmaf(a.x, s, b.x), fmaf(a.y, s, b.y), fmaf(a.z, s, b.z)};
}
__device__ __forceinline__ float4 withXyz(float4 base, V3 p) {
    base.x = p.x;
    base.y = p.y;
    base.z = p.z;
    return base;
}
__device__ __forceinline__ V3 mix3(V3 a, V3 b, float t) {
    // a + (b - a) t, the form used for every interpolation in the solver.
    return {a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t};
}
__device__ __forceinline__ V3 safeNormalized(V3 v) {
    const float m = length(v);
    return m > 1.0e-14f ? v / m : v3(1.0f, 0.0f, 0.0f);
}
// v + 2w (q x v) + q x (2 q x v)
__device__ __forceinline__ V3 rotate(const mr_float4& q, V3 v) {
    const V3 qv = v3(q.x, q.y, q.z);
    const V3 t = cross(qv, v) * 2.0f;
    return v + t * q.w + cross(qv, t);
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) {
    return fminf(fmaxf(v, lo), hi);
}

// Kinematic drive of a pinned knot at time t: m = (amplitude, omega, phase),
// offset A (sin(omega t + phase) - sin(phase)) along z, so a driven knot
// starts where it was created. Evaluated in double, rounded to float.
__device__ __forceinline__ float driveOffset(const float4& m, double t) {
    return static_cast<float>(static_cast<double>(m.x) *
                              (sin(static_cast<double>(m.y) * t + static_cast<double>(m.z)) - sin(static_cast<double>(m.z))));
}

__device__ __forceinline__ void recordFailure(uint32_t* failure, uint32_t bit) {
    atomicOr(failure, bit);
}

}  // namespace numi_dev
