#pragma once

// Small vector helpers shared by the oracle's translation units (test
// infrastructure only). Every operation is written out so that the FP32
// twin evaluates in the device kernels' order; fma3 is the only fused form.

#include <algorithm>
#include <cmath>

#include "numi_cloth_hip.h"

namespace oracle_math {

template <typename R>
struct V {
    R x = 0, y = 0, z = 0;
};
template <typename R> V<R> operator+(V<R> a, V<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> V<R> operator-(V<R> a, V<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> V<R> operator-(V<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> V<R> operator*(V<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> V<R> operator/(V<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename R> R dot(V<R> a, V<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> V<R> cross(V<R> a, V<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> R norm(V<R> a) { return std::sqrt(dot(a, a)); }
template <typename R> V<R> fma3(V<R> a, R s, V<R> b) {
    return {std::fma(a.x, s, b.x), std::fma(a.y, s, b.y), std::fma(a.z, s, b.z)};
}
template <typename R> V<R> mix3(V<R> a, V<R> b, R t) {
    return {a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t};
}
template <typename R> V<R> v3(const mr_float4& v) { return {R(v.x), R(v.y), R(v.z)}; }
template <typename R> R clampR(R v, R lo, R hi) { return std::min(std::max(v, lo), hi); }

template <typename R>
struct Quat {
    R x, y, z, w;
};
template <typename R> Quat<R> q4(const mr_float4& v) { return {R(v.x), R(v.y), R(v.z), R(v.w)}; }
// v + 2w(q x v) + q x (2 q x v)
template <typename R> V<R> rotate(Quat<R> q, V<R> v) {
    const V<R> u{q.x, q.y, q.z};
    const V<R> t = cross(u, v) * R(2);
    return v + t * q.w + cross(u, t);
}


}  // namespace oracle_math

using namespace oracle_math;  // NOLINT: the oracle sources use the short names
