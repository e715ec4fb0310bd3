#pragma once

// Core XPBD kernels of the cloth substep on gfx950: configuration check,
// predict, distance / crossing-knot / chord-bend / grip projections, ground
// and static-sphere contact, the unilateral strain limiter and finalize with
// cloth/ground Coulomb friction.
//
// Device layout (structure of arrays, one buffer per field, 16-byte records
// where the field is a vector so every access is one coalesced dwordx4):
//   pos[P]  float4  position xyz, inverse mass w
//   prev[P] float4  substep-start position xyz, mass w
//   vel[P]  float4  velocity xyz, predicted vz (during the substep) or
//                   accumulated ground normal impulse (after finalize) w
//   distances: ij[D] uint2, restCompliance[D] float2, lambda[D] float,
//              maxLength[D] float = rest * (1 + strain limit)
//   bends:     ends[B] uint2 (first, third), chordCompliance[B] float2,
//              lambda[B] float
//   knots:     quad[K] uint4, cosineCompliance[K] float2, lambda[K] float
// Colour batches are contiguous index ranges, so one launch per colour reads
// each record exactly once.
//
// Every kernel gates on ok[s], the per-substep result of validConfig
// (reference src/metal/ClothBag.metal:16-55), which the reference
// re-evaluates in every thread of every dispatch.

#include "device_math.hpp"

namespace numi_dev {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kMaxWaves = 16;

// --------------------------------------------------------------- predict --

struct BeginArgs {
    const NumiClothBagGPUConfig* cfg;
    const uint32_t* ok;
    float4* pos;
    float4* prev;
    float4* vel;
    const float4* anchor;       // kinematic drive anchors, or null
    const float4* motion;       // per knot (amplitude, omega, phase, damping), or null
    float4 drive;               // (amplitude, omega, phase, -) when motion is null
    const double* time;         // time at the end of the substep
    float dampingDt;            // timestep the per-knot damping factors use
    float* distanceLambda;
    float* bendLambda;
    float* knotLambda;
    NumiClothBagGPUGrip* grips;
    uint32_t particles, distances, bends, knots, grips_;
    float dampingFactor;        // 1 / (1 + c dt), 1 when undamped
    uint32_t advance;           // 1: also x += v dt (nothing runs in between)
    uint32_t* failure;
};

// begin_substep (reference src/metal/ClothBag.metal:491-575) for knots and
// the constraint multipliers, optionally fused with advance_positions
// (:907-938) when no grip re-attachment or air load runs between them.
__global__ void k_begin(BeginArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (*a.ok == 0u) return;
    if (i < a.particles) {
        float4 p = a.pos[i];
        float4 v = a.vel[i];
        float4 q = a.prev[i];
        const V3 x = xyz(p);
        const V3 u = xyz(v);
        if (!finite3(x) || !finite3(u) || !isfinite(p.w) || p.w < 0.0f) {
            recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        } else {
            q = withXyz(q, x);
            v.w = 0.0f;
            const float dt = a.cfg->gravityAndTimestep.w;
            if (p.w > 0.0f) {
                // Damping acts on the carried velocity only, so a resting
                // knot still receives the full gravity increment.
                float factor = a.dampingFactor;
                if (a.motion != nullptr) {
                    const float c = a.motion[i].w;
                    factor = c > 0.0f ? 1.0f / (1.0f + c * a.dampingDt) : 1.0f;
                }
                const V3 carried = factor != 1.0f ? u * factor : u;
                const V3 nu = fma3(xyz(a.cfg->gravityAndTimestep), dt, carried);
                v = withXyz(v, nu);
                v.w = v.z;
                if (a.advance) p = withXyz(p, fma3(nu, dt, x));
            } else if (a.anchor != nullptr) {
                // Kinematic drive, evaluated in double:
                // z = anchor.z + A (sin(omega t + phase) - sin(phase)).
                const float4 anchor = a.anchor[i];
                const float4 m = a.motion != nullptr ? a.motion[i] : a.drive;
                const float offset = driveOffset(m, *a.time);
                p = withXyz(p, v3(anchor.x, anchor.y, anchor.z + offset));
            }
            a.prev[i] = q;
            a.vel[i] = v;
            a.pos[i] = p;
        }
    }
    if (i < a.distances) a.distanceLambda[i] = 0.0f;
    if (i < a.bends) a.bendLambda[i] = 0.0f;
    if (i < a.knots) a.knotLambda[i] = 0.0f;
    if (i < a.grips_) {
        a.grips[i].lambda.x = 0.0f;
        a.grips[i].lambda.y = 0.0f;
        a.grips[i].lambda.z = 0.0f;
    }
}

// advance_positions (reference src/metal/ClothBag.metal:907-938), cloth part.
__global__ void k_advance(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                          float4* __restrict__ pos, const float4* __restrict__ vel, uint32_t particles) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= particles || *ok == 0u) return;
    float4 p = pos[i];
    if (p.w > 0.0f) {
        const float4 v = vel[i];
        pos[i] = withXyz(p, fma3(xyz(v), cfg->gravityAndTimestep.w, xyz(p)));
    }
}

// ------------------------------------------------------------ distances --

// One XPBD distance projection (reference solve_distance,
// src/metal/ClothBag.metal:1148-1213) on values already loaded: updates a, b
// and lambda in place; returns false when nothing is written.
// invDt2 = 1 / (dt * dt) of the substep: alpha = compliance * invDt2.
// The arithmetic without branches: every value is computed and the result
// selected, failures are or-ed into `fail` (so that several independent
// projections can be interleaved by the compiler).
__device__ __forceinline__ bool projectDistanceCore(float4& a, float4& b, float2 rc, float& lam, float invDt2,
                                                    uint32_t& fail) {
    const V3 d = xyz(b) - xyz(a);
    const float len = length(d);
    const float alpha = rc.y * invDt2;
    const float denom = a.w + b.w + alpha;
    const bool bad = !isfinite(len) || !isfinite(alpha) || !isfinite(denom) || !isfinite(rc.x) || !isfinite(lam);
    const bool go = !bad && len > 1.0e-12f && denom > 0.0f;
    const float value = len - rc.x;
    const float dl = (-value - alpha * lam) / denom;
    const V3 corr = d * (dl / len);
    const float nl = lam + dl;
    const V3 na = xyz(a) - corr * a.w;
    const V3 nb = xyz(b) + corr * b.w;
    const bool badOut = go && (!finite3(na) || !finite3(nb) || !isfinite(nl));
    fail |= bad || badOut ? NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE : 0u;
    const bool apply = go && !badOut;
    if (apply) {
        a = withXyz(a, na);
        b = withXyz(b, nb);
        lam = nl;
    }
    return apply;
}

__device__ __forceinline__ bool projectDistance(float4& a, float4& b, float2 rc, float& lam, float invDt2,
                                                uint32_t* failure) {
    uint32_t fail = 0u;
    const bool apply = projectDistanceCore(a, b, rc, lam, invDt2, fail);
    if (fail != 0u) recordFailure(failure, fail);
    return apply;
}

// solve_distance over one colour batch [first, first + count).
__global__ void k_distance(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                           uint32_t first, uint32_t count, float4* __restrict__ pos,
                           const uint2* __restrict__ ij, const float2* __restrict__ restCompliance,
                           float* __restrict__ lambda, uint32_t* __restrict__ failure) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count || *ok == 0u) return;
    const uint32_t c = first + t;
    const uint2 e = ij[c];
    float lam = lambda[c];
    float4 a = pos[e.x];
    float4 b = pos[e.y];
    const float dt = cfg->gravityAndTimestep.w;
    if (!projectDistance(a, b, restCompliance[c], lam, 1.0f / (dt * dt), failure)) return;
    pos[e.x] = a;
    pos[e.y] = b;
    lambda[c] = lam;
}

// One unilateral strain limit (reference limit_strain,
// src/metal/ClothBag.metal:1616-1673) on loaded values.
__device__ __forceinline__ bool limitStrainPair(float4& a, float4& b, float limit, uint32_t* failure) {
    const V3 d = xyz(b) - xyz(a);
    const float len = length(d);
    if (!isfinite(len) || !isfinite(limit)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return false;
    }
    if (!(len > limit) || !(len > 1.0e-12f)) return false;
    const float denom = a.w + b.w;
    if (!(denom > 0.0f) || !isfinite(denom)) return false;
    const float excess = len - limit;
    const V3 dir = d * (1.0f / len);
    const float share = excess / denom;
    a = withXyz(a, xyz(a) + dir * (a.w * share));
    b = withXyz(b, xyz(b) - dir * (b.w * share));
    return true;
}

__global__ void k_strain(const uint32_t* __restrict__ ok, uint32_t first, uint32_t count,
                         float4* __restrict__ pos, const uint2* __restrict__ ij,
                         const float* __restrict__ maxLength, uint32_t* __restrict__ failure) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count || *ok == 0u) return;
    const uint32_t c = first + t;
    const uint2 e = ij[c];
    float4 a = pos[e.x];
    float4 b = pos[e.y];
    if (!limitStrainPair(a, b, maxLength[c], failure)) return;
    pos[e.x] = a;
    pos[e.y] = b;
}

// ----------------------------------------------------------------- knots --

// One crossing-angle cosine constraint on the warp and weft chords through a
// knot (reference solve_knot, src/metal/ClothBag.metal:1215-1313).
__device__ __forceinline__ bool projectKnot(float4& p0, float4& p1, float4& p2, float4& p3, float2 cc, float& lam,
                                            float dt, uint32_t* failure) {
    const V3 warpV = xyz(p1) - xyz(p0);
    const V3 weftV = xyz(p3) - xyz(p2);
    const float warpL = length(warpV);
    const float weftL = length(weftV);
    if (!isfinite(warpL) || !isfinite(weftL) || !isfinite(cc.x) || !isfinite(cc.y) || !isfinite(lam)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return false;
    }
    if (!(warpL > 1.0e-12f) || !(weftL > 1.0e-12f)) return false;
    const V3 warp = warpV / warpL;
    const V3 weft = weftV / weftL;
    const float cosine = clampf(dot(warp, weft), -1.0f, 1.0f);
    const float value = cosine - cc.x;
    const V3 gWarp = (weft - warp * cosine) / warpL;
    const V3 gWeft = (warp - weft * cosine) / weftL;
    const V3 g[4] = {-gWarp, gWarp, -gWeft, gWeft};
    const float w[4] = {p0.w, p1.w, p2.w, p3.w};
    const float alpha = cc.y / (dt * dt);
    float denom = alpha;
    for (int k = 0; k < 4; ++k) denom += w[k] * dot(g[k], g[k]);
    if (!(denom > 0.0f) || !isfinite(denom)) return false;
    const float dl = (-value - alpha * lam) / denom;
    const float nl = lam + dl;
    const V3 n0 = xyz(p0) + g[0] * (w[0] * dl);
    const V3 n1 = xyz(p1) + g[1] * (w[1] * dl);
    const V3 n2 = xyz(p2) + g[2] * (w[2] * dl);
    const V3 n3 = xyz(p3) + g[3] * (w[3] * dl);
    if (!finite3(n0) || !finite3(n1) || !finite3(n2) || !finite3(n3) || !isfinite(nl)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return false;
    }
    p0 = withXyz(p0, n0);
    p1 = withXyz(p1, n1);
    p2 = withXyz(p2, n2);
    p3 = withXyz(p3, n3);
    lam = nl;
    return true;
}

__global__ void k_knot(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                       uint32_t first, uint32_t count, float4* __restrict__ pos,
                       const uint4* __restrict__ quad, const float2* __restrict__ cosineCompliance,
                       float* __restrict__ lambda, uint32_t* __restrict__ failure) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count || *ok == 0u) return;
    const uint32_t c = first + t;
    const uint4 q = quad[c];
    float lam = lambda[c];
    float4 p0 = pos[q.x], p1 = pos[q.y], p2 = pos[q.z], p3 = pos[q.w];
    if (!projectKnot(p0, p1, p2, p3, cosineCompliance[c], lam, cfg->gravityAndTimestep.w, failure)) return;
    pos[q.x] = p0;
    pos[q.y] = p1;
    pos[q.z] = p2;
    pos[q.w] = p3;
    lambda[c] = lam;
}

// ----------------------------------------------------------------- bends --

// One chord bend (reference solve_bend, src/metal/ClothBag.metal:1315-1447):
// chord-length XPBD between the outer knots with a ground active set (a
// grounded endpoint that would be pushed down keeps its height) and a step
// fraction that stops free endpoints at the ground.
// invDt2 = 1 / (dt * dt) of the substep: alpha = compliance * invDt2.
// Without branches, as projectDistanceCore: every value computed, the
// result selected, failures or-ed into `fail`.
__device__ __forceinline__ bool projectBendCore(float4& a, float4& b, float2 cc, float& lam, float invDt2, bool ground,
                                                float floorZ, uint32_t& fail) {
    const V3 d = xyz(b) - xyz(a);
    const float chord = length(d);
    const bool bad = !isfinite(chord) || !isfinite(cc.x) || !isfinite(cc.y) || !isfinite(lam);
    bool go = !bad && chord > 1.0e-12f;
    const float value = chord - cc.x;
    const V3 dir = d * (1.0f / chord);
    const V3 g0 = -dir, g1 = dir;
    const float w0 = a.w, w1 = b.w;
    const float alpha = cc.y * invDt2;
    float freeDenom = alpha;
    freeDenom += w0 * dot(g0, g0);
    freeDenom += w1 * dot(g1, g1);
    go = go && freeDenom >= 1.0e-16f && isfinite(freeDenom);
    const float numerator = -value - alpha * lam;
    const float freeDl = numerator / freeDenom;
    const bool active0 = ground && a.z <= floorZ + 1.0e-9f && g0.z * freeDl < 0.0f;
    const bool active1 = ground && b.z <= floorZ + 1.0e-9f && g1.z * freeDl < 0.0f;
    float denom = alpha;
    denom += w0 * (dot(g0, g0) - (active0 ? g0.z * g0.z : 0.0f));
    denom += w1 * (dot(g1, g1) - (active1 ? g1.z * g1.z : 0.0f));
    go = go && denom >= 1.0e-16f && isfinite(denom);
    // Without an active ground knot denom is freeDenom to the bit (the same
    // sums with zero subtracted), so the free multiplier is reused.
    const float dl = active0 || active1 ? numerator / denom : freeDl;
    float fraction = 1.0f;
    if (ground) {
        const float down0 = g0.z * w0 * dl;
        if (!active0 && down0 < 0.0f) fraction = fminf(fraction, fmaxf(0.0f, a.z - floorZ) / -down0);
        const float down1 = g1.z * w1 * dl;
        if (!active1 && down1 < 0.0f) fraction = fminf(fraction, fmaxf(0.0f, b.z - floorZ) / -down1);
    }
    const float applied = dl * fraction;
    const float nl = lam + applied;
    V3 ca = g0 * (w0 * applied);
    V3 cb = g1 * (w1 * applied);
    if (active0) ca.z = 0.0f;
    if (active1) cb.z = 0.0f;
    V3 na = xyz(a) + ca;
    V3 nb = xyz(b) + cb;
    if (ground) {
        na.z = fmaxf(na.z, floorZ);
        nb.z = fmaxf(nb.z, floorZ);
    }
    const bool badOut = go && (!finite3(na) || !finite3(nb) || !isfinite(nl));
    fail |= bad || badOut ? NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE : 0u;
    const bool apply = go && !badOut;
    if (apply) {
        a = withXyz(a, na);
        b = withXyz(b, nb);
        lam = nl;
    }
    return apply;
}

__device__ __forceinline__ bool projectBend(float4& a, float4& b, float2 cc, float& lam, float invDt2, bool ground,
                                            float floorZ, uint32_t* failure) {
    uint32_t fail = 0u;
    const bool apply = projectBendCore(a, b, cc, lam, invDt2, ground, floorZ, fail);
    if (fail != 0u) recordFailure(failure, fail);
    return apply;
}

__global__ void k_bend(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                       uint32_t first, uint32_t count, float4* __restrict__ pos,
                       const uint2* __restrict__ ends, const float2* __restrict__ chordCompliance,
                       float* __restrict__ lambda, uint32_t* __restrict__ failure) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count || *ok == 0u) return;
    const uint32_t c = first + t;
    const uint2 e = ends[c];
    float lam = lambda[c];
    float4 a = pos[e.x];
    float4 b = pos[e.y];
    const float dt = cfg->gravityAndTimestep.w;
    if (!projectBend(a, b, chordCompliance[c], lam, 1.0f / (dt * dt), cfg->constraintCounts.z != 0u,
                     cfg->clothMaterial.x, failure)) {
        return;
    }
    pos[e.x] = a;
    pos[e.y] = b;
    lambda[c] = lam;
}

// ------------------------------------------------------------------ grip --

// update_grip_attachment (reference src/metal/ClothBag.metal:577-673).
__global__ void k_grip_attach(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                              const float4* __restrict__ pos, NumiClothBagGPUGrip* __restrict__ grips,
                              uint32_t gripCount, const uint32_t* __restrict__ cuff, uint32_t particles,
                              uint32_t* __restrict__ failure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (*ok == 0u || i >= gripCount || i >= cfg->control.w || !(cfg->gripTargetAndActive.w > 0.0f)) return;
    NumiClothBagGPUGrip grip = grips[i];
    if (grip.particle.y == cfg->gripControl.x) return;
    if (grip.particle.y == 0xffffffffu || grip.particle.y + 1u != cfg->gripControl.x) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_GRIP_CAPTURE);
        return;
    }
    const V3 target = xyz(cfg->gripTargetAndActive);
    uint32_t particle = grip.particle.x;
    if (cfg->gripControl.y == 1u) {
        const uint32_t rings = cfg->gripControl.z;
        uint32_t centre = 0u;
        float nearest = INFINITY;
        for (uint32_t ring = 0u; ring < rings; ++ring) {
            const uint32_t candidate = cuff[rings + ring];
            if (candidate >= particles) {
                recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_RANGE);
                return;
            }
            const V3 sep = xyz(pos[candidate]) - target;
            const float d2 = dot(sep, sep);
            if (d2 < nearest) {
                nearest = d2;
                centre = ring;
            }
        }
        const uint32_t width = 2u * cfg->gripControl.w + 1u;
        const uint32_t row = i / width;
        const uint32_t slot = i % width;
        const uint32_t ring = (centre + rings + slot - cfg->gripControl.w) % rings;
        particle = cuff[row * rings + ring];
        grip.particle.x = particle;
        grip.particle.w = centre;
    }
    if (particle >= particles) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_RANGE);
        return;
    }
    const V3 x = xyz(pos[particle]);
    const V3 sep = x - target;
    const float capture = length(sep);
    if (!finite3(x) || !isfinite(capture)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    if (capture > cfg->gripMaterial.x) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_GRIP_CAPTURE);
        return;
    }
    const mr_float4 q = cfg->gripOrientation;
    const mr_float4 inverse = {-q.x, -q.y, -q.z, q.w};
    const V3 local = rotate(inverse, sep);
    const V3 rebuilt = target + rotate(q, local);
    const float error = length(x - rebuilt);
    grip.targetOffsetAndCompliance.x = local.x;
    grip.targetOffsetAndCompliance.y = local.y;
    grip.targetOffsetAndCompliance.z = local.z;
    grip.particle.y = cfg->gripControl.x;
    grip.particle.z = __float_as_uint(fmaxf(__uint_as_float(grip.particle.z), capture));
    grip.lambda.x = 0.0f;
    grip.lambda.y = 0.0f;
    grip.lambda.z = 0.0f;
    grip.lambda.w = fmaxf(grip.lambda.w, error);
    if (!finite4(grip.targetOffsetAndCompliance) || !finite4(grip.lambda)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    grips[i] = grip;
}

// solve_grip (reference src/metal/ClothBag.metal:1449-1501): vector XPBD
// towards the handle-frame target g + q o.
// Grip i of solve_grip (the caller has checked the config is valid).
__device__ __forceinline__ void projectGrip(const NumiClothBagGPUConfig* __restrict__ cfg, float4* __restrict__ pos,
                                            NumiClothBagGPUGrip* __restrict__ grips, uint32_t i, uint32_t gripCount,
                                            uint32_t particles, uint32_t* __restrict__ failure) {
    if (i >= gripCount || i >= cfg->control.w || !(cfg->gripTargetAndActive.w > 0.0f)) return;
    NumiClothBagGPUGrip grip = grips[i];
    if (grip.particle.y != cfg->gripControl.x) return;
    const uint32_t index = grip.particle.x;
    if (index >= particles) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_RANGE);
        return;
    }
    float4 p = pos[index];
    const float dt = cfg->gravityAndTimestep.w;
    const float alpha = grip.targetOffsetAndCompliance.w / (dt * dt);
    const float denom = p.w + alpha;
    if (!isfinite(alpha) || !isfinite(denom) || !finite4(grip.targetOffsetAndCompliance) ||
        !finite4(grip.lambda)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    if (!(denom > 0.0f)) return;
    const V3 target = xyz(cfg->gripTargetAndActive) + rotate(cfg->gripOrientation, xyz(grip.targetOffsetAndCompliance));
    const V3 value = xyz(p) - target;
    const V3 lam = xyz(grip.lambda);
    const V3 dl = (-value - lam * alpha) / denom;
    const V3 nl = lam + dl;
    const V3 np = xyz(p) + dl * p.w;
    if (!finite3(np) || !finite3(nl)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    grip.lambda.x = nl.x;
    grip.lambda.y = nl.y;
    grip.lambda.z = nl.z;
    pos[index] = withXyz(p, np);
    grips[i] = grip;
}

__global__ void k_grip(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                       float4* __restrict__ pos, NumiClothBagGPUGrip* __restrict__ grips, uint32_t gripCount,
                       uint32_t particles, uint32_t* __restrict__ failure) {
    if (*ok == 0u) return;
    projectGrip(cfg, pos, grips, blockIdx.x * blockDim.x + threadIdx.x, gripCount, particles, failure);
}

// ------------------------------------------------------ ground + spheres --

// Cloth part of solve_ground (reference src/metal/ClothBag.metal:1575-1614)
// followed by static sphere contact, which the reference does not have
// (its spheres must be dynamic fruits): a dynamic knot inside a sphere grown
// by the yarn radius is moved radially onto its surface. Returns true when
// the knot moved.
__device__ __forceinline__ bool collideKnot(float4& p, bool ground, float r, const float4* spheres,
                                            uint32_t sphereCount) {
    bool changed = false;
    if (ground && !(p.z >= r)) {
        p.z = fmaxf(p.z, r);
        changed = true;
    }
    if (p.w > 0.0f) {
        for (uint32_t k = 0; k < sphereCount; ++k) {
            const float4 s = spheres[k];
            const V3 d = xyz(p) - xyz(s);
            const float dist = length(d);
            const float reach = s.w + r;
            if (dist < reach && dist > 1.0e-12f) {
                p = withXyz(p, xyz(s) + d * (reach / dist));
                changed = true;
            }
        }
    }
    return changed;
}

// ------------------------------------------------------ sphere windows --
//
// The spheres that can act on the knots of one workgroup. In a pass a knot
// is tested against the spheres in their list order at a sequence of
// positions (the start, then each push). The positions start inside the
// workgroup's box B; while they stay inside B grown by pad, a sphere whose
// centre lies more than reach + 2 pad outside B on some axis is farther than
// reach + pad from every one of them along that axis (pad also covers the
// rounding of the coordinate differences), so its test fails each time and
// leaving it out changes nothing. The window keeps the other spheres in
// list order; a knot that leaves B + pad is redone with the whole list.
// The result is therefore the same bytes as collideKnot over all spheres.
struct SphereWindow {
    float4 lo, hi;  // B grown by pad
    uint32_t count;
};

constexpr uint32_t kMaxBlockWaves = 16;

// Lists of fewer than this many spheres are tested whole (collideKnot): the
// window's box reductions and barriers cost more than testing a few spheres
// per knot (measured on gfx950: ≈3.3 µs per tile pass for the window against
// ≈0.8 µs for a one-sphere list). Both give the same bytes; the solver passes
// the threshold to the kernels (NUMI_CLOTH_HIP_SPHERE_WINDOW overrides it at
// create, so tests can run either path on any scene).
constexpr uint32_t kSphereWindowMin = 5;

// Every thread of the workgroup calls this with the box of its own knots
// (+inf / -inf when it has none). spheres is the workgroup's LDS copy of the
// list, complete before the call (wave 0 reads it after the first barrier,
// while the other waves wait: a global list would put two memory round
// trips on every pass); culled must hold sphereCount entries; scratch holds
// 2 * kMaxBlockWaves. Ends with a workgroup barrier.
__device__ void buildSphereWindow(V3 lo, V3 hi, float radius, const float4* __restrict__ spheres,
                                  uint32_t sphereCount, float4* culled, SphereWindow& w, float4* scratch) {
    for (int offset = 32; offset > 0; offset >>= 1) {
        lo.x = fminf(lo.x, __shfl_xor(lo.x, offset));
        lo.y = fminf(lo.y, __shfl_xor(lo.y, offset));
        lo.z = fminf(lo.z, __shfl_xor(lo.z, offset));
        hi.x = fmaxf(hi.x, __shfl_xor(hi.x, offset));
        hi.y = fmaxf(hi.y, __shfl_xor(hi.y, offset));
        hi.z = fmaxf(hi.z, __shfl_xor(hi.z, offset));
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
        scratch[wave] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        scratch[kMaxBlockWaves + wave] = make_float4(hi.x, hi.y, hi.z, 0.0f);
    }
    __syncthreads();
    if (wave == 0) {
        const uint32_t waves = blockDim.x >> 6;
        float4 l = lane < waves ? scratch[lane] : make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
        float4 h = lane < waves ? scratch[kMaxBlockWaves + lane] : make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
        float reach = 0.0f;
        for (uint32_t k = lane; k < sphereCount; k += 64) reach = fmaxf(reach, spheres[k].w + radius);
        for (int offset = 32; offset > 0; offset >>= 1) {
            l.x = fminf(l.x, __shfl_xor(l.x, offset));
            l.y = fminf(l.y, __shfl_xor(l.y, offset));
            l.z = fminf(l.z, __shfl_xor(l.z, offset));
            h.x = fmaxf(h.x, __shfl_xor(h.x, offset));
            h.y = fmaxf(h.y, __shfl_xor(h.y, offset));
            h.z = fmaxf(h.z, __shfl_xor(h.z, offset));
            reach = fmaxf(reach, __shfl_xor(reach, offset));
        }
        const float extent = fmaxf(fmaxf(fmaxf(fabsf(l.x), fabsf(l.y)), fmaxf(fabsf(l.z), fabsf(h.x))),
                                   fmaxf(fabsf(h.y), fabsf(h.z)));
        const float pad = 0.25f * reach + 1.0e-5f * extent;
        // An empty or non-finite box keeps every sphere and an unbounded window.
        const bool bounded = isfinite(pad) && l.x <= h.x && l.y <= h.y && l.z <= h.z;
        uint32_t count = 0;
        for (uint32_t base = 0; base < sphereCount; base += 64) {
            const uint32_t k = base + lane;
            bool keep = false;
            float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (k < sphereCount) {
                c = spheres[k];
                const float g = c.w + radius + 2.0f * pad;
                keep = !bounded || (c.x >= l.x - g && c.x <= h.x + g && c.y >= l.y - g && c.y <= h.y + g &&
                                    c.z >= l.z - g && c.z <= h.z + g);
            }
            const unsigned long long ballot = __ballot(keep);
            const uint32_t before = static_cast<uint32_t>(__popcll(ballot & ((1ull << lane) - 1ull)));
            if (keep) culled[count + before] = c;
            count += static_cast<uint32_t>(__popcll(ballot));
        }
        if (lane == 0) {
            w.count = count;
            if (bounded) {
                w.lo = make_float4(l.x - pad, l.y - pad, l.z - pad, 0.0f);
                w.hi = make_float4(h.x + pad, h.y + pad, h.z + pad, 0.0f);
            } else {
                w.lo = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
                w.hi = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ bool insideWindow(const float4& p, const SphereWindow& w) {
    return p.x >= w.lo.x && p.x <= w.hi.x && p.y >= w.lo.y && p.y <= w.hi.y && p.z >= w.lo.z && p.z <= w.hi.z;
}

// collideKnot with the window's spheres; the same result as collideKnot
// over all spheres (see SphereWindow).
__device__ __forceinline__ bool collideKnotWindow(float4& p, bool ground, float r, const SphereWindow& w,
                                                  const float4* culled, const float4* spheres,
                                                  uint32_t sphereCount) {
    const float4 start = p;
    bool changed = false;
    if (ground && !(p.z >= r)) {
        p.z = fmaxf(p.z, r);
        changed = true;
    }
    if (p.w > 0.0f) {
        bool inside = insideWindow(p, w);
        for (uint32_t k = 0; k < w.count; ++k) {
            const float4 s = culled[k];
            const V3 d = xyz(p) - xyz(s);
            const float dist = length(d);
            const float reach = s.w + r;
            if (dist < reach && dist > 1.0e-12f) {
                p = withXyz(p, xyz(s) + d * (reach / dist));
                changed = true;
                inside = inside && insideWindow(p, w);
            }
        }
        if (!inside) {
            p = start;
            return collideKnot(p, ground, r, spheres, sphereCount);
        }
    }
    return changed;
}

// Box of a thread's knot for buildSphereWindow (identity when it has none).
__device__ __forceinline__ void knotBox(const float4& p, bool has, V3& lo, V3& hi) {
    lo = has ? xyz(p) : v3(INFINITY, INFINITY, INFINITY);
    hi = has ? xyz(p) : v3(-INFINITY, -INFINITY, -INFINITY);
}

// Fruit part of solve_ground: z >= radius with the impulse that removes the
// penetration over the substep.
__device__ __forceinline__ void fruitGround(const NumiClothBagGPUConfig* __restrict__ cfg,
                                            NumiClothBagGPUFruit* __restrict__ fruits, uint32_t i,
                                            uint32_t* __restrict__ failure) {
    NumiClothBagGPUFruit f = fruits[i];
    const float penetration = f.previousAndRadius.w - f.positionAndInverseMass.z;
    if (!isfinite(penetration)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    if (penetration > 0.0f) {
        f.velocityAndGroundImpulse.w += penetration / (f.positionAndInverseMass.w * cfg->gravityAndTimestep.w);
        f.positionAndInverseMass.z = f.previousAndRadius.w;
        fruits[i] = f;
    }
}

// The pending self-contact corrections of one knot (the mean of its
// fixed-point sum, when its touched bit is set in `word`) and the
// accumulator cleared for reuse; live = false clears without moving.
struct FixedAccumulator;
__device__ __forceinline__ void applyPendingCorrection(float4& p, uint32_t i, uint32_t word, bool live,
                                                       FixedAccumulator* __restrict__ acc,
                                                       uint32_t* __restrict__ failure);

// Ground and static spheres for every knot. Dynamic shared memory: the
// spheres, then the window's list (2 * sphereCount float4). With fruits
// (fruitCount > 0) the first workgroup's first threads also run the fruit
// part of the ground pass (the schedule's next launch, on the fruits only).
// With selfAcc the self-contact corrections of the solve before it are
// applied first (its accumulator and touched mask cleared), as the
// schedule's apply pass would.
__global__ void k_collide(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                          float4* __restrict__ pos, uint32_t particles, const float4* __restrict__ spheres,
                          uint32_t sphereCount, uint32_t windowMin, uint32_t* __restrict__ failure,
                          NumiClothBagGPUFruit* __restrict__ fruits = nullptr, uint32_t fruitCount = 0u,
                          FixedAccumulator* __restrict__ selfAcc = nullptr, uint32_t* __restrict__ selfTouched = nullptr) {
    extern __shared__ float4 sphereCache[];
    if (blockIdx.x == 0 && threadIdx.x < fruitCount && *ok != 0u && cfg->constraintCounts.z != 0u)
        fruitGround(cfg, fruits, threadIdx.x, failure);
    __shared__ SphereWindow window;
    __shared__ float4 scratch[2 * kMaxBlockWaves];
    float4* culled = sphereCache + sphereCount;
    for (uint32_t k = threadIdx.x; k < sphereCount; k += blockDim.x) sphereCache[k] = spheres[k];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = *ok != 0u;
    const bool has = live && i < particles;
    float4 p0 = has ? pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool corrected = false;
    if (selfAcc != nullptr && i < particles) {
        // Every lane of the 32-knot word reads it before its first lane
        // clears it (the lanes of one wave, in order).
        const uint32_t word = selfTouched[i >> 5];
        if ((word >> (i & 31u)) & 1u) {
            applyPendingCorrection(p0, i, word, live, selfAcc, failure);
            corrected = live;
        }
        if ((i & 31u) == 0u && word != 0u) selfTouched[i >> 5] = 0u;
    }
    const float r = cfg->clothMaterial.x;
    const bool windowed = sphereCount >= windowMin;
    if (windowed) {
        V3 lo, hi;
        knotBox(p0, has && isfinite(p0.z), lo, hi);
        buildSphereWindow(lo, hi, r, sphereCache, sphereCount, culled, window, scratch);
    } else {
        __syncthreads();
    }
    if (!has) return;
    float4 p = p0;
    if (!isfinite(p.z)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    const bool ground = cfg->constraintCounts.z != 0u;
    const bool moved = windowed ? collideKnotWindow(p, ground, r, window, culled, sphereCache, sphereCount)
                                : collideKnot(p, ground, r, sphereCache, sphereCount);
    if (moved || corrected) pos[i] = p;
}

// ------------------------------------------------------------- finalize --

// finalize_substep (reference src/metal/ClothBag.metal:2890-2921) fused with
// apply_cloth_ground_friction (:2976-3022); nothing runs between them on the
// cloth knots. A grid-stride loop over a bounded grid: the friction contact
// count and the largest cone ratio are reduced per wavefront, then per
// workgroup in LDS, and each workgroup adds one integer atomic pair (a
// draped sheet has most knots in ground contact; one atomic pair per wave on
// the same two words serialised ~16k waves).
constexpr uint32_t kFinalizeMaxBlocks = 1024;

__global__ void k_finalize(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                           float4* __restrict__ pos, const float4* __restrict__ prev, float4* __restrict__ vel,
                           uint32_t particles, NumiClothBagGPUFrictionStatus* __restrict__ friction,
                           uint32_t* __restrict__ failure) {
    __shared__ uint32_t blockContacts;
    __shared__ uint32_t blockRatio;
    if (threadIdx.x == 0) {
        blockContacts = 0u;
        blockRatio = 0u;
    }
    __syncthreads();
    if (*ok == 0u) return;
    const float inverseDt = 1.0f / cfg->gravityAndTimestep.w;
    const bool ground = cfg->constraintCounts.z != 0u;
    const float mu = cfg->clothMaterial.z;
    const float contactHeight = cfg->clothMaterial.x + 1.0e-6f;
    uint32_t contacts = 0u;
    float ratio = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < particles; i += gridDim.x * blockDim.x) {
        const float4 p = pos[i];
        const float4 v = vel[i];
        const float4 q = prev[i];
        const float predicted = v.w;
        V3 u = (xyz(p) - xyz(q)) * inverseDt;
        float jn = 0.0f;
        if (ground && p.w > 0.0f && p.z <= contactHeight) {
            jn = fmaxf(0.0f, (u.z - predicted) / p.w);
        }
        bool good = finite3(u);
        if (good && ground) {
            if (!(mu >= 0.0f) || !isfinite(mu)) {
                good = false;
            } else if (jn > 0.0f && p.w > 0.0f) {
                const V3 tangent = v3(u.x, u.y, 0.0f);
                const float slip = length(tangent);
                if (slip > 1.0e-10f) {
                    const float limit = mu * jn;
                    const float jt = fminf(slip / p.w, limit);
                    if (jt > 0.0f) {
                        u = u - tangent * (jt * p.w / slip);
                        contacts += 1u;
                        ratio = fmaxf(ratio, limit > 0.0f ? jt / limit : 0.0f);
                        good = finite3(u) && isfinite(jn);
                    }
                }
            }
        }
        if (!good) {
            recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        } else {
            vel[i] = make_float4(u.x, u.y, u.z, jn);
        }
    }
    for (int offset = 32; offset > 0; offset >>= 1) {
        contacts += __shfl_xor(contacts, offset);
        ratio = fmaxf(ratio, __shfl_xor(ratio, offset));
    }
    if ((threadIdx.x & 63u) == 0u && contacts != 0u) {
        atomicAdd(&blockContacts, contacts);
        atomicMax(&blockRatio, __float_as_uint(ratio));
    }
    __syncthreads();
    if (threadIdx.x == 0 && blockContacts != 0u) {
        atomicAdd(&friction->counters.z, blockContacts);
        atomicMax(&friction->metrics.x, blockRatio);
    }
}

}  // namespace numi_dev
