#pragma once

// Produce-bag families: rigid fruit spheres, fruit/yarn and fruit/fruit
// contact, air loads, friction, orientation and mouth release.
// Reference kernels (src/metal/ClothBag.metal):
//   begin_substep (fruit part) :491-575   apply_fruit_aerodynamics :844-905
//   clear/accumulate/reduce/apply_yarn_aerodynamics :675-842
//   advance_positions (fruit part) :907-938
//   update_released_fruit :1025-1146      solve_fruit_pair :1503-1573
//   solve_ground (fruit part) :1575-1614  solve_yarn_batch :1675-1786
//   applyYarnCorrection :352-462          build_yarn_contacts :2733-2888
//   finalize_fruit :2923-2942             integrate_fruit_orientation :2944-2974
//   apply_yarn_friction :3024-3158        apply_fruit_pair_friction :3160-3248
//   apply_fruit_ground_friction :3250-3335
//
// The fruit/yarn phase loops keep the reference's order exactly: fruit f
// visits segment (phase + f) mod n of the colour batch, one phase at a time.
// They run as a single workgroup with every knot staged in LDS and all
// colour batches of a call in one launch, so a phase costs one workgroup
// barrier instead of a dispatch.

#include "device_math.hpp"

namespace numi_dev {

__device__ __forceinline__ float fruitInverseInertia(const NumiClothBagGPUFruit& f) {
    const float r = f.previousAndRadius.w;
    return 2.5f * f.positionAndInverseMass.w / (r * r);
}

__device__ __forceinline__ void applyFruitImpulse(NumiClothBagGPUFruit& f, V3 impulse, V3 offset) {
    const V3 dv = impulse * f.positionAndInverseMass.w;
    f.velocityAndGroundImpulse.x += dv.x;
    f.velocityAndGroundImpulse.y += dv.y;
    f.velocityAndGroundImpulse.z += dv.z;
    const V3 dw = cross(offset, impulse) * fruitInverseInertia(f);
    f.angularVelocity.x += dw.x;
    f.angularVelocity.y += dw.y;
    f.angularVelocity.z += dw.z;
}

__device__ __forceinline__ void recordFriction(NumiClothBagGPUFrictionStatus* status, uint32_t counter, float jt,
                                               float limit) {
    atomicAdd(&(&status->counters.x)[counter], 1u);
    if (limit > 0.0f) atomicMax(&status->metrics.x, __float_as_uint(jt / limit));
}

__device__ __forceinline__ void setXyz(mr_float4& v, V3 p) {
    v.x = p.x;
    v.y = p.y;
    v.z = p.z;
}

// ------------------------------------------------------------- predict --

// Fruit part of begin_substep, plus the per-substep reset of the fruit-pair
// and fruit/yarn impulse accumulators.
__global__ void k_fruit_begin(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                              NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                              NumiClothBagGPUFruitPair* __restrict__ pairs, uint32_t pairCount,
                              NumiClothBagGPUYarnContact* __restrict__ contacts, uint32_t contactCount,
                              uint32_t* __restrict__ failure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (*ok == 0u) return;
    if (i < fruitCount) {
        NumiClothBagGPUFruit f = fruits[i];
        const V3 x = xyz(f.positionAndInverseMass);
        const V3 v = xyz(f.velocityAndGroundImpulse);
        const float w = f.positionAndInverseMass.w, r = f.previousAndRadius.w;
        if (!finite3(x) || !finite3(v) || !finite4(f.angularVelocity) || !finite4(f.orientation) || !isfinite(w) ||
            !isfinite(r) || !(w > 0.0f) || !(r > 0.0f)) {
            recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        } else {
            const float dt = cfg->gravityAndTimestep.w;
            setXyz(f.previousAndRadius, x);
            setXyz(f.velocityAndGroundImpulse, fma3(xyz(cfg->gravityAndTimestep), dt, v));
            f.velocityAndGroundImpulse.w = 0.0f;
            fruits[i] = f;
        }
    }
    if (i < pairCount) pairs[i].contact = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
    if (i < contactCount) {
        contacts[i].control.w = 0u;
        contacts[i].fruitNormalAndImpulse = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
        contacts[i].segmentImpulse = mr_float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
}

__global__ void k_fruit_advance(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                                NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fruitCount || *ok == 0u) return;
    NumiClothBagGPUFruit f = fruits[i];
    setXyz(f.positionAndInverseMass,
           fma3(xyz(f.velocityAndGroundImpulse), cfg->gravityAndTimestep.w, xyz(f.positionAndInverseMass)));
    fruits[i] = f;
}

// ----------------------------------------------------------- air loads --

__global__ void k_aero_clear(const uint32_t* __restrict__ ok, float4* __restrict__ forces, uint32_t particles) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < particles && *ok != 0u) forces[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Crossflow and axial skin drag of one yarn segment, split to its knots
// (one colour batch: the two knots of a segment belong to no other segment
// of the batch, so the accumulation is race free).
__global__ void k_aero_accumulate(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                                  uint32_t first, uint32_t count, const float4* __restrict__ pos,
                                  const float4* __restrict__ vel, const uint2* __restrict__ ij,
                                  float4* __restrict__ forces, NumiClothBagGPUAerodynamicsStatus* __restrict__ status,
                                  uint32_t* __restrict__ failure) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count || *ok == 0u) return;
    const uint2 e = ij[first + t];
    const V3 span = xyz(pos[e.y]) - xyz(pos[e.x]);
    const float len = length(span);
    if (!(len > 1.0e-12f)) return;
    const V3 axis = span / len;
    const V3 air = xyz(cfg->airVelocityAndDensity);
    const V3 rel = (xyz(vel[e.x]) + xyz(vel[e.y])) * 0.5f - air;
    const V3 axial = axis * dot(rel, axis);
    const V3 crossflow = rel - axial;
    const float density = cfg->airVelocityAndDensity.w;
    const float pi = 3.14159265358979323846f;
    const float crossCoeff = 0.5f * density * cfg->aerodynamicCoefficients.x * (2.0f * cfg->clothMaterial.x * len);
    const float skinCoeff = 0.5f * density * cfg->aerodynamicCoefficients.y * (2.0f * pi * cfg->clothMaterial.x * len);
    const V3 force = crossflow * (-crossCoeff * length(crossflow)) - axial * (skinCoeff * length(axial));
    if (!finite3(force)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    float4 fa = forces[e.x], fb = forces[e.y];
    fa.x += 0.5f * force.x;
    fa.y += 0.5f * force.y;
    fa.z += 0.5f * force.z;
    fb.x += 0.5f * force.x;
    fb.y += 0.5f * force.y;
    fb.z += 0.5f * force.z;
    forces[e.x] = fa;
    forces[e.y] = fb;
    atomicMax(&status->metrics.x, __float_as_uint(length(force)));
}

// One 256-lane workgroup: total drag power and sum of w|F|^2 in a fixed
// order (strided per-lane sums, then a pairwise tree), and the attenuation
// 1 / (1 + dt sum(w|F|^2) / -P) that keeps the explicit drag dissipative.
__global__ void k_aero_reduce(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                              const float4* __restrict__ pos, const float4* __restrict__ vel,
                              const float4* __restrict__ forces, uint32_t particles, float4* __restrict__ reduction) {
    __shared__ float power[256];
    __shared__ float weighted[256];
    if (*ok == 0u) return;
    const uint32_t lane = threadIdx.x;
    float p = 0.0f, w = 0.0f;
    const V3 air = xyz(cfg->airVelocityAndDensity);
    for (uint32_t i = lane; i < particles; i += 256) {
        const float4 x = pos[i];
        if (x.w > 0.0f) {
            const V3 f = xyz(forces[i]);
            p += dot(f, xyz(vel[i]) - air);
            w += x.w * dot(f, f);
        }
    }
    power[lane] = p;
    weighted[lane] = w;
    __syncthreads();
    for (uint32_t stride = 128; stride > 0; stride >>= 1) {
        if (lane < stride) {
            power[lane] += power[lane + stride];
            weighted[lane] += weighted[lane + stride];
        }
        __syncthreads();
    }
    if (lane == 0) {
        float attenuation = 1.0f;
        if (weighted[0] > 0.0f) {
            attenuation = power[0] < 0.0f ? 1.0f / (1.0f + cfg->gravityAndTimestep.w * weighted[0] / -power[0]) : 0.0f;
        }
        reduction[0] = make_float4(power[0], weighted[0], attenuation, 0.0f);
    }
}

__global__ void k_aero_apply(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                             const float4* __restrict__ pos, float4* __restrict__ vel, const float4* __restrict__ forces,
                             const float4* __restrict__ reduction, uint32_t particles, uint32_t* __restrict__ failure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= particles || *ok == 0u) return;
    const float w = pos[i].w;
    float4 v = vel[i];
    if (w > 0.0f) {
        const V3 f = xyz(forces[i]);
        const V3 nv = xyz(v) + f * (reduction[0].z * w * cfg->gravityAndTimestep.w);
        v = withXyz(v, nv);
        v.w = v.z;
    }
    if (!finite3(xyz(v))) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    vel[i] = v;
}

// Exact quadratic translational / rotational drag decay of a fruit.
__global__ void k_fruit_aero(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                             NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                             NumiClothBagGPUAerodynamicsStatus* __restrict__ status, uint32_t* __restrict__ failure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fruitCount || *ok == 0u) return;
    NumiClothBagGPUFruit f = fruits[i];
    const float pi = 3.14159265358979323846f;
    const float density = cfg->airVelocityAndDensity.w;
    const float dt = cfg->gravityAndTimestep.w;
    const float w = f.positionAndInverseMass.w;
    const float r = f.previousAndRadius.w;
    const V3 air = xyz(cfg->airVelocityAndDensity);
    const V3 rel = xyz(f.velocityAndGroundImpulse) - air;
    const float speed = length(rel);
    const float transCoeff = 0.5f * density * cfg->aerodynamicCoefficients.z * pi * r * r;
    const float transAtt = 1.0f / (1.0f + w * transCoeff * speed * dt);
    setXyz(f.velocityAndGroundImpulse, air + rel * transAtt);
    const V3 omega = xyz(f.angularVelocity);
    const float spin = length(omega);
    const float rotCoeff = (3.0f * pi * pi / 8.0f) * density * cfg->aerodynamicCoefficients.w * r * r * r * r * r;
    const float rotAtt = 1.0f / (1.0f + fruitInverseInertia(f) * rotCoeff * spin * dt);
    setXyz(f.angularVelocity, omega * rotAtt);
    atomicMax(&status->metrics.y, __float_as_uint(transCoeff * speed * speed));
    atomicMax(&status->metrics.z, __float_as_uint(rotCoeff * spin * spin));
    if (!finite3(xyz(f.velocityAndGroundImpulse)) || !finite3(xyz(f.angularVelocity))) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    fruits[i] = f;
}

// ------------------------------------------------------------- contacts --

// Positional sphere/sphere separation of one colour batch of fruit pairs.
// One sphere/sphere separation (reference solve_fruit_pair,
// src/metal/ClothBag.metal:1503-1573) on fruits held by the caller.
__device__ __forceinline__ void projectFruitPair(NumiClothBagGPUFruit& a, NumiClothBagGPUFruit& b,
                                                 NumiClothBagGPUFruitPair& pair, float dt, uint32_t* failure,
                                                 bool& wrote) {
    wrote = false;
    const V3 d = xyz(b.positionAndInverseMass) - xyz(a.positionAndInverseMass);
    const float len = length(d);
    const float target = a.previousAndRadius.w + b.previousAndRadius.w;
    if (!isfinite(len) || !isfinite(target) || !finite4(pair.contact)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    if (!(len < target) || !(len > 1.0e-12f)) return;
    const float denom = a.positionAndInverseMass.w + b.positionAndInverseMass.w;
    if (!(denom > 0.0f) || !isfinite(denom)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    const float lambda = (target - len) / denom;
    const V3 n = d / len;
    const V3 c = n * lambda;
    NumiClothBagGPUFruit na = a, nb = b;
    setXyz(na.positionAndInverseMass, xyz(a.positionAndInverseMass) - c * a.positionAndInverseMass.w);
    setXyz(nb.positionAndInverseMass, xyz(b.positionAndInverseMass) + c * b.positionAndInverseMass.w);
    const float impulse = lambda / dt;
    NumiClothBagGPUFruitPair np = pair;
    setXyz(np.contact, xyz(pair.contact) + n * impulse);
    np.contact.w += impulse;
    if (!finite3(xyz(na.positionAndInverseMass)) || !finite3(xyz(nb.positionAndInverseMass)) ||
        !finite4(np.contact)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    a = na;
    b = nb;
    pair = np;
    wrote = true;
}

// All fruit-pair colours of one solve pass in one workgroup (the reference
// launches one dispatch per colour, :4519-4528); every thread calls it. The
// fruits are staged in LDS (`local`, at least fruitCount records), a
// colour's pairs share no fruit, and a barrier separates colours; the
// fruits are written back and the workgroup synchronised.
// A pair acts only when its fruits overlap, and until one acts the fruits
// stay as they start: the colours before the first colour holding a pair
// that overlaps at the start act on nothing and are skipped (the non-finite
// checks they would make see the start state and are made up front). With
// at most one pair per thread, thread g keeps pair g (colourOf[g]: its
// colour) in registers for the whole pass.
__device__ void fruitPairColours(const NumiClothBagGPUConfig* __restrict__ cfg, const uint2* __restrict__ colours,
                                 const uint32_t* __restrict__ colourOf, uint32_t colourCount,
                                 NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                                 NumiClothBagGPUFruitPair* __restrict__ pairs, uint32_t pairCount,
                                 uint32_t* __restrict__ failure, NumiClothBagGPUFruit* local) {
    __shared__ uint32_t firstColour;
    const uint32_t tid = threadIdx.x, threads = blockDim.x;
    for (uint32_t i = tid; i < fruitCount; i += threads) local[i] = fruits[i];
    if (tid == 0) firstColour = colourCount;
    const bool held = pairCount <= threads;
    NumiClothBagGPUFruitPair mine{};
    uint32_t myColour = ~0u;
    if (held && tid < pairCount) {
        mine = pairs[tid];
        myColour = colourOf[tid];
    }
    __syncthreads();
    for (uint32_t g = tid; g < pairCount; g += threads) {
        const uint32_t c = held ? myColour : colourOf[g];
        if (c >= colourCount) continue;
        const NumiClothBagGPUFruitPair pair = held ? mine : pairs[g];
        const NumiClothBagGPUFruit& a = local[pair.fruitsAndColor.x];
        const NumiClothBagGPUFruit& b = local[pair.fruitsAndColor.y];
        const float len = length(xyz(b.positionAndInverseMass) - xyz(a.positionAndInverseMass));
        const float target = a.previousAndRadius.w + b.previousAndRadius.w;
        if (!isfinite(len) || !isfinite(target) || !finite4(pair.contact)) {
            recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        } else if (len < target && len > 1.0e-12f) {
            atomicMin(&firstColour, c);
        }
    }
    __syncthreads();
    const float dt = cfg->gravityAndTimestep.w;
    for (uint32_t c = firstColour; c < colourCount; ++c) {
        if (held) {
            if (myColour == c) {
                NumiClothBagGPUFruit a = local[mine.fruitsAndColor.x];
                NumiClothBagGPUFruit b = local[mine.fruitsAndColor.y];
                bool wrote = false;
                projectFruitPair(a, b, mine, dt, failure, wrote);
                if (wrote) {
                    local[mine.fruitsAndColor.x] = a;
                    local[mine.fruitsAndColor.y] = b;
                    pairs[tid] = mine;
                }
            }
        } else {
            const uint2 range = colours[c];
            for (uint32_t t = tid; t < range.y; t += threads) {
                NumiClothBagGPUFruitPair pair = pairs[range.x + t];
                NumiClothBagGPUFruit a = local[pair.fruitsAndColor.x];
                NumiClothBagGPUFruit b = local[pair.fruitsAndColor.y];
                bool wrote = false;
                projectFruitPair(a, b, pair, dt, failure, wrote);
                if (wrote) {
                    local[pair.fruitsAndColor.x] = a;
                    local[pair.fruitsAndColor.y] = b;
                    pairs[range.x + t] = pair;
                }
            }
        }
        __syncthreads();
    }
    for (uint32_t i = tid; i < fruitCount; i += threads) fruits[i] = local[i];
    __syncthreads();
}

constexpr uint32_t kFruitPairThreads = 256;  // k_fruit_pairs workgroup: up to this many pairs held in registers

__global__ void __launch_bounds__(kFruitPairThreads)
k_fruit_pairs(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                              const uint2* __restrict__ colours, const uint32_t* __restrict__ colourOf,
                              uint32_t colourCount, NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                              NumiClothBagGPUFruitPair* __restrict__ pairs, uint32_t pairCount,
                              uint32_t* __restrict__ failure) {
    __shared__ NumiClothBagGPUFruit local[32];
    if (*ok == 0u) return;
    fruitPairColours(cfg, colours, colourOf, colourCount, fruits, fruitCount, pairs, pairCount, failure, local);
}

// Fruit part of solve_ground: z >= radius with the impulse that removes the
// penetration over the substep.
__global__ void k_fruit_ground(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                               NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                               uint32_t* __restrict__ failure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fruitCount || *ok == 0u || cfg->constraintCounts.z == 0u) return;
    fruitGround(cfg, fruits, i, failure);
}

struct PointSample {
    V3 closest;
    float weight, distance;
};

__device__ __forceinline__ PointSample samplePointSegment(V3 point, V3 a, V3 b) {
    const V3 d = b - a;
    const float l2 = dot(d, d);
    const float w = l2 > 1.0e-20f ? clampf(dot(d, point - a) / l2, 0.0f, 1.0f) : 0.0f;
    const V3 c = fma3(d, w, a);
    return {c, w, length(c - point)};
}

__device__ __forceinline__ V3 degenerateYarnNormal(V3 a, V3 b) {
    const V3 dir = safeNormalized(b - a);
    const V3 axis = fabsf(dir.z) < 0.9f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
    return safeNormalized(cross(dir, axis));
}

// Present sample and swept conservative-advancement impact of every
// (fruit, segment) record.
__global__ void k_yarn_contacts_build(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                                      const float4* __restrict__ pos, const float4* __restrict__ prev,
                                      const uint2* __restrict__ ij, uint32_t segments,
                                      const NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                                      NumiClothBagGPUYarnContact* __restrict__ contacts, uint32_t* __restrict__ failure) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= segments * fruitCount || *ok == 0u) return;
    const uint32_t fi = index / segments;
    const uint32_t si = index - fi * segments;
    const uint2 e = ij[si];
    const NumiClothBagGPUFruit fruit = fruits[fi];
    const V3 a = xyz(pos[e.x]), b = xyz(pos[e.y]);
    const V3 a0 = xyz(prev[e.x]), b0 = xyz(prev[e.y]);
    const V3 c = xyz(fruit.positionAndInverseMass), c0 = xyz(fruit.previousAndRadius);
    const float target = fruit.previousAndRadius.w + cfg->clothMaterial.x;
    const PointSample now = samplePointSegment(c, a, b);
    const bool degenerate = now.distance < 1.0e-12f;
    const V3 normal = degenerate ? degenerateYarnNormal(a, b) : (now.closest - c) / now.distance;
    float time = 0.0f, sweptWeight = 0.0f, removed = 0.0f;
    V3 sweptNormal = v3(0.0f, 0.0f, 0.0f);
    bool impact = false;
    const float bound = length(c - c0) + length(a - a0) + length(b - b0);
    if (bound >= 1.0e-14f) {
        const float tolerance = 1.0e-9f;
        auto at = [&](float t) { return samplePointSegment(mix3(c0, c, t), mix3(a0, a, t), mix3(b0, b, t)); };
        PointSample hit = at(0.0f);
        bool found = hit.distance <= target + tolerance;
        if (!found) {
            for (int step = 0; step < 80; ++step) {
                hit = at(time);
                const float gap = hit.distance - target;
                if (gap <= tolerance) {
                    found = true;
                    break;
                }
                const float advance = 0.9f * gap / bound;
                if (!isfinite(advance) || !(advance > 0.0f) || time + advance >= 1.0f) break;
                time += fmaxf(advance, 1.0e-10f);
            }
        }
        if (found && hit.distance >= 1.0e-12f) {
            const V3 ct = mix3(c0, c, time), at_ = mix3(a0, a, time), bt = mix3(b0, b, time);
            sweptNormal = (hit.closest - ct) / hit.distance;
            sweptWeight = hit.weight;
            const V3 segRest = (a - at_) * (1.0f - sweptWeight) + (b - bt) * sweptWeight;
            const V3 ballRest = c - ct;
            removed = dot(ballRest - segRest, sweptNormal);
            impact = removed > 0.0f;
            if (!impact) removed = 0.0f;
        }
    }
    NumiClothBagGPUYarnContact r = contacts[index];
    r.identity = mr_uint4{fi, si, e.x, e.y};
    r.currentNormalAndSeparation = mr_float4{normal.x, normal.y, normal.z, now.distance - target};
    r.sweptNormalAndAdvance = mr_float4{sweptNormal.x, sweptNormal.y, sweptNormal.z, removed};
    r.weightsAndTime = mr_float4{now.weight, sweptWeight, time, target};
    r.control.x = now.distance < target ? 1u : 0u;
    r.control.y = impact ? 1u : 0u;
    r.control.z = degenerate ? 1u : 0u;
    if (!finite4(r.currentNormalAndSeparation) || !finite4(r.sweptNormalAndAdvance) || !finite4(r.weightsAndTime)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    contacts[index] = r;
}

// applyYarnCorrection (reference :352-462): normal response of a fruit and a
// yarn segment with a ground active set of at most three steps. Knots are in
// LDS; returns the accumulated multiplier.
__device__ __forceinline__ float yarnCorrection(float4& first, float4& second, NumiClothBagGPUFruit& fruit, float wa, float wb, V3 n,
                                float distance, bool groundEnabled, float floorZ, uint32_t* failure, bool& wrote) {
    wrote = false;
    float4 a = first, b = second;
    bool aActive = groundEnabled && n.z < 0.0f && wa > 0.0f && a.z <= floorZ + 1.0e-9f;
    bool bActive = groundEnabled && n.z < 0.0f && wb > 0.0f && b.z <= floorZ + 1.0e-9f;
    if (aActive) a.z = floorZ;
    if (bActive) b.z = floorZ;
    float remaining = distance;
    float total = 0.0f;
    for (int k = 0; k < 3 && remaining > 1.0e-14f; ++k) {
        V3 ra = n * a.w;
        V3 rb = n * b.w;
        if (aActive) ra.z = 0.0f;
        if (bActive) rb.z = 0.0f;
        const float denom = fruit.positionAndInverseMass.w + dot(n, ra) * wa * wa + dot(n, rb) * wb * wb;
        if (!(denom > 0.0f) || !isfinite(denom)) break;
        const float free = remaining / denom;
        float step = free;
        if (groundEnabled && n.z < 0.0f) {
            const float va = ra.z * wa;
            if (!aActive && va < 0.0f) step = fminf(step, fmaxf(0.0f, a.z - floorZ) / -va);
            const float vb = rb.z * wb;
            if (!bActive && vb < 0.0f) step = fminf(step, fmaxf(0.0f, b.z - floorZ) / -vb);
        }
        setXyz(fruit.positionAndInverseMass, xyz(fruit.positionAndInverseMass) - n * (fruit.positionAndInverseMass.w * step));
        a = withXyz(a, xyz(a) + ra * (wa * step));
        b = withXyz(b, xyz(b) + rb * (wb * step));
        total += step;
        remaining = fmaxf(0.0f, remaining - denom * step);
        if (step >= free - 1.0e-14f) break;
        bool activated = false;
        if (!aActive && wa > 0.0f && a.z <= floorZ + 1.0e-9f) {
            a.z = floorZ;
            aActive = true;
            activated = true;
        }
        if (!bActive && wb > 0.0f && b.z <= floorZ + 1.0e-9f) {
            b.z = floorZ;
            bActive = true;
            activated = true;
        }
        if (!activated) break;
    }
    if (!finite3(xyz(fruit.positionAndInverseMass)) || !finite3(xyz(a)) || !finite3(xyz(b)) || !isfinite(total)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return 0.0f;
    }
    first = a;
    second = b;
    wrote = true;
    return total;
}

// ------------------------------------------------------------- finalize --

__global__ void k_fruit_finalize(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                                 NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                                 uint32_t* __restrict__ failure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fruitCount || *ok == 0u) return;
    NumiClothBagGPUFruit f = fruits[i];
    setXyz(f.velocityAndGroundImpulse,
           (xyz(f.positionAndInverseMass) - xyz(f.previousAndRadius)) / cfg->gravityAndTimestep.w);
    if (!finite4(f.velocityAndGroundImpulse)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    fruits[i] = f;
}

// Coulomb friction of one fruit pair (reference apply_fruit_pair_friction,
// src/metal/ClothBag.metal:3160-3248) on fruits held by the caller.
__device__ __forceinline__ void fruitPairFriction(const NumiClothBagGPUConfig* cfg, NumiClothBagGPUFruit& a,
                                                  NumiClothBagGPUFruit& b, const NumiClothBagGPUFruitPair& pair,
                                                  NumiClothBagGPUFrictionStatus* status, uint32_t* failure,
                                                  bool& wrote) {
    wrote = false;
    const float mu = cfg->fruitMaterial.x;
    const float jn = pair.contact.w;
    const float nl = length(xyz(pair.contact));
    if (!(mu >= 0.0f) || !isfinite(mu)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    if (!(jn > 0.0f) || !(nl > 1.0e-10f)) return;
    const V3 n = xyz(pair.contact) / nl;
    const V3 oa = n * a.previousAndRadius.w;
    const V3 ob = n * -b.previousAndRadius.w;
    const V3 va = xyz(a.velocityAndGroundImpulse) + cross(xyz(a.angularVelocity), oa);
    const V3 vb = xyz(b.velocityAndGroundImpulse) + cross(xyz(b.angularVelocity), ob);
    const V3 rel = vb - va;
    const V3 tv = rel - n * dot(rel, n);
    const float slip = length(tv);
    if (!(slip > 1.0e-10f)) return;
    const V3 tangent = tv / slip;
    const V3 la = cross(oa, tangent);
    const V3 lb = cross(ob, tangent);
    const float denom = a.positionAndInverseMass.w + b.positionAndInverseMass.w +
                        fruitInverseInertia(a) * dot(la, la) + fruitInverseInertia(b) * dot(lb, lb);
    if (!(denom > 0.0f) || !isfinite(denom)) return;
    const float limit = mu * jn;
    const float jt = fminf(slip / denom, limit);
    if (!(jt > 0.0f)) return;
    NumiClothBagGPUFruit na = a, nb = b;
    const V3 onB = tangent * -jt;
    applyFruitImpulse(nb, onB, ob);
    applyFruitImpulse(na, -onB, oa);
    if (!finite4(na.velocityAndGroundImpulse) || !finite4(nb.velocityAndGroundImpulse) ||
        !finite4(na.angularVelocity) || !finite4(nb.angularVelocity)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    a = na;
    b = nb;
    wrote = true;
    recordFriction(status, 0, jt, limit);
}

// All fruit-pair friction colours in one workgroup, fruits in LDS.
__global__ void k_fruit_pairs_friction(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                                       const uint2* __restrict__ colours, uint32_t colourCount,
                                       NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                                       const NumiClothBagGPUFruitPair* __restrict__ pairs,
                                       NumiClothBagGPUFrictionStatus* __restrict__ status,
                                       uint32_t* __restrict__ failure) {
    __shared__ NumiClothBagGPUFruit local[32];
    if (*ok == 0u) return;
    for (uint32_t i = threadIdx.x; i < fruitCount; i += blockDim.x) local[i] = fruits[i];
    __syncthreads();
    for (uint32_t c = 0; c < colourCount; ++c) {
        const uint2 range = colours[c];
        for (uint32_t t = threadIdx.x; t < range.y; t += blockDim.x) {
            const NumiClothBagGPUFruitPair pair = pairs[range.x + t];
            NumiClothBagGPUFruit a = local[pair.fruitsAndColor.x];
            NumiClothBagGPUFruit b = local[pair.fruitsAndColor.y];
            bool wrote = false;
            fruitPairFriction(cfg, a, b, pair, status, failure, wrote);
            if (wrote) {
                local[pair.fruitsAndColor.x] = a;
                local[pair.fruitsAndColor.y] = b;
            }
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < fruitCount; i += blockDim.x) fruits[i] = local[i];
}

__global__ void k_fruit_ground_friction(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                                        NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                                        NumiClothBagGPUFrictionStatus* __restrict__ status,
                                        uint32_t* __restrict__ failure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fruitCount || *ok == 0u || cfg->constraintCounts.z == 0u) return;
    NumiClothBagGPUFruit f = fruits[i];
    const float jn = f.velocityAndGroundImpulse.w;
    const float mu = cfg->fruitMaterial.y;
    const float rolling = cfg->fruitMaterial.z;
    if (!(mu >= 0.0f) || !(rolling >= 0.0f) || !isfinite(mu) || !isfinite(rolling)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    if (!(jn > 0.0f)) return;
    const V3 n = v3(0.0f, 0.0f, 1.0f);
    const V3 offset = v3(0.0f, 0.0f, -f.previousAndRadius.w);
    const V3 cv = xyz(f.velocityAndGroundImpulse) + cross(xyz(f.angularVelocity), offset);
    const V3 tv = cv - n * dot(cv, n);
    const float slip = length(tv);
    if (slip > 1.0e-10f) {
        const V3 t = tv / slip;
        const V3 lever = cross(offset, t);
        const float denom = f.positionAndInverseMass.w + fruitInverseInertia(f) * dot(lever, lever);
        const float limit = mu * jn;
        const float jt = fminf(slip / denom, limit);
        if (jt > 0.0f) {
            applyFruitImpulse(f, t * -jt, offset);
            recordFriction(status, 3, jt, limit);
        }
    }
    const V3 roll = v3(f.angularVelocity.x, f.angularVelocity.y, 0.0f);
    const float rollSpeed = length(roll);
    if (rollSpeed > 1.0e-12f && rolling > 0.0f) {
        const float inertia = fruitInverseInertia(f);
        const float required = rollSpeed / inertia;
        const float limit = rolling * f.previousAndRadius.w * jn;
        const float angular = fminf(required, limit);
        setXyz(f.angularVelocity, xyz(f.angularVelocity) - roll * (angular * inertia / rollSpeed));
        atomicAdd(&status->metrics.z, 1u);
        if (limit > 0.0f) atomicMax(&status->metrics.y, __float_as_uint(angular / limit));
    }
    if (!finite4(f.velocityAndGroundImpulse) || !finite4(f.angularVelocity)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    fruits[i] = f;
}

// q += dt/2 (w x q) then normalise (reference :2944-2974).
__global__ void k_fruit_orientation(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                                    NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                                    uint32_t* __restrict__ failure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fruitCount || *ok == 0u) return;
    NumiClothBagGPUFruit f = fruits[i];
    const V3 w = xyz(f.angularVelocity);
    const mr_float4 q = f.orientation;
    const V3 qv = v3(q.x, q.y, q.z);
    const V3 dv = w * q.w + cross(w, qv);
    const float dw = -dot(w, qv);
    const float h = 0.5f * cfg->gravityAndTimestep.w;
    const float4 adv = make_float4(fmaf(dv.x, h, q.x), fmaf(dv.y, h, q.y), fmaf(dv.z, h, q.z), fmaf(dw, h, q.w));
    const float m2 = adv.x * adv.x + adv.y * adv.y + adv.z * adv.z + adv.w * adv.w;
    if (!isfinite(adv.x) || !isfinite(adv.y) || !isfinite(adv.z) || !isfinite(adv.w) || !(m2 > 1.0e-20f)) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    const float s = 1.0f / sqrtf(m2);
    f.orientation = mr_float4{adv.x * s, adv.y * s, adv.z * s, adv.w * s};
    fruits[i] = f;
}

// --------------------------------------------------------- mouth release --

__device__ bool insideMouth(const float4* __restrict__ pos, const uint32_t* __restrict__ rim, uint32_t n, V3 centre,
                            V3 tangent, V3 bitangent, V3 point, float clearance) {
    const V3 rp = point - centre;
    const float px = dot(rp, tangent), py = dot(rp, bitangent);
    bool inside = false;
    float nearest = INFINITY;
    uint32_t second = n - 1;
    for (uint32_t first = 0; first < n; ++first) {
        const V3 a = xyz(pos[rim[first]]) - centre;
        const V3 b = xyz(pos[rim[second]]) - centre;
        const float ax = dot(a, tangent), ay = dot(a, bitangent);
        const float bx = dot(b, tangent), by = dot(b, bitangent);
        const float ex = bx - ax, ey = by - ay;
        const float l2 = ex * ex + ey * ey;
        const float t = l2 > 1.0e-20f ? clampf(((px - ax) * ex + (py - ay) * ey) / l2, 0.0f, 1.0f) : 0.0f;
        const float sx = px - (ax + t * ex), sy = py - (ay + t * ey);
        nearest = fminf(nearest, sx * sx + sy * sy);
        if ((ay > py) != (by > py)) {
            const float cx = ax + (bx - ax) * (py - ay) / (by - ay);
            if (px < cx) inside = !inside;
        }
        second = first;
    }
    return inside || nearest <= clearance * clearance;
}

__device__ float rimDistance(const float4* __restrict__ pos, const uint32_t* __restrict__ rim, uint32_t n, V3 p) {
    float nearest = INFINITY;
    for (uint32_t k = 0; k < n; ++k) {
        const V3 a = xyz(pos[rim[k]]);
        const V3 e = xyz(pos[rim[(k + 1) % n]]) - a;
        const float l2 = dot(e, e);
        const float t = l2 > 1.0e-20f ? clampf(dot(p - a, e) / l2, 0.0f, 1.0f) : 0.0f;
        const V3 s = p - (a + e * t);
        nearest = fminf(nearest, dot(s, s));
    }
    return sqrtf(nearest);
}

// One workgroup, one lane per fruit (update_released_fruit,
// reference :1025-1146): mouth plane from the rim, polygon test in the mouth
// plane, release when the fruit clears the cap or the rim by the hysteresis.
__global__ void k_release(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                          const float4* __restrict__ pos, const uint32_t* __restrict__ rim,
                          const NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                          NumiClothBagGPUReleaseStatus* __restrict__ status, uint32_t* __restrict__ maximumClearance,
                          uint32_t* __restrict__ failure) {
    const uint32_t i = threadIdx.x;
    if (*ok == 0u || i >= fruitCount || cfg->mouthControl.x == 0u) return;
    const uint32_t n = cfg->mouthControl.x;
    if (n < 3u || i >= 32u) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_RANGE);
        return;
    }
    V3 centre = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < n; ++k) centre = centre + xyz(pos[rim[k]]);
    centre = centre / static_cast<float>(n);
    V3 area = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < n; ++k) area = area + cross(xyz(pos[rim[k]]) - centre, xyz(pos[rim[(k + 1) % n]]) - centre);
    if (!(length(area) > 1.0e-8f)) return;
    V3 normal = area / length(area);
    if (dot(normal, centre - xyz(pos[cfg->mouthControl.y])) < 0.0f) normal = -normal;
    V3 tangent = xyz(pos[rim[0]]) - centre;
    tangent = tangent - normal * dot(tangent, normal);
    if (!(length(tangent) > 1.0e-8f)) {
        tangent = cross(normal, fabsf(normal.z) < 0.9f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f));
    }
    tangent = tangent / length(tangent);
    V3 bitangent = cross(normal, tangent);
    bitangent = bitangent / length(bitangent);
    const NumiClothBagGPUFruit f = fruits[i];
    const V3 p = xyz(f.positionAndInverseMass);
    const float required = f.previousAndRadius.w + cfg->clothMaterial.x;
    const bool insideProjection = insideMouth(pos, rim, n, centre, tangent, bitangent, p, 0.0f);
    const bool nearProjection = insideMouth(pos, rim, n, centre, tangent, bitangent, p, required);
    const float outward = dot(p - centre, normal);
    const uint32_t mask = 1u << i;
    if (insideProjection && outward > 0.0f) atomicOr(&status->masks.x, mask);
    const float axial = outward - required;
    const float rimClear = rimDistance(pos, rim, n, p) - required;
    const bool throughCap = nearProjection && axial > cfg->mouthMaterial.x;
    const uint32_t candidates = atomicOr(&status->masks.x, 0u);
    const bool aroundRim = (candidates & mask) != 0u && !insideProjection && outward > 0.0f &&
                           rimClear > cfg->mouthMaterial.x;
    const float clearance = throughCap ? axial : (aroundRim ? rimClear : axial);
    atomicMax(maximumClearance + i, __float_as_uint(fmaxf(0.0f, clearance)));
    if (throughCap || aroundRim) atomicOr(&status->masks.y, mask);
}

}  // namespace numi_dev
