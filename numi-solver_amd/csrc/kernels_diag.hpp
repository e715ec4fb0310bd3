#pragma once

// Substep bookkeeping and residual diagnostics.
//
// k_prepare plays the role of prepare_trajectory_substep (reference
// src/metal/ClothBag.metal:466-489): it copies the next per-substep
// configuration into the live slot every other kernel reads. The index comes
// from a device cursor that k_prepare advances itself, so the same captured
// substep graph is replayed for every substep of a step() call.
//
// Diagnostics reduce maxima with wavefront shuffles (64 lanes), then one
// partial per workgroup, then a single fixed-order workgroup pass: no float
// atomics, identical bits on every run.

#include "device_math.hpp"

namespace numi_dev {

struct SubstepSlot {
    NumiClothBagGPUConfig config;   // live configuration
    uint32_t ok;                    // validConfig result
    uint32_t cursor;                // next substep index
    uint32_t substep;               // index of the live substep
    uint32_t reserved;
    double time;                    // time at the end of the live substep
};

__device__ bool configValid(const NumiClothBagGPUConfig& c, uint32_t* failure) {
    if (c.control.x != NUMI_CLOTH_BAG_GPU_ABI_VERSION) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_ABI);
        return false;
    }
    const mr_float4& q = c.gripOrientation;
    const float qn = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    bool bad = !(c.gravityAndTimestep.w > 0.0f) || !isfinite(c.gravityAndTimestep.w) ||
               !finite3(xyz(c.gravityAndTimestep)) || !finite3(xyz(c.gripTargetAndActive)) || !finite4(q);
    bad = bad || (c.gripTargetAndActive.w > 0.0f &&
                  (fabsf(qn - 1.0f) > 1.0e-4f || c.gripControl.x == 0u || !(c.gripMaterial.x > 0.0f)));
    bad = bad || c.gripControl.y > 1u;
    bad = bad || (c.gripControl.y == 1u &&
                  (c.gripControl.z < 2u * c.gripControl.w + 1u ||
                   c.gripControl.z > NUMI_CLOTH_BAG_GPU_MOUTH_RIM_CAPACITY || c.gripControl.w > 16u ||
                   c.control.w != 2u * (2u * c.gripControl.w + 1u)));
    bad = bad || !finite4(c.gripMaterial) || c.gripMaterial.x < 0.0f;
    bad = bad || !finite4(c.clothMaterial) || c.clothMaterial.x < 0.0f;
    bad = bad || !finite4(c.fruitMaterial);
    bad = bad || !finite4(c.airVelocityAndDensity) || c.airVelocityAndDensity.w < 0.0f;
    bad = bad || !finite4(c.aerodynamicCoefficients) || c.aerodynamicCoefficients.x < 0.0f ||
          c.aerodynamicCoefficients.y < 0.0f || c.aerodynamicCoefficients.z < 0.0f ||
          c.aerodynamicCoefficients.w < 0.0f;
    bad = bad || c.mouthControl.x > NUMI_CLOTH_BAG_GPU_MOUTH_RIM_CAPACITY ||
          (c.mouthControl.x != 0u && c.mouthControl.y >= c.control.y);
    bad = bad || !finite4(c.mouthMaterial) || c.mouthMaterial.x < 0.0f;
    if (bad) recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
    return !bad;
}

// One thread. An ABI-mismatched config is not copied (the previous live
// configuration stays), exactly as the reference's prepare kernel does.
__global__ void k_prepare(const NumiClothBagGPUConfig* __restrict__ configs, const double* __restrict__ times,
                          SubstepSlot* __restrict__ slot, uint32_t* __restrict__ failure) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint32_t s = slot->cursor;
    const NumiClothBagGPUConfig next = configs[s];
    if (next.control.x != NUMI_CLOTH_BAG_GPU_ABI_VERSION) {
        recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_ABI);
    } else {
        slot->config = next;
    }
    slot->ok = configValid(slot->config, failure) ? 1u : 0u;
    slot->time = times[s];
    slot->substep = s;
    slot->cursor = s + 1u;
}

__global__ void k_reset_cursor(SubstepSlot* slot) {
    if (blockIdx.x == 0 && threadIdx.x == 0) slot->cursor = 0u;
}

__global__ void k_fill_u32(uint32_t* __restrict__ data, uint32_t count, uint32_t value) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) data[i] = value;
}

__device__ __forceinline__ float waveMax(float v) {
    for (int offset = 32; offset > 0; offset >>= 1) v = fmaxf(v, __shfl_xor(v, offset));
    return v;
}

// Block maximum of up to four values per thread; lane 0 of wave 0 writes.
__device__ void blockMax4(float v[4], float4* out) {
    __shared__ float partial[4][kMaxWaves];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    for (int k = 0; k < 4; ++k) {
        const float m = waveMax(v[k]);
        if (lane == 0) partial[k][wave] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint32_t waves = (blockDim.x + 63u) >> 6;
        for (int k = 0; k < 4; ++k)
            for (uint32_t w = 0; w < waves; ++w) r[k] = fmaxf(r[k], partial[k][w]);
        *out = make_float4(r[0], r[1], r[2], r[3]);
    }
}

// Per-workgroup partial maxima:
//  x strain violation over distances, y ground penetration over knots,
//  z static-sphere penetration over knots, w unused here.
__global__ void k_diag_partial(const float4* __restrict__ pos, uint32_t particles,
                               const uint2* __restrict__ ij, const float* __restrict__ maxLength,
                               uint32_t distances, const float4* __restrict__ spheres, uint32_t sphereCount,
                               float yarnRadius, float4* __restrict__ partials) {
    const uint32_t stride = gridDim.x * blockDim.x;
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < distances; c += stride) {
        const uint2 e = ij[c];
        const float len = length(xyz(pos[e.y]) - xyz(pos[e.x]));
        v[0] = fmaxf(v[0], len - maxLength[c]);
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < particles; i += stride) {
        const float4 p = pos[i];
        v[1] = fmaxf(v[1], yarnRadius - p.z);
        if (p.w > 0.0f) {
            for (uint32_t k = 0; k < sphereCount; ++k) {
                const float4 s = spheres[k];
                v[2] = fmaxf(v[2], s.w + yarnRadius - length(xyz(p) - xyz(s)));
            }
        }
    }
    blockMax4(v, partials + blockIdx.x);
}

__global__ void k_diag_final(const float4* __restrict__ partials, uint32_t count, float4* __restrict__ out) {
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
        const float4 p = partials[i];
        v[0] = fmaxf(v[0], p.x);
        v[1] = fmaxf(v[1], p.y);
        v[2] = fmaxf(v[2], p.z);
        v[3] = fmaxf(v[3], p.w);
    }
    blockMax4(v, out);
}

}  // namespace numi_dev
