#pragma once

// Two passes of the independent CPU reference's schedule (tools/cloth_bag.cpp)
// that the Metal path lacks, enabled per scene (NumiClothHipSceneExtension
// sweptGround / pointSelfCollision).
//
// Swept ground (reference sweepGroundPrediction, :2692-2721): after the
// advance, a knot or fruit that started the substep at or above contact
// height and is predicted below it goes back to contact height; a fruit adds
// the removed advance / (w dt) to its ground impulse.
//
// Point self-collision (reference solveSelfCollision, :4249-4315): knots i < j
// with 1e-12 <= |x_j - x_i| < 2r, j outside i's two-hop neighbourhood and
// w_i + w_j > 0, each pair corrected by d * (2r - |d|) / (|d| (w_i + w_j)),
// -w_i of it on i and +w_j on j. Every pair is evaluated on one position
// snapshot; corrections are summed per knot in 2^-32 m fixed point (integer
// atomics, order independent) and each knot moves by their mean. The pairs
// are found on a hashed uniform grid of cells a little over 2r wide, so a
// pair within 2r always lies in neighbouring cells despite FP32 rounding of
// the cell coordinates; the grid only accelerates the pair set, which the
// distance test defines.

#include "kernels_self.hpp"

namespace numi_dev {

constexpr float kPointCellScale = 1.0009765625f;  // cell = 2r (1 + 2^-10)

__global__ void k_swept_ground(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                               float4* __restrict__ pos, const float4* __restrict__ prev, uint32_t particles,
                               NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (*ok == 0u || cfg->constraintCounts.z == 0u) return;
    const float r = cfg->clothMaterial.x;
    if (i < particles) {
        float4 x = pos[i];
        if (prev[i].z >= r && x.z < r) {
            x.z = r;
            pos[i] = x;
        }
    }
    if (i < fruitCount) {
        NumiClothBagGPUFruit f = fruits[i];
        const float radius = f.previousAndRadius.w;
        if (f.previousAndRadius.z >= radius && f.positionAndInverseMass.z < radius) {
            const float removed = radius - f.positionAndInverseMass.z;
            f.velocityAndGroundImpulse.w += removed / (f.positionAndInverseMass.w * cfg->gravityAndTimestep.w);
            f.positionAndInverseMass.z = radius;
            fruits[i] = f;
        }
    }
}

struct PointCell {
    int x, y, z;
    bool valid;
};

__device__ __forceinline__ PointCell pointCell(const float4& p, float inverseCell) {
    const float fx = floorf(p.x * inverseCell), fy = floorf(p.y * inverseCell), fz = floorf(p.z * inverseCell);
    // cells beyond +-2^30 (or non-finite knots) take no part
    const bool valid = fabsf(fx) < 1073741824.0f && fabsf(fy) < 1073741824.0f && fabsf(fz) < 1073741824.0f;
    return {valid ? static_cast<int>(fx) : 0, valid ? static_cast<int>(fy) : 0, valid ? static_cast<int>(fz) : 0,
            valid};
}

__device__ __forceinline__ uint32_t pointBucket(int x, int y, int z, uint32_t mask) {
    return ((static_cast<uint32_t>(x) * 73856093u) ^ (static_cast<uint32_t>(y) * 19349663u) ^
            (static_cast<uint32_t>(z) * 83492791u)) &
           mask;
}

__device__ __forceinline__ float pointInverseCell(const NumiClothBagGPUConfig* cfg) {
    return 1.0f / (2.0f * cfg->clothMaterial.x * kPointCellScale);
}

// Clears the bucket counts and the contact counter of the pass.
__global__ void k_point_clear(uint32_t* __restrict__ counts, uint32_t buckets, uint32_t* __restrict__ contacts) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < buckets; b += gridDim.x * blockDim.x) counts[b] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) *contacts = 0u;
}

__global__ void k_point_count(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                              const float4* __restrict__ pos, uint32_t particles, uint32_t mask,
                              uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= particles || *ok == 0u) return;
    const PointCell c = pointCell(pos[i], pointInverseCell(cfg));
    if (c.valid) atomicAdd(&counts[pointBucket(c.x, c.y, c.z, mask)], 1u);
}

// After the exclusive scan of the counts: each knot takes a slot of its
// bucket (any order; the solve does not depend on it). Afterwards ends[b] is
// the end of bucket b.
__global__ void k_point_place(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                              const float4* __restrict__ pos, uint32_t particles, uint32_t mask,
                              uint32_t* __restrict__ ends, uint32_t* __restrict__ ids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= particles || *ok == 0u) return;
    const PointCell c = pointCell(pos[i], pointInverseCell(cfg));
    if (c.valid) ids[atomicAdd(&ends[pointBucket(c.x, c.y, c.z, mask)], 1u)] = i;
}

// Every pair (i < j) within 2r, from knot i: the knots of the 27 cells around
// i's cell whose own cell is that neighbour (two cells may share a bucket).
__global__ void k_point_solve(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                              const float4* __restrict__ pos, uint32_t particles, uint32_t mask,
                              const uint32_t* __restrict__ ends, const uint32_t* __restrict__ ids,
                              const uint32_t* __restrict__ localStart, const uint32_t* __restrict__ localList,
                              FixedAccumulator* __restrict__ acc, uint32_t* __restrict__ touched,
                              uint32_t* __restrict__ contacts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= particles || *ok == 0u) return;
    const float inverseCell = pointInverseCell(cfg);
    const float4 a = pos[i];
    const PointCell c = pointCell(a, inverseCell);
    if (!c.valid) return;
    const float target = 2.0f * cfg->clothMaterial.x;
    const uint32_t l0 = localStart[i], l1 = localStart[i + 1];
    uint32_t found = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = c.x + dx, ny = c.y + dy, nz = c.z + dz;
                const uint32_t b = pointBucket(nx, ny, nz, mask);
                const uint32_t first = b == 0u ? 0u : ends[b - 1u], last = ends[b];
                for (uint32_t k = first; k < last; ++k) {
                    const uint32_t j = ids[k];
                    if (j <= i) continue;
                    const float4 q = pos[j];
                    const PointCell cj = pointCell(q, inverseCell);
                    if (cj.x != nx || cj.y != ny || cj.z != nz) continue;
                    const V3 d = xyz(q) - xyz(a);
                    const float len = length(d);
                    if (!(len < target) || len < 1.0e-12f) continue;
                    const float denom = a.w + q.w;
                    if (!(denom > 0.0f)) continue;
                    if (inList(localList, l0, l1, j)) continue;
                    const V3 correction = d * ((target - len) / (len * denom));
                    addFixed(acc, touched, i, correction * -a.w);
                    addFixed(acc, touched, j, correction * q.w);
                    ++found;
                }
            }
    if (found != 0u) atomicAdd(contacts, found);
}

}  // namespace numi_dev
