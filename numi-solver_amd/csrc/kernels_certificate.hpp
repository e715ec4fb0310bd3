#pragma once

// Residuals of the data-dependent certificate loop that the independent CPU
// reference runs at the last substep of a frame (tools/cloth_bag.cpp:
// 5143-5185): after each certificate pass it measures
//   x  strain-limit violation, max(len - rest (1 + limit)) over distances
//      (measureStrainLimitViolation, :1765-1779);
//   y  ground penetration, max(r - z) over knots and max(radius - z) over
//      fruits, 0 without the ground (measureGroundPenetration, :1880-1894);
//   z  fruit/yarn penetration, max(radius + r - |centre - segment|) over
//      fruits and distance segments (measureBallYarnPenetration, :3111-3137;
//      its bounding-box test skips only pairs whose term is negative);
//   w  segment self-penetration, max(2r - segment distance) over the self
//      candidates of the substep's broad phase (measurePrimitiveSelfPenetration,
//      :4182-4245; pairs outside the broad phase's margin cannot be within 2r);
// each clamped below at 0 as the reference starts its maxima at 0. Partial
// maxima per workgroup (fmaxf drops NaN as std::max does), reduced by
// k_diag_final.

#include "kernels_bag.hpp"
#include "kernels_diag.hpp"
#include "kernels_self.hpp"

namespace numi_dev {

__global__ void k_certificate_partial(const NumiClothBagGPUConfig* __restrict__ cfg, const float4* __restrict__ pos,
                                      uint32_t particles, const uint2* __restrict__ ij,
                                      const float* __restrict__ maxLength, uint32_t distances,
                                      const NumiClothBagGPUFruit* __restrict__ fruits, uint32_t fruitCount,
                                      const SelfGrid* __restrict__ grid, const uint2* __restrict__ pairs,
                                      float4* __restrict__ partials) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
    const float r = cfg->clothMaterial.x;
    const bool ground = cfg->constraintCounts.z != 0u;
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t c = first; c < distances; c += stride) {
        const uint2 e = ij[c];
        v[0] = fmaxf(v[0], length(xyz(pos[e.y]) - xyz(pos[e.x])) - maxLength[c]);
    }
    if (ground) {
        for (uint32_t i = first; i < particles; i += stride) v[1] = fmaxf(v[1], r - pos[i].z);
        for (uint32_t f = first; f < fruitCount; f += stride)
            v[1] = fmaxf(v[1], fruits[f].previousAndRadius.w - fruits[f].positionAndInverseMass.z);
    }
    const uint64_t visits = static_cast<uint64_t>(fruitCount) * distances;
    for (uint64_t k = first; k < visits; k += stride) {
        const uint32_t f = static_cast<uint32_t>(k / distances), s = static_cast<uint32_t>(k % distances);
        const NumiClothBagGPUFruit& fruit = fruits[f];
        const uint2 e = ij[s];
        const float target = fruit.previousAndRadius.w + r;
        const PointSample ps = samplePointSegment(xyz(fruit.positionAndInverseMass), xyz(pos[e.x]), xyz(pos[e.y]));
        v[2] = fmaxf(v[2], target - ps.distance);
    }
    const uint32_t pairCount = grid != nullptr && grid->valid ? grid->pairCount : 0u;
    for (uint32_t k = first; k < pairCount; k += stride) {
        const uint2 pr = pairs[k];
        const uint2 a = ij[pr.x], b = ij[pr.y];
        const SegmentSample c = sampleSegments(xyz(pos[a.x]), xyz(pos[a.y]), xyz(pos[b.x]), xyz(pos[b.y]));
        v[3] = fmaxf(v[3], 2.0f * r - c.distance);
    }
    blockMax4(v, partials + blockIdx.x);
}

}  // namespace numi_dev
