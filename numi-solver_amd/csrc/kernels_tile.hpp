#pragma once

// LDS-resident tile programs (see tile_plan.hpp for the plan and its
// exactness argument). One workgroup per tile:
//   1. gather the tile's loaded knots (owned + halo) from posIn into LDS,
//   2. run every pass of the program on LDS, one item per thread slot, with a
//      workgroup barrier between colours; multipliers are read from the
//      input buffer and written to the output buffer by the owning tile only,
//   3. run the trailing per-knot ground/sphere passes,
//   4. write the owned knots to posOut.
// Inputs and outputs are separate buffers (ping-pong across launches), so a
// tile never reads a value another tile of the same launch has written.
// The per-constraint arithmetic is the same device function the per-colour
// kernels call, so results are bit-identical to the per-colour schedule.

#include "kernels_core.hpp"

namespace numi_dev {

constexpr uint32_t kMaxTilePasses = 64;
constexpr uint32_t kTileDistanceKind = 0;  // values of numi_hip::TilePassKind
constexpr uint32_t kTileKnotKind = 1;
constexpr uint32_t kTileBendKind = 2;
constexpr uint32_t kTileStrainKind = 3;

struct TileProgramArgs {
    const NumiClothBagGPUConfig* cfg;
    const uint32_t* ok;
    const float4* posIn;
    float4* posOut;
    const uint32_t* loadStart;
    const uint32_t* loadIds;
    const uint32_t* ownedCount;
    const uint32_t* itemStart;
    const uint4* items;
    const float2* distanceRestCompliance;
    const float* distanceMaxLength;
    const float* distanceLambdaIn;
    float* distanceLambdaOut;
    const float2* bendChordCompliance;
    const float* bendLambdaIn;
    float* bendLambdaOut;
    const float2* knotCosineCompliance;
    const float* knotLambdaIn;
    float* knotLambdaOut;
    const float4* spheres;
    uint32_t sphereCount;
    uint32_t passes;
    uint32_t trailingCollides;
    uint32_t* failure;
    uint32_t kinds[kMaxTilePasses];
};

__global__ void k_tile_program(TileProgramArgs a) {
    extern __shared__ float4 lds[];
    const uint32_t tile = blockIdx.x;
    const uint32_t base = a.loadStart[tile];
    const uint32_t loaded = a.loadStart[tile + 1] - base;
    const uint32_t owned = a.ownedCount[tile];
    float4* sphereCache = lds + loaded;
    for (uint32_t i = threadIdx.x; i < loaded; i += blockDim.x) lds[i] = a.posIn[a.loadIds[base + i]];
    if (a.trailingCollides != 0)
        for (uint32_t k = threadIdx.x; k < a.sphereCount; k += blockDim.x) sphereCache[k] = a.spheres[k];
    __syncthreads();
    const bool live = *a.ok != 0u;
    const float dt = a.cfg->gravityAndTimestep.w;
    const bool ground = a.cfg->constraintCounts.z != 0u;
    const float floorZ = a.cfg->clothMaterial.x;
    for (uint32_t k = 0; k < a.passes; ++k) {
        const uint32_t kind = a.kinds[k];
        const uint32_t first = a.itemStart[tile * a.passes + k];
        const uint32_t last = a.itemStart[tile * a.passes + k + 1];
        for (uint32_t j = first + threadIdx.x; j < last; j += blockDim.x) {
            const uint4 item = a.items[j];
            const uint32_t la = item.x & 0xffffu, lb = item.x >> 16;
            if (kind == kTileKnotKind) {
                const uint32_t c = item.z & 0x7fffffffu;
                float lam = a.knotLambdaIn[c];
                if (live) {
                    const uint32_t lc = item.y & 0xffffu, ld = item.y >> 16;
                    float4 p0 = lds[la], p1 = lds[lb], p2 = lds[lc], p3 = lds[ld];
                    if (projectKnot(p0, p1, p2, p3, a.knotCosineCompliance[c], lam, dt, a.failure)) {
                        lds[la] = p0;
                        lds[lb] = p1;
                        lds[lc] = p2;
                        lds[ld] = p3;
                    }
                }
                if ((item.z >> 31) != 0u) a.knotLambdaOut[c] = lam;
                continue;
            }
            const uint32_t c = item.y & 0x7fffffffu;
            const bool owner = (item.y >> 31) != 0u;
            if (kind == kTileDistanceKind) {
                float lam = a.distanceLambdaIn[c];
                if (live) {
                    float4 p = lds[la], q = lds[lb];
                    const float2 rc = make_float2(__uint_as_float(item.z), __uint_as_float(item.w));
                    if (projectDistance(p, q, rc, lam, dt, a.failure)) {
                        lds[la] = p;
                        lds[lb] = q;
                    }
                }
                if (owner) a.distanceLambdaOut[c] = lam;
            } else if (kind == kTileBendKind) {
                float lam = a.bendLambdaIn[c];
                if (live) {
                    float4 p = lds[la], q = lds[lb];
                    const float2 cc = make_float2(__uint_as_float(item.z), __uint_as_float(item.w));
                    if (projectBend(p, q, cc, lam, dt, ground, floorZ, a.failure)) {
                        lds[la] = p;
                        lds[lb] = q;
                    }
                }
                if (owner) a.bendLambdaOut[c] = lam;
            } else if (live) {  // strain limit
                float4 p = lds[la], q = lds[lb];
                if (limitStrainPair(p, q, __uint_as_float(item.z), a.failure)) {
                    lds[la] = p;
                    lds[lb] = q;
                }
            }
        }
        __syncthreads();
    }
    if (live) {
        for (uint32_t pass = 0; pass < a.trailingCollides; ++pass) {
            for (uint32_t i = threadIdx.x; i < loaded; i += blockDim.x) {
                float4 p = lds[i];
                if (!isfinite(p.z)) {
                    if (i < owned) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
                    continue;
                }
                if (collideKnot(p, ground, floorZ, sphereCache, a.sphereCount)) lds[i] = p;
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < owned; i += blockDim.x) a.posOut[a.loadIds[base + i]] = lds[i];
}

__global__ void k_copy_float4(const float4* __restrict__ from, float4* __restrict__ to, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) to[i] = from[i];
}

__global__ void k_copy_float(const float* __restrict__ from, float* __restrict__ to, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) to[i] = from[i];
}

}  // namespace numi_dev
