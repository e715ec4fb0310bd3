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
    uint32_t leadingCollide;    // 1: the ground/sphere pass before the first colour pass
    uint32_t trailingCollides;
    uint32_t* failure;
    uint32_t kinds[kMaxTilePasses];
};

// Ground/sphere pass over every loaded knot (owned and halo: the pass acts on
// one knot at a time, so halo knots stay equal to their owners' values).
// Non-finite knots are skipped and reported by their owning tile only.
__device__ void tileCollide(float4* lds, uint32_t loaded, uint32_t owned, const TileProgramArgs& a, bool ground,
                            float radius, const float4* sphereCache, float4* culled, SphereWindow& window,
                            float4* scratch) {
    if (a.sphereCount != 0) {
        V3 lo = v3(INFINITY, INFINITY, INFINITY), hi = v3(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t i = threadIdx.x; i < loaded; i += blockDim.x) {
            const float4 p = lds[i];
            if (!isfinite(p.z)) continue;
            lo = v3(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z));
            hi = v3(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z));
        }
        buildSphereWindow(lo, hi, radius, a.spheres, a.sphereCount, culled, window, scratch);
    }
    for (uint32_t i = threadIdx.x; i < loaded; i += blockDim.x) {
        float4 p = lds[i];
        if (!isfinite(p.z)) {
            if (i < owned) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
            continue;
        }
        const bool moved = a.sphereCount != 0
                               ? collideKnotWindow(p, ground, radius, window, culled, sphereCache, a.sphereCount)
                               : collideKnot(p, ground, radius, sphereCache, 0u);
        if (moved) lds[i] = p;
    }
    __syncthreads();
}

constexpr uint32_t kTilePrefetch = 4;  // item slots per thread and pass held in registers

__device__ __forceinline__ uint32_t tileItemConstraint(uint32_t kind, const uint4& item) {
    return (kind == kTileKnotKind ? item.z : item.y) & 0x7fffffffu;
}

__device__ __forceinline__ bool tileItemOwner(uint32_t kind, const uint4& item) {
    return ((kind == kTileKnotKind ? item.z : item.y) >> 31) != 0u && kind != kTileStrainKind;
}

__device__ __forceinline__ float tileItemLambda(const TileProgramArgs& a, uint32_t kind, const uint4& item) {
    const uint32_t c = tileItemConstraint(kind, item);
    if (kind == kTileDistanceKind) return a.distanceLambdaIn[c];
    if (kind == kTileBendKind) return a.bendLambdaIn[c];
    if (kind == kTileKnotKind) return a.knotLambdaIn[c];
    return 0.0f;
}

// One constraint of a pass on the LDS knots; lam is its multiplier from the
// input buffer, written to the output buffer when this tile owns it.
__device__ __forceinline__ void tileItem(float4* lds, const TileProgramArgs& a, uint32_t kind, const uint4& item,
                                         float lam, bool live, float dt, bool ground, float floorZ) {
    const uint32_t la = item.x & 0xffffu, lb = item.x >> 16;
    const uint32_t c = tileItemConstraint(kind, item);
    if (kind == kTileKnotKind) {
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
        if (tileItemOwner(kind, item)) a.knotLambdaOut[c] = lam;
    } else if (kind == kTileDistanceKind) {
        if (live) {
            float4 p = lds[la], q = lds[lb];
            const float2 rc = make_float2(__uint_as_float(item.z), __uint_as_float(item.w));
            if (projectDistance(p, q, rc, lam, dt, a.failure)) {
                lds[la] = p;
                lds[lb] = q;
            }
        }
        if (tileItemOwner(kind, item)) a.distanceLambdaOut[c] = lam;
    } else if (kind == kTileBendKind) {
        if (live) {
            float4 p = lds[la], q = lds[lb];
            const float2 cc = make_float2(__uint_as_float(item.z), __uint_as_float(item.w));
            if (projectBend(p, q, cc, lam, dt, ground, floorZ, a.failure)) {
                lds[la] = p;
                lds[lb] = q;
            }
        }
        if (tileItemOwner(kind, item)) a.bendLambdaOut[c] = lam;
    } else if (live) {  // strain limit
        float4 p = lds[la], q = lds[lb];
        if (limitStrainPair(p, q, __uint_as_float(item.z), a.failure)) {
            lds[la] = p;
            lds[lb] = q;
        }
    }
}

// The first kTilePrefetch item slots of a thread in every pass are software
// pipelined: the items of pass k + 2 and the multipliers of pass k + 1 are
// requested before pass k runs, so their latency overlaps the LDS work and
// the barrier instead of following each other. Slots beyond those (tiles
// with more than kTilePrefetch x workgroup items in one pass) load in place.
__global__ void k_tile_program(TileProgramArgs a) {
    extern __shared__ float4 lds[];
    __shared__ SphereWindow window;
    __shared__ float4 scratch[2 * kMaxBlockWaves];
    __shared__ uint32_t starts[kMaxTilePasses + 1];
    const uint32_t tile = blockIdx.x;
    const uint32_t base = a.loadStart[tile];
    const uint32_t loaded = a.loadStart[tile + 1] - base;
    const uint32_t owned = a.ownedCount[tile];
    const uint32_t passes = a.passes;
    float4* sphereCache = lds + loaded;
    float4* culled = sphereCache + a.sphereCount;
    const bool collides = a.leadingCollide != 0 || a.trailingCollides != 0;
    for (uint32_t k = threadIdx.x; k <= passes; k += blockDim.x) starts[k] = a.itemStart[tile * passes + k];
    // Items of passes 0 and 1, multipliers of pass 0 (issued before the knot
    // gather so the two overlap).
    const uint32_t t = threadIdx.x, stride = blockDim.x;
    uint4 cur[kTilePrefetch], next[kTilePrefetch], ahead[kTilePrefetch];
    float lam[kTilePrefetch], nextLam[kTilePrefetch];
    auto loadItems = [&](uint32_t k, uint4* dst) {
        if (k >= passes) return;
        const uint32_t first = a.itemStart[tile * passes + k], last = a.itemStart[tile * passes + k + 1];
#pragma unroll
        for (uint32_t m = 0; m < kTilePrefetch; ++m) {
            const uint32_t j = first + t + m * stride;
            dst[m] = j < last ? a.items[j] : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto loadLambdas = [&](uint32_t k, const uint4* src, float* dst) {
        if (k >= passes) return;
        const uint32_t kind = a.kinds[k];
        const uint32_t first = a.itemStart[tile * passes + k], last = a.itemStart[tile * passes + k + 1];
#pragma unroll
        for (uint32_t m = 0; m < kTilePrefetch; ++m)
            dst[m] = first + t + m * stride < last ? tileItemLambda(a, kind, src[m]) : 0.0f;
    };
    loadItems(0, cur);
    loadItems(1, next);
    loadLambdas(0, cur, lam);
    for (uint32_t i = t; i < loaded; i += stride) lds[i] = a.posIn[a.loadIds[base + i]];
    if (collides)
        for (uint32_t k = t; k < a.sphereCount; k += stride) sphereCache[k] = a.spheres[k];
    __syncthreads();
    const bool live = *a.ok != 0u;
    const float dt = a.cfg->gravityAndTimestep.w;
    const bool ground = a.cfg->constraintCounts.z != 0u;
    const float floorZ = a.cfg->clothMaterial.x;
    if (live && a.leadingCollide != 0)
        tileCollide(lds, loaded, owned, a, ground, floorZ, sphereCache, culled, window, scratch);
    for (uint32_t k = 0; k < passes; ++k) {
        loadLambdas(k + 1, next, nextLam);
        loadItems(k + 2, ahead);
        const uint32_t kind = a.kinds[k];
        const uint32_t first = starts[k], last = starts[k + 1];
#pragma unroll
        for (uint32_t m = 0; m < kTilePrefetch; ++m)
            if (first + t + m * stride < last) tileItem(lds, a, kind, cur[m], lam[m], live, dt, ground, floorZ);
        for (uint32_t j = first + t + kTilePrefetch * stride; j < last; j += stride) {
            const uint4 item = a.items[j];
            tileItem(lds, a, kind, item, tileItemLambda(a, kind, item), live, dt, ground, floorZ);
        }
        __syncthreads();
#pragma unroll
        for (uint32_t m = 0; m < kTilePrefetch; ++m) {
            cur[m] = next[m];
            next[m] = ahead[m];
            lam[m] = nextLam[m];
        }
    }
    if (live)
        for (uint32_t pass = 0; pass < a.trailingCollides; ++pass)
            tileCollide(lds, loaded, owned, a, ground, floorZ, sphereCache, culled, window, scratch);
    for (uint32_t i = t; i < owned; i += stride) a.posOut[a.loadIds[base + i]] = lds[i];
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
