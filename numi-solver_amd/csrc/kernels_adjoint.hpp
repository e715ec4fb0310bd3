#pragma once

// Specimen kernels: a whole substep of one contact-free specimen (hanging
// strip) in one workgroup, every knot and multiplier resident in LDS, and
// its reverse-mode adjoint.
//
// Forward (k_specimen_forward): for each substep of the launch: begin
// (damping, gravity, drive) + advance, `iterations` x [distance colours,
// bend colours on even iterations], the strain-limit invocations of the
// reference schedule (1 + 8 + 9 when sweeps > 0), finalize; the loss
// sum_k |x_k - target_k|^2 of the tracked knots after the substep. The
// per-constraint arithmetic is the device functions the per-colour kernels
// use (projectDistance, projectBend without ground, limitStrainPair), in the
// same colour order, so the state equals the plain schedule's.
// In record mode every distance and bend pass first writes the positions it
// reads (the tape) and every iteration the multipliers it starts from. A
// strain pass in which a constraint acts tapes only what it moves: a flag
// per constraint and, for each constraint that moves, its two endpoints as
// the pass found them (a colour's constraints share no knot, so the knots'
// slots of the position tape hold them); the reverse reads the flags and the
// endpoints of flagged constraints only.
//
// Reverse (k_specimen_reverse): replays a segment of substeps backwards from
// the tape with the hand-derived adjoint of each projection (same formulas as
// the CPU oracle's cloth_adjoint.cpp) and of begin / advance / finalize,
// accumulating dL/d(compliance) per constraint, dL/dw per knot and dL/dc
// (damping) per knot; the state adjoint (dL/dx, dL/dv) is carried between
// segments in global memory.

#include "kernels_core.hpp"

namespace numi_dev {

constexpr uint32_t kSpecimenThreads = 1024;
constexpr uint32_t kSpecimenOwn = 5;  // knots per thread: specimens of <= 5120 knots

struct SpecimenInfo {
    uint32_t knotFirst, knotCount;
    uint32_t distanceCount, bendCount;  // the specimen's constraints (all colours)
    uint32_t trackedFirst, trackedCount;
    uint32_t pad[2];
};

struct SpecimenArgs {
    const SpecimenInfo* specimens;
    const uint4* distanceRanges;   // specimen x distanceColours: (first, count, local offset, -)
    const uint4* bendRanges;       // specimen x bendColours
    uint32_t distanceColours, bendColours;
    const NumiClothBagGPUConfig* configs;  // per substep of the trajectory
    const double* times;                   // per substep: time at its end
    uint32_t firstSubstep, substeps;       // range of this launch
    uint32_t iterations, sweeps;
    float4* pos;
    float4* prev;
    float4* vel;
    const float4* anchor;
    const float4* motion;          // per knot (amplitude, omega, phase, damping) or null
    float4 drive;                  // scene drive when motion is null
    float dampingFactor;           // scene damping factor when motion is null
    float dampingDt;
    const uint2* ij;
    const float2* restCompliance;
    const float* maxLength;
    float* dLambda;
    const uint2* bEnds;
    const float2* chordCompliance;
    float* bLambda;
    const uint32_t* tracked;       // merged knot indices, grouped by specimen
    const float* targets;          // substep x trackedTotal x 3 (weights when linear)
    uint32_t trackedTotal;
    uint32_t linear;               // 1: loss = sum w . x, 0: sum |x - target|^2
    double* loss;                  // per specimen
    uint32_t record;               // 0: loss only, 1: tape only, 2: tape and loss
    uint32_t passes;               // passes per substep (tape stride)
    uint32_t particles, distances, bends;
    float4* tapeX;                 // [substep in launch][pass][particles]
    float* tapeLambdaD;            // [substep in launch][iteration][distances]
    float* tapeLambdaB;            // [substep in launch][even iteration][bends]
    float4* tapeV0;                // [substep in launch][particles]
    uint32_t* tapeActive;          // [substep in launch][pass][specimen]
    uint8_t* tapeStrainActs;       // [substep in launch][strain pass][distances]: the constraint moved
    uint32_t* strainTaped;         // record mode: strain passes taped (acting), summed over specimens
    float* lossGrad;               // [substep in launch][trackedTotal][3]
    uint32_t specimenCount;
    uint32_t* failure;
    // reverse
    float4* gxState;               // dL/dx carried between segments
    float4* gvState;               // dL/dv
    float* gradInverseMass;        // dL/dw per knot (accumulated)
    float* gradDamping;            // dL/dc per knot (accumulated)
    float* gradDistance;           // dL/dc per distance (accumulated)
    float* gradBend;               // dL/dc per bend (accumulated)
};

// begin_substep + advance_positions for one knot, the arithmetic of k_begin.
__device__ __forceinline__ void specimenBegin(const SpecimenArgs& a, uint32_t i, float4& x, float4& v, V3 g, float dt,
                                              double endTime) {
    const V3 u = xyz(v);
    v.w = 0.0f;
    if (x.w > 0.0f) {
        float factor = a.dampingFactor;
        if (a.motion != nullptr) {
            const float c = a.motion[i].w;
            factor = c > 0.0f ? 1.0f / (1.0f + c * a.dampingDt) : 1.0f;
        }
        const V3 carried = factor != 1.0f ? u * factor : u;
        const V3 nu = fma3(g, dt, carried);
        v = withXyz(v, nu);
        v.w = nu.z;
        x = withXyz(x, fma3(nu, dt, xyz(x)));
    } else if (a.anchor != nullptr) {
        const float4 anchor = a.anchor[i];
        const float4 m = a.motion != nullptr ? a.motion[i] : a.drive;
        const float offset = driveOffset(m, endTime);
        x = withXyz(x, v3(anchor.x, anchor.y, anchor.z + offset));
    }
}

__device__ __forceinline__ uint32_t strainInvocations(uint32_t sweeps) { return sweeps == 0 ? 0u : 18u; }

__global__ void __launch_bounds__(kSpecimenThreads) k_specimen_forward(SpecimenArgs a) {
    extern __shared__ float4 specimenLds[];
    __shared__ double lossParts[kSpecimenThreads / 64];
    const uint32_t spec = blockIdx.x;
    const SpecimenInfo info = a.specimens[spec];
    const uint32_t tid = threadIdx.x, threads = blockDim.x;
    const uint32_t n = info.knotCount, k0 = info.knotFirst;
    float4* x = specimenLds;
    float* lamD = reinterpret_cast<float*>(x + n);
    float* lamB = lamD + info.distanceCount;
    const uint4* dRanges = a.distanceRanges + static_cast<size_t>(spec) * a.distanceColours;
    const uint4* bRanges = a.bendRanges + static_cast<size_t>(spec) * a.bendColours;
    float4 v[kSpecimenOwn];
    V3 prevOwn[kSpecimenOwn];
#pragma unroll
    for (uint32_t m = 0; m < kSpecimenOwn; ++m) {
        const uint32_t i = tid + m * threads;
        if (i < n) {
            x[i] = a.pos[k0 + i];
            v[m] = a.vel[k0 + i];
        }
    }
    __syncthreads();
    for (uint32_t j = 0; j < a.substeps; ++j) {
        const uint32_t s = a.firstSubstep + j;
        const NumiClothBagGPUConfig& cfg = a.configs[s];
        const float dt = cfg.gravityAndTimestep.w;
        const float invDt2 = 1.0f / (dt * dt);
        const V3 g = xyz(cfg.gravityAndTimestep);
        const double endTime = a.times[s];
#pragma unroll
        for (uint32_t m = 0; m < kSpecimenOwn; ++m) {
            const uint32_t i = tid + m * threads;
            if (i < n) {
                float4 p = x[i];
                if (a.record) a.tapeV0[static_cast<size_t>(j) * a.particles + k0 + i] = v[m];
                prevOwn[m] = xyz(p);
                specimenBegin(a, k0 + i, p, v[m], g, dt, endTime);
                x[i] = p;
            }
        }
        for (uint32_t t = tid; t < info.distanceCount; t += threads) lamD[t] = 0.0f;
        for (uint32_t t = tid; t < info.bendCount; t += threads) lamB[t] = 0.0f;
        __syncthreads();
        uint32_t pass = 0;
        auto snapshot = [&](uint32_t p) {
            float4* out = a.tapeX + (static_cast<size_t>(j) * a.passes + p) * a.particles + k0;
            for (uint32_t i = tid; i < n; i += threads) out[i] = x[i];
        };
        for (uint32_t it = 0; it < a.iterations; ++it) {
            if (a.record) {
                float* out = a.tapeLambdaD + (static_cast<size_t>(j) * a.iterations + it) * a.distances;
                for (uint32_t c = 0; c < a.distanceColours; ++c) {
                    const uint4 r = dRanges[c];
                    for (uint32_t t = tid; t < r.y; t += threads) out[r.x + t] = lamD[r.z + t];
                }
            }
            for (uint32_t c = 0; c < a.distanceColours; ++c, ++pass) {
                if (a.record) {
                    snapshot(pass);
                    __syncthreads();
                }
                const uint4 r = dRanges[c];
                for (uint32_t t = tid; t < r.y; t += threads) {
                    const uint32_t k = r.x + t;
                    const uint2 e = a.ij[k];
                    float lam = lamD[r.z + t];
                    float4 pa = x[e.x - k0], pb = x[e.y - k0];
                    if (projectDistance(pa, pb, a.restCompliance[k], lam, invDt2, a.failure)) {
                        x[e.x - k0] = pa;
                        x[e.y - k0] = pb;
                        lamD[r.z + t] = lam;
                    }
                }
                __syncthreads();
            }
            if (it % 2 == 0) {
                if (a.record) {
                    float* out = a.tapeLambdaB + (static_cast<size_t>(j) * ((a.iterations + 1) / 2) + it / 2) * a.bends;
                    for (uint32_t c = 0; c < a.bendColours; ++c) {
                        const uint4 r = bRanges[c];
                        for (uint32_t t = tid; t < r.y; t += threads) out[r.x + t] = lamB[r.z + t];
                    }
                }
                for (uint32_t c = 0; c < a.bendColours; ++c, ++pass) {
                    if (a.record) {
                        snapshot(pass);
                        __syncthreads();
                    }
                    const uint4 r = bRanges[c];
                    for (uint32_t t = tid; t < r.y; t += threads) {
                        const uint32_t k = r.x + t;
                        const uint2 e = a.bEnds[k];
                        float lam = lamB[r.z + t];
                        float4 pa = x[e.x - k0], pb = x[e.y - k0];
                        if (projectBend(pa, pb, a.chordCompliance[k], lam, invDt2, false, 0.0f, a.failure)) {
                            x[e.x - k0] = pa;
                            x[e.y - k0] = pb;
                            lamB[r.z + t] = lam;
                        }
                    }
                    __syncthreads();
                }
            }
        }
        const uint32_t invocations = strainInvocations(a.sweeps);
        const uint32_t strainPasses = invocations * a.sweeps * a.distanceColours;
        for (uint32_t inv = 0; inv < invocations; ++inv) {
            for (uint32_t sweep = 0; sweep < a.sweeps; ++sweep) {
                bool moved = false;
                for (uint32_t c = 0; c < a.distanceColours; ++c, ++pass) {
                    const uint4 r = dRanges[c];
                    uint8_t* acts = a.record ? a.tapeStrainActs + (static_cast<size_t>(j) * strainPasses +
                                                                   (pass - (a.passes - strainPasses))) * a.distances
                                             : nullptr;
                    float4* xs = a.record ? a.tapeX + (static_cast<size_t>(j) * a.passes + pass) * a.particles : nullptr;
                    bool taped = false;
                    for (uint32_t t = tid; t < r.y; t += threads) {
                        const uint32_t k = r.x + t;
                        const uint2 e = a.ij[k];
                        float4 pa = x[e.x - k0], pb = x[e.y - k0];
                        if (a.record) {
                            // The reverse's test of whether the projection moves the pair.
                            const float len = length(xyz(pb) - xyz(pa));
                            const bool moves = len > a.maxLength[k] && len > 1.0e-12f && pa.w + pb.w > 0.0f;
                            acts[k] = moves ? 1u : 0u;
                            if (moves) {
                                xs[e.x] = pa;
                                xs[e.y] = pb;
                                taped = true;
                            }
                        }
                        if (limitStrainPair(pa, pb, a.maxLength[k], a.failure)) {
                            x[e.x - k0] = pa;
                            x[e.y - k0] = pb;
                            moved = true;
                        }
                    }
                    if (a.record) {
                        // The reverse visits only the passes in which a pair moved.
                        const bool act = __syncthreads_or(taped ? 1 : 0) != 0;
                        if (tid == 0) {
                            a.tapeActive[(static_cast<size_t>(j) * a.passes + pass) * a.specimenCount + spec] = act;
                            if (act) atomicAdd(a.strainTaped, 1u);
                        }
                    } else {
                        __syncthreads();
                    }
                }
                // A sweep that moved no knot leaves the positions the next
                // sweep of this invocation would check: the remaining sweeps
                // are no-ops (the strain tile program stops the same way), and
                // their passes are marked untaped for the reverse.
                if (__syncthreads_or(moved ? 1 : 0) == 0) {
                    const uint32_t skipped = (a.sweeps - 1 - sweep) * a.distanceColours;
                    if (a.record && tid == 0)
                        for (uint32_t q = 0; q < skipped; ++q)
                            a.tapeActive[(static_cast<size_t>(j) * a.passes + pass + q) * a.specimenCount + spec] = 0u;
                    pass += skipped;
                    break;
                }
            }
        }
        // finalize (no ground): v = (x - prev) / dt
        const float inverseDt = 1.0f / dt;
#pragma unroll
        for (uint32_t m = 0; m < kSpecimenOwn; ++m) {
            const uint32_t i = tid + m * threads;
            if (i < n) {
                const V3 u = (xyz(x[i]) - prevOwn[m]) * inverseDt;
                if (!finite3(u)) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
                v[m] = make_float4(u.x, u.y, u.z, 0.0f);
            }
        }
        // loss of the tracked knots after the substep (fixed-order sums)
        double part = 0.0;
        for (uint32_t t = tid; t < info.trackedCount; t += threads) {
            const uint32_t q = info.trackedFirst + t;
            const uint32_t i = a.tracked[q] - k0;
            const float* target = a.targets + (static_cast<size_t>(s) * a.trackedTotal + q) * 3;
            float gx0, gx1, gx2;
            if (a.linear) {
                gx0 = target[0];
                gx1 = target[1];
                gx2 = target[2];
                part += static_cast<double>(x[i].x) * gx0 + static_cast<double>(x[i].y) * gx1 +
                        static_cast<double>(x[i].z) * gx2;
            } else {
                const float rx = x[i].x - target[0], ry = x[i].y - target[1], rz = x[i].z - target[2];
                part += static_cast<double>(rx) * rx + static_cast<double>(ry) * ry + static_cast<double>(rz) * rz;
                gx0 = 2.0f * rx;
                gx1 = 2.0f * ry;
                gx2 = 2.0f * rz;
            }
            if (a.record) {
                float* out = a.lossGrad + (static_cast<size_t>(j) * a.trackedTotal + q) * 3;
                out[0] = gx0;
                out[1] = gx1;
                out[2] = gx2;
            }
        }
        if (a.record != 1u) {
            for (int offset = 32; offset > 0; offset >>= 1) part += __shfl_xor(part, offset);
            if ((tid & 63u) == 0u) lossParts[tid / 64] = part;
        }
        __syncthreads();
        if (a.record != 1u && tid == 0) {
            double sum = 0.0;
            for (uint32_t w = 0; w < threads / 64; ++w) sum += lossParts[w];
            a.loss[spec] += sum;
        }
    }
    // State after the launch, laid out as the plain solver keeps it.
#pragma unroll
    for (uint32_t m = 0; m < kSpecimenOwn; ++m) {
        const uint32_t i = tid + m * threads;
        if (i < n) {
            a.pos[k0 + i] = x[i];
            a.prev[k0 + i] = make_float4(prevOwn[m].x, prevOwn[m].y, prevOwn[m].z, a.prev[k0 + i].w);
            a.vel[k0 + i] = v[m];
        }
    }
    for (uint32_t c = 0; c < a.distanceColours; ++c) {
        const uint4 r = dRanges[c];
        for (uint32_t t = tid; t < r.y; t += threads) a.dLambda[r.x + t] = lamD[r.z + t];
    }
    for (uint32_t c = 0; c < a.bendColours; ++c) {
        const uint4 r = bRanges[c];
        for (uint32_t t = tid; t < r.y; t += threads) a.bLambda[r.x + t] = lamB[r.z + t];
    }
}

// Adjoint of one XPBD projection given its inputs (see the forward in
// projectDistance / projectBend): ga, gb are dL/d(outputs) on entry and
// dL/d(inputs) on return; glam likewise; galpha, gwa, gwb accumulate.
__device__ __forceinline__ void projectionAdjoint(V3 xa, V3 xb, float rest, float alpha, float lam, float wa, float wb,
                                                  V3& ga, V3& gb, float& glam, float& galpha, float& gwa,
                                                  float& gwb) {
    const V3 d = xb - xa;
    const float len = length(d);
    const V3 n = d / len;
    const float denom = wa + wb + alpha;
    const float numerator = -(len - rest) - alpha * lam;
    const float dl = numerator / denom;
    const V3 corr = n * dl;
    const V3 gcorr = gb * wb - ga * wa;
    gwa += -dot(corr, ga);
    gwb += dot(corr, gb);
    const float gdl = glam + dot(gcorr, n);
    const V3 gn = gcorr * dl;
    const float gnum = gdl / denom;
    const float gden = -gdl * dl / denom;
    const float glen = -gnum;
    galpha += -lam * gnum + gden;
    glam += -alpha * gnum;
    gwa += gden;
    gwb += gden;
    const V3 gd = (gn - n * dot(gn, n)) / len + n * glen;
    ga = ga - gd;
    gb = gb + gd;
}

__device__ __forceinline__ void strainAdjoint(V3 xa, V3 xb, float limit, float wa, float wb, V3& ga, V3& gb,
                                              float& gwa, float& gwb) {
    const V3 d = xb - xa;
    const float len = length(d);
    const V3 n = d / len;
    const float W = wa + wb;
    const float ex = len - limit;
    const float sa = wa * ex / W, sb = wb * ex / W;
    const V3 gn = ga * sa - gb * sb;
    const float gsa = dot(ga, n), gsb = -dot(gb, n);
    const float gex = (gsa * wa + gsb * wb) / W;
    const float gW = -(gsa * sa + gsb * sb) / W;
    gwa += gsa * ex / W + gW;
    gwb += gsb * ex / W + gW;
    const V3 gd = (gn - n * dot(gn, n)) / len + n * gex;
    ga = ga - gd;
    gb = gb + gd;
}

__global__ void __launch_bounds__(kSpecimenThreads) k_specimen_reverse(SpecimenArgs a) {
    extern __shared__ float4 specimenLds[];
    const uint32_t spec = blockIdx.x;
    const SpecimenInfo info = a.specimens[spec];
    const uint32_t tid = threadIdx.x, threads = blockDim.x;
    const uint32_t n = info.knotCount, k0 = info.knotFirst;
    float4* gx = specimenLds;  // xyz: dL/dx, w: dL/dw accumulated
    float* glD = reinterpret_cast<float*>(gx + n);
    float* glB = glD + info.distanceCount;
    const uint4* dRanges = a.distanceRanges + static_cast<size_t>(spec) * a.distanceColours;
    const uint4* bRanges = a.bendRanges + static_cast<size_t>(spec) * a.bendColours;
    V3 gv[kSpecimenOwn];
    float gdamp[kSpecimenOwn];
#pragma unroll
    for (uint32_t m = 0; m < kSpecimenOwn; ++m) {
        const uint32_t i = tid + m * threads;
        gdamp[m] = 0.0f;
        if (i < n) {
            const float4 s = a.gxState[k0 + i];
            gx[i] = make_float4(s.x, s.y, s.z, 0.0f);
            gv[m] = xyz(a.gvState[k0 + i]);
        }
    }
    __syncthreads();
    const uint32_t halfIterations = (a.iterations + 1) / 2;
    const uint32_t strainPasses = strainInvocations(a.sweeps) * a.sweeps * a.distanceColours;
    const uint32_t iterationPasses = a.passes - strainPasses;
    for (uint32_t j = a.substeps; j-- > 0;) {
        const uint32_t s = a.firstSubstep + j;
        const NumiClothBagGPUConfig& cfg = a.configs[s];
        const float dt = cfg.gravityAndTimestep.w;
        const float inverseDt = 1.0f / dt;
        const float alphaScale = 1.0f / (dt * dt);
        for (uint32_t t = tid; t < info.trackedCount; t += threads) {
            const uint32_t q = info.trackedFirst + t;
            const uint32_t i = a.tracked[q] - k0;
            const float* gl = a.lossGrad + (static_cast<size_t>(j) * a.trackedTotal + q) * 3;
            float4 g = gx[i];
            g.x += gl[0];
            g.y += gl[1];
            g.z += gl[2];
            gx[i] = g;
        }
        __syncthreads();
        V3 gprev[kSpecimenOwn];
#pragma unroll
        for (uint32_t m = 0; m < kSpecimenOwn; ++m) {
            const uint32_t i = tid + m * threads;
            if (i < n) {
                gx[i] = withXyz(gx[i], xyz(gx[i]) + gv[m] * inverseDt);
                gprev[m] = gv[m] * -inverseDt;
            }
        }
        for (uint32_t t = tid; t < info.distanceCount; t += threads) glD[t] = 0.0f;
        for (uint32_t t = tid; t < info.bendCount; t += threads) glB[t] = 0.0f;
        __syncthreads();
        auto tapeAt = [&](uint32_t pass) {
            return a.tapeX + (static_cast<size_t>(j) * a.passes + pass) * a.particles;
        };
        // strain passes, last first
        for (uint32_t p = strainPasses; p-- > 0;) {
            const uint32_t pass = iterationPasses + p;
            if (a.tapeActive[(static_cast<size_t>(j) * a.passes + pass) * a.specimenCount + spec] == 0u) continue;
            const uint4 r = dRanges[p % a.distanceColours];
            const float4* xs = tapeAt(pass);
            const uint8_t* acts = a.tapeStrainActs + (static_cast<size_t>(j) * strainPasses + p) * a.distances;
            for (uint32_t t = tid; t < r.y; t += threads) {
                const uint32_t k = r.x + t;
                if (acts[k] == 0u) continue;  // the projection left the pair as it was
                const uint2 e = a.ij[k];
                const float4 pa = xs[e.x], pb = xs[e.y];
                const V3 d = xyz(pb) - xyz(pa);
                const float len = length(d);
                const float limit = a.maxLength[k];
                if (!(len > limit) || !(len > 1.0e-12f) || !(pa.w + pb.w > 0.0f)) continue;
                float4 ga = gx[e.x - k0], gb = gx[e.y - k0];
                V3 va = xyz(ga), vb = xyz(gb);
                strainAdjoint(xyz(pa), xyz(pb), limit, pa.w, pb.w, va, vb, ga.w, gb.w);
                gx[e.x - k0] = withXyz(ga, va);
                gx[e.y - k0] = withXyz(gb, vb);
            }
            __syncthreads();
        }
        // iterations, last first
        uint32_t pass = iterationPasses;
        for (uint32_t it = a.iterations; it-- > 0;) {
            if (it % 2 == 0) {
                const float* lamBefore =
                    a.tapeLambdaB + (static_cast<size_t>(j) * halfIterations + it / 2) * a.bends;
                for (uint32_t c = a.bendColours; c-- > 0;) {
                    --pass;
                    const uint4 r = bRanges[c];
                    const float4* xs = tapeAt(pass);
                    for (uint32_t t = tid; t < r.y; t += threads) {
                        const uint32_t k = r.x + t;
                        const uint2 e = a.bEnds[k];
                        const float4 pa = xs[e.x], pb = xs[e.y];
                        const float2 cc = a.chordCompliance[k];
                        const float alpha = cc.y * alphaScale;  // as the forward computes it
                        const float chord = length(xyz(pb) - xyz(pa));
                        if (!(chord > 1.0e-12f) || !(alpha + pa.w + pb.w >= 1.0e-16f)) continue;
                        float4 ga = gx[e.x - k0], gb = gx[e.y - k0];
                        V3 va = xyz(ga), vb = xyz(gb);
                        float galpha = 0.0f;
                        projectionAdjoint(xyz(pa), xyz(pb), cc.x, alpha, lamBefore[k], pa.w, pb.w, va, vb,
                                          glB[r.z + t], galpha, ga.w, gb.w);
                        gx[e.x - k0] = withXyz(ga, va);
                        gx[e.y - k0] = withXyz(gb, vb);
                        a.gradBend[k] += galpha * alphaScale;
                    }
                    __syncthreads();
                }
            }
            const float* lamBefore = a.tapeLambdaD + (static_cast<size_t>(j) * a.iterations + it) * a.distances;
            for (uint32_t c = a.distanceColours; c-- > 0;) {
                --pass;
                const uint4 r = dRanges[c];
                const float4* xs = tapeAt(pass);
                for (uint32_t t = tid; t < r.y; t += threads) {
                    const uint32_t k = r.x + t;
                    const uint2 e = a.ij[k];
                    const float4 pa = xs[e.x], pb = xs[e.y];
                    const float2 rc = a.restCompliance[k];
                    const float alpha = rc.y * alphaScale;
                    const float len = length(xyz(pb) - xyz(pa));
                    if (!(len > 1.0e-12f) || !(pa.w + pb.w + alpha > 0.0f)) continue;
                    float4 ga = gx[e.x - k0], gb = gx[e.y - k0];
                    V3 va = xyz(ga), vb = xyz(gb);
                    float galpha = 0.0f;
                    projectionAdjoint(xyz(pa), xyz(pb), rc.x, alpha, lamBefore[k], pa.w, pb.w, va, vb,
                                      glD[r.z + t], galpha, ga.w, gb.w);
                    gx[e.x - k0] = withXyz(ga, va);
                    gx[e.y - k0] = withXyz(gb, vb);
                    a.gradDistance[k] += galpha * alphaScale;
                }
                __syncthreads();
            }
        }
        // begin + advance: x = x0 + v' dt, v' = f v0 + g dt, prev = x0
#pragma unroll
        for (uint32_t m = 0; m < kSpecimenOwn; ++m) {
            const uint32_t i = tid + m * threads;
            if (i < n) {
                const float4 g = gx[i];
                const float w = a.pos[k0 + i].w;
                if (w > 0.0f) {
                    float c = 0.0f, f = a.dampingFactor;
                    if (a.motion != nullptr) {
                        c = a.motion[k0 + i].w;
                        f = c > 0.0f ? 1.0f / (1.0f + c * a.dampingDt) : 1.0f;
                    }
                    const V3 gvp = xyz(g) * dt;
                    gx[i] = withXyz(g, xyz(g) + gprev[m]);
                    gv[m] = gvp * f;
                    if (c > 0.0f) {
                        const V3 v0 = xyz(a.tapeV0[static_cast<size_t>(j) * a.particles + k0 + i]);
                        gdamp[m] += dot(gvp, v0) * (-a.dampingDt * f * f);
                    }
                } else {
                    gx[i] = withXyz(g, v3(0.0f, 0.0f, 0.0f));
                    gv[m] = v3(0.0f, 0.0f, 0.0f);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (uint32_t m = 0; m < kSpecimenOwn; ++m) {
        const uint32_t i = tid + m * threads;
        if (i < n) {
            const float4 g = gx[i];
            a.gxState[k0 + i] = make_float4(g.x, g.y, g.z, 0.0f);
            a.gvState[k0 + i] = make_float4(gv[m].x, gv[m].y, gv[m].z, 0.0f);
            a.gradInverseMass[k0 + i] += g.w;
            a.gradDamping[k0 + i] += gdamp[m];
        }
    }
}

}  // namespace numi_dev
