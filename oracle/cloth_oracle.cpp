// CPU oracle of the XPBD cloth substep loop, in FP64 and as an FP32 twin.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the
// product library.
//
// It restates, in the same colour-batch order, the reference's
// schedule-identical oracle runOracle (reference
// tools/cloth_bag_metal.mm:3432-3902) generalised from one substep to a
// trajectory of substeps, using the per-kernel arithmetic of reference
// src/metal/ClothBag.metal:
//   begin_substep          :491-575     advance_positions    :907-938
//   update_grip_attachment :577-673     solve_distance       :1148-1213
//   solve_knot             :1215-1313   solve_bend           :1315-1447
//   solve_grip             :1449-1501   solve_ground         :1575-1614
//   limit_strain           :1616-1673   finalize_substep     :2890-2921
//   apply_cloth_ground_friction :2976-3022
//   self contact (sampleSegments :135-211, edgeBoundsOverlap :213-274,
//                 solve_self_contact :2125-2448)
// and the pass order of runGPU (reference tools/cloth_bag_metal.mm:4211-4709,
// pass counts :47-49). Scene facts the reference has no record for (static
// spheres, kinematic drive, damping, hash self-contact) follow the
// definitions in include/numi_cloth_hip.h.
//
// Two instances of the same template:
//   Real = double  the FP64 oracle (inputs widened once, state kept in FP64);
//   Real = float   an FP32 twin that evaluates every expression in the order
//                  the device kernels do (fma where they use fma), used to
//                  show that the device path differs from FP64 only by FP32
//                  rounding.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>

#include "numi_cloth_hip.h"
#include "oracle_math.hpp"

namespace {

template <typename R>
struct Particle {
    V<R> x, prev, v;
    R w = 0, predictedVz = 0, groundImpulse = 0;
};

// Self-contact accumulator: FP64 sums exactly in double; the FP32 twin uses
// the device's 2^-32 m fixed point so that sums are order independent.
struct Fixed {
    long long x = 0, y = 0, z = 0;  // sum of the corrections, 2^-32 m units
    long long m = 0;                // sum of their lengths, same units
    float maxLength = 0.0f;         // the longest correction
    uint32_t count = 0;
};

template <typename R>
struct Oracle {
    const NumiClothHipSceneDesc& d;
    uint32_t P, D, B, K, G;
    std::vector<Particle<R>> p;
    std::vector<R> dLambda, bLambda, kLambda, maxLength;
    std::vector<V<R>> gripLambda, gripOffset, anchor;
    std::vector<uint32_t> gripParticle, gripGeneration, gripCentre;
    std::vector<R> gripCapture, gripError;
    std::vector<uint32_t> localStart, localList;
    std::vector<std::pair<uint32_t, uint32_t>> candidates;
    std::vector<V<R>> deltaSum;
    std::vector<Fixed> fixed;
    std::vector<uint32_t> deltaCount;
    std::vector<R> deltaLengthSum, deltaLengthMax;
    uint32_t failure = 0;
    uint32_t frictionContacts = 0;
    uint32_t presentContacts = 0, sweptContacts = 0;
    uint32_t certificatePasses = 0;  // passes of the last certificate loop
    uint32_t selfFrictionContacts = 0;
    uint32_t pointContacts = 0;  // pairs of the last point self-collision pass
    double maximumConeRatio = 0;
    // Impulse aggregate of each candidate over the substep's self-contact
    // calls (reference aggregate_self_impulses, src/metal/ClothBag.metal:
    // 2503-2553: sum of -normal * impulse, impulse, endpoint shares, count).
    struct SelfImpulse {
        V<R> n{};
        R impulse = 0;
        R e[4] = {0, 0, 0, 0};
        uint32_t count = 0;
    };
    std::vector<SelfImpulse> selfImpulses;
    double time = 0;

    explicit Oracle(const NumiClothHipSceneDesc& scene)
        : d(scene), P(scene.particleCount), D(scene.distanceCount), B(scene.bendCount), K(scene.knotCount),
          G(scene.gripCount) {
        p.resize(P);
        anchor.resize(P);
        for (uint32_t i = 0; i < P; ++i) {
            const NumiClothBagGPUParticle& s = d.particles[i];
            p[i].x = v3<R>(s.positionAndInverseMass);
            p[i].prev = v3<R>(s.previousAndMass);
            p[i].v = v3<R>(s.velocity);
            p[i].w = R(s.positionAndInverseMass.w);
            anchor[i] = p[i].x;
        }
        dLambda.assign(D, R(0));
        bLambda.assign(B, R(0));
        kLambda.assign(K, R(0));
        maxLength.resize(D);
        for (uint32_t k = 0; k < D; ++k) {
            const mr_float4 m = d.distances[k].material;
            // The device precomputes rest * (1 + limit) in FP32.
            maxLength[k] = sizeof(R) == sizeof(float) ? R(m.x * (1.0f + m.w)) : R(m.x) * (R(1) + R(m.w));
        }
        gripLambda.assign(G, V<R>{});
        gripOffset.resize(G);
        gripParticle.resize(G);
        gripGeneration.resize(G);
        gripCentre.resize(G);
        gripCapture.assign(G, R(0));
        gripError.assign(G, R(0));
        for (uint32_t i = 0; i < G; ++i) {
            const NumiClothBagGPUGrip& g = d.grips[i];
            gripOffset[i] = v3<R>(g.targetOffsetAndCompliance);
            gripParticle[i] = g.particle.x;
            gripGeneration[i] = g.particle.y;
            gripCentre[i] = g.particle.w;
        }
        if (d.extension.selfContact || d.extension.pointSelfCollision) buildLocality();
        initBag();
    }

    // Two-hop neighbourhoods (including the knot) over the distance graph.
    void buildLocality() {
        std::vector<std::vector<uint32_t>> direct(P);
        for (uint32_t k = 0; k < D; ++k) {
            const mr_uint4 e = d.distances[k].particlesAndColor;
            direct[e.x].push_back(e.y);
            direct[e.y].push_back(e.x);
        }
        localStart.assign(P + 1, 0);
        for (uint32_t i = 0; i < P; ++i) {
            std::vector<uint32_t> near{i};
            for (uint32_t a : direct[i]) {
                near.push_back(a);
                for (uint32_t b : direct[a]) near.push_back(b);
            }
            std::sort(near.begin(), near.end());
            near.erase(std::unique(near.begin(), near.end()), near.end());
            localStart[i + 1] = localStart[i] + static_cast<uint32_t>(near.size());
            localList.insert(localList.end(), near.begin(), near.end());
        }
    }

    bool near(uint32_t a, uint32_t b) const {
        return std::binary_search(localList.begin() + localStart[a], localList.begin() + localStart[a + 1], b);
    }

    // ------------------------------------------------------------ begin --
    // Damping factor 1 / (1 + c dt0) of a damping coefficient c, in the
    // arithmetic of each precision (FP32 twin: float operations).
    static R dampingFactor(float damping, float dt0) {
        if (!(damping > 0.0f)) return R(1);
        return sizeof(R) == sizeof(float) ? R(1.0f / (1.0f + damping * dt0)) : R(1) / (R(1) + R(damping) * R(dt0));
    }

    // Kinematic drive offset A (sin(omega t + phase) - sin(phase)), evaluated
    // in double and rounded to float as the device does.
    static R driveOffset(float amplitude, float omega, float phase, double time) {
        return R(static_cast<float>(static_cast<double>(amplitude) *
                                    (std::sin(static_cast<double>(omega) * time + static_cast<double>(phase)) -
                                     std::sin(static_cast<double>(phase)))));
    }

    void begin(const NumiClothBagGPUConfig& c, double endTime) {
        const R dt = R(c.gravityAndTimestep.w);
        const V<R> g = v3<R>(c.gravityAndTimestep);
        const NumiClothHipSceneExtension& e = d.extension;
        const float dt0 = d.config->gravityAndTimestep.w;
        const R factor = dampingFactor(e.damping, dt0);
        const bool driven = e.driveAmplitude != 0.0f || e.particleMotion != nullptr;
        const R globalOffset = e.driveAmplitude != 0.0f
                                   ? driveOffset(e.driveAmplitude, e.driveOmega, e.drivePhase, endTime)
                                   : R(0);
        for (uint32_t i = 0; i < P; ++i) {
            Particle<R>& q = p[i];
            q.prev = q.x;
            q.groundImpulse = 0;
            q.predictedVz = 0;
            if (q.w > R(0)) {
                const R f = e.particleMotion != nullptr ? dampingFactor(e.particleMotion[i].w, dt0) : factor;
                const V<R> carried = f != R(1) ? q.v * f : q.v;
                q.v = fma3(g, dt, carried);
                q.predictedVz = q.v.z;
            } else if (driven) {
                const R offset = e.particleMotion != nullptr
                                     ? driveOffset(e.particleMotion[i].x, e.particleMotion[i].y,
                                                   e.particleMotion[i].z, endTime)
                                     : globalOffset;
                q.x = {anchor[i].x, anchor[i].y, anchor[i].z + offset};
            }
        }
        std::fill(dLambda.begin(), dLambda.end(), R(0));
        std::fill(bLambda.begin(), bLambda.end(), R(0));
        std::fill(kLambda.begin(), kLambda.end(), R(0));
        std::fill(gripLambda.begin(), gripLambda.end(), V<R>{});
    }

    void attachGrips(const NumiClothBagGPUConfig& c) {
        if (!(c.gripTargetAndActive.w > 0.0f)) return;
        const V<R> target = v3<R>(c.gripTargetAndActive);
        const Quat<R> q = q4<R>(c.gripOrientation);
        for (uint32_t i = 0; i < G && i < c.control.w; ++i) {
            if (gripGeneration[i] == c.gripControl.x) continue;
            if (gripGeneration[i] == 0xffffffffu || gripGeneration[i] + 1u != c.gripControl.x) {
                failure |= NUMI_CLOTH_BAG_GPU_FAILURE_GRIP_CAPTURE;
                continue;
            }
            uint32_t particle = gripParticle[i];
            if (c.gripControl.y == 1u) {
                const uint32_t rings = c.gripControl.z;
                uint32_t centre = 0;
                R nearest = R(INFINITY);
                for (uint32_t ring = 0; ring < rings; ++ring) {
                    const V<R> sep = p[d.cuffParticles[rings + ring]].x - target;
                    const R d2 = dot(sep, sep);
                    if (d2 < nearest) {
                        nearest = d2;
                        centre = ring;
                    }
                }
                const uint32_t width = 2u * c.gripControl.w + 1u;
                const uint32_t ring = (centre + rings + i % width - c.gripControl.w) % rings;
                particle = d.cuffParticles[(i / width) * rings + ring];
                gripParticle[i] = particle;
                gripCentre[i] = centre;
            }
            const V<R> sep = p[particle].x - target;
            const R capture = norm(sep);
            if (capture > R(c.gripMaterial.x)) {
                failure |= NUMI_CLOTH_BAG_GPU_FAILURE_GRIP_CAPTURE;
                continue;
            }
            const V<R> local = rotate(Quat<R>{-q.x, -q.y, -q.z, q.w}, sep);
            const R error = norm(p[particle].x - (target + rotate(q, local)));
            gripOffset[i] = local;
            gripGeneration[i] = c.gripControl.x;
            gripCapture[i] = std::max(gripCapture[i], capture);
            gripLambda[i] = V<R>{};
            gripError[i] = std::max(gripError[i], error);
        }
    }

    void advance(R dt) {
        for (Particle<R>& q : p)
            if (q.w > R(0)) q.x = fma3(q.v, dt, q.x);
    }

    // Swept ground pass after the advance (reference sweepGroundPrediction,
    // tools/cloth_bag.cpp:2692-2721): knots and fruits that crossed contact
    // height go back to it; a fruit's removed advance becomes ground impulse.
    void sweptGround(const NumiClothBagGPUConfig& c) {
        if (c.constraintCounts.z == 0u) return;
        const R r = R(c.clothMaterial.x);
        for (Particle<R>& q : p)
            if (q.prev.z >= r && q.x.z < r) q.x.z = r;
        for (Fruit& f : fruit)
            if (f.prev.z >= f.radius && f.x.z < f.radius) {
                const R removed = f.radius - f.x.z;
                f.groundImpulse += removed / (f.invMass * R(c.gravityAndTimestep.w));
                f.x.z = f.radius;
            }
    }

    // Point self-collision (reference solveSelfCollision, tools/cloth_bag.cpp:
    // 4249-4315) with every pair on one snapshot and each knot moved by the
    // mean of its corrections (FP64: exact sums; FP32 twin: the device's
    // 2^-32 m fixed point). The pair set is every i < j within 2r (and at
    // least 1e-12), outside i's two-hop list, with w_i + w_j > 0; the grid
    // here (cells of 2r, neighbours only) just finds them.
    void pointSelf(const NumiClothBagGPUConfig& c) {
        const R target = R(2) * R(c.clothMaterial.x);
        const double cellSize = 2.0 * static_cast<double>(c.clothMaterial.x) * 1.0009765625;
        std::map<std::array<long long, 3>, std::vector<uint32_t>> cells;
        auto cellOf = [&](uint32_t i) {
            return std::array<long long, 3>{static_cast<long long>(std::floor(static_cast<double>(p[i].x.x) / cellSize)),
                                            static_cast<long long>(std::floor(static_cast<double>(p[i].x.y) / cellSize)),
                                            static_cast<long long>(std::floor(static_cast<double>(p[i].x.z) / cellSize))};
        };
        for (uint32_t i = 0; i < P; ++i)
            if (std::isfinite(static_cast<double>(p[i].x.x)) && std::isfinite(static_cast<double>(p[i].x.y)) &&
                std::isfinite(static_cast<double>(p[i].x.z)))
                cells[cellOf(i)].push_back(i);
        clearAccumulators();
        pointContacts = 0;
        for (const auto& [cell, members] : cells)
            for (uint32_t i : members)
                for (long long dz = -1; dz <= 1; ++dz)
                    for (long long dy = -1; dy <= 1; ++dy)
                        for (long long dx = -1; dx <= 1; ++dx) {
                            const auto found = cells.find({cell[0] + dx, cell[1] + dy, cell[2] + dz});
                            if (found == cells.end()) continue;
                            for (uint32_t j : found->second) {
                                if (j <= i) continue;
                                const V<R> dd = p[j].x - p[i].x;
                                const R len = norm(dd);
                                if (!(len < target) || len < R(1.0e-12)) continue;
                                const R denom = p[i].w + p[j].w;
                                if (!(denom > R(0))) continue;
                                if (near(i, j)) continue;
                                const V<R> correction = dd * ((target - len) / (len * denom));
                                accumulate(i, correction * -p[i].w);
                                accumulate(j, correction * p[j].w);
                                ++pointContacts;
                            }
                        }
        for (uint32_t i = 0; i < P; ++i) {
            V<R> move;
            if (blendedMove(i, move)) p[i].x = p[i].x + move;
        }
    }

    // ------------------------------------------------------- constraints --
    void distances(R dt) {
        for (uint32_t b = 0; b < d.distanceBatchCount; ++b) {
            const mr_uint4 batch = d.distanceBatches[b].control;
            for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
                const NumiClothBagGPUDistance& r = d.distances[k];
                Particle<R>& a = p[r.particlesAndColor.x];
                Particle<R>& c = p[r.particlesAndColor.y];
                const V<R> diff = c.x - a.x;
                const R len = norm(diff);
                // alpha = compliance * (1 / dt^2), the device's form (one
                // reciprocal per substep instead of a division per constraint).
                const R alpha = R(r.material.y) * (R(1) / (dt * dt));
                const R denom = a.w + c.w + alpha;
                if (!(len > R(1.0e-12)) || !(denom > R(0))) continue;
                const R dl = (-(len - R(r.material.x)) - alpha * dLambda[k]) / denom;
                const V<R> corr = diff * (dl / len);
                dLambda[k] += dl;
                a.x = a.x - corr * a.w;
                c.x = c.x + corr * c.w;
            }
        }
    }

    void knots(R dt) {
        for (uint32_t b = 0; b < d.knotBatchCount; ++b) {
            const mr_uint4 batch = d.knotBatches[b].control;
            for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
                const NumiClothBagGPUKnot& r = d.knots[k];
                Particle<R>* q[4] = {&p[r.particles.x], &p[r.particles.y], &p[r.particles.z], &p[r.particles.w]};
                const V<R> warpV = q[1]->x - q[0]->x;
                const V<R> weftV = q[3]->x - q[2]->x;
                const R warpL = norm(warpV);
                const R weftL = norm(weftV);
                if (!(warpL > R(1.0e-12)) || !(weftL > R(1.0e-12))) continue;
                const V<R> warp = warpV / warpL;
                const V<R> weft = weftV / weftL;
                const R cosine = clampR(dot(warp, weft), R(-1), R(1));
                const V<R> gWarp = (weft - warp * cosine) / warpL;
                const V<R> gWeft = (warp - weft * cosine) / weftL;
                const V<R> g[4] = {-gWarp, gWarp, -gWeft, gWeft};
                const R alpha = R(r.material.y) / (dt * dt);
                R denom = alpha;
                for (int m = 0; m < 4; ++m) denom += q[m]->w * dot(g[m], g[m]);
                if (!(denom > R(0))) continue;
                const R dl = (-(cosine - R(r.material.x)) - alpha * kLambda[k]) / denom;
                kLambda[k] += dl;
                for (int m = 0; m < 4; ++m) q[m]->x = q[m]->x + g[m] * (q[m]->w * dl);
            }
        }
    }

    void bends(const NumiClothBagGPUConfig& c, R dt) {
        const bool ground = c.constraintCounts.z != 0u;
        const R floorZ = R(c.clothMaterial.x);
        for (uint32_t b = 0; b < d.bendBatchCount; ++b) {
            const mr_uint4 batch = d.bendBatches[b].control;
            for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
                const NumiClothBagGPUBend& r = d.bends[k];
                Particle<R>* q[2] = {&p[r.particlesAndColor.x], &p[r.particlesAndColor.z]};
                const V<R> diff = q[1]->x - q[0]->x;
                const R chord = norm(diff);
                if (!(chord > R(1.0e-12))) continue;
                const V<R> dir = diff * (R(1) / chord);
                const V<R> g[2] = {-dir, dir};
                const R alpha = R(r.material.z) * (R(1) / (dt * dt));
                R freeDenom = alpha;
                for (int m = 0; m < 2; ++m) freeDenom += q[m]->w * dot(g[m], g[m]);
                if (!(freeDenom >= R(1.0e-16))) continue;
                const R numerator = -(chord - R(r.material.x)) - alpha * bLambda[k];
                const R freeDl = numerator / freeDenom;
                bool active[2];
                R denom = alpha;
                for (int m = 0; m < 2; ++m) {
                    active[m] = ground && q[m]->x.z <= floorZ + R(1.0e-9) && g[m].z * freeDl < R(0);
                    denom += q[m]->w * (dot(g[m], g[m]) - (active[m] ? g[m].z * g[m].z : R(0)));
                }
                if (!(denom >= R(1.0e-16))) continue;
                const R dl = numerator / denom;
                R fraction = R(1);
                if (ground) {
                    for (int m = 0; m < 2; ++m) {
                        if (active[m]) continue;
                        const R down = g[m].z * q[m]->w * dl;
                        if (down < R(0)) fraction = std::min(fraction, std::max(R(0), q[m]->x.z - floorZ) / -down);
                    }
                }
                const R applied = dl * fraction;
                bLambda[k] += applied;
                for (int m = 0; m < 2; ++m) {
                    V<R> corr = g[m] * (q[m]->w * applied);
                    if (active[m]) corr.z = R(0);
                    q[m]->x = q[m]->x + corr;
                    if (ground) q[m]->x.z = std::max(q[m]->x.z, floorZ);
                }
            }
        }
    }

    void grips(const NumiClothBagGPUConfig& c, R dt) {
        if (!(c.gripTargetAndActive.w > 0.0f)) return;
        const V<R> target = v3<R>(c.gripTargetAndActive);
        const Quat<R> q = q4<R>(c.gripOrientation);
        for (uint32_t i = 0; i < G && i < c.control.w; ++i) {
            if (gripGeneration[i] != c.gripControl.x) continue;
            Particle<R>& x = p[gripParticle[i]];
            const R alpha = R(d.grips[i].targetOffsetAndCompliance.w) / (dt * dt);
            const R denom = x.w + alpha;
            if (!(denom > R(0))) continue;
            const V<R> value = x.x - (target + rotate(q, gripOffset[i]));
            const V<R> dl = (-value - gripLambda[i] * alpha) / denom;
            gripLambda[i] = gripLambda[i] + dl;
            x.x = x.x + dl * x.w;
        }
    }

    void collide(const NumiClothBagGPUConfig& c) {
        const bool ground = c.constraintCounts.z != 0u;
        const R r = R(c.clothMaterial.x);
        const NumiClothHipSceneExtension& e = d.extension;
        for (Particle<R>& q : p) {
            if (ground) q.x.z = std::max(q.x.z, r);
            if (q.w > R(0)) {
                for (uint32_t k = 0; k < e.staticSphereCount; ++k) {
                    const mr_float4& s = e.staticSpheres[k];
                    const V<R> diff = q.x - v3<R>(s);
                    const R dist = norm(diff);
                    const R reach = R(s.w) + r;
                    if (dist < reach && dist > R(1.0e-12)) q.x = v3<R>(s) + diff * (reach / dist);
                }
            }
        }
    }

    void strain(uint32_t sweeps) {
        for (uint32_t sweep = 0; sweep < sweeps; ++sweep) {
            for (uint32_t b = 0; b < d.distanceBatchCount; ++b) {
                const mr_uint4 batch = d.distanceBatches[b].control;
                for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
                    const NumiClothBagGPUDistance& r = d.distances[k];
                    Particle<R>& a = p[r.particlesAndColor.x];
                    Particle<R>& c = p[r.particlesAndColor.y];
                    const V<R> diff = c.x - a.x;
                    const R len = norm(diff);
                    const R limit = maxLength[k];
                    if (!(len > limit) || !(len > R(1.0e-12))) continue;
                    const R denom = a.w + c.w;
                    if (!(denom > R(0))) continue;
                    const V<R> dir = diff * (R(1) / len);
                    const R excess = len - limit;
                    const R share = excess / denom;
                    a.x = a.x + dir * (a.w * share);
                    c.x = c.x - dir * (c.w * share);
                }
            }
        }
    }

    // ------------------------------------------------------ self contact --
    struct Box {
        V<R> lo, hi;
    };
    std::vector<Box> boxes;

    static long cellOf(R v, R origin, R cell) { return static_cast<long>(std::floor((v - origin) / cell)); }

    // Broad phase from the current (prev, x) snapshot; see include/numi_cloth_hip.h.
    void broadPhase(const NumiClothBagGPUConfig& c) {
        const R expand = sizeof(R) == sizeof(float) ? R(c.clothMaterial.x + d.extension.selfContactMargin)
                                                    : R(c.clothMaterial.x) + R(d.extension.selfContactMargin);
        const R cell = R(c.clothMaterial.y);
        boxes.resize(D);
        V<R> origin{R(INFINITY), R(INFINITY), R(INFINITY)};
        for (uint32_t s = 0; s < D; ++s) {
            const mr_uint4 e = d.distances[s].particlesAndColor;
            const V<R> pts[4] = {p[e.x].x, p[e.y].x, p[e.x].prev, p[e.y].prev};
            Box b{pts[0], pts[0]};
            for (int k = 1; k < 4; ++k) {
                b.lo = {std::min(b.lo.x, pts[k].x), std::min(b.lo.y, pts[k].y), std::min(b.lo.z, pts[k].z)};
                b.hi = {std::max(b.hi.x, pts[k].x), std::max(b.hi.y, pts[k].y), std::max(b.hi.z, pts[k].z)};
            }
            b.lo = {b.lo.x - expand, b.lo.y - expand, b.lo.z - expand};
            b.hi = {b.hi.x + expand, b.hi.y + expand, b.hi.z + expand};
            boxes[s] = b;
            origin = {std::min(origin.x, b.lo.x), std::min(origin.y, b.lo.y), std::min(origin.z, b.lo.z)};
        }
        candidates.clear();
        selfImpulses.clear();
        if (!(cell > R(0))) return;
        std::unordered_map<uint64_t, std::vector<uint32_t>> cells;
        auto key = [](long x, long y, long z) {
            return (static_cast<uint64_t>(x & 0x1fffff) << 42) | (static_cast<uint64_t>(y & 0x1fffff) << 21) |
                   static_cast<uint64_t>(z & 0x1fffff);
        };
        for (uint32_t s = 0; s < D; ++s) {
            const Box& b = boxes[s];
            for (long z = cellOf(b.lo.z, origin.z, cell); z <= cellOf(b.hi.z, origin.z, cell); ++z)
                for (long y = cellOf(b.lo.y, origin.y, cell); y <= cellOf(b.hi.y, origin.y, cell); ++y)
                    for (long x = cellOf(b.lo.x, origin.x, cell); x <= cellOf(b.hi.x, origin.x, cell); ++x)
                        cells[key(x, y, z)].push_back(s);
        }
        for (uint32_t s = 0; s < D; ++s) {
            const Box& a = boxes[s];
            const mr_uint4 es = d.distances[s].particlesAndColor;
            for (long z = cellOf(a.lo.z, origin.z, cell); z <= cellOf(a.hi.z, origin.z, cell); ++z)
                for (long y = cellOf(a.lo.y, origin.y, cell); y <= cellOf(a.hi.y, origin.y, cell); ++y)
                    for (long x = cellOf(a.lo.x, origin.x, cell); x <= cellOf(a.hi.x, origin.x, cell); ++x) {
                        for (uint32_t t : cells[key(x, y, z)]) {
                            if (t <= s) continue;
                            const Box& b = boxes[t];
                            if (!(a.lo.x <= b.hi.x && a.lo.y <= b.hi.y && a.lo.z <= b.hi.z && a.hi.x >= b.lo.x &&
                                  a.hi.y >= b.lo.y && a.hi.z >= b.lo.z)) {
                                continue;
                            }
                            if (cellOf(std::max(a.lo.x, b.lo.x), origin.x, cell) != x ||
                                cellOf(std::max(a.lo.y, b.lo.y), origin.y, cell) != y ||
                                cellOf(std::max(a.lo.z, b.lo.z), origin.z, cell) != z) {
                                continue;
                            }
                            const mr_uint4 et = d.distances[t].particlesAndColor;
                            if (near(es.x, et.x) || near(es.x, et.y) || near(es.y, et.x) || near(es.y, et.y)) continue;
                            candidates.emplace_back(s, t);
                        }
                    }
        }
        std::sort(candidates.begin(), candidates.end());
    }

    struct Sample {
        V<R> first, second;
        R wa, wb, distance;
    };

    static Sample segments(V<R> a0, V<R> a1, V<R> b0, V<R> b1) {
        const V<R> da = a1 - a0, db = b1 - b0, off = a0 - b0;
        const R la = dot(da, da), lb = dot(db, db), pb = dot(db, off);
        R wa = 0, wb = 0;
        const R tiny = R(1.0e-20);
        if (la <= tiny && lb <= tiny) return {a0, b0, R(0), R(0), norm(b0 - a0)};
        if (la <= tiny) {
            wb = clampR(pb / lb, R(0), R(1));
        } else {
            const R pa = dot(da, off);
            if (lb <= tiny) {
                wa = clampR(-pa / la, R(0), R(1));
            } else {
                const R cr = dot(da, db);
                const R den = la * lb - cr * cr;
                if (den > tiny) wa = clampR((cr * pb - pa * lb) / den, R(0), R(1));
                wb = (cr * wa + pb) / lb;
                if (wb < R(0)) {
                    wb = R(0);
                    wa = clampR(-pa / la, R(0), R(1));
                } else if (wb > R(1)) {
                    wb = R(1);
                    wa = clampR((cr - pa) / la, R(0), R(1));
                }
            }
        }
        const V<R> pA = fma3(da, wa, a0), pB = fma3(db, wb, b0);
        return {pA, pB, wa, wb, norm(pB - pA)};
    }

    static V<R> normalOf(V<R> sep, V<R> da, V<R> db, V<R> previousOffset) {
        const R len = norm(sep);
        if (len > R(1.0e-12)) return sep / len;
        const V<R> c = cross(da, db);
        const R m = norm(c);
        V<R> n = m > R(1.0e-14) ? c / m : V<R>{R(1), R(0), R(0)};
        if (dot(previousOffset, n) > R(0)) n = -n;
        return n;
    }

    void accumulate(uint32_t i, V<R> delta) {
        if (sizeof(R) == sizeof(float)) {
            Fixed& f = fixed[i];
            const float length = static_cast<float>(norm(delta));
            f.x += std::llrint(static_cast<double>(delta.x) * 4294967296.0);
            f.y += std::llrint(static_cast<double>(delta.y) * 4294967296.0);
            f.z += std::llrint(static_cast<double>(delta.z) * 4294967296.0);
            f.m += std::llrint(static_cast<double>(length) * 4294967296.0);
            f.maxLength = std::max(f.maxLength, length);
            f.count += 1;
        } else {
            deltaSum[i] = deltaSum[i] + delta;
            const R length = norm(delta);
            deltaLengthSum[i] = deltaLengthSum[i] + length;
            deltaLengthMax[i] = std::max(deltaLengthMax[i], length);
            deltaCount[i] += 1;
        }
    }

    void clearAccumulators() {
        fixed.assign(P, Fixed{});
        deltaSum.assign(P, V<R>{});
        deltaCount.assign(P, 0u);
        deltaLengthSum.assign(P, R(0));
        deltaLengthMax.assign(P, R(0));
    }

    // The move of knot i from the corrections accumulated in one Jacobi pass:
    // their sum scaled by (longest length) / (sum of lengths). One correction
    // moves the knot by itself, n equal ones by one of them (the mean), and a
    // correction that shrinks to zero drops out continuously (a count-based
    // mean would jump when a contact at the threshold appears or vanishes).
    // FP32: the device's 2^-32 m fixed-point sums, combined in double.
    bool blendedMove(uint32_t i, V<R>& move) const {
        if (sizeof(R) == sizeof(float)) {
            const Fixed& f = fixed[i];
            if (f.count == 0 || f.m <= 0) return false;
            const double scale = static_cast<double>(f.maxLength) / static_cast<double>(f.m);
            move = V<R>{R(static_cast<double>(f.x) * scale), R(static_cast<double>(f.y) * scale),
                        R(static_cast<double>(f.z) * scale)};
            return true;
        }
        if (deltaCount[i] == 0 || !(deltaLengthSum[i] > R(0))) return false;
        move = deltaSum[i] * (deltaLengthMax[i] / deltaLengthSum[i]);
        return true;
    }

    // Jacobi self-contact solve over the candidates, one snapshot.
    void selfContact(const NumiClothBagGPUConfig& c, uint32_t mode) {
        if (!d.extension.selfContact) return;
        const R r = R(c.clothMaterial.x);
        const R target = R(2) * r;
        const R tolerance = R(1.0e-9);
        clearAccumulators();
        selfImpulses.resize(candidates.size());
        for (size_t k = 0; k < candidates.size(); ++k) {
            const auto& pr = candidates[k];
            const mr_uint4 ea = d.distances[pr.first].particlesAndColor;
            const mr_uint4 eb = d.distances[pr.second].particlesAndColor;
            const Particle<R>& A0 = p[ea.x];
            const Particle<R>& A1 = p[ea.y];
            const Particle<R>& B0 = p[eb.x];
            const Particle<R>& B1 = p[eb.y];
            R aLo[3] = {std::min(A0.x.x, A1.x.x), std::min(A0.x.y, A1.x.y), std::min(A0.x.z, A1.x.z)};
            R aHi[3] = {std::max(A0.x.x, A1.x.x), std::max(A0.x.y, A1.x.y), std::max(A0.x.z, A1.x.z)};
            R bLo[3] = {std::min(B0.x.x, B1.x.x), std::min(B0.x.y, B1.x.y), std::min(B0.x.z, B1.x.z)};
            R bHi[3] = {std::max(B0.x.x, B1.x.x), std::max(B0.x.y, B1.x.y), std::max(B0.x.z, B1.x.z)};
            if (mode == 1) {
                const R qa[2][3] = {{A0.prev.x, A0.prev.y, A0.prev.z}, {A1.prev.x, A1.prev.y, A1.prev.z}};
                const R qb[2][3] = {{B0.prev.x, B0.prev.y, B0.prev.z}, {B1.prev.x, B1.prev.y, B1.prev.z}};
                for (int a = 0; a < 3; ++a) {
                    aLo[a] = std::min(aLo[a], std::min(qa[0][a], qa[1][a]));
                    aHi[a] = std::max(aHi[a], std::max(qa[0][a], qa[1][a]));
                    bLo[a] = std::min(bLo[a], std::min(qb[0][a], qb[1][a]));
                    bHi[a] = std::max(bHi[a], std::max(qb[0][a], qb[1][a]));
                }
            }
            bool overlap = true;
            for (int a = 0; a < 3; ++a) overlap = overlap && (aLo[a] - r <= bHi[a] + r) && (aHi[a] + r >= bLo[a] - r);
            if (!overlap) continue;
            R wa = 0, wb = 0, correction = 0;
            V<R> n{};
            if (mode == 0) {
                const Sample s = segments(A0.x, A1.x, B0.x, B1.x);
                if (s.distance < target) {
                    wa = s.wa;
                    wb = s.wb;
                    n = normalOf(s.second - s.first, A1.x - A0.x, B1.x - B0.x,
                                 mix3(A0.prev, A1.prev, wa) - mix3(B0.prev, B1.prev, wb));
                    correction = target - s.distance;
                }
            } else {
                const R bound = norm(A0.x - A0.prev) + norm(A1.x - A1.prev) + norm(B0.x - B0.prev) +
                                norm(B1.x - B1.prev);
                if (bound >= R(1.0e-14)) {
                    auto at = [&](R t) {
                        return segments(mix3(A0.prev, A0.x, t), mix3(A1.prev, A1.x, t), mix3(B0.prev, B0.x, t),
                                        mix3(B1.prev, B1.x, t));
                    };
                    R t = 0;
                    Sample hit = at(R(0));
                    bool found = hit.distance <= target + tolerance;
                    if (!found) {
                        for (int step = 0; step < 80; ++step) {
                            hit = at(t);
                            const R gap = hit.distance - target;
                            if (gap <= tolerance) {
                                found = true;
                                break;
                            }
                            const R adv = R(0.9) * gap / bound;
                            if (!std::isfinite(adv) || !(adv > R(0)) || t + adv >= R(1)) break;
                            t += std::max(adv, R(1.0e-10));
                        }
                    }
                    if (found) {
                        wa = hit.wa;
                        wb = hit.wb;
                        const V<R> a0t = mix3(A0.prev, A0.x, t), a1t = mix3(A1.prev, A1.x, t);
                        const V<R> b0t = mix3(B0.prev, B0.x, t), b1t = mix3(B1.prev, B1.x, t);
                        n = normalOf(hit.second - hit.first, a1t - a0t, b1t - b0t,
                                     mix3(A0.prev, A1.prev, wa) - mix3(B0.prev, B1.prev, wb));
                        const V<R> restA = (A0.x - a0t) * (R(1) - wa) + (A1.x - a1t) * wa;
                        const V<R> restB = (B0.x - b0t) * (R(1) - wb) + (B1.x - b1t) * wb;
                        correction = dot(restA - restB, n);
                        if (!(correction > R(0))) correction = 0;
                    }
                }
            }
            if (!(correction > R(0))) continue;
            const R ua = R(1) - wa, ub = R(1) - wb;
            const R den = A0.w * ua * ua + A1.w * wa * wa + B0.w * ub * ub + B1.w * wb * wb;
            if (!(den > R(0))) continue;
            const R lambda = correction / den;
            accumulate(ea.x, n * -(A0.w * ua * lambda));
            accumulate(ea.y, n * -(A1.w * wa * lambda));
            accumulate(eb.x, n * (B0.w * ub * lambda));
            accumulate(eb.y, n * (B1.w * wb * lambda));
            (mode == 0 ? presentContacts : sweptContacts) += 1;
            // Impulse record (reference :2391-2407) into the aggregate.
            const R impulse = lambda / R(c.gravityAndTimestep.w);
            SelfImpulse& a = selfImpulses[k];
            a.count += 1;
            a.n = {a.n.x + -n.x * impulse, a.n.y + -n.y * impulse, a.n.z + -n.z * impulse};
            a.impulse = a.impulse + impulse;
            a.e[0] = a.e[0] + ua * impulse;
            a.e[1] = a.e[1] + wa * impulse;
            a.e[2] = a.e[2] + ub * impulse;
            a.e[3] = a.e[3] + wb * impulse;
        }
        for (uint32_t i = 0; i < P; ++i) {
            V<R> move;
            if (blendedMove(i, move)) p[i].x = p[i].x + move;
        }
    }

    // Self-contact friction (reference apply_self_friction,
    // src/metal/ClothBag.metal:2580-2731) on every candidate that made contact
    // in the substep, Jacobi from one velocity snapshot, each touched knot
    // taking the mean of its velocity changes (the hash redesign's rule, the
    // same as the normal response).
    void selfFriction(const NumiClothBagGPUConfig& c) {
        if (!d.extension.selfContact) return;
        const R mu = R(c.clothMaterial.w);
        clearAccumulators();
        for (size_t k = 0; k < selfImpulses.size(); ++k) {
            const SelfImpulse& a = selfImpulses[k];
            if (a.count == 0) continue;
            const mr_uint4 ea = d.distances[candidates[k].first].particlesAndColor;
            const mr_uint4 eb = d.distances[candidates[k].second].particlesAndColor;
            const R normalImpulse = a.impulse;
            const R normalLength = norm(a.n);
            const R firstSum = a.e[0] + a.e[1];
            const R secondSum = a.e[2] + a.e[3];
            if (!(normalImpulse > R(0) && normalLength > R(1.0e-10) && firstSum > R(1.0e-12) &&
                  secondSum > R(1.0e-12)))
                continue;
            const V<R> normal = a.n / normalLength;
            const R f0 = a.e[0] / firstSum, f1 = a.e[1] / firstSum;
            const R s0 = a.e[2] / secondSum, s1 = a.e[3] / secondSum;
            const Particle<R>&A0 = p[ea.x], &A1 = p[ea.y], &B0 = p[eb.x], &B1 = p[eb.y];
            const V<R> firstVelocity = A0.v * f0 + A1.v * f1;
            const V<R> secondVelocity = B0.v * s0 + B1.v * s1;
            const V<R> relative = firstVelocity - secondVelocity;
            const V<R> tangentVelocity = relative - normal * dot(relative, normal);
            const R slip = norm(tangentVelocity);
            if (!(slip > R(1.0e-10))) continue;
            const V<R> tangent = tangentVelocity / slip;
            const R denominator = A0.w * f0 * f0 + A1.w * f1 * f1 + B0.w * s0 * s0 + B1.w * s1 * s1;
            if (!(denominator > R(0))) continue;
            const R limit = mu * normalImpulse;
            const R tangential = std::min(slip / denominator, limit);
            if (!(tangential > R(0))) continue;
            const V<R> onFirst = tangent * -tangential;
            accumulate(ea.x, onFirst * (A0.w * f0));
            accumulate(ea.y, onFirst * (A1.w * f1));
            accumulate(eb.x, -(onFirst * (B0.w * s0)));
            accumulate(eb.y, -(onFirst * (B1.w * s1)));
            selfFrictionContacts += 1;
            if (limit > R(0)) maximumConeRatio = std::max(maximumConeRatio, static_cast<double>(tangential / limit));
        }
        for (uint32_t i = 0; i < P; ++i) {
            V<R> move;
            if (blendedMove(i, move)) p[i].v = p[i].v + move;
        }
    }

    // ---------------------------------------------------------- finalize --
    void finalize(const NumiClothBagGPUConfig& c) {
        const bool ground = c.constraintCounts.z != 0u;
        const R inverseDt = R(1) / R(c.gravityAndTimestep.w);
        const R mu = R(c.clothMaterial.z);
        for (Particle<R>& q : p) {
            q.v = (q.x - q.prev) * inverseDt;
            q.groundImpulse = 0;
            if (ground && q.w > R(0) && q.x.z <= R(c.clothMaterial.x) + R(1.0e-6)) {
                q.groundImpulse = std::max(R(0), (q.v.z - q.predictedVz) / q.w);
            }
            if (!ground || !(q.groundImpulse > R(0)) || !(q.w > R(0))) continue;
            const V<R> tangent{q.v.x, q.v.y, R(0)};
            const R slip = norm(tangent);
            if (!(slip > R(1.0e-10))) continue;
            const R limit = mu * q.groundImpulse;
            const R jt = std::min(slip / q.w, limit);
            if (!(jt > R(0))) continue;
            q.v = q.v - tangent * (jt * q.w / slip);
            frictionContacts += 1;
            if (limit > R(0)) maximumConeRatio = std::max(maximumConeRatio, static_cast<double>(jt / limit));
        }
    }

    // ============================================================== bag ==
    // Fruit spheres, fruit/yarn records and air loads (reference
    // src/metal/ClothBag.metal:491-575, :675-905, :1025-1146, :1503-1614,
    // :1675-1786, :2733-2974, :3024-3335), in the order of the device
    // kernels in numi-solver_amd/csrc/kernels_bag.hpp.
    struct Fruit {
        V<R> x, prev, v, w;
        R q[4];
        R invMass = 0, radius = 0, groundImpulse = 0;
    };
    struct PairContact {
        V<R> n;
        R j = 0;
    };
    struct Yarn {
        V<R> normal, sweptNormal, fruitN;
        R separation = 0, advance = 0, wCurrent = 0, wSwept = 0, time = 0, target = 0;
        R fruitJ = 0, segA = 0, segB = 0, firstTime = 0, sweptAdvance = 0;
        uint32_t overlap = 0, impact = 0, degenerate = 0, responses = 0;
    };
    std::vector<Fruit> fruit;
    std::vector<PairContact> pairContact;
    std::vector<Yarn> yarn;
    std::vector<V<R>> airForce;
    bool yarnAir = false;
    uint32_t candidateMask = 0, releasedMask = 0;
    uint32_t pairFrictionContacts = 0, yarnFrictionContacts = 0, fruitGroundContacts = 0, rollingContacts = 0;
    double maximumRollingRatio = 0;
    double maximumYarnAeroForce = 0, maximumFruitAeroForce = 0, maximumFruitAeroTorque = 0;
    std::vector<double> mouthClearance;

    void initBag() {
        const uint32_t F = d.fruitCount;
        fruit.resize(F);
        for (uint32_t i = 0; i < F; ++i) {
            const NumiClothBagGPUFruit& s = d.fruits[i];
            Fruit& f = fruit[i];
            f.x = v3<R>(s.positionAndInverseMass);
            f.prev = v3<R>(s.previousAndRadius);
            f.v = v3<R>(s.velocityAndGroundImpulse);
            f.w = v3<R>(s.angularVelocity);
            f.q[0] = R(s.orientation.x);
            f.q[1] = R(s.orientation.y);
            f.q[2] = R(s.orientation.z);
            f.q[3] = R(s.orientation.w);
            f.invMass = R(s.positionAndInverseMass.w);
            f.radius = R(s.previousAndRadius.w);
            f.groundImpulse = R(s.velocityAndGroundImpulse.w);
        }
        pairContact.assign(d.fruitPairCount, PairContact{});
        yarn.assign(d.yarnContactCount, Yarn{});
        const NumiClothBagGPUConfig& c = *d.config;
        yarnAir = c.airVelocityAndDensity.w > 0.0f &&
                  (c.aerodynamicCoefficients.x > 0.0f || c.aerodynamicCoefficients.y > 0.0f) && D != 0;
        candidateMask = d.initialCandidateMask;
        releasedMask = d.initialReleasedMask;
        mouthClearance.assign(F, 0.0);
    }

    R inverseInertia(const Fruit& f) const { return R(2.5) * f.invMass / (f.radius * f.radius); }

    void applyImpulse(Fruit& f, V<R> impulse, V<R> offset) {
        f.v = f.v + impulse * f.invMass;
        f.w = f.w + cross(offset, impulse) * inverseInertia(f);
    }

    void fruitBegin(const NumiClothBagGPUConfig& c) {
        const R dt = R(c.gravityAndTimestep.w);
        for (Fruit& f : fruit) {
            f.prev = f.x;
            f.v = fma3(v3<R>(c.gravityAndTimestep), dt, f.v);
            f.groundImpulse = 0;
        }
        for (PairContact& p : pairContact) p = PairContact{};
        for (Yarn& y : yarn) {
            y.responses = 0;
            y.fruitN = V<R>{};
            y.fruitJ = 0;
            y.segA = y.segB = y.firstTime = y.sweptAdvance = 0;
        }
    }

    void yarnAirLoads(const NumiClothBagGPUConfig& c) {
        if (!yarnAir) return;
        airForce.assign(P, V<R>{});
        const V<R> air = v3<R>(c.airVelocityAndDensity);
        const R density = R(c.airVelocityAndDensity.w);
        const R pi = R(3.14159265358979323846);
        const R radius = R(c.clothMaterial.x);
        for (uint32_t b = 0; b < d.distanceBatchCount; ++b) {
            const mr_uint4 batch = d.distanceBatches[b].control;
            for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
                const mr_uint4 e = d.distances[k].particlesAndColor;
                const V<R> span = p[e.y].x - p[e.x].x;
                const R len = norm(span);
                if (!(len > R(1.0e-12))) continue;
                const V<R> axis = span / len;
                const V<R> rel = (p[e.x].v + p[e.y].v) * R(0.5) - air;
                const V<R> axial = axis * dot(rel, axis);
                const V<R> crossflow = rel - axial;
                const R crossCoeff = R(0.5) * density * R(c.aerodynamicCoefficients.x) * (R(2) * radius * len);
                const R skinCoeff = R(0.5) * density * R(c.aerodynamicCoefficients.y) * (R(2) * pi * radius * len);
                const V<R> force = crossflow * (-crossCoeff * norm(crossflow)) - axial * (skinCoeff * norm(axial));
                maximumYarnAeroForce = std::max(maximumYarnAeroForce, static_cast<double>(norm(force)));
                airForce[e.x] = airForce[e.x] + force * R(0.5);
                airForce[e.y] = airForce[e.y] + force * R(0.5);
            }
        }
        // 256 strided lane sums, then a pairwise tree (the device order).
        R power[256] = {}, weighted[256] = {};
        for (uint32_t lane = 0; lane < 256; ++lane) {
            for (uint32_t i = lane; i < P; i += 256) {
                if (p[i].w > R(0)) {
                    power[lane] += dot(airForce[i], p[i].v - air);
                    weighted[lane] += p[i].w * dot(airForce[i], airForce[i]);
                }
            }
        }
        for (uint32_t stride = 128; stride > 0; stride >>= 1)
            for (uint32_t lane = 0; lane < stride; ++lane) {
                power[lane] += power[lane + stride];
                weighted[lane] += weighted[lane + stride];
            }
        R attenuation = 1;
        if (weighted[0] > R(0)) {
            attenuation = power[0] < R(0) ? R(1) / (R(1) + R(c.gravityAndTimestep.w) * weighted[0] / -power[0]) : R(0);
        }
        const R dt = R(c.gravityAndTimestep.w);
        for (uint32_t i = 0; i < P; ++i) {
            if (p[i].w > R(0)) {
                p[i].v = p[i].v + airForce[i] * (attenuation * p[i].w * dt);
                p[i].predictedVz = p[i].v.z;
            }
        }
    }

    void fruitAirLoads(const NumiClothBagGPUConfig& c) {
        const R pi = R(3.14159265358979323846);
        const R density = R(c.airVelocityAndDensity.w);
        const R dt = R(c.gravityAndTimestep.w);
        const V<R> air = v3<R>(c.airVelocityAndDensity);
        for (Fruit& f : fruit) {
            const V<R> rel = f.v - air;
            const R speed = norm(rel);
            const R transCoeff = R(0.5) * density * R(c.aerodynamicCoefficients.z) * pi * f.radius * f.radius;
            const R transAtt = R(1) / (R(1) + f.invMass * transCoeff * speed * dt);
            f.v = air + rel * transAtt;
            const R spin = norm(f.w);
            const R r = f.radius;
            const R rotCoeff = (R(3) * pi * pi / R(8)) * density * R(c.aerodynamicCoefficients.w) * r * r * r * r * r;
            const R rotAtt = R(1) / (R(1) + inverseInertia(f) * rotCoeff * spin * dt);
            f.w = f.w * rotAtt;
            maximumFruitAeroForce = std::max(maximumFruitAeroForce, static_cast<double>(transCoeff * speed * speed));
            maximumFruitAeroTorque = std::max(maximumFruitAeroTorque, static_cast<double>(rotCoeff * spin * spin));
        }
    }

    void fruitAdvance(R dt) {
        for (Fruit& f : fruit) f.x = fma3(f.v, dt, f.x);
    }

    void fruitPairs(const NumiClothBagGPUConfig& c) {
        for (uint32_t b = 0; b < d.fruitPairBatchCount; ++b) {
            const mr_uint4 batch = d.fruitPairBatches[b].control;
            for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
                const mr_uint4 pr = d.fruitPairs[k].fruitsAndColor;
                Fruit& a = fruit[pr.x];
                Fruit& bf = fruit[pr.y];
                const V<R> diff = bf.x - a.x;
                const R len = norm(diff);
                const R target = a.radius + bf.radius;
                if (!(len < target) || !(len > R(1.0e-12))) continue;
                const R denom = a.invMass + bf.invMass;
                if (!(denom > R(0))) continue;
                const R lambda = (target - len) / denom;
                const V<R> n = diff / len;
                const V<R> corr = n * lambda;
                a.x = a.x - corr * a.invMass;
                bf.x = bf.x + corr * bf.invMass;
                const R impulse = lambda / R(c.gravityAndTimestep.w);
                pairContact[k].n = pairContact[k].n + n * impulse;
                pairContact[k].j += impulse;
            }
        }
    }

    void fruitGround(const NumiClothBagGPUConfig& c) {
        if (c.constraintCounts.z == 0u) return;
        for (Fruit& f : fruit) {
            const R penetration = f.radius - f.x.z;
            if (penetration > R(0)) {
                f.groundImpulse += penetration / (f.invMass * R(c.gravityAndTimestep.w));
                f.x.z = f.radius;
            }
        }
    }

    struct PointSample {
        V<R> closest;
        R weight, distance;
    };

    static PointSample pointSegment(V<R> point, V<R> a, V<R> b) {
        const V<R> dd = b - a;
        const R l2 = dot(dd, dd);
        const R w = l2 > R(1.0e-20) ? clampR(dot(dd, point - a) / l2, R(0), R(1)) : R(0);
        const V<R> cl = fma3(dd, w, a);
        return {cl, w, norm(cl - point)};
    }

    static V<R> safeUnit(V<R> v) {
        const R m = norm(v);
        return m > R(1.0e-14) ? v / m : V<R>{R(1), R(0), R(0)};
    }

    static V<R> degenerateNormal(V<R> a, V<R> b) {
        const V<R> dir = safeUnit(b - a);
        const V<R> axis = std::abs(dir.z) < R(0.9) ? V<R>{R(0), R(0), R(1)} : V<R>{R(1), R(0), R(0)};
        return safeUnit(cross(dir, axis));
    }

    void buildYarnContacts(const NumiClothBagGPUConfig& c) {
        const uint32_t F = d.fruitCount;
        if (yarn.empty()) return;
        for (uint32_t fi = 0; fi < F; ++fi) {
            const Fruit& f = fruit[fi];
            for (uint32_t si = 0; si < D; ++si) {
                Yarn& y = yarn[fi * D + si];
                const mr_uint4 e = d.distances[si].particlesAndColor;
                const V<R> a = p[e.x].x, b = p[e.y].x, a0 = p[e.x].prev, b0 = p[e.y].prev;
                const R target = f.radius + R(c.clothMaterial.x);
                const PointSample now = pointSegment(f.x, a, b);
                const bool degenerate = now.distance < R(1.0e-12);
                const V<R> normal = degenerate ? degenerateNormal(a, b) : (now.closest - f.x) / now.distance;
                R time = 0, sweptWeight = 0, removed = 0;
                V<R> sweptNormal{};
                bool impact = false;
                const R bound = norm(f.x - f.prev) + norm(a - a0) + norm(b - b0);
                if (bound >= R(1.0e-14)) {
                    auto at = [&](R t) { return pointSegment(mix3(f.prev, f.x, t), mix3(a0, a, t), mix3(b0, b, t)); };
                    PointSample hit = at(R(0));
                    bool found = hit.distance <= target + R(1.0e-9);
                    if (!found) {
                        for (int step = 0; step < 80; ++step) {
                            hit = at(time);
                            const R gap = hit.distance - target;
                            if (gap <= R(1.0e-9)) {
                                found = true;
                                break;
                            }
                            const R adv = R(0.9) * gap / bound;
                            if (!std::isfinite(adv) || !(adv > R(0)) || time + adv >= R(1)) break;
                            time += std::max(adv, R(1.0e-10));
                        }
                    }
                    if (found && hit.distance >= R(1.0e-12)) {
                        const V<R> ct = mix3(f.prev, f.x, time), at_ = mix3(a0, a, time), bt = mix3(b0, b, time);
                        sweptNormal = (hit.closest - ct) / hit.distance;
                        sweptWeight = hit.weight;
                        const V<R> segRest = (a - at_) * (R(1) - sweptWeight) + (b - bt) * sweptWeight;
                        removed = dot((f.x - ct) - segRest, sweptNormal);
                        impact = removed > R(0);
                        if (!impact) removed = 0;
                    }
                }
                y.normal = normal;
                y.separation = now.distance - target;
                y.sweptNormal = sweptNormal;
                y.advance = removed;
                y.wCurrent = now.weight;
                y.wSwept = sweptWeight;
                y.time = time;
                y.target = target;
                y.overlap = now.distance < target ? 1u : 0u;
                y.impact = impact ? 1u : 0u;
                y.degenerate = degenerate ? 1u : 0u;
            }
        }
    }

    R yarnCorrection(Particle<R>& first, Particle<R>& second, Fruit& f, R wa, R wb, V<R> n, R distance,
                     bool groundEnabled, R floorZ) {
        V<R> a = first.x, b = second.x;
        const R ia = first.w, ib = second.w;
        bool aActive = groundEnabled && n.z < R(0) && wa > R(0) && a.z <= floorZ + R(1.0e-9);
        bool bActive = groundEnabled && n.z < R(0) && wb > R(0) && b.z <= floorZ + R(1.0e-9);
        if (aActive) a.z = floorZ;
        if (bActive) b.z = floorZ;
        R remaining = distance, total = 0;
        for (int k = 0; k < 3 && remaining > R(1.0e-14); ++k) {
            V<R> ra = n * ia, rb = n * ib;
            if (aActive) ra.z = 0;
            if (bActive) rb.z = 0;
            const R denom = f.invMass + dot(n, ra) * wa * wa + dot(n, rb) * wb * wb;
            if (!(denom > R(0))) break;
            const R freeStep = remaining / denom;
            R step = freeStep;
            if (groundEnabled && n.z < R(0)) {
                const R va = ra.z * wa;
                if (!aActive && va < R(0)) step = std::min(step, std::max(R(0), a.z - floorZ) / -va);
                const R vb = rb.z * wb;
                if (!bActive && vb < R(0)) step = std::min(step, std::max(R(0), b.z - floorZ) / -vb);
            }
            f.x = f.x - n * (f.invMass * step);
            a = a + ra * (wa * step);
            b = b + rb * (wb * step);
            total += step;
            remaining = std::max(R(0), remaining - denom * step);
            if (step >= freeStep - R(1.0e-14)) break;
            bool activated = false;
            if (!aActive && wa > R(0) && a.z <= floorZ + R(1.0e-9)) {
                a.z = floorZ;
                aActive = true;
                activated = true;
            }
            if (!bActive && wb > R(0) && b.z <= floorZ + R(1.0e-9)) {
                b.z = floorZ;
                bActive = true;
                activated = true;
            }
            if (!activated) break;
        }
        first.x = a;
        second.x = b;
        return total;
    }

    // Phase order of the reference: fruit f visits segment (phase + f) mod n;
    // within a phase the fruits touch disjoint segments of one colour.
    void yarnBatches(const NumiClothBagGPUConfig& c, uint32_t mode) {
        if (yarn.empty()) return;
        const uint32_t F = d.fruitCount;
        const bool ground = c.constraintCounts.z != 0u;
        const R floorZ = R(c.clothMaterial.x);
        const R dt = R(c.gravityAndTimestep.w);
        for (uint32_t bi = 0; bi < d.distanceBatchCount; ++bi) {
            const mr_uint4 batch = d.distanceBatches[bi].control;
            for (uint32_t phase = 0; phase < batch.y; ++phase) {
                for (uint32_t fi = 0; fi < F; ++fi) {
                    const uint32_t segment = batch.x + (phase + fi) % batch.y;
                    const mr_uint4 e = d.distances[segment].particlesAndColor;
                    Yarn& y = yarn[fi * D + segment];
                    Fruit& f = fruit[fi];
                    V<R> n{};
                    R wa = 0, wb = 0, correction = 0;
                    if (mode == 1) {
                        if (y.impact) {
                            n = y.sweptNormal;
                            wb = y.wSwept;
                            wa = R(1) - wb;
                            correction = y.advance;
                        }
                    } else {
                        const PointSample s = pointSegment(f.x, p[e.x].x, p[e.y].x);
                        const R target = f.radius + R(c.clothMaterial.x);
                        if (s.distance < target) {
                            n = s.distance < R(1.0e-12) ? degenerateNormal(p[e.x].x, p[e.y].x)
                                                        : (s.closest - f.x) / s.distance;
                            wb = s.weight;
                            wa = R(1) - wb;
                            correction = target - s.distance;
                        }
                    }
                    if (!(correction > R(0))) continue;
                    const R lambda = yarnCorrection(p[e.x], p[e.y], f, wa, wb, n, correction, ground, floorZ);
                    if (lambda > R(0)) {
                        const R impulse = lambda / dt;
                        y.fruitN = y.fruitN - n * impulse;
                        y.fruitJ += impulse;
                        y.segA += wa * impulse;
                        y.segB += wb * impulse;
                        if (mode == 1) {
                            y.firstTime = y.time;
                            y.sweptAdvance += correction;
                        }
                        y.responses += 1;
                    }
                }
            }
        }
    }

    void finalizeFruit(const NumiClothBagGPUConfig& c) {
        for (Fruit& f : fruit) f.v = (f.x - f.prev) / R(c.gravityAndTimestep.w);
    }

    void yarnFriction(const NumiClothBagGPUConfig& c) {
        if (yarn.empty()) return;
        const uint32_t F = d.fruitCount;
        const R mu = R(c.fruitMaterial.w);
        const bool ground = c.constraintCounts.z != 0u;
        const R band = R(c.clothMaterial.x) + R(1.0e-6);
        for (uint32_t bi = 0; bi < d.distanceBatchCount; ++bi) {
            const mr_uint4 batch = d.distanceBatches[bi].control;
            for (uint32_t phase = 0; phase < batch.y; ++phase) {
                for (uint32_t fi = 0; fi < F; ++fi) {
                    const uint32_t segment = batch.x + (phase + fi) % batch.y;
                    const mr_uint4 e = d.distances[segment].particlesAndColor;
                    const Yarn& y = yarn[fi * D + segment];
                    Fruit& f = fruit[fi];
                    const R jn = y.fruitJ;
                    const R nl = norm(y.fruitN);
                    if (!(jn > R(0)) || !(nl > R(1.0e-10))) continue;
                    const V<R> n = y.fruitN / nl;
                    R wx = y.segA / jn, wy = y.segB / jn;
                    const R ws = wx + wy;
                    if (!(ws > R(1.0e-12))) continue;
                    wx = wx / ws;
                    wy = wy / ws;
                    Particle<R>& a = p[e.x];
                    Particle<R>& b = p[e.y];
                    const V<R> yarnV = a.v * wx + b.v * wy;
                    const V<R> offset = n * -f.radius;
                    const V<R> rel = (f.v + cross(f.w, offset)) - yarnV;
                    const V<R> tv = rel - n * dot(rel, n);
                    const R slip = norm(tv);
                    if (!(slip > R(1.0e-10))) continue;
                    const V<R> t = tv / slip;
                    const V<R> lever = cross(offset, t);
                    R denom = f.invMass + inverseInertia(f) * dot(lever, lever);
                    V<R> ra = t * a.w, rb = t * b.w;
                    if (ground) {
                        if (a.x.z <= band && ra.z < R(0)) ra.z = 0;
                        if (b.x.z <= band && rb.z < R(0)) rb.z = 0;
                    }
                    denom += dot(t, ra) * wx * wx + dot(t, rb) * wy * wy;
                    if (!(denom > R(0))) continue;
                    const R limit = mu * jn;
                    const R jt = std::min(slip / denom, limit);
                    if (!(jt > R(0))) continue;
                    applyImpulse(f, t * -jt, offset);
                    a.v = a.v + ra * (jt * wx);
                    b.v = b.v + rb * (jt * wy);
                    yarnFrictionContacts += 1;
                    if (limit > R(0)) maximumConeRatio = std::max(maximumConeRatio, static_cast<double>(jt / limit));
                }
            }
        }
    }

    void pairFriction(const NumiClothBagGPUConfig& c) {
        const R mu = R(c.fruitMaterial.x);
        for (uint32_t bi = 0; bi < d.fruitPairBatchCount; ++bi) {
            const mr_uint4 batch = d.fruitPairBatches[bi].control;
            for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
                const mr_uint4 pr = d.fruitPairs[k].fruitsAndColor;
                const R jn = pairContact[k].j;
                const R nl = norm(pairContact[k].n);
                if (!(jn > R(0)) || !(nl > R(1.0e-10))) continue;
                Fruit& a = fruit[pr.x];
                Fruit& b = fruit[pr.y];
                const V<R> n = pairContact[k].n / nl;
                const V<R> oa = n * a.radius, ob = n * -b.radius;
                const V<R> rel = (b.v + cross(b.w, ob)) - (a.v + cross(a.w, oa));
                const V<R> tv = rel - n * dot(rel, n);
                const R slip = norm(tv);
                if (!(slip > R(1.0e-10))) continue;
                const V<R> t = tv / slip;
                const V<R> la = cross(oa, t), lb = cross(ob, t);
                const R denom = a.invMass + b.invMass + inverseInertia(a) * dot(la, la) + inverseInertia(b) * dot(lb, lb);
                if (!(denom > R(0))) continue;
                const R limit = mu * jn;
                const R jt = std::min(slip / denom, limit);
                if (!(jt > R(0))) continue;
                const V<R> onB = t * -jt;
                applyImpulse(b, onB, ob);
                applyImpulse(a, -onB, oa);
                pairFrictionContacts += 1;
                if (limit > R(0)) maximumConeRatio = std::max(maximumConeRatio, static_cast<double>(jt / limit));
            }
        }
    }

    void fruitGroundFriction(const NumiClothBagGPUConfig& c) {
        if (c.constraintCounts.z == 0u) return;
        const R mu = R(c.fruitMaterial.y), rolling = R(c.fruitMaterial.z);
        for (Fruit& f : fruit) {
            const R jn = f.groundImpulse;
            if (!(jn > R(0))) continue;
            const V<R> n{R(0), R(0), R(1)};
            const V<R> offset{R(0), R(0), -f.radius};
            const V<R> cv = f.v + cross(f.w, offset);
            const V<R> tv = cv - n * dot(cv, n);
            const R slip = norm(tv);
            if (slip > R(1.0e-10)) {
                const V<R> t = tv / slip;
                const V<R> lever = cross(offset, t);
                const R denom = f.invMass + inverseInertia(f) * dot(lever, lever);
                const R limit = mu * jn;
                const R jt = std::min(slip / denom, limit);
                if (jt > R(0)) {
                    applyImpulse(f, t * -jt, offset);
                    fruitGroundContacts += 1;
                    if (limit > R(0)) maximumConeRatio = std::max(maximumConeRatio, static_cast<double>(jt / limit));
                }
            }
            const V<R> roll{f.w.x, f.w.y, R(0)};
            const R rollSpeed = norm(roll);
            if (rollSpeed > R(1.0e-12) && rolling > R(0)) {
                const R inertia = inverseInertia(f);
                const R limit = rolling * f.radius * jn;
                const R angular = std::min(rollSpeed / inertia, limit);
                f.w = f.w - roll * (angular * inertia / rollSpeed);
                rollingContacts += 1;
                if (limit > R(0)) maximumRollingRatio = std::max(maximumRollingRatio, static_cast<double>(angular / limit));
            }
        }
    }

    void orientation(const NumiClothBagGPUConfig& c) {
        const R h = R(0.5) * R(c.gravityAndTimestep.w);
        for (Fruit& f : fruit) {
            const V<R> qv{f.q[0], f.q[1], f.q[2]};
            const V<R> dv = f.w * f.q[3] + cross(f.w, qv);
            const R dw = -dot(f.w, qv);
            const R a[4] = {std::fma(dv.x, h, f.q[0]), std::fma(dv.y, h, f.q[1]), std::fma(dv.z, h, f.q[2]),
                            std::fma(dw, h, f.q[3])};
            const R m2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
            const R s = R(1) / std::sqrt(m2);
            for (int k = 0; k < 4; ++k) f.q[k] = a[k] * s;
        }
    }

    bool insideMouth(V<R> centre, V<R> tangent, V<R> bitangent, V<R> point, R clearance) const {
        const uint32_t n = d.config->mouthControl.x;
        const V<R> rp = point - centre;
        const R px = dot(rp, tangent), py = dot(rp, bitangent);
        bool inside = false;
        R nearest = R(INFINITY);
        uint32_t second = n - 1;
        for (uint32_t first = 0; first < n; ++first) {
            const V<R> a = p[d.mouthRimParticles[first]].x - centre;
            const V<R> b = p[d.mouthRimParticles[second]].x - centre;
            const R ax = dot(a, tangent), ay = dot(a, bitangent), bx = dot(b, tangent), by = dot(b, bitangent);
            const R ex = bx - ax, ey = by - ay;
            const R l2 = ex * ex + ey * ey;
            const R t = l2 > R(1.0e-20) ? clampR(((px - ax) * ex + (py - ay) * ey) / l2, R(0), R(1)) : R(0);
            const R sx = px - (ax + t * ex), sy = py - (ay + t * ey);
            nearest = std::min(nearest, sx * sx + sy * sy);
            if ((ay > py) != (by > py)) {
                const R cx = ax + (bx - ax) * (py - ay) / (by - ay);
                if (px < cx) inside = !inside;
            }
            second = first;
        }
        return inside || nearest <= clearance * clearance;
    }

    R rimDistance(V<R> point) const {
        const uint32_t n = d.config->mouthControl.x;
        R nearest = R(INFINITY);
        for (uint32_t k = 0; k < n; ++k) {
            const V<R> a = p[d.mouthRimParticles[k]].x;
            const V<R> e = p[d.mouthRimParticles[(k + 1) % n]].x - a;
            const R l2 = dot(e, e);
            const R t = l2 > R(1.0e-20) ? clampR(dot(point - a, e) / l2, R(0), R(1)) : R(0);
            const V<R> s = point - (a + e * t);
            nearest = std::min(nearest, dot(s, s));
        }
        return std::sqrt(nearest);
    }

    void release(const NumiClothBagGPUConfig& c) {
        const uint32_t n = c.mouthControl.x;
        if (n == 0 || d.mouthRimCount == 0) return;
        V<R> centre{};
        for (uint32_t k = 0; k < n; ++k) centre = centre + p[d.mouthRimParticles[k]].x;
        centre = centre / R(n);
        V<R> area{};
        for (uint32_t k = 0; k < n; ++k)
            area = area + cross(p[d.mouthRimParticles[k]].x - centre, p[d.mouthRimParticles[(k + 1) % n]].x - centre);
        if (!(norm(area) > R(1.0e-8))) return;
        V<R> normal = area / norm(area);
        if (dot(normal, centre - p[c.mouthControl.y].x) < R(0)) normal = -normal;
        V<R> tangent = p[d.mouthRimParticles[0]].x - centre;
        tangent = tangent - normal * dot(tangent, normal);
        if (!(norm(tangent) > R(1.0e-8))) {
            tangent = cross(normal, std::abs(normal.z) < R(0.9) ? V<R>{R(0), R(0), R(1)} : V<R>{R(1), R(0), R(0)});
        }
        tangent = tangent / norm(tangent);
        V<R> bitangent = cross(normal, tangent);
        bitangent = bitangent / norm(bitangent);
        // Candidate bits set this substep are visible to the rim test of the
        // same fruit only (each fruit reads its own bit).
        for (uint32_t i = 0; i < fruit.size(); ++i) {
            const Fruit& f = fruit[i];
            const R required = f.radius + R(c.clothMaterial.x);
            const bool insideProjection = insideMouth(centre, tangent, bitangent, f.x, R(0));
            const bool nearProjection = insideMouth(centre, tangent, bitangent, f.x, required);
            const R outward = dot(f.x - centre, normal);
            const uint32_t mask = 1u << i;
            if (insideProjection && outward > R(0)) candidateMask |= mask;
            const R axial = outward - required;
            const R rimClear = rimDistance(f.x) - required;
            const bool throughCap = nearProjection && axial > R(c.mouthMaterial.x);
            const bool aroundRim = (candidateMask & mask) != 0u && !insideProjection && outward > R(0) &&
                                   rimClear > R(c.mouthMaterial.x);
            const R clearance = throughCap ? axial : (aroundRim ? rimClear : axial);
            mouthClearance[i] = std::max(mouthClearance[i], static_cast<double>(std::max(R(0), clearance)));
            if (throughCap || aroundRim) releasedMask |= mask;
        }
    }

    void ground(const NumiClothBagGPUConfig& c, bool colliders) {
        if (colliders) collide(c);
        fruitGround(c);
    }

    void finalContacts(const NumiClothBagGPUConfig& c, bool colliders) {
        fruitPairs(c);
        yarnBatches(c, 0);
        ground(c, colliders);
    }

    // Residuals of the independent CPU reference's certificate loop, each
    // max(0, ...) in the device's arithmetic: strain-limit violation
    // (measureStrainLimitViolation, tools/cloth_bag.cpp:1765-1779), ground
    // penetration of knots and fruits (measureGroundPenetration, :1880-1894),
    // fruit/yarn penetration (measureBallYarnPenetration, :3111-3137), segment
    // self-penetration over the self candidates (measurePrimitiveSelfPenetration,
    // :4182-4245); true when all four fall below the loop's thresholds
    // (:5171-5174). residuals() returns them in that order.
    std::array<R, 4> residuals(const NumiClothBagGPUConfig& c) const {
        const R r = R(c.clothMaterial.x);
        R strainMax = R(0), groundMax = R(0), yarnMax = R(0), selfMax = R(0);
        for (uint32_t k = 0; k < D; ++k) {
            const mr_uint4 e = d.distances[k].particlesAndColor;
            strainMax = std::max(strainMax, norm(p[e.y].x - p[e.x].x) - maxLength[k]);
        }
        if (c.constraintCounts.z != 0u) {
            for (const Particle<R>& q : p) groundMax = std::max(groundMax, r - q.x.z);
            for (const Fruit& f : fruit) groundMax = std::max(groundMax, f.radius - f.x.z);
        }
        for (const Fruit& f : fruit) {
            const R target = f.radius + r;
            for (uint32_t k = 0; k < D; ++k) {
                const mr_uint4 e = d.distances[k].particlesAndColor;
                yarnMax = std::max(yarnMax, target - pointSegment(f.x, p[e.x].x, p[e.y].x).distance);
            }
        }
        if (d.extension.selfContact) {
            for (const auto& pr : candidates) {
                const mr_uint4 ea = d.distances[pr.first].particlesAndColor;
                const mr_uint4 eb = d.distances[pr.second].particlesAndColor;
                selfMax = std::max(selfMax, R(2) * r - segments(p[ea.x].x, p[ea.y].x, p[eb.x].x, p[eb.y].x).distance);
            }
        }
        return {strainMax, groundMax, yarnMax, selfMax};
    }

    bool certified(const NumiClothBagGPUConfig& c) const {
        const std::array<R, 4> v = residuals(c);
        return static_cast<double>(v[3]) < 1.0e-6 && static_cast<double>(v[2]) < 1.0e-6 &&
               static_cast<double>(v[1]) < 1.0e-9 && static_cast<double>(v[0]) < 1.0e-6;
    }

    // certify: the last substep of a call, with extension.certificatePasses
    // set, runs the data-dependent certificate loop after the schedule's
    // certificate rounds (tools/cloth_bag.cpp:5143-5185).
    void substep(const NumiClothBagGPUConfig& c, uint32_t iterations, uint32_t sweeps, bool certify = false) {
        const R dt = R(c.gravityAndTimestep.w);
        time += static_cast<double>(c.gravityAndTimestep.w);
        const NumiClothHipSceneExtension& e = d.extension;
        begin(c, time);
        fruitBegin(c);
        attachGrips(c);
        yarnAirLoads(c);
        fruitAirLoads(c);
        advance(dt);
        fruitAdvance(dt);
        if (e.sweptGround) sweptGround(c);
        const bool self = e.selfContact != 0;
        if (self) {
            broadPhase(c);
            selfContact(c, 1);
        }
        buildYarnContacts(c);
        yarnBatches(c, 1);
        const bool colliders = d.config->constraintCounts.z != 0u || e.staticSphereCount != 0u;
        for (uint32_t it = 0; it < iterations; ++it) {
            distances(dt);
            knots(dt);
            if (it % 2 == 0) bends(c, dt);
            grips(c, dt);
            fruitPairs(c);
            yarnBatches(c, 0);
            if (self) selfContact(c, 0);
            if (e.pointSelfCollision) pointSelf(c);
            ground(c, colliders);
        }
        strain(sweeps);
        if (sweeps != 0) {
            for (int pass = 0; pass < 8; ++pass) {
                fruitPairs(c);
                yarnBatches(c, 0);
                strain(sweeps);
                ground(c, colliders);
            }
            for (int round = 0; round <= 8; ++round) {
                if (self) selfContact(c, 0);
                strain(sweeps);
                for (int pass = 0; pass < 2; ++pass) finalContacts(c, colliders);
            }
            if (certify) {
                for (uint32_t pass = 0; pass < e.certificatePasses; ++pass) {
                    if (self) selfContact(c, 0);
                    strain(sweeps);
                    for (int final = 0; final < 2; ++final) finalContacts(c, colliders);
                    certificatePasses = pass + 1;
                    if (certified(c)) break;
                }
            }
        }
        buildYarnContacts(c);
        finalize(c);
        selfFriction(c);
        finalizeFruit(c);
        yarnFriction(c);
        pairFriction(c);
        fruitGroundFriction(c);
        orientation(c);
        release(c);
    }
};

}  // namespace

extern "C" {

// Output arrays (any may be NULL). Positions/velocities/prev are P x 3.
typedef struct NumiClothOracleOutput {
    double* positions;
    double* previous;
    double* velocities;
    double* groundImpulse;    /* P */
    double* distanceLambda;   /* D */
    double* bendLambda;       /* B */
    double* knotLambda;       /* K */
    double* gripLambda;       /* G x 3 */
    uint32_t* gripParticle;   /* G */
    uint32_t failure;
    uint32_t frictionContacts;
    double maximumConeRatio;
    uint32_t presentContacts;
    uint32_t sweptContacts;
    uint32_t candidateCount;  /* candidates of the last broad phase */
    uint32_t certificatePasses; /* passes of the certificate loop at the run's last substep */
    double* fruits;           /* F x 20: the five float4 fields of a fruit record */
    double* fruitPairs;       /* FP x 4: contact (normal impulse vector, impulse) */
    double* yarnContacts;     /* Y x 24: the six float4/uint4 fields after identity */
    double* mouthClearance;   /* F: largest release clearance over the run */
    double maximumYarnAeroForce;
    double maximumFruitAeroForce;
    double maximumFruitAeroTorque;
    uint32_t candidateMask;
    uint32_t releasedMask;
    uint32_t pairFrictionContacts;
    uint32_t yarnFrictionContacts;
    uint32_t fruitGroundContacts;
    uint32_t rollingContacts;
    double maximumRollingRatio;
    uint32_t selfFrictionContacts;
    uint32_t pointSelfContacts;  /* pairs of the last point self-collision pass */
    /* Self-contact impulse aggregates of the last substep, one row of 11 per
     * candidate (s, t, records, sum of -n * impulse xyz, sum of impulses,
     * endpoint shares e0..e3), up to selfImpulseCapacity rows; the count of
     * candidates is written to selfImpulseCount. NULL skips them. */
    double* selfImpulses;
    uint32_t selfImpulseCapacity;
    uint32_t selfImpulseCount;
    /* Per grip, as the ABI grip record holds them: particle, attachment
     * generation, largest capture distance (float bits in the FP32 twin,
     * rounded to float in FP64) and the selected cuff-patch centre ring;
     * then the largest capture error. NULL skips them. */
    uint32_t* gripRecords;      /* G x 4 */
    double* gripCaptureError;   /* G */
} NumiClothOracleOutput;

}  // extern "C"

template <typename R>
static int runOracle(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs, uint32_t substeps,
                     uint32_t iterations, uint32_t sweeps, NumiClothOracleOutput* out) {
    if (scene == nullptr || configs == nullptr || out == nullptr) return 1;
    Oracle<R> o(*scene);
    for (uint32_t k = 0; k < substeps; ++k)
        o.substep(configs[k], iterations, sweeps, k + 1 == substeps && scene->extension.certificatePasses != 0);
    for (uint32_t i = 0; i < o.P; ++i) {
        const Particle<R>& q = o.p[i];
        if (out->positions) {
            out->positions[3 * i] = q.x.x;
            out->positions[3 * i + 1] = q.x.y;
            out->positions[3 * i + 2] = q.x.z;
        }
        if (out->previous) {
            out->previous[3 * i] = q.prev.x;
            out->previous[3 * i + 1] = q.prev.y;
            out->previous[3 * i + 2] = q.prev.z;
        }
        if (out->velocities) {
            out->velocities[3 * i] = q.v.x;
            out->velocities[3 * i + 1] = q.v.y;
            out->velocities[3 * i + 2] = q.v.z;
        }
        if (out->groundImpulse) out->groundImpulse[i] = q.groundImpulse;
    }
    for (uint32_t k = 0; k < o.D && out->distanceLambda; ++k) out->distanceLambda[k] = o.dLambda[k];
    for (uint32_t k = 0; k < o.B && out->bendLambda; ++k) out->bendLambda[k] = o.bLambda[k];
    for (uint32_t k = 0; k < o.K && out->knotLambda; ++k) out->knotLambda[k] = o.kLambda[k];
    for (uint32_t i = 0; i < o.G; ++i) {
        if (out->gripLambda) {
            out->gripLambda[3 * i] = o.gripLambda[i].x;
            out->gripLambda[3 * i + 1] = o.gripLambda[i].y;
            out->gripLambda[3 * i + 2] = o.gripLambda[i].z;
        }
        if (out->gripParticle) out->gripParticle[i] = o.gripParticle[i];
    }
    out->failure = o.failure;
    out->frictionContacts = o.frictionContacts;
    out->maximumConeRatio = o.maximumConeRatio;
    out->presentContacts = o.presentContacts;
    out->sweptContacts = o.sweptContacts;
    out->candidateCount = static_cast<uint32_t>(o.candidates.size());
    out->certificatePasses = o.certificatePasses;
    for (uint32_t i = 0; i < o.fruit.size() && out->fruits; ++i) {
        const auto& f = o.fruit[i];
        const double r[20] = {f.x.x,    f.x.y,    f.x.z,    f.invMass, f.prev.x, f.prev.y, f.prev.z,
                              f.radius, f.v.x,    f.v.y,    f.v.z,     f.groundImpulse, f.w.x, f.w.y,
                              f.w.z,    0.0,      f.q[0],   f.q[1],    f.q[2],   f.q[3]};
        std::copy(r, r + 20, out->fruits + 20 * i);
    }
    for (uint32_t k = 0; k < o.pairContact.size() && out->fruitPairs; ++k) {
        const auto& c = o.pairContact[k];
        const double r[4] = {c.n.x, c.n.y, c.n.z, c.j};
        std::copy(r, r + 4, out->fruitPairs + 4 * k);
    }
    for (uint32_t k = 0; k < o.yarn.size() && out->yarnContacts; ++k) {
        const auto& y = o.yarn[k];
        const double r[24] = {y.normal.x, y.normal.y, y.normal.z, y.separation,
                              y.sweptNormal.x, y.sweptNormal.y, y.sweptNormal.z, y.advance,
                              y.wCurrent, y.wSwept, y.time, y.target,
                              double(y.overlap), double(y.impact), double(y.degenerate), double(y.responses),
                              y.fruitN.x, y.fruitN.y, y.fruitN.z, y.fruitJ,
                              y.segA, y.segB, y.firstTime, y.sweptAdvance};
        std::copy(r, r + 24, out->yarnContacts + 24 * k);
    }
    for (uint32_t i = 0; i < o.mouthClearance.size() && out->mouthClearance; ++i)
        out->mouthClearance[i] = o.mouthClearance[i];
    out->maximumYarnAeroForce = o.maximumYarnAeroForce;
    out->maximumFruitAeroForce = o.maximumFruitAeroForce;
    out->maximumFruitAeroTorque = o.maximumFruitAeroTorque;
    out->candidateMask = o.candidateMask;
    out->releasedMask = o.releasedMask;
    out->pairFrictionContacts = o.pairFrictionContacts;
    out->yarnFrictionContacts = o.yarnFrictionContacts;
    out->fruitGroundContacts = o.fruitGroundContacts;
    out->rollingContacts = o.rollingContacts;
    out->maximumRollingRatio = o.maximumRollingRatio;
    out->selfFrictionContacts = o.selfFrictionContacts;
    out->pointSelfContacts = o.pointContacts;
    for (uint32_t i = 0; i < o.G; ++i) {
        if (out->gripRecords) {
            const float capture = static_cast<float>(o.gripCapture[i]);
            uint32_t bits;
            std::memcpy(&bits, &capture, sizeof bits);
            const uint32_t r[4] = {o.gripParticle[i], o.gripGeneration[i], bits, o.gripCentre[i]};
            std::copy(r, r + 4, out->gripRecords + 4 * i);
        }
        if (out->gripCaptureError) out->gripCaptureError[i] = o.gripError[i];
    }
    out->selfImpulseCount = static_cast<uint32_t>(o.selfImpulses.size());
    for (uint32_t k = 0; k < o.selfImpulses.size() && k < out->selfImpulseCapacity && out->selfImpulses; ++k) {
        const auto& a = o.selfImpulses[k];
        const double r[11] = {double(o.candidates[k].first), double(o.candidates[k].second), double(a.count),
                              a.n.x, a.n.y, a.n.z, a.impulse, a.e[0], a.e[1], a.e[2], a.e[3]};
        std::copy(r, r + 11, out->selfImpulses + 11 * k);
    }
    return 0;
}

extern "C" {

int numi_cloth_oracle_run(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs,
                          uint32_t substeps, uint32_t iterations, uint32_t sweeps, NumiClothOracleOutput* out) {
    return runOracle<double>(scene, configs, substeps, iterations, sweeps, out);
}

int numi_cloth_oracle_run_fp32(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs,
                               uint32_t substeps, uint32_t iterations, uint32_t sweeps, NumiClothOracleOutput* out) {
    return runOracle<float>(scene, configs, substeps, iterations, sweeps, out);
}

// FP32 broad phase on a given snapshot (the scene's arrays replaced by
// `particles`), for bit-exact comparison with the device candidate list.
// Writes up to `capacity` sorted (s, t) pairs; returns the total count.
int64_t numi_cloth_oracle_self_candidates_fp32(const NumiClothHipSceneDesc* scene,
                                               const NumiClothBagGPUParticle* particles,
                                               const NumiClothBagGPUConfig* config, uint32_t* pairs,
                                               uint32_t capacity) {
    if (scene == nullptr || particles == nullptr || config == nullptr) return -1;
    NumiClothHipSceneDesc copy = *scene;
    copy.particles = particles;
    copy.extension.selfContact = 1;
    Oracle<float> o(copy);
    o.broadPhase(*config);
    for (std::size_t k = 0; k < o.candidates.size() && k < capacity; ++k) {
        pairs[2 * k] = o.candidates[k].first;
        pairs[2 * k + 1] = o.candidates[k].second;
    }
    return static_cast<int64_t>(o.candidates.size());
}

// The certificate loop's four residuals (strain-limit violation, ground,
// fruit/yarn and segment self-penetration; each max(0, ...)) of the scene's
// own state under `config`, the self candidates from a broad phase on that
// state; precision 0 = FP64, 1 = FP32 twin. Returns 0, 1 on a null argument.
int numi_cloth_oracle_certificate_residuals(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* config,
                                            uint32_t precision, double* out) {
    if (scene == nullptr || config == nullptr || out == nullptr) return 1;
    auto measure = [&](auto oracle) {
        if (scene->extension.selfContact) oracle.broadPhase(*config);
        const auto v = oracle.residuals(*config);
        for (int k = 0; k < 4; ++k) out[k] = static_cast<double>(v[k]);
    };
    if (precision == 1) measure(Oracle<float>(*scene));
    else measure(Oracle<double>(*scene));
    return 0;
}

}  // extern "C"
