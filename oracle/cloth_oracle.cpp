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
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "numi_cloth_hip.h"

namespace {

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

template <typename R>
struct Particle {
    V<R> x, prev, v;
    R w = 0, predictedVz = 0, groundImpulse = 0;
};

// Self-contact accumulator: FP64 sums exactly in double; the FP32 twin uses
// the device's 2^-32 m fixed point so that sums are order independent.
struct Fixed {
    long long x = 0, y = 0, z = 0;
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
    uint32_t failure = 0;
    uint32_t frictionContacts = 0;
    uint32_t presentContacts = 0, sweptContacts = 0;
    double maximumConeRatio = 0;
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
        if (d.extension.selfContact) buildLocality();
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
    void begin(const NumiClothBagGPUConfig& c, R driveOffset) {
        const R dt = R(c.gravityAndTimestep.w);
        const V<R> g = v3<R>(c.gravityAndTimestep);
        const float damping = d.extension.damping;
        const float dt0 = d.config->gravityAndTimestep.w;
        const R factor = damping > 0.0f
                             ? (sizeof(R) == sizeof(float) ? R(1.0f / (1.0f + damping * dt0))
                                                           : R(1) / (R(1) + R(damping) * R(dt0)))
                             : R(1);
        const bool driven = d.extension.driveAmplitude != 0.0f;
        for (uint32_t i = 0; i < P; ++i) {
            Particle<R>& q = p[i];
            q.prev = q.x;
            q.groundImpulse = 0;
            q.predictedVz = 0;
            if (q.w > R(0)) {
                const V<R> carried = factor != R(1) ? q.v * factor : q.v;
                q.v = fma3(g, dt, carried);
                q.predictedVz = q.v.z;
            } else if (driven) {
                q.x = {anchor[i].x, anchor[i].y, anchor[i].z + driveOffset};
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
                const R alpha = R(r.material.y) / (dt * dt);
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
                const V<R> dir = diff / chord;
                const V<R> g[2] = {-dir, dir};
                const R alpha = R(r.material.z) / (dt * dt);
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
                    const V<R> dir = diff / len;
                    const R excess = len - limit;
                    a.x = a.x + dir * (a.w * excess / denom);
                    c.x = c.x - dir * (c.w * excess / denom);
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
            f.x += std::llrint(static_cast<double>(delta.x) * 4294967296.0);
            f.y += std::llrint(static_cast<double>(delta.y) * 4294967296.0);
            f.z += std::llrint(static_cast<double>(delta.z) * 4294967296.0);
            f.count += 1;
        } else {
            deltaSum[i] = deltaSum[i] + delta;
            deltaCount[i] += 1;
        }
    }

    // Jacobi self-contact solve over the candidates, one snapshot.
    void selfContact(const NumiClothBagGPUConfig& c, uint32_t mode) {
        if (!d.extension.selfContact) return;
        const R r = R(c.clothMaterial.x);
        const R target = R(2) * r;
        const R tolerance = R(1.0e-9);
        fixed.assign(P, Fixed{});
        deltaSum.assign(P, V<R>{});
        deltaCount.assign(P, 0u);
        for (const auto& pr : candidates) {
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
        }
        for (uint32_t i = 0; i < P; ++i) {
            if (sizeof(R) == sizeof(float)) {
                const Fixed& f = fixed[i];
                if (f.count == 0) continue;
                const double scale = (1.0 / 4294967296.0) / static_cast<double>(f.count);
                p[i].x = p[i].x + V<R>{R(static_cast<double>(f.x) * scale), R(static_cast<double>(f.y) * scale),
                                       R(static_cast<double>(f.z) * scale)};
            } else {
                if (deltaCount[i] == 0) continue;
                p[i].x = p[i].x + deltaSum[i] / R(deltaCount[i]);
            }
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

    void substep(const NumiClothBagGPUConfig& c, uint32_t iterations, uint32_t sweeps) {
        const R dt = R(c.gravityAndTimestep.w);
        time += static_cast<double>(c.gravityAndTimestep.w);
        const NumiClothHipSceneExtension& e = d.extension;
        const R driveOffset = e.driveAmplitude != 0.0f
                                  ? R(static_cast<float>(e.driveAmplitude * std::sin(e.driveOmega * time + e.drivePhase)))
                                  : R(0);
        begin(c, driveOffset);
        attachGrips(c);
        advance(dt);
        const bool self = e.selfContact != 0;
        if (self) {
            broadPhase(c);
            selfContact(c, 1);
        }
        const bool colliders = d.config->constraintCounts.z != 0u || e.staticSphereCount != 0u;
        for (uint32_t it = 0; it < iterations; ++it) {
            distances(dt);
            knots(dt);
            if (it % 2 == 0) bends(c, dt);
            grips(c, dt);
            if (self) selfContact(c, 0);
            if (colliders) collide(c);
        }
        strain(sweeps);
        if (sweeps != 0) {
            for (int pass = 0; pass < 8; ++pass) {
                strain(sweeps);
                if (colliders) collide(c);
            }
            if (self) selfContact(c, 0);
            strain(sweeps);
            for (int pass = 0; pass < 2; ++pass)
                if (colliders) collide(c);
            for (int certificate = 0; certificate < 8; ++certificate) {
                if (self) selfContact(c, 0);
                strain(sweeps);
                for (int pass = 0; pass < 2; ++pass)
                    if (colliders) collide(c);
            }
        }
        finalize(c);
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
    uint32_t reserved;
} NumiClothOracleOutput;

}  // extern "C"

template <typename R>
static int runOracle(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs, uint32_t substeps,
                     uint32_t iterations, uint32_t sweeps, NumiClothOracleOutput* out) {
    if (scene == nullptr || configs == nullptr || out == nullptr) return 1;
    if (scene->fruitCount != 0 || scene->yarnContactCount != 0) return 5;
    Oracle<R> o(*scene);
    for (uint32_t k = 0; k < substeps; ++k) o.substep(configs[k], iterations, sweeps);
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

}  // extern "C"
