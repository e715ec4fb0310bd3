// CPU FP64 oracle of the XPBD cloth substep loop.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the
// product library.
//
// It restates, in double precision and in the same colour-batch order, the
// reference's schedule-identical oracle runOracle
// (reference tools/cloth_bag_metal.mm:3432-3902) generalised from one substep
// to a trajectory of substeps, using the per-kernel arithmetic of
// reference src/metal/ClothBag.metal:
//   begin_substep          :491-575     advance_positions    :907-938
//   update_grip_attachment :577-673     solve_distance       :1148-1213
//   solve_knot             :1215-1313   solve_bend           :1315-1447
//   solve_grip             :1449-1501   solve_ground         :1575-1614
//   limit_strain           :1616-1673   finalize_substep     :2890-2921
//   apply_cloth_ground_friction :2976-3022
// and the pass order of runGPU (reference tools/cloth_bag_metal.mm:4211-4709,
// pass counts :47-49). Scene facts the reference has no record for (static
// spheres, kinematic drive, damping) follow the definitions in
// include/numi_cloth_hip.h.
//
// Inputs are the ABI v13 arrays (FP32), widened to double once; state stays
// in double across substeps.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "numi_cloth_hip.h"

namespace {

struct D3 {
    double x = 0, y = 0, z = 0;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator-(D3 a) { return {-a.x, -a.y, -a.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
D3 operator/(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(D3 a) { return std::sqrt(dot(a, a)); }
D3 d3(const mr_float4& v) { return {v.x, v.y, v.z}; }

struct Quat {
    double x, y, z, w;
};
Quat q4(const mr_float4& v) { return {v.x, v.y, v.z, v.w}; }
// v + 2w(q x v) + q x (2 q x v)
D3 rotate(Quat q, D3 v) {
    const D3 u{q.x, q.y, q.z};
    const D3 t = cross(u, v) * 2.0;
    return v + t * q.w + cross(u, t);
}

struct Particle {
    D3 x, prev, v;
    double w = 0, mass = 0, predictedVz = 0, groundImpulse = 0;
};

struct State {
    std::vector<Particle> p;
    std::vector<double> dLambda, bLambda, kLambda;
    std::vector<D3> gripLambda, gripOffset;
    std::vector<uint32_t> gripParticle, gripGeneration, gripCentre;
    std::vector<double> gripCapture, gripError;
    std::vector<D3> anchor;
    uint32_t failure = 0;
    uint32_t frictionContacts = 0;
    double maximumConeRatio = 0;
    double time = 0;
};

struct Scene {
    const NumiClothHipSceneDesc* d;
    uint32_t P, D, B, K, G;
};

void beginSubstep(const Scene& s, State& st, const NumiClothBagGPUConfig& c, double driveOffset) {
    const double dt = c.gravityAndTimestep.w;
    const D3 g = d3(c.gravityAndTimestep);
    const double damping = s.d->extension.damping;
    const double factor = damping > 0.0 ? 1.0 / (1.0 + damping * static_cast<double>(s.d->config->gravityAndTimestep.w)) : 1.0;
    const bool driven = s.d->extension.driveAmplitude != 0.0f;
    for (uint32_t i = 0; i < s.P; ++i) {
        Particle& q = st.p[i];
        q.prev = q.x;
        q.groundImpulse = 0.0;
        q.predictedVz = 0.0;
        if (q.w > 0.0) {
            q.v = (factor != 1.0 ? q.v * factor : q.v) + g * dt;
            q.predictedVz = q.v.z;
        } else if (driven) {
            q.x = {st.anchor[i].x, st.anchor[i].y, st.anchor[i].z + driveOffset};
        }
    }
    std::fill(st.dLambda.begin(), st.dLambda.end(), 0.0);
    std::fill(st.bLambda.begin(), st.bLambda.end(), 0.0);
    std::fill(st.kLambda.begin(), st.kLambda.end(), 0.0);
    std::fill(st.gripLambda.begin(), st.gripLambda.end(), D3{});
}

void attachGrips(const Scene& s, State& st, const NumiClothBagGPUConfig& c) {
    if (!(c.gripTargetAndActive.w > 0.0f)) return;
    const D3 target = d3(c.gripTargetAndActive);
    const Quat q = q4(c.gripOrientation);
    for (uint32_t i = 0; i < s.G && i < c.control.w; ++i) {
        if (st.gripGeneration[i] == c.gripControl.x) continue;
        if (st.gripGeneration[i] == 0xffffffffu || st.gripGeneration[i] + 1u != c.gripControl.x) {
            st.failure |= NUMI_CLOTH_BAG_GPU_FAILURE_GRIP_CAPTURE;
            continue;
        }
        uint32_t particle = st.gripParticle[i];
        if (c.gripControl.y == 1u) {
            const uint32_t rings = c.gripControl.z;
            uint32_t centre = 0;
            double nearest = INFINITY;
            for (uint32_t ring = 0; ring < rings; ++ring) {
                const D3 sep = st.p[s.d->cuffParticles[rings + ring]].x - target;
                const double d2 = dot(sep, sep);
                if (d2 < nearest) {
                    nearest = d2;
                    centre = ring;
                }
            }
            const uint32_t width = 2u * c.gripControl.w + 1u;
            const uint32_t ring = (centre + rings + i % width - c.gripControl.w) % rings;
            particle = s.d->cuffParticles[(i / width) * rings + ring];
            st.gripParticle[i] = particle;
            st.gripCentre[i] = centre;
        }
        const D3 sep = st.p[particle].x - target;
        const double capture = norm(sep);
        if (capture > c.gripMaterial.x) {
            st.failure |= NUMI_CLOTH_BAG_GPU_FAILURE_GRIP_CAPTURE;
            continue;
        }
        const D3 local = rotate({-q.x, -q.y, -q.z, q.w}, sep);
        const double error = norm(st.p[particle].x - (target + rotate(q, local)));
        st.gripOffset[i] = local;
        st.gripGeneration[i] = c.gripControl.x;
        st.gripCapture[i] = std::max(st.gripCapture[i], capture);
        st.gripLambda[i] = D3{};
        st.gripError[i] = std::max(st.gripError[i], error);
    }
}

void advance(const Scene& s, State& st, double dt) {
    for (Particle& q : st.p)
        if (q.w > 0.0) q.x = q.x + q.v * dt;
}

void solveDistances(const Scene& s, State& st, double dt) {
    const NumiClothHipSceneDesc& d = *s.d;
    for (uint32_t b = 0; b < d.distanceBatchCount; ++b) {
        const mr_uint4 batch = d.distanceBatches[b].control;
        for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
            const NumiClothBagGPUDistance& r = d.distances[k];
            Particle& a = st.p[r.particlesAndColor.x];
            Particle& c = st.p[r.particlesAndColor.y];
            const D3 diff = c.x - a.x;
            const double len = norm(diff);
            const double alpha = static_cast<double>(r.material.y) / (dt * dt);
            const double denom = a.w + c.w + alpha;
            if (!(len > 1.0e-12) || !(denom > 0.0)) continue;
            const double dl = (-(len - r.material.x) - alpha * st.dLambda[k]) / denom;
            const D3 corr = diff * (dl / len);
            st.dLambda[k] += dl;
            a.x = a.x - corr * a.w;
            c.x = c.x + corr * c.w;
        }
    }
}

void solveKnots(const Scene& s, State& st, double dt) {
    const NumiClothHipSceneDesc& d = *s.d;
    for (uint32_t b = 0; b < d.knotBatchCount; ++b) {
        const mr_uint4 batch = d.knotBatches[b].control;
        for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
            const NumiClothBagGPUKnot& r = d.knots[k];
            Particle* q[4] = {&st.p[r.particles.x], &st.p[r.particles.y], &st.p[r.particles.z], &st.p[r.particles.w]};
            const D3 warpV = q[1]->x - q[0]->x;
            const D3 weftV = q[3]->x - q[2]->x;
            const double warpL = norm(warpV);
            const double weftL = norm(weftV);
            if (!(warpL > 1.0e-12) || !(weftL > 1.0e-12)) continue;
            const D3 warp = warpV / warpL;
            const D3 weft = weftV / weftL;
            const double cosine = std::clamp(dot(warp, weft), -1.0, 1.0);
            const D3 gWarp = (weft - warp * cosine) / warpL;
            const D3 gWeft = (warp - weft * cosine) / weftL;
            const D3 g[4] = {-gWarp, gWarp, -gWeft, gWeft};
            const double alpha = static_cast<double>(r.material.y) / (dt * dt);
            double denom = alpha;
            for (int m = 0; m < 4; ++m) denom += q[m]->w * dot(g[m], g[m]);
            if (!(denom > 0.0)) continue;
            const double dl = (-(cosine - r.material.x) - alpha * st.kLambda[k]) / denom;
            st.kLambda[k] += dl;
            for (int m = 0; m < 4; ++m) q[m]->x = q[m]->x + g[m] * (q[m]->w * dl);
        }
    }
}

void solveBends(const Scene& s, State& st, const NumiClothBagGPUConfig& c, double dt) {
    const NumiClothHipSceneDesc& d = *s.d;
    const bool ground = c.constraintCounts.z != 0u;
    const double floorZ = c.clothMaterial.x;
    for (uint32_t b = 0; b < d.bendBatchCount; ++b) {
        const mr_uint4 batch = d.bendBatches[b].control;
        for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
            const NumiClothBagGPUBend& r = d.bends[k];
            Particle* q[2] = {&st.p[r.particlesAndColor.x], &st.p[r.particlesAndColor.z]};
            const D3 diff = q[1]->x - q[0]->x;
            const double chord = norm(diff);
            if (!(chord > 1.0e-12)) continue;
            const D3 dir = diff / chord;
            const D3 g[2] = {-dir, dir};
            const double alpha = static_cast<double>(r.material.z) / (dt * dt);
            double freeDenom = alpha;
            for (int m = 0; m < 2; ++m) freeDenom += q[m]->w * dot(g[m], g[m]);
            if (!(freeDenom >= 1.0e-16)) continue;
            const double numerator = -(chord - r.material.x) - alpha * st.bLambda[k];
            const double freeDl = numerator / freeDenom;
            bool active[2];
            double denom = alpha;
            for (int m = 0; m < 2; ++m) {
                active[m] = ground && q[m]->x.z <= floorZ + 1.0e-9 && g[m].z * freeDl < 0.0;
                denom += q[m]->w * (dot(g[m], g[m]) - (active[m] ? g[m].z * g[m].z : 0.0));
            }
            if (!(denom >= 1.0e-16)) continue;
            const double dl = numerator / denom;
            double fraction = 1.0;
            if (ground) {
                for (int m = 0; m < 2; ++m) {
                    if (active[m]) continue;
                    const double down = g[m].z * q[m]->w * dl;
                    if (down < 0.0) fraction = std::min(fraction, std::max(0.0, q[m]->x.z - floorZ) / -down);
                }
            }
            const double applied = dl * fraction;
            st.bLambda[k] += applied;
            for (int m = 0; m < 2; ++m) {
                D3 corr = g[m] * (q[m]->w * applied);
                if (active[m]) corr.z = 0.0;
                q[m]->x = q[m]->x + corr;
                if (ground) q[m]->x.z = std::max(q[m]->x.z, floorZ);
            }
        }
    }
}

void solveGrips(const Scene& s, State& st, const NumiClothBagGPUConfig& c, double dt) {
    if (!(c.gripTargetAndActive.w > 0.0f)) return;
    const D3 target = d3(c.gripTargetAndActive);
    const Quat q = q4(c.gripOrientation);
    for (uint32_t i = 0; i < s.G && i < c.control.w; ++i) {
        if (st.gripGeneration[i] != c.gripControl.x) continue;
        Particle& p = st.p[st.gripParticle[i]];
        const double alpha = static_cast<double>(s.d->grips[i].targetOffsetAndCompliance.w) / (dt * dt);
        const double denom = p.w + alpha;
        if (!(denom > 0.0)) continue;
        const D3 value = p.x - (target + rotate(q, st.gripOffset[i]));
        const D3 dl = (-value - st.gripLambda[i] * alpha) / denom;
        st.gripLambda[i] = st.gripLambda[i] + dl;
        p.x = p.x + dl * p.w;
    }
}

void collide(const Scene& s, State& st, const NumiClothBagGPUConfig& c) {
    const bool ground = c.constraintCounts.z != 0u;
    const double r = c.clothMaterial.x;
    const NumiClothHipSceneExtension& e = s.d->extension;
    for (Particle& q : st.p) {
        if (ground) q.x.z = std::max(q.x.z, r);
        if (q.w > 0.0) {
            for (uint32_t k = 0; k < e.staticSphereCount; ++k) {
                const mr_float4& sphere = e.staticSpheres[k];
                const D3 diff = q.x - d3(sphere);
                const double dist = norm(diff);
                const double reach = static_cast<double>(sphere.w) + r;
                if (dist < reach && dist > 1.0e-12) q.x = d3(sphere) + diff * (reach / dist);
            }
        }
    }
}

void limitStrain(const Scene& s, State& st, uint32_t sweeps) {
    const NumiClothHipSceneDesc& d = *s.d;
    for (uint32_t sweep = 0; sweep < sweeps; ++sweep) {
        for (uint32_t b = 0; b < d.distanceBatchCount; ++b) {
            const mr_uint4 batch = d.distanceBatches[b].control;
            for (uint32_t k = batch.x; k < batch.x + batch.y; ++k) {
                const NumiClothBagGPUDistance& r = d.distances[k];
                Particle& a = st.p[r.particlesAndColor.x];
                Particle& c = st.p[r.particlesAndColor.y];
                const D3 diff = c.x - a.x;
                const double len = norm(diff);
                const double limit = static_cast<double>(r.material.x) * (1.0 + static_cast<double>(r.material.w));
                if (!(len > limit) || !(len > 1.0e-12)) continue;
                const double denom = a.w + c.w;
                if (!(denom > 0.0)) continue;
                const D3 dir = diff / len;
                const double excess = len - limit;
                a.x = a.x + dir * (a.w * excess / denom);
                c.x = c.x - dir * (c.w * excess / denom);
            }
        }
    }
}

void finalize(const Scene& s, State& st, const NumiClothBagGPUConfig& c, double dt) {
    const bool ground = c.constraintCounts.z != 0u;
    for (Particle& q : st.p) {
        q.v = (q.x - q.prev) / dt;
        q.groundImpulse = 0.0;
        if (ground && q.w > 0.0 && q.x.z <= static_cast<double>(c.clothMaterial.x) + 1.0e-6) {
            q.groundImpulse = std::max(0.0, (q.v.z - q.predictedVz) / q.w);
        }
        if (!ground || !(q.groundImpulse > 0.0) || !(q.w > 0.0)) continue;
        const D3 tangent{q.v.x, q.v.y, 0.0};
        const double slip = norm(tangent);
        if (!(slip > 1.0e-10)) continue;
        const double limit = static_cast<double>(c.clothMaterial.z) * q.groundImpulse;
        const double jt = std::min(slip / q.w, limit);
        if (!(jt > 0.0)) continue;
        q.v = q.v - tangent * (jt * q.w / slip);
        st.frictionContacts += 1;
        if (limit > 0.0) st.maximumConeRatio = std::max(st.maximumConeRatio, jt / limit);
    }
}

void substep(const Scene& s, State& st, const NumiClothBagGPUConfig& c, uint32_t iterations, uint32_t sweeps) {
    const double dt = c.gravityAndTimestep.w;
    st.time += dt;
    const NumiClothHipSceneExtension& e = s.d->extension;
    const double driveOffset = e.driveAmplitude != 0.0f
                                   ? static_cast<double>(static_cast<float>(
                                         e.driveAmplitude * std::sin(e.driveOmega * st.time + e.drivePhase)))
                                   : 0.0;
    beginSubstep(s, st, c, driveOffset);
    attachGrips(s, st, c);
    advance(s, st, dt);
    const bool colliders = s.d->config->constraintCounts.z != 0u || e.staticSphereCount != 0u;
    for (uint32_t it = 0; it < iterations; ++it) {
        solveDistances(s, st, dt);
        solveKnots(s, st, dt);
        if (it % 2 == 0) solveBends(s, st, c, dt);
        solveGrips(s, st, c, dt);
        if (colliders) collide(s, st, c);
    }
    limitStrain(s, st, sweeps);
    if (sweeps != 0) {
        for (int pass = 0; pass < 8; ++pass) {
            limitStrain(s, st, sweeps);
            if (colliders) collide(s, st, c);
        }
        limitStrain(s, st, sweeps);
        for (int pass = 0; pass < 2; ++pass)
            if (colliders) collide(s, st, c);
        for (int certificate = 0; certificate < 8; ++certificate) {
            limitStrain(s, st, sweeps);
            for (int pass = 0; pass < 2; ++pass)
                if (colliders) collide(s, st, c);
        }
    }
    finalize(s, st, c, dt);
}

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
} NumiClothOracleOutput;

// Runs `substeps` substeps (one config each) from the scene's initial state.
int numi_cloth_oracle_run(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs,
                          uint32_t substeps, uint32_t iterations, uint32_t sweeps, NumiClothOracleOutput* out) {
    if (scene == nullptr || configs == nullptr || out == nullptr) return 1;
    if (scene->fruitCount != 0 || scene->yarnContactCount != 0 || scene->extension.selfContact != 0) return 5;
    Scene s{scene, scene->particleCount, scene->distanceCount, scene->bendCount, scene->knotCount, scene->gripCount};
    State st;
    st.p.resize(s.P);
    st.anchor.resize(s.P);
    for (uint32_t i = 0; i < s.P; ++i) {
        const NumiClothBagGPUParticle& src = scene->particles[i];
        Particle& q = st.p[i];
        q.x = d3(src.positionAndInverseMass);
        q.prev = d3(src.previousAndMass);
        q.v = d3(src.velocity);
        q.w = src.positionAndInverseMass.w;
        q.mass = src.previousAndMass.w;
        st.anchor[i] = q.x;
    }
    st.dLambda.assign(s.D, 0.0);
    st.bLambda.assign(s.B, 0.0);
    st.kLambda.assign(s.K, 0.0);
    st.gripLambda.assign(s.G, D3{});
    st.gripOffset.resize(s.G);
    st.gripParticle.resize(s.G);
    st.gripGeneration.resize(s.G);
    st.gripCentre.resize(s.G);
    st.gripCapture.assign(s.G, 0.0);
    st.gripError.assign(s.G, 0.0);
    for (uint32_t i = 0; i < s.G; ++i) {
        const NumiClothBagGPUGrip& g = scene->grips[i];
        st.gripOffset[i] = d3(g.targetOffsetAndCompliance);
        st.gripParticle[i] = g.particle.x;
        st.gripGeneration[i] = g.particle.y;
        st.gripCentre[i] = g.particle.w;
    }
    for (uint32_t k = 0; k < substeps; ++k) substep(s, st, configs[k], iterations, sweeps);
    for (uint32_t i = 0; i < s.P; ++i) {
        const Particle& q = st.p[i];
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
    if (out->distanceLambda) std::copy(st.dLambda.begin(), st.dLambda.end(), out->distanceLambda);
    if (out->bendLambda) std::copy(st.bLambda.begin(), st.bLambda.end(), out->bendLambda);
    if (out->knotLambda) std::copy(st.kLambda.begin(), st.kLambda.end(), out->knotLambda);
    for (uint32_t i = 0; i < s.G; ++i) {
        if (out->gripLambda) {
            out->gripLambda[3 * i] = st.gripLambda[i].x;
            out->gripLambda[3 * i + 1] = st.gripLambda[i].y;
            out->gripLambda[3 * i + 2] = st.gripLambda[i].z;
        }
        if (out->gripParticle) out->gripParticle[i] = st.gripParticle[i];
    }
    out->failure = st.failure;
    out->frictionContacts = st.frictionContacts;
    out->maximumConeRatio = st.maximumConeRatio;
    return 0;
}

}  // extern "C"
