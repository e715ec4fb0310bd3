// FP64 reverse-mode adjoint of the cloth substep loop for contact-free
// specimens (hanging strips): test infrastructure only, like the rest of
// oracle/.
//
// Forward: the same substep schedule and FP64 arithmetic as the oracle's
// Oracle<double> for scenes made of distances, chord bends and strain limits
// with gravity, damping and kinematically driven pinned knots (reference
// begin_substep src/metal/ClothBag.metal:491-575, advance_positions
// :907-938, solve_distance :1148-1213, solve_bend :1315-1447, limit_strain
// :1616-1673, finalize_substep :2890-2921; pass order
// tools/cloth_bag_metal.mm:4211-4709). Every applied projection is recorded
// (endpoint positions and multiplier before it).
//
// Loss: L = sum over substeps s and tracked knots k of |x_k(s) - T_k(s)|^2,
// x_k(s) the position after substep s.
//
// Reverse: the recorded projections are replayed backwards with the
// hand-derived adjoint of each projection; begin/advance/finalize are
// differentiated per knot. Outputs dL/d(compliance) per distance and bend
// (through alpha = c / dt^2), dL/dw per knot (inverse mass) and dL/dc per
// knot (damping coefficient in 1 / (1 + c dt)).
//
// Parameter overrides (compliance scale per constraint, inverse-mass scale
// and damping per knot, all in double) let a test take central finite
// differences of the FP64 loss without FP32 rounding of the inputs.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "numi_cloth_hip.h"
#include "oracle_math.hpp"

namespace {


template <typename R>
struct Entry {
    uint8_t kind;  // 0 distance, 1 bend, 2 strain
    uint32_t k, a, b;
    V<R> xa, xb;
    R lambda;
};

// State before a substep: the reverse pass re-runs the substep from it with
// the tape on (one substep's tape is held at a time).
template <typename R>
struct SubstepRecord {
    std::vector<V<R>> x0, v0;
    double time0 = 0;
    R dt = 0;
};

struct Problem {
    const uint32_t* tracked = nullptr;
    uint32_t trackedCount = 0;
    const double* targets = nullptr;
    const double* distanceComplianceScale = nullptr;
    const double* bendComplianceScale = nullptr;
    const double* inverseMassScale = nullptr;
    const double* damping = nullptr;
};

template <typename R>
struct Model {
    using Vec = V<R>;
    static constexpr bool kTwin = sizeof(R) == sizeof(float);
    const NumiClothHipSceneDesc& d;
    uint32_t P, D, B;
    std::vector<Vec> x, v, prev, anchor;
    std::vector<R> w, damping, dCompliance, bCompliance, maxLength, dLambda, bLambda;
    std::vector<Entry<R>> tape;
    std::vector<SubstepRecord<R>> substeps;
    bool record = false;
    double time = 0;
    float dt0 = 0;

    Model(const NumiClothHipSceneDesc& scene, const Problem& pr)
        : d(scene), P(scene.particleCount), D(scene.distanceCount), B(scene.bendCount) {
        x.resize(P);
        v.resize(P);
        prev.resize(P);
        anchor.resize(P);
        w.resize(P);
        damping.resize(P);
        const NumiClothHipSceneExtension& e = d.extension;
        for (uint32_t i = 0; i < P; ++i) {
            const NumiClothBagGPUParticle& s = d.particles[i];
            x[i] = v3<R>(s.positionAndInverseMass);
            prev[i] = v3<R>(s.previousAndMass);
            v[i] = v3<R>(s.velocity);
            w[i] = R(double(s.positionAndInverseMass.w) * (pr.inverseMassScale ? pr.inverseMassScale[i] : 1.0));
            anchor[i] = x[i];
            const float c = e.particleMotion != nullptr ? e.particleMotion[i].w : e.damping;
            damping[i] = R(pr.damping != nullptr ? pr.damping[i] : double(c));
        }
        dCompliance.resize(D);
        maxLength.resize(D);
        for (uint32_t k = 0; k < D; ++k) {
            const mr_float4 m = d.distances[k].material;
            dCompliance[k] = R(double(m.y) * (pr.distanceComplianceScale ? pr.distanceComplianceScale[k] : 1.0));
            // The device precomputes rest * (1 + limit) in FP32.
            maxLength[k] = kTwin ? R(m.x * (1.0f + m.w)) : R(m.x) * (R(1) + R(m.w));
        }
        bCompliance.resize(B);
        for (uint32_t k = 0; k < B; ++k) {
            bCompliance[k] = R(double(d.bends[k].material.z) * (pr.bendComplianceScale ? pr.bendComplianceScale[k] : 1.0));
        }
        dLambda.assign(D, R(0));
        bLambda.assign(B, R(0));
        dt0 = d.config->gravityAndTimestep.w;
    }

    R factor(uint32_t i) const {
        if (!(damping[i] > R(0))) return R(1);
        return kTwin ? R(1.0f / (1.0f + float(damping[i]) * dt0)) : R(1) / (R(1) + damping[i] * R(dt0));
    }

    R driveOffset(uint32_t i, double endTime) const {
        const NumiClothHipSceneExtension& e = d.extension;
        float a = e.driveAmplitude, om = e.driveOmega, ph = e.drivePhase;
        if (e.particleMotion != nullptr) {
            a = e.particleMotion[i].x;
            om = e.particleMotion[i].y;
            ph = e.particleMotion[i].z;
        }
        return R(static_cast<float>(double(a) * (std::sin(double(om) * endTime + double(ph)) - std::sin(double(ph)))));
    }

    void distancePass(const NumiClothBagGPUBatch& batch, R dt) {
        for (uint32_t k = batch.control.x; k < batch.control.x + batch.control.y; ++k) {
            const mr_uint4 e = d.distances[k].particlesAndColor;
            const Vec diff = x[e.y] - x[e.x];
            const R len = norm(diff);
            const R alpha = dCompliance[k] * (R(1) / (dt * dt));
            const R denom = w[e.x] + w[e.y] + alpha;
            if (!(len > R(1.0e-12)) || !(denom > R(0))) continue;
            if (record) tape.push_back({0, k, e.x, e.y, x[e.x], x[e.y], dLambda[k]});
            const R dl = (-(len - R(d.distances[k].material.x)) - alpha * dLambda[k]) / denom;
            const Vec corr = diff * (dl / len);
            dLambda[k] += dl;
            x[e.x] = x[e.x] - corr * w[e.x];
            x[e.y] = x[e.y] + corr * w[e.y];
        }
    }

    void bendPass(const NumiClothBagGPUBatch& batch, R dt) {
        for (uint32_t k = batch.control.x; k < batch.control.x + batch.control.y; ++k) {
            const mr_uint4 e = d.bends[k].particlesAndColor;
            const uint32_t a = e.x, c = e.z;
            const Vec diff = x[c] - x[a];
            const R chord = norm(diff);
            if (!(chord > R(1.0e-12))) continue;
            const Vec dir = diff * (R(1) / chord);
            const Vec g[2] = {-dir, dir};
            const R alpha = bCompliance[k] * (R(1) / (dt * dt));
            R denom = alpha + w[a] * dot(g[0], g[0]) + w[c] * dot(g[1], g[1]);
            if (!(denom >= R(1.0e-16))) continue;
            if (record) tape.push_back({1, k, a, c, x[a], x[c], bLambda[k]});
            const R dl = (-(chord - R(d.bends[k].material.x)) - alpha * bLambda[k]) / denom;
            bLambda[k] += dl;
            x[a] = x[a] + g[0] * (w[a] * dl);
            x[c] = x[c] + g[1] * (w[c] * dl);
        }
    }

    void strainPass(const NumiClothBagGPUBatch& batch) {
        for (uint32_t k = batch.control.x; k < batch.control.x + batch.control.y; ++k) {
            const mr_uint4 e = d.distances[k].particlesAndColor;
            const Vec diff = x[e.y] - x[e.x];
            const R len = norm(diff);
            const R limit = maxLength[k];
            if (!(len > limit) || !(len > R(1.0e-12))) continue;
            const R denom = w[e.x] + w[e.y];
            if (!(denom > R(0))) continue;
            if (record) tape.push_back({2, k, e.x, e.y, x[e.x], x[e.y], R(0)});
            const R excess = len - limit;
            const Vec dir = diff * (R(1) / len);
            const R share = excess / denom;
            x[e.x] = x[e.x] + dir * (w[e.x] * share);
            x[e.y] = x[e.y] - dir * (w[e.y] * share);
        }
    }

    void strain(uint32_t sweeps) {
        for (uint32_t sweep = 0; sweep < sweeps; ++sweep)
            for (uint32_t b = 0; b < d.distanceBatchCount; ++b) strainPass(d.distanceBatches[b]);
    }

    SubstepRecord<R> snapshot(const NumiClothBagGPUConfig& c) const {
        return {x, v, time, R(c.gravityAndTimestep.w)};
    }

    void restore(const SubstepRecord<R>& rec) {
        x = rec.x0;
        v = rec.v0;
        time = rec.time0;
    }

    void substep(const NumiClothBagGPUConfig& c, uint32_t iterations, uint32_t sweeps) {
        const R dt = R(c.gravityAndTimestep.w);
        time += static_cast<double>(c.gravityAndTimestep.w);
        const Vec g = v3<R>(c.gravityAndTimestep);
        const bool driven = d.extension.driveAmplitude != 0.0f || d.extension.particleMotion != nullptr;
        for (uint32_t i = 0; i < P; ++i) {
            prev[i] = x[i];
            if (w[i] > R(0)) {
                const R f = factor(i);
                const Vec carried = f != R(1) ? v[i] * f : v[i];
                v[i] = fma3(g, dt, carried);
                x[i] = fma3(v[i], dt, x[i]);
            } else if (driven) {
                x[i] = {anchor[i].x, anchor[i].y, anchor[i].z + driveOffset(i, time)};
            }
        }
        std::fill(dLambda.begin(), dLambda.end(), R(0));
        std::fill(bLambda.begin(), bLambda.end(), R(0));
        for (uint32_t it = 0; it < iterations; ++it) {
            for (uint32_t b = 0; b < d.distanceBatchCount; ++b) distancePass(d.distanceBatches[b], dt);
            if (it % 2 == 0)
                for (uint32_t b = 0; b < d.bendBatchCount; ++b) bendPass(d.bendBatches[b], dt);
        }
        strain(sweeps);
        if (sweeps != 0) {
            for (int pass = 0; pass < 8; ++pass) strain(sweeps);
            for (int round = 0; round <= 8; ++round) strain(sweeps);
        }
        const R inverseDt = R(1) / dt;
        for (uint32_t i = 0; i < P; ++i) v[i] = (x[i] - prev[i]) * inverseDt;
    }
};

// Adjoint of one XPBD projection x_a -= w_a n dl, x_b += w_b n dl,
// lambda += dl with dl = (-(len - rest) - alpha lambda) / (w_a + w_b + alpha).
template <typename R>
void projectionAdjoint(const Entry<R>& e, R rest, R alpha, R wa, R wb, V<R>& ga, V<R>& gb, R& glambda, R& galpha,
                       R& gwa, R& gwb) {
    using Vec = V<R>;
    const Vec d = e.xb - e.xa;
    const R len = norm(d);
    const Vec n = d / len;
    const R denom = wa + wb + alpha;
    const R numerator = -(len - rest) - alpha * e.lambda;
    const R dl = numerator / denom;
    // corr = n dl; x_a' = x_a - w_a corr; x_b' = x_b + w_b corr
    const Vec corr = n * dl;
    const Vec gcorr = gb * wb - ga * wa;
    gwa += -dot(corr, ga);
    gwb += dot(corr, gb);
    const R gdl = glambda + dot(gcorr, n);
    Vec gn = gcorr * dl;
    const R gnum = gdl / denom;
    const R gden = -gdl * dl / denom;
    R glen = -gnum;
    galpha += -e.lambda * gnum + gden;
    glambda += -alpha * gnum;
    gwa += gden;
    gwb += gden;
    // n = d / len
    const Vec gd = (gn - n * dot(gn, n)) / len + n * glen;
    ga = ga - gd;
    gb = gb + gd;
}

// Adjoint of one strain limit x_a += w_a e/W n, x_b -= w_b e/W n,
// e = len - maxLength, W = w_a + w_b.
template <typename R>
void strainAdjoint(const Entry<R>& e, R maxLength, R wa, R wb, V<R>& ga, V<R>& gb, R& gwa, R& gwb) {
    using Vec = V<R>;
    const Vec d = e.xb - e.xa;
    const R len = norm(d);
    const Vec n = d / len;
    const R W = wa + wb;
    const R ex = len - maxLength;
    const R sa = wa * ex / W, sb = wb * ex / W;
    const Vec gn = ga * sa - gb * sb;
    const R gsa = dot(ga, n), gsb = -dot(gb, n);
    const R gex = (gsa * wa + gsb * wb) / W;
    const R gW = -(gsa * sa + gsb * sb) / W;
    gwa += gsa * ex / W + gW;
    gwb += gsb * ex / W + gW;
    const R glen = gex;
    const Vec gd = (gn - n * dot(gn, n)) / len + n * glen;
    ga = ga - gd;
    gb = gb + gd;
}

bool supported(const NumiClothHipSceneDesc& d) {
    const NumiClothBagGPUConfig& c = *d.config;
    return d.gripCount == 0 && d.knotCount == 0 && d.fruitCount == 0 && d.fruitPairCount == 0 &&
           d.yarnContactCount == 0 && d.extension.staticSphereCount == 0 && d.extension.selfContact == 0 &&
           d.extension.pointSelfCollision == 0 && d.extension.sweptGround == 0 && d.extension.certificatePasses == 0 &&
           c.constraintCounts.z == 0u && !(c.airVelocityAndDensity.w > 0.0f) && d.mouthRimCount == 0;
}

}  // namespace

extern "C" {

typedef struct NumiClothOracleAdjointProblem {
    const uint32_t* tracked;       /* knot indices in the loss */
    uint32_t trackedCount;
    uint32_t observable;           /* 0: sum |x - target|^2, 1: sum target . x (weights) */
    const double* targets;         /* substeps x trackedCount x 3 */
    const double* distanceComplianceScale;  /* D, or NULL */
    const double* bendComplianceScale;      /* B, or NULL */
    const double* inverseMassScale;         /* P, or NULL */
    const double* damping;                  /* P damping coefficients, or NULL */
} NumiClothOracleAdjointProblem;

typedef struct NumiClothOracleGradients {
    double loss;
    double* distanceCompliance;    /* D: dL/dc, or NULL */
    double* bendCompliance;        /* B */
    double* inverseMass;           /* P: dL/dw */
    double* damping;               /* P: dL/dc_damping */
    double* positions;             /* P x 3 after the run, or NULL */
    double* trajectory;            /* substeps x trackedCount x 3 tracked positions, or NULL */
    double* trajectoryLow;         /* the same, position - trajectory (nonzero only in extended precision) */
} NumiClothOracleGradients;

}  // extern "C"

namespace {

// R = double: the FP64 adjoint. R = float: its twin, every expression in the
// order of the device kernels (k_specimen_forward / k_specimen_reverse).
template <typename R>
int runAdjoint(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs, uint32_t substeps,
               uint32_t iterations, uint32_t sweeps, const NumiClothOracleAdjointProblem* problem,
               NumiClothOracleGradients* out) {
    using Vec = V<R>;
    if (scene == nullptr || configs == nullptr || problem == nullptr || out == nullptr) return 1;
    if (!supported(*scene)) return 5;
    Problem pr;
    pr.tracked = problem->tracked;
    pr.trackedCount = problem->trackedCount;
    pr.targets = problem->targets;
    pr.distanceComplianceScale = problem->distanceComplianceScale;
    pr.bendComplianceScale = problem->bendComplianceScale;
    pr.inverseMassScale = problem->inverseMassScale;
    pr.damping = problem->damping;
    Model<R> m(*scene, pr);
    const bool gradients = out->distanceCompliance || out->bendCompliance || out->inverseMass || out->damping;
    const uint32_t T = pr.trackedCount;
    if (problem->observable > 1u) return 1;
    const bool linear = problem->observable == 1u;
    std::vector<std::vector<Vec>> lossGrad(gradients ? substeps : 0);
    double loss = 0.0;
    for (uint32_t s = 0; s < substeps; ++s) {
        if (gradients) m.substeps.push_back(m.snapshot(configs[s]));
        m.substep(configs[s], iterations, sweeps);
        if (gradients) lossGrad[s].resize(T);
        for (uint32_t t = 0; t < T; ++t) {
            const double* target = pr.targets + 3 * (size_t(s) * T + t);
            const Vec w{R(target[0]), R(target[1]), R(target[2])};
            const Vec& x = m.x[pr.tracked[t]];
            if (out->trajectory != nullptr) {
                double* row = out->trajectory + 3 * (size_t(s) * T + t);
                row[0] = double(x.x), row[1] = double(x.y), row[2] = double(x.z);
                if (out->trajectoryLow != nullptr) {
                    double* low = out->trajectoryLow + 3 * (size_t(s) * T + t);
                    low[0] = double(x.x - R(row[0])), low[1] = double(x.y - R(row[1])), low[2] = double(x.z - R(row[2]));
                }
            }
            if (linear) {
                loss += double(x.x) * double(w.x) + double(x.y) * double(w.y) + double(x.z) * double(w.z);
                if (gradients) lossGrad[s][t] = w;
            } else {
                const Vec r = x - w;
                loss += double(r.x) * double(r.x) + double(r.y) * double(r.y) + double(r.z) * double(r.z);
                if (gradients) lossGrad[s][t] = Vec{R(2) * r.x, R(2) * r.y, R(2) * r.z};
            }
        }
    }
    out->loss = loss;
    if (out->positions != nullptr)
        for (uint32_t i = 0; i < m.P; ++i) {
            out->positions[3 * i] = m.x[i].x;
            out->positions[3 * i + 1] = m.x[i].y;
            out->positions[3 * i + 2] = m.x[i].z;
        }
    if (!gradients) return 0;

    const uint32_t P = m.P;
    std::vector<Vec> gx(P), gv(P);
    std::vector<R> gw(P, R(0)), gdamp(P, R(0)), gcd(m.D, R(0)), gcb(m.B, R(0)), gld(m.D), glb(m.B);
    m.record = true;
    for (uint32_t s = substeps; s-- > 0;) {
        const SubstepRecord<R>& rec = m.substeps[s];
        const R dt = rec.dt;
        m.restore(rec);
        m.tape.clear();
        m.substep(configs[s], iterations, sweeps);
        for (uint32_t t = 0; t < T; ++t) gx[pr.tracked[t]] = gx[pr.tracked[t]] + lossGrad[s][t];
        // finalize: v = (x - prev) / dt, prev = x0
        const R inverseDt = R(1) / dt;
        std::vector<Vec> gprev(P);
        for (uint32_t i = 0; i < P; ++i) {
            gx[i] = gx[i] + gv[i] * inverseDt;
            gprev[i] = gv[i] * -inverseDt;
        }
        std::fill(gld.begin(), gld.end(), R(0));
        std::fill(glb.begin(), glb.end(), R(0));
        const R alphaScale = R(1) / (dt * dt);
        for (size_t k = m.tape.size(); k-- > 0;) {
            const Entry<R>& e = m.tape[k];
            if (e.kind == 2) {
                strainAdjoint(e, m.maxLength[e.k], m.w[e.a], m.w[e.b], gx[e.a], gx[e.b], gw[e.a], gw[e.b]);
                continue;
            }
            const bool bend = e.kind == 1;
            const R compliance = bend ? m.bCompliance[e.k] : m.dCompliance[e.k];
            const R rest = bend ? R(m.d.bends[e.k].material.x) : R(m.d.distances[e.k].material.x);
            const R alpha = compliance * alphaScale;  // as the forward computes it
            R galpha = R(0);
            R& gl = bend ? glb[e.k] : gld[e.k];
            projectionAdjoint(e, rest, alpha, m.w[e.a], m.w[e.b], gx[e.a], gx[e.b], gl, galpha, gw[e.a], gw[e.b]);
            (bend ? gcb[e.k] : gcd[e.k]) += galpha * alphaScale;
        }
        // begin + advance: x = x0 + v' dt, v' = f v0 + g dt, prev = x0.
        for (uint32_t i = 0; i < P; ++i) {
            if (m.w[i] > R(0)) {
                const R f = m.factor(i);
                const Vec gvp = gx[i] * dt;
                gx[i] = gx[i] + gprev[i];
                gv[i] = gvp * f;
                if (m.damping[i] > R(0)) gdamp[i] += dot(gvp, rec.v0[i]) * (-R(m.dt0) * f * f);
            } else {
                gx[i] = Vec{};
                gv[i] = Vec{};
            }
        }
    }
    auto copyOut = [](double* dst, const std::vector<R>& src) {
        if (dst != nullptr)
            for (size_t k = 0; k < src.size(); ++k) dst[k] = double(src[k]);
    };
    copyOut(out->distanceCompliance, gcd);
    copyOut(out->bendCompliance, gcb);
    copyOut(out->inverseMass, gw);
    copyOut(out->damping, gdamp);
    return 0;
}

}  // namespace

extern "C" {

// Returns 0, 1 on a null argument, 5 when the scene has a family the
// adjoint does not cover (contacts, grips, knots, fruits, air).
int numi_cloth_oracle_adjoint(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs,
                              uint32_t substeps, uint32_t iterations, uint32_t sweeps,
                              const NumiClothOracleAdjointProblem* problem, NumiClothOracleGradients* out) {
    return runAdjoint<double>(scene, configs, substeps, iterations, sweeps, problem, out);
}

// The same algorithm in x87 extended precision (64-bit significand): finite
// differences of its loss resolve parameters whose effect on the loss lies
// below the FP64 trajectory's rounding noise at small steps.
int numi_cloth_oracle_adjoint_extended(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs,
                                       uint32_t substeps, uint32_t iterations, uint32_t sweeps,
                                       const NumiClothOracleAdjointProblem* problem, NumiClothOracleGradients* out) {
    return runAdjoint<long double>(scene, configs, substeps, iterations, sweeps, problem, out);
}

// The FP32 twin (parameter overrides must be NULL or exact in FP32).
int numi_cloth_oracle_adjoint_fp32(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs,
                                   uint32_t substeps, uint32_t iterations, uint32_t sweeps,
                                   const NumiClothOracleAdjointProblem* problem, NumiClothOracleGradients* out) {
    return runAdjoint<float>(scene, configs, substeps, iterations, sweeps, problem, out);
}

}  // extern "C"
