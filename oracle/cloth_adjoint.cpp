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

using Vec = V<double>;

struct Entry {
    uint8_t kind;  // 0 distance, 1 bend, 2 strain
    uint32_t k, a, b;
    Vec xa, xb;
    double lambda;
};

// State before a substep: the reverse pass re-runs the substep from it with
// the tape on (one substep's tape is held at a time).
struct SubstepRecord {
    std::vector<Vec> x0, v0;
    double time0 = 0, dt = 0;
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

struct Model {
    const NumiClothHipSceneDesc& d;
    uint32_t P, D, B;
    std::vector<Vec> x, v, prev, anchor;
    std::vector<double> w, damping, dCompliance, bCompliance, maxLength, dLambda, bLambda;
    std::vector<Entry> tape;
    std::vector<SubstepRecord> substeps;
    bool record = false;
    double time = 0;
    double dt0 = 0;

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
            x[i] = v3<double>(s.positionAndInverseMass);
            prev[i] = v3<double>(s.previousAndMass);
            v[i] = v3<double>(s.velocity);
            w[i] = double(s.positionAndInverseMass.w) * (pr.inverseMassScale ? pr.inverseMassScale[i] : 1.0);
            anchor[i] = x[i];
            const float c = e.particleMotion != nullptr ? e.particleMotion[i].w : e.damping;
            damping[i] = pr.damping != nullptr ? pr.damping[i] : double(c);
        }
        dCompliance.resize(D);
        maxLength.resize(D);
        for (uint32_t k = 0; k < D; ++k) {
            const mr_float4 m = d.distances[k].material;
            dCompliance[k] = double(m.y) * (pr.distanceComplianceScale ? pr.distanceComplianceScale[k] : 1.0);
            maxLength[k] = double(m.x) * (1.0 + double(m.w));
        }
        bCompliance.resize(B);
        for (uint32_t k = 0; k < B; ++k) {
            bCompliance[k] = double(d.bends[k].material.z) * (pr.bendComplianceScale ? pr.bendComplianceScale[k] : 1.0);
        }
        dLambda.assign(D, 0.0);
        bLambda.assign(B, 0.0);
        dt0 = d.config->gravityAndTimestep.w;
    }

    double factor(uint32_t i) const { return damping[i] > 0.0 ? 1.0 / (1.0 + damping[i] * dt0) : 1.0; }

    double driveOffset(uint32_t i, double endTime) const {
        const NumiClothHipSceneExtension& e = d.extension;
        float a = e.driveAmplitude, om = e.driveOmega, ph = e.drivePhase;
        if (e.particleMotion != nullptr) {
            a = e.particleMotion[i].x;
            om = e.particleMotion[i].y;
            ph = e.particleMotion[i].z;
        }
        return double(static_cast<float>(double(a) * std::sin(double(om) * endTime + double(ph))));
    }

    void distancePass(const NumiClothBagGPUBatch& batch, double dt) {
        for (uint32_t k = batch.control.x; k < batch.control.x + batch.control.y; ++k) {
            const mr_uint4 e = d.distances[k].particlesAndColor;
            const Vec diff = x[e.y] - x[e.x];
            const double len = norm(diff);
            const double alpha = dCompliance[k] / (dt * dt);
            const double denom = w[e.x] + w[e.y] + alpha;
            if (!(len > 1.0e-12) || !(denom > 0.0)) continue;
            if (record) tape.push_back({0, k, e.x, e.y, x[e.x], x[e.y], dLambda[k]});
            const double dl = (-(len - double(d.distances[k].material.x)) - alpha * dLambda[k]) / denom;
            const Vec corr = diff * (dl / len);
            dLambda[k] += dl;
            x[e.x] = x[e.x] - corr * w[e.x];
            x[e.y] = x[e.y] + corr * w[e.y];
        }
    }

    void bendPass(const NumiClothBagGPUBatch& batch, double dt) {
        for (uint32_t k = batch.control.x; k < batch.control.x + batch.control.y; ++k) {
            const mr_uint4 e = d.bends[k].particlesAndColor;
            const uint32_t a = e.x, c = e.z;
            const Vec diff = x[c] - x[a];
            const double chord = norm(diff);
            if (!(chord > 1.0e-12)) continue;
            const Vec dir = diff / chord;
            const Vec g[2] = {-dir, dir};
            const double alpha = bCompliance[k] / (dt * dt);
            double denom = alpha + w[a] * dot(g[0], g[0]) + w[c] * dot(g[1], g[1]);
            if (!(denom >= 1.0e-16)) continue;
            if (record) tape.push_back({1, k, a, c, x[a], x[c], bLambda[k]});
            const double dl = (-(chord - double(d.bends[k].material.x)) - alpha * bLambda[k]) / denom;
            bLambda[k] += dl;
            x[a] = x[a] + g[0] * (w[a] * dl);
            x[c] = x[c] + g[1] * (w[c] * dl);
        }
    }

    void strainPass(const NumiClothBagGPUBatch& batch) {
        for (uint32_t k = batch.control.x; k < batch.control.x + batch.control.y; ++k) {
            const mr_uint4 e = d.distances[k].particlesAndColor;
            const Vec diff = x[e.y] - x[e.x];
            const double len = norm(diff);
            const double limit = maxLength[k];
            if (!(len > limit) || !(len > 1.0e-12)) continue;
            const double denom = w[e.x] + w[e.y];
            if (!(denom > 0.0)) continue;
            if (record) tape.push_back({2, k, e.x, e.y, x[e.x], x[e.y], 0.0});
            const double excess = len - limit;
            const Vec dir = diff / len;
            x[e.x] = x[e.x] + dir * (w[e.x] * excess / denom);
            x[e.y] = x[e.y] - dir * (w[e.y] * excess / denom);
        }
    }

    void strain(uint32_t sweeps) {
        for (uint32_t sweep = 0; sweep < sweeps; ++sweep)
            for (uint32_t b = 0; b < d.distanceBatchCount; ++b) strainPass(d.distanceBatches[b]);
    }

    SubstepRecord snapshot(const NumiClothBagGPUConfig& c) const {
        return {x, v, time, double(c.gravityAndTimestep.w)};
    }

    void restore(const SubstepRecord& rec) {
        x = rec.x0;
        v = rec.v0;
        time = rec.time0;
    }

    void substep(const NumiClothBagGPUConfig& c, uint32_t iterations, uint32_t sweeps) {
        const double dt = c.gravityAndTimestep.w;
        time += static_cast<double>(c.gravityAndTimestep.w);
        const Vec g = v3<double>(c.gravityAndTimestep);
        const bool driven = d.extension.driveAmplitude != 0.0f || d.extension.particleMotion != nullptr;
        for (uint32_t i = 0; i < P; ++i) {
            prev[i] = x[i];
            if (w[i] > 0.0) {
                const double f = factor(i);
                const Vec carried = f != 1.0 ? v[i] * f : v[i];
                v[i] = fma3(g, dt, carried);
                x[i] = fma3(v[i], dt, x[i]);
            } else if (driven) {
                x[i] = {anchor[i].x, anchor[i].y, anchor[i].z + driveOffset(i, time)};
            }
        }
        std::fill(dLambda.begin(), dLambda.end(), 0.0);
        std::fill(bLambda.begin(), bLambda.end(), 0.0);
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
        const double inverseDt = 1.0 / dt;
        for (uint32_t i = 0; i < P; ++i) v[i] = (x[i] - prev[i]) * inverseDt;
    }
};

// Adjoint of one XPBD projection x_a -= w_a n dl, x_b += w_b n dl,
// lambda += dl with dl = (-(len - rest) - alpha lambda) / (w_a + w_b + alpha).
void projectionAdjoint(const Entry& e, double rest, double alpha, double wa, double wb, Vec& ga, Vec& gb,
                       double& glambda, double& galpha, double& gwa, double& gwb) {
    const Vec d = e.xb - e.xa;
    const double len = norm(d);
    const Vec n = d / len;
    const double denom = wa + wb + alpha;
    const double numerator = -(len - rest) - alpha * e.lambda;
    const double dl = numerator / denom;
    // corr = n dl; x_a' = x_a - w_a corr; x_b' = x_b + w_b corr
    const Vec corr = n * dl;
    const Vec gcorr = gb * wb - ga * wa;
    gwa += -dot(corr, ga);
    gwb += dot(corr, gb);
    const double gdl = glambda + dot(gcorr, n);
    Vec gn = gcorr * dl;
    const double gnum = gdl / denom;
    const double gden = -gdl * dl / denom;
    double glen = -gnum;
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
void strainAdjoint(const Entry& e, double maxLength, double wa, double wb, Vec& ga, Vec& gb, double& gwa,
                   double& gwb) {
    const Vec d = e.xb - e.xa;
    const double len = norm(d);
    const Vec n = d / len;
    const double W = wa + wb;
    const double ex = len - maxLength;
    const double sa = wa * ex / W, sb = wb * ex / W;
    const Vec gn = ga * sa - gb * sb;
    const double gsa = dot(ga, n), gsb = -dot(gb, n);
    const double gex = (gsa * wa + gsb * wb) / W;
    const double gW = -(gsa * sa + gsb * sb) / W;
    gwa += gsa * ex / W + gW;
    gwb += gsb * ex / W + gW;
    const double glen = gex;
    const Vec gd = (gn - n * dot(gn, n)) / len + n * glen;
    ga = ga - gd;
    gb = gb + gd;
}

bool supported(const NumiClothHipSceneDesc& d) {
    const NumiClothBagGPUConfig& c = *d.config;
    return d.gripCount == 0 && d.knotCount == 0 && d.fruitCount == 0 && d.fruitPairCount == 0 &&
           d.yarnContactCount == 0 && d.extension.staticSphereCount == 0 && d.extension.selfContact == 0 &&
           c.constraintCounts.z == 0u && !(c.airVelocityAndDensity.w > 0.0f) && d.mouthRimCount == 0;
}

}  // namespace

extern "C" {

typedef struct NumiClothOracleAdjointProblem {
    const uint32_t* tracked;       /* knot indices in the loss */
    uint32_t trackedCount;
    uint32_t reserved;
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
} NumiClothOracleGradients;

// Returns 0, 1 on a null argument, 5 when the scene has a family the
// adjoint does not cover (contacts, grips, knots, fruits, air).
int numi_cloth_oracle_adjoint(const NumiClothHipSceneDesc* scene, const NumiClothBagGPUConfig* configs,
                              uint32_t substeps, uint32_t iterations, uint32_t sweeps,
                              const NumiClothOracleAdjointProblem* problem, NumiClothOracleGradients* out) {
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
    Model m(*scene, pr);
    const bool gradients = out->distanceCompliance || out->bendCompliance || out->inverseMass || out->damping;
    const uint32_t T = pr.trackedCount;
    std::vector<std::vector<Vec>> lossGrad(gradients ? substeps : 0);
    double loss = 0.0;
    for (uint32_t s = 0; s < substeps; ++s) {
        if (gradients) m.substeps.push_back(m.snapshot(configs[s]));
        m.substep(configs[s], iterations, sweeps);
        if (gradients) lossGrad[s].resize(T);
        for (uint32_t t = 0; t < T; ++t) {
            const double* target = pr.targets + 3 * (size_t(s) * T + t);
            const Vec r = m.x[pr.tracked[t]] - Vec{target[0], target[1], target[2]};
            loss += dot(r, r);
            if (gradients) lossGrad[s][t] = r * 2.0;
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
    std::vector<double> gw(P, 0.0), gdamp(P, 0.0), gcd(m.D, 0.0), gcb(m.B, 0.0), gld(m.D), glb(m.B);
    m.record = true;
    for (uint32_t s = substeps; s-- > 0;) {
        const SubstepRecord& rec = m.substeps[s];
        const double dt = rec.dt;
        m.restore(rec);
        m.tape.clear();
        m.substep(configs[s], iterations, sweeps);
        for (uint32_t t = 0; t < T; ++t) gx[pr.tracked[t]] = gx[pr.tracked[t]] + lossGrad[s][t];
        // finalize: v = (x - prev) / dt, prev = x0
        const double inverseDt = 1.0 / dt;
        std::vector<Vec> gprev(P);
        for (uint32_t i = 0; i < P; ++i) {
            gx[i] = gx[i] + gv[i] * inverseDt;
            gprev[i] = gv[i] * -inverseDt;
        }
        std::fill(gld.begin(), gld.end(), 0.0);
        std::fill(glb.begin(), glb.end(), 0.0);
        for (size_t k = m.tape.size(); k-- > 0;) {
            const Entry& e = m.tape[k];
            if (e.kind == 2) {
                strainAdjoint(e, m.maxLength[e.k], m.w[e.a], m.w[e.b], gx[e.a], gx[e.b], gw[e.a], gw[e.b]);
                continue;
            }
            const bool bend = e.kind == 1;
            const double compliance = bend ? m.bCompliance[e.k] : m.dCompliance[e.k];
            const double rest = bend ? double(m.d.bends[e.k].material.x) : double(m.d.distances[e.k].material.x);
            const double alpha = compliance / (dt * dt);
            double galpha = 0.0;
            double& gl = bend ? glb[e.k] : gld[e.k];
            projectionAdjoint(e, rest, alpha, m.w[e.a], m.w[e.b], gx[e.a], gx[e.b], gl, galpha, gw[e.a], gw[e.b]);
            (bend ? gcb[e.k] : gcd[e.k]) += galpha / (dt * dt);
        }
        // begin + advance: x = x0 + v' dt, v' = f v0 + g dt, prev = x0.
        for (uint32_t i = 0; i < P; ++i) {
            if (m.w[i] > 0.0) {
                const double f = m.factor(i);
                const Vec gvp = gx[i] * dt;
                gx[i] = gx[i] + gprev[i];
                gv[i] = gvp * f;
                if (m.damping[i] > 0.0) gdamp[i] += dot(gvp, rec.v0[i]) * (-m.dt0 * f * f);
            } else {
                gx[i] = Vec{};
                gv[i] = Vec{};
            }
        }
    }
    if (out->distanceCompliance) std::memcpy(out->distanceCompliance, gcd.data(), m.D * sizeof(double));
    if (out->bendCompliance) std::memcpy(out->bendCompliance, gcb.data(), m.B * sizeof(double));
    if (out->inverseMass) std::memcpy(out->inverseMass, gw.data(), P * sizeof(double));
    if (out->damping) std::memcpy(out->damping, gdamp.data(), P * sizeof(double));
    return 0;
}

}  // extern "C"
