// Frame driver of the produce-bag trajectory replays (declared in
// include/numi_cloth_replay.h): pose generators, per-frame substep configs,
// the scenario's initial state, FNV-1a frame hashes, handle lag, OBJ dumps.
// Host code; the stepping goes through the solver's C ABI.

#include "numi_cloth_replay.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "common.hpp"
#include "scene_internal.hpp"

namespace numi_hip {
namespace {

constexpr double kSpinLift = 0.52;  // reference kSpinAirborneLift, tools/cloth_bag_metal.mm:86
constexpr double kPi = 3.14159265358979323846;

double smooth01(double v) {
    const double x = std::min(1.0, std::max(0.0, v));
    return x * x * (3.0 - 2.0 * x);
}

// Grip base point: the authored position of cuff knot (top level, ring 0).
void cuffBase(double base[3]) { bagAuthoredPosition(kBagLevels - 1, 0, base); }

// Reference pickupGripTarget (tools/cloth_bag_metal.mm:361-371): lift over
// 0.8 s, a snap down and sideways at 1.0 s, a recovery at 1.45 s.
void pickupTarget(double t, double out[3]) {
    double base[3];
    cuffBase(base);
    const double lift = smooth01(t / 0.80);
    const double snap = smooth01((t - 1.00) / 0.23);
    const double recovery = smooth01((t - 1.45) / 0.50);
    out[0] = base[0] + -0.10 * snap;
    out[1] = base[1] + 0.04 * std::sin(kPi * snap);
    out[2] = base[2] + (1.25 * lift - 0.85 * snap + 0.75 * recovery);
}

// Reference spinGripTarget (:373-388): a 0.28 m circle whose angular speed
// ramps to 4.8 rad/s with time constant 0.18 s, around the lifted base.
void spinTarget(double t, double out[3]) {
    double base[3];
    cuffBase(base);
    base[2] += kSpinLift;
    constexpr double radius = 0.28, speed = 4.8, ramp = 0.18;
    const double angle = speed * (t - ramp * (1.0 - std::exp(-t / ramp)));
    out[0] = base[0] + radius * (std::cos(angle) - 1.0);
    out[1] = base[1] + radius * std::sin(angle);
    out[2] = base[2] + 0.035 * std::sin(0.5 * angle);
}

NumiClothReplayPose poseAt(NumiClothReplayScenario scenario, double t, const NumiClothGripTrajectory* trajectory) {
    NumiClothReplayPose pose{};
    pose.orientation[3] = 1.0;
    pose.attachmentGeneration = 1;
    switch (scenario) {
        case NUMI_CLOTH_REPLAY_GROUNDED:
            cuffBase(pose.target);
            pose.active = 0;
            return pose;
        case NUMI_CLOTH_REPLAY_SPIN:
            spinTarget(t, pose.target);
            pose.active = 1;
            return pose;
        case NUMI_CLOTH_REPLAY_PICKUP:
            pickupTarget(t, pose.target);
            pose.active = 1;
            return pose;
        case NUMI_CLOTH_REPLAY_RECORDED: {
            if (trajectory == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "recorded replay needs a trajectory");
            NumiClothGripPose sample{};
            if (numi_cloth_grip_trajectory_sample(trajectory, t, &sample) != NUMI_CLOTH_HIP_OK)
                fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, numi_cloth_hip_last_error());
            cuffBase(pose.target);
            for (int k = 0; k < 3; ++k) pose.target[k] = pose.target[k] + sample.translation[k];
            for (int k = 0; k < 4; ++k) pose.orientation[k] = sample.orientation[k];
            pose.active = sample.active;
            pose.attachmentGeneration = sample.attachmentGeneration;
            return pose;
        }
    }
    fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "unknown replay scenario");
}

bool nearestPatch(const NumiClothGripTrajectory* trajectory) {
    if (trajectory == nullptr) return false;
    NumiClothGripTrajectoryInfo info{};
    if (numi_cloth_grip_trajectory_info(trajectory, &info) != NUMI_CLOTH_HIP_OK) return false;
    return info.selectNearestCuffPatch != 0;
}

// The grip fields of a config from a pose (reference gripQuaternion4 and the
// config updates of :5190-5203).
void applyPose(NumiClothBagGPUConfig& c, const NumiClothReplayPose& pose, bool nearest) {
    c.gripTargetAndActive = mr_float4{static_cast<float>(pose.target[0]), static_cast<float>(pose.target[1]),
                                      static_cast<float>(pose.target[2]), pose.active ? 1.0f : 0.0f};
    c.gripOrientation = mr_float4{static_cast<float>(pose.orientation[0]), static_cast<float>(pose.orientation[1]),
                                  static_cast<float>(pose.orientation[2]), static_cast<float>(pose.orientation[3])};
    c.gripControl.x = pose.attachmentGeneration;
    c.gripControl.y = nearest ? 1u : 0u;
}

void frameConfigs(const NumiClothBagGPUConfig& base, NumiClothReplayScenario scenario,
                  const NumiClothGripTrajectory* trajectory, uint32_t frame, uint32_t perFrame,
                  NumiClothBagGPUConfig* out) {
    const bool nearest = nearestPatch(trajectory);
    for (uint32_t s = 0; s < perFrame; ++s) {
        const uint64_t completed = static_cast<uint64_t>(frame) * perFrame + s + 1;
        out[s] = base;
        applyPose(out[s], poseAt(scenario, static_cast<double>(completed) * NUMI_CLOTH_REPLAY_TIMESTEP, trajectory),
                  nearest);
    }
}

struct Fnv {
    uint64_t value = 1469598103934665603ull;
    void bytes(const void* data, size_t count) {
        if (count == 0) return;
        const auto* b = static_cast<const uint8_t*>(data);
        for (size_t k = 0; k < count; ++k) {
            value ^= b[k];
            value *= 1099511628211ull;
        }
    }
};

uint64_t stateHash(const NumiClothHipSceneDesc& d, const NumiClothHipState& s) {
    const bool missing = (d.particleCount && !s.particles) || (d.distanceCount && !s.distances) ||
                         (d.gripCount && !s.grips) || (d.knotCount && !s.knots) || (d.bendCount && !s.bends) ||
                         (d.fruitCount && (!s.fruits || !s.mouthClearance)) || (d.fruitPairCount && !s.fruitPairs) ||
                         (d.yarnContactCount && !s.yarnContacts) || !s.selfStatus || !s.frictionStatus ||
                         !s.aerodynamicsStatus || !s.releaseStatus || !s.failure;
    if (missing) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "state hash needs every non-empty state array");
    Fnv h;
    h.bytes(s.particles, sizeof(NumiClothBagGPUParticle) * d.particleCount);
    h.bytes(s.distances, sizeof(NumiClothBagGPUDistance) * d.distanceCount);
    h.bytes(s.grips, sizeof(NumiClothBagGPUGrip) * d.gripCount);
    h.bytes(s.knots, sizeof(NumiClothBagGPUKnot) * d.knotCount);
    h.bytes(s.bends, sizeof(NumiClothBagGPUBend) * d.bendCount);
    h.bytes(s.fruits, sizeof(NumiClothBagGPUFruit) * d.fruitCount);
    h.bytes(s.fruitPairs, sizeof(NumiClothBagGPUFruitPair) * d.fruitPairCount);
    h.bytes(s.yarnContacts, sizeof(NumiClothBagGPUYarnContact) * d.yarnContactCount);
    h.bytes(s.selfStatus, sizeof(NumiClothBagGPUSelfStatus));
    h.bytes(s.frictionStatus, sizeof(NumiClothBagGPUFrictionStatus));
    h.bytes(s.aerodynamicsStatus, sizeof(NumiClothBagGPUAerodynamicsStatus));
    h.bytes(s.releaseStatus, sizeof(NumiClothBagGPUReleaseStatus));
    h.bytes(s.mouthClearance, sizeof(float) * d.fruitCount);
    h.value ^= *s.failure;
    h.value *= 1099511628211ull;
    return h.value;
}

void writeObj(const std::string& path, const NumiClothBagGPUParticle* particles, uint32_t particleCount,
              const NumiClothBagGPUFruit* fruits, uint32_t fruitCount, const NumiClothBagGPUGrip* grips,
              uint32_t gripCount, const NumiClothBagGPUConfig& c) {
    if (particleCount != kBagKnots)
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "OBJ export covers the produce-bag topology (1465 knots)");
    if (particles == nullptr || (fruitCount != 0 && fruits == nullptr) || (gripCount != 0 && grips == nullptr))
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null OBJ input");
    std::ofstream out(path);
    if (!out) fail(NUMI_CLOTH_HIP_ERROR_IO, "failed to open Metal cloth OBJ: " + path);
    out << std::setprecision(9);
    out << "# Numi Solver explicit-yarn Metal cloth bag\n";
    out << "# vertices " << particleCount << " render_triangles "
        << (2u * kBagAround * (kBagLevels - 1u) + 2u * (kBagGrid - 1u) * (kBagGrid - 1u)) << '\n';
    for (uint32_t i = 0; i < particleCount; ++i) {
        const mr_float4& x = particles[i].positionAndInverseMass;
        out << "v " << x.x << ' ' << x.y << ' ' << x.z << '\n';
    }
    for (uint32_t level = 0; level + 1 < kBagLevels; ++level) {
        for (uint32_t ring = 0; ring < kBagAround; ++ring) {
            const uint32_t next = (ring + 1) % kBagAround;
            const uint32_t a = bagWallIndex(level, ring) + 1, b = bagWallIndex(level, next) + 1;
            const uint32_t cc = bagWallIndex(level + 1, ring) + 1, d = bagWallIndex(level + 1, next) + 1;
            out << "f " << a << ' ' << b << ' ' << cc << '\n';
            out << "f " << b << ' ' << d << ' ' << cc << '\n';
        }
    }
    for (uint32_t row = 0; row + 1 < kBagGrid; ++row) {
        for (uint32_t column = 0; column + 1 < kBagGrid; ++column) {
            const uint32_t a = bagGridIndex(row, column) + 1, b = bagGridIndex(row, column + 1) + 1;
            const uint32_t cc = bagGridIndex(row + 1, column) + 1, d = bagGridIndex(row + 1, column + 1) + 1;
            if ((row + column) % 2 == 0) {
                out << "f " << a << ' ' << cc << ' ' << b << '\n';
                out << "f " << b << ' ' << cc << ' ' << d << '\n';
            } else {
                out << "f " << a << ' ' << d << ' ' << b << '\n';
                out << "f " << a << ' ' << cc << ' ' << d << '\n';
            }
        }
    }
    out << "# grip center " << c.gripTargetAndActive.x << ' ' << c.gripTargetAndActive.y << ' '
        << c.gripTargetAndActive.z << " active " << (c.gripTargetAndActive.w > 0.0f ? 1 : 0) << " orientation "
        << c.gripOrientation.w << ' ' << c.gripOrientation.x << ' ' << c.gripOrientation.y << ' '
        << c.gripOrientation.z << " patch_center " << (gripCount == 0 ? 0u : grips[0].particle.w) << '\n';
    for (uint32_t k = 0; k < fruitCount; ++k) {
        const NumiClothBagGPUFruit& f = fruits[k];
        out << "# ball " << k << " center " << f.positionAndInverseMass.x << ' ' << f.positionAndInverseMass.y << ' '
            << f.positionAndInverseMass.z << " radius " << f.previousAndRadius.w << " appearance " << f.identity.x
            << " orientation " << f.orientation.w << ' ' << f.orientation.x << ' ' << f.orientation.y << ' '
            << f.orientation.z << " angular_velocity " << f.angularVelocity.x << ' ' << f.angularVelocity.y << ' '
            << f.angularVelocity.z << '\n';
    }
    if (!out) fail(NUMI_CLOTH_HIP_ERROR_IO, "failed to write Metal cloth OBJ: " + path);
}

// q (xyzw) applied to v (reference rotateGripVector,
// include/numi/grip_trajectory.h:161-175).
void rotate(const mr_float4& q, const double v[3], double out[3]) {
    const double qx = q.x, qy = q.y, qz = q.z, qw = q.w;
    const double tx = 2.0 * (qy * v[2] - qz * v[1]);
    const double ty = 2.0 * (qz * v[0] - qx * v[2]);
    const double tz = 2.0 * (qx * v[1] - qy * v[0]);
    out[0] = v[0] + qw * tx + qy * tz - qz * ty;
    out[1] = v[1] + qw * ty + qz * tx - qx * tz;
    out[2] = v[2] + qw * tz + qx * ty - qy * tx;
}

struct HostState {
    std::vector<NumiClothBagGPUParticle> particles;
    std::vector<NumiClothBagGPUDistance> distances;
    std::vector<NumiClothBagGPUGrip> grips;
    std::vector<NumiClothBagGPUKnot> knots;
    std::vector<NumiClothBagGPUBend> bends;
    std::vector<NumiClothBagGPUFruit> fruits;
    std::vector<NumiClothBagGPUFruitPair> fruitPairs;
    std::vector<NumiClothBagGPUYarnContact> yarnContacts;
    NumiClothBagGPUSelfStatus selfStatus{};
    NumiClothBagGPUFrictionStatus frictionStatus{};
    NumiClothBagGPUAerodynamicsStatus aerodynamicsStatus{};
    NumiClothBagGPUReleaseStatus releaseStatus{};
    uint32_t failure = 0;
    std::vector<float> mouthClearance;

    explicit HostState(const NumiClothHipSceneDesc& d)
        : particles(d.particleCount), distances(d.distanceCount), grips(d.gripCount), knots(d.knotCount),
          bends(d.bendCount), fruits(d.fruitCount), fruitPairs(d.fruitPairCount), yarnContacts(d.yarnContactCount),
          mouthClearance(d.fruitCount) {}

    // Non-null pointers even for empty families (the hash reads 0 bytes).
    template <typename T>
    static T* at(std::vector<T>& v) {
        static T dummy{};
        return v.empty() ? &dummy : v.data();
    }

    NumiClothHipState view() {
        return NumiClothHipState{at(particles), at(distances), at(grips), at(knots), at(bends), at(fruits),
                                 at(fruitPairs), at(yarnContacts), &selfStatus, &frictionStatus,
                                 &aerodynamicsStatus, &releaseStatus, &failure, at(mouthClearance)};
    }
};

void check(int status) {
    if (status != NUMI_CLOTH_HIP_OK) fail(status, numi_cloth_hip_last_error());
}

void runReplay(NumiClothHipSolver* solver, const NumiClothHipSceneDesc& d, NumiClothReplayScenario scenario,
               const NumiClothGripTrajectory* trajectory, uint32_t frames, uint32_t iterations, uint32_t sweeps,
               uint32_t dumpEvery, const char* prefix, NumiClothReplayResult& r) {
    if (solver == nullptr || d.config == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null solver or scene");
    const uint32_t perFrame = NUMI_CLOTH_REPLAY_SUBSTEPS_PER_FRAME;
    if (prefix != nullptr) {
        if (dumpEvery == 0) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "OBJ dumps need dumpEvery > 0");
        writeObj(std::string(prefix) + "-0.obj", d.particles, d.particleCount, d.fruits, d.fruitCount, d.grips,
                 d.gripCount, *d.config);
    }
    HostState state(d);
    std::vector<NumiClothBagGPUConfig> configs(perFrame);
    r.framesRun = 0;
    r.failureFree = 1;
    r.maximumHandleLag = 0.0;
    r.stepSeconds = 0.0;
    for (uint32_t frame = 0; frame < frames; ++frame) {
        frameConfigs(*d.config, scenario, trajectory, frame, perFrame, configs.data());
        const auto start = std::chrono::steady_clock::now();
        check(numi_cloth_hip_step(solver, configs.data(), perFrame, iterations, sweeps));
        check(numi_cloth_hip_synchronize(solver));
        r.stepSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        NumiClothHipState view = state.view();
        check(numi_cloth_hip_read(solver, &view));
        ++r.framesRun;
        if (r.frameHashes != nullptr) r.frameHashes[frame] = stateHash(d, view);
        if (state.failure != NUMI_CLOTH_BAG_GPU_FAILURE_NONE) r.failureFree = 0;
        const NumiClothBagGPUConfig& last = configs.back();
        if (scenario != NUMI_CLOTH_REPLAY_GROUNDED && last.gripTargetAndActive.w > 0.0f) {
            for (const NumiClothBagGPUGrip& g : state.grips) {
                if (g.particle.x >= d.particleCount) continue;
                const mr_float4& x = state.particles[g.particle.x].positionAndInverseMass;
                const double offset[3] = {g.targetOffsetAndCompliance.x, g.targetOffsetAndCompliance.y,
                                          g.targetOffsetAndCompliance.z};
                double rotated[3];
                rotate(last.gripOrientation, offset, rotated);
                const double dx = x.x - (last.gripTargetAndActive.x + rotated[0]);
                const double dy = x.y - (last.gripTargetAndActive.y + rotated[1]);
                const double dz = x.z - (last.gripTargetAndActive.z + rotated[2]);
                r.maximumHandleLag = std::max(r.maximumHandleLag, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
        r.releasedMask = state.releaseStatus.masks.y;
        const uint32_t completed = frame + 1;
        if (prefix != nullptr && (completed % dumpEvery == 0 || completed == frames)) {
            writeObj(std::string(prefix) + "-" + std::to_string(completed) + ".obj", state.particles.data(),
                     d.particleCount, state.fruits.data(), d.fruitCount, state.grips.data(), d.gripCount, last);
        }
        if (!r.failureFree) break;
    }
}

}  // namespace
}  // namespace numi_hip

using namespace numi_hip;

extern "C" {

int numi_cloth_replay_pose(NumiClothReplayScenario scenario, double timeSeconds,
                           const NumiClothGripTrajectory* trajectory, NumiClothReplayPose* out) {
    return guarded([&] {
        if (out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null pose");
        *out = poseAt(scenario, timeSeconds, trajectory);
    });
}

int numi_cloth_replay_prepare(NumiClothHipScene* scene, NumiClothReplayScenario scenario,
                              const NumiClothGripTrajectory* trajectory) {
    return guarded([&] {
        if (scene == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null scene");
        if (scene->particles.size() != kBagKnots)
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "replays drive the produce-bag scene");
        if (scenario == NUMI_CLOTH_REPLAY_SPIN) {
            const float lift = static_cast<float>(kSpinLift);
            for (NumiClothBagGPUParticle& p : scene->particles) {
                p.positionAndInverseMass.z += lift;
                p.previousAndMass.z += lift;
            }
            for (NumiClothBagGPUFruit& f : scene->fruits) {
                f.positionAndInverseMass.z += lift;
                f.previousAndRadius.z += lift;
            }
        }
        scene->config.constraintCounts.z = scenario == NUMI_CLOTH_REPLAY_SPIN ? 0u : 1u;
        applyPose(scene->config, poseAt(scenario, 0.0, trajectory), nearestPatch(trajectory));
        scene->publish();
    });
}

int numi_cloth_replay_frame_configs(const NumiClothBagGPUConfig* base, NumiClothReplayScenario scenario,
                                    const NumiClothGripTrajectory* trajectory, uint32_t frame,
                                    uint32_t substepsPerFrame, NumiClothBagGPUConfig* out) {
    return guarded([&] {
        if (base == nullptr || out == nullptr || substepsPerFrame == 0)
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null config or no substeps");
        frameConfigs(*base, scenario, trajectory, frame, substepsPerFrame, out);
    });
}

int numi_cloth_state_hash(const NumiClothHipSceneDesc* scene, const NumiClothHipState* state, uint64_t* out) {
    return guarded([&] {
        if (scene == nullptr || state == nullptr || out == nullptr)
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null hash argument");
        *out = stateHash(*scene, *state);
    });
}

int numi_cloth_obj_write(const char* path, const NumiClothBagGPUParticle* particles, uint32_t particleCount,
                         const NumiClothBagGPUFruit* fruits, uint32_t fruitCount, const NumiClothBagGPUGrip* grips,
                         uint32_t gripCount, const NumiClothBagGPUConfig* config) {
    return guarded([&] {
        if (path == nullptr || config == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null OBJ argument");
        writeObj(path, particles, particleCount, fruits, fruitCount, grips, gripCount, *config);
    });
}

int numi_cloth_replay_run(NumiClothHipSolver* solver, const NumiClothHipSceneDesc* scene,
                          NumiClothReplayScenario scenario, const NumiClothGripTrajectory* trajectory,
                          uint32_t frames, uint32_t iterations, uint32_t strainSweeps, uint32_t dumpEvery,
                          const char* dumpPrefix, NumiClothReplayResult* out) {
    return guarded([&] {
        if (scene == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null replay argument");
        runReplay(solver, *scene, scenario, trajectory, frames, iterations, strainSweeps, dumpEvery, dumpPrefix,
                  *out);
    });
}

}  // extern "C"
