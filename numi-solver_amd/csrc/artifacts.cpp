// Input artifacts of the cloth path: the numi.cloth.material.v1 material file
// (reference include/numi/cloth_material.h:15-192) and the
// numi.grip.trajectory v1/v2/v3 CSV (reference
// include/numi/grip_trajectory.h:18-362). Field names, bounds, accepted
// syntax and error messages are kept identical because the reference's
// comparator scripts match them (reference cmake/ExpectMaterialFailure.cmake,
// cmake/ExpectGripTrajectoryFailure.cmake).

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "numi_cloth_scene.h"

namespace numi_hip {

thread_local std::string gLastError;

void setLastError(const std::string& message) { gLastError = message; }

namespace {

std::string trimmed(const std::string& text) {
    std::size_t begin = 0;
    std::size_t end = text.size();
    auto space = [](char c) {
        return c == ' ' || c == '\t' || c == '\r' || c == '\n';
    };
    while (begin < end && space(text[begin])) ++begin;
    while (end > begin && space(text[end - 1])) --end;
    return text.substr(begin, end - begin);
}

// std::stod with the "whole token, finite" rule of the reference parsers.
bool parseWholeNumber(const std::string& text, double* value) {
    std::size_t used = 0;
    try {
        *value = std::stod(text, &used);
    } catch (const std::exception&) {
        return false;
    }
    return used == text.size();
}

void copyText(char* destination, std::size_t capacity, const std::string& text) {
    std::snprintf(destination, capacity, "%s", text.c_str());
}

// ---- material -------------------------------------------------------------

struct MaterialField {
    const char* name;
    double NumiClothMaterialValues::*member;
    double minimum;
    double maximum;
};

const MaterialField kMaterialFields[] = {
    {"ordinary_node_mass_kg", &NumiClothMaterialValues::ordinaryNodeMassKg, 0.000010, 0.000200},
    {"hem_node_mass_kg", &NumiClothMaterialValues::hemNodeMassKg, 0.000020, 0.000400},
    {"yarn_radius_m", &NumiClothMaterialValues::yarnRadiusM, 0.0005, 0.010},
    {"axial_body_compliance_m_per_n", &NumiClothMaterialValues::axialBodyComplianceMPerN, 1.0e-11, 1.0e-4},
    {"axial_cuff_compliance_m_per_n", &NumiClothMaterialValues::axialCuffComplianceMPerN, 1.0e-12, 1.0e-5},
    {"knot_compliance", &NumiClothMaterialValues::knotCompliance, 1.0e-10, 1.0e-2},
    {"bend_body_compliance_m_per_n", &NumiClothMaterialValues::bendBodyComplianceMPerN, 1.0e-5, 100.0},
    {"bend_cuff_compliance_m_per_n", &NumiClothMaterialValues::bendCuffComplianceMPerN, 1.0e-12, 1.0e-3},
    {"grip_compliance_m_per_n", &NumiClothMaterialValues::gripComplianceMPerN, 1.0e-8, 1.0},
    {"cloth_ground_friction", &NumiClothMaterialValues::clothGroundFriction, 0.01, 2.0},
    {"cloth_self_friction", &NumiClothMaterialValues::clothSelfFriction, 0.01, 2.0},
    {"fruit_yarn_friction", &NumiClothMaterialValues::fruitYarnFriction, 0.01, 2.0},
    {"fruit_ground_friction", &NumiClothMaterialValues::fruitGroundFriction, 0.01, 2.0},
    {"fruit_pair_friction", &NumiClothMaterialValues::fruitPairFriction, 0.01, 2.0},
    {"fruit_rolling_resistance", &NumiClothMaterialValues::fruitRollingResistance, 0.0001, 0.20},
    {"yarn_crossflow_drag", &NumiClothMaterialValues::yarnCrossflowDrag, 0.10, 3.0},
    {"yarn_skin_friction", &NumiClothMaterialValues::yarnSkinFriction, 0.0001, 0.20},
    {"fruit_drag", &NumiClothMaterialValues::fruitDrag, 0.05, 2.0},
    {"fruit_rotational_drag", &NumiClothMaterialValues::fruitRotationalDrag, 0.0001, 0.20},
};

void setMaterialDefaults(NumiClothMaterialArtifact* out) {
    std::memset(out, 0, sizeof(*out));
    NumiClothMaterialValues& v = out->values;
    v.ordinaryNodeMassKg = 0.000050;
    v.hemNodeMassKg = 0.000100;
    v.yarnRadiusM = 0.004;
    v.axialBodyComplianceMPerN = 1.0e-8;
    v.axialCuffComplianceMPerN = 1.0e-9;
    v.knotCompliance = 2.0e-6;
    v.bendBodyComplianceMPerN = 8.0e-2;
    v.bendCuffComplianceMPerN = 1.0e-8;
    v.gripComplianceMPerN = 2.0e-4;
    v.clothGroundFriction = 0.45;
    v.clothSelfFriction = 0.34;
    v.fruitYarnFriction = 0.36;
    v.fruitGroundFriction = 0.42;
    v.fruitPairFriction = 0.30;
    v.fruitRollingResistance = 0.015;
    v.yarnCrossflowDrag = 1.10;
    v.yarnSkinFriction = 0.010;
    v.fruitDrag = 0.47;
    v.fruitRotationalDrag = 0.010;
    out->loaded = 0;
    copyText(out->parametersHash, sizeof(out->parametersHash), "authored-defaults");
    copyText(out->observationsHash, sizeof(out->observationsHash), "none");
}

void loadMaterial(const std::string& path, NumiClothMaterialArtifact* out) {
    std::ifstream stream(path);
    if (!stream) {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE, "cannot open cloth material artifact: " + path);
    }
    setMaterialDefaults(out);
    out->loaded = 1;
    std::set<std::string> seen;
    std::string schema;
    std::string line;
    while (std::getline(stream, line)) {
        line = trimmed(line);
        if (line.empty() || line[0] == '#') continue;
        const std::size_t equals = line.find('=');
        if (equals == std::string::npos) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE, "cloth material line must contain '='");
        }
        const std::string name = trimmed(line.substr(0, equals));
        const std::string text = trimmed(line.substr(equals + 1));
        if (name.empty() || text.empty() || !seen.insert(name).second) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                 "cloth material fields must be nonempty and unique");
        }
        if (name == "schema") {
            schema = text;
            continue;
        }
        if (name == "parameters_hash") {
            copyText(out->parametersHash, sizeof(out->parametersHash), text);
            continue;
        }
        if (name == "observations_hash") {
            copyText(out->observationsHash, sizeof(out->observationsHash), text);
            continue;
        }
        const MaterialField* field = nullptr;
        for (const MaterialField& candidate : kMaterialFields) {
            if (name == candidate.name) field = &candidate;
        }
        if (field == nullptr) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE, "unknown cloth material field: " + name);
        }
        double value = 0.0;
        if (!parseWholeNumber(text, &value)) {
            // A token std::stod rejects outright is "invalid"; a partially
            // consumed or non-finite one is "non-finite", as in the reference.
            std::size_t used = 0;
            bool parsed = true;
            try {
                (void)std::stod(text, &used);
            } catch (const std::exception&) {
                parsed = false;
            }
            fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                 std::string(parsed ? "non-finite cloth material value for "
                                    : "invalid cloth material value for ") + name);
        }
        if (!std::isfinite(value)) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE, "non-finite cloth material value for " + name);
        }
        out->values.*(field->member) = value;
    }
    if (schema != "numi.cloth.material.v1") {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE, "cloth material schema mismatch");
    }
    if (!seen.count("parameters_hash") || !seen.count("observations_hash")) {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE, "cloth material provenance hashes are required");
    }
    for (const MaterialField& field : kMaterialFields) {
        if (!seen.count(field.name)) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                 std::string("cloth material artifact is missing: ") + field.name);
        }
    }
    for (const MaterialField& field : kMaterialFields) {
        const double value = out->values.*(field.member);
        if (!(value >= field.minimum && value <= field.maximum)) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                 std::string("cloth material parameter outside contract bounds: ") +
                     field.name);
        }
    }
}

// ---- grip trajectory ------------------------------------------------------

struct Quat {
    double x, y, z, w;
};

double quatDot(const Quat& a, const Quat& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}

Quat quatNormalized(const Quat& q) {
    const double n = std::sqrt(quatDot(q, q));
    if (!(n > 1.0e-14)) {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE, "grip trajectory quaternion has zero length");
    }
    return {q.x / n, q.y / n, q.z / n, q.w / n};
}

// Shortest-arc slerp with a normalised-lerp fallback above cos 0.9995.
Quat quatSlerp(const Quat& a, Quat b, double t) {
    double c = quatDot(a, b);
    if (c < 0.0) {
        b = {-b.x, -b.y, -b.z, -b.w};
        c = -c;
    }
    c = std::min(1.0, std::max(-1.0, c));
    if (c > 0.9995) {
        return quatNormalized({a.x + t * (b.x - a.x), a.y + t * (b.y - a.y),
                               a.z + t * (b.z - a.z), a.w + t * (b.w - a.w)});
    }
    const double angle = std::acos(c);
    const double s = std::sin(angle);
    const double wa = std::sin((1.0 - t) * angle) / s;
    const double wb = std::sin(t * angle) / s;
    return quatNormalized({wa * a.x + wb * b.x, wa * a.y + wb * b.y,
                           wa * a.z + wb * b.z, wa * a.w + wb * b.w});
}

std::vector<std::string> splitComma(const std::string& line) {
    std::vector<std::string> fields;
    std::size_t begin = 0;
    while (true) {
        const std::size_t comma = line.find(',', begin);
        fields.push_back(trimmed(line.substr(
            begin, comma == std::string::npos ? std::string::npos : comma - begin)));
        if (comma == std::string::npos) break;
        begin = comma + 1;
    }
    return fields;
}

double trajectoryNumber(const std::string& text, const char* field) {
    std::size_t used = 0;
    double value = 0.0;
    try {
        value = std::stod(text, &used);
    } catch (const std::exception&) {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE,
             std::string("invalid grip trajectory number for ") + field);
    }
    if (used != text.size() || !std::isfinite(value)) {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE,
             std::string("non-finite grip trajectory number for ") + field);
    }
    return value;
}

}  // namespace

uint64_t fnv1a(const void* data, std::size_t bytes, uint64_t hash) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (std::size_t i = 0; i < bytes; ++i) {
        hash ^= p[i];
        hash *= 1099511628211ull;
    }
    return hash;
}

}  // namespace numi_hip

struct NumiClothGripTrajectory {
    std::vector<NumiClothGripPose> poses;
    std::string schema;
    std::string fingerprint;
    bool nearestPatch = false;
};

namespace numi_hip {
namespace {

void loadTrajectory(const std::string& path, NumiClothGripTrajectory* result) {
    std::ifstream stream(path, std::ios::binary);
    if (!stream) {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE, "cannot open grip trajectory: " + path);
    }
    std::ostringstream buffer;
    buffer << stream.rdbuf();
    const std::string source = buffer.str();
    char fingerprint[32];
    std::snprintf(fingerprint, sizeof(fingerprint), "0x%llx",
                  static_cast<unsigned long long>(
                      fnv1a(source.data(), source.size(), 1469598103934665603ull)));
    result->fingerprint = fingerprint;

    const std::vector<std::string> header{
        "time_s", "translation_x_m", "translation_y_m", "translation_z_m",
        "quaternion_x", "quaternion_y", "quaternion_z", "quaternion_w", "active"};
    std::istringstream lines(source);
    std::string line;
    std::size_t lineNumber = 0;
    bool haveSchema = false;
    bool haveHeader = false;
    bool reactivation = false;
    bool previousActive = true;
    uint32_t generation = 1;
    while (std::getline(lines, line)) {
        ++lineNumber;
        line = trimmed(line);
        if (line.empty() || line[0] == '#') continue;
        if (!haveSchema) {
            if (line != "schema=numi.grip.trajectory.v1" &&
                line != "schema=numi.grip.trajectory.v2" &&
                line != "schema=numi.grip.trajectory.v3") {
                fail(NUMI_CLOTH_HIP_ERROR_PARSE, "grip trajectory schema mismatch");
            }
            result->schema = line.substr(7);
            reactivation = result->schema != "numi.grip.trajectory.v1";
            result->nearestPatch = result->schema == "numi.grip.trajectory.v3";
            haveSchema = true;
            continue;
        }
        const std::vector<std::string> fields = splitComma(line);
        if (!haveHeader) {
            if (fields != header) {
                fail(NUMI_CLOTH_HIP_ERROR_PARSE, "grip trajectory CSV header mismatch");
            }
            haveHeader = true;
            continue;
        }
        if (fields.size() != header.size()) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                 "grip trajectory row has wrong field count at " +
                     std::to_string(lineNumber));
        }
        NumiClothGripPose pose{};
        pose.timeSeconds = trajectoryNumber(fields[0], "time_s");
        pose.translation[0] = trajectoryNumber(fields[1], "translation_x_m");
        pose.translation[1] = trajectoryNumber(fields[2], "translation_y_m");
        pose.translation[2] = trajectoryNumber(fields[3], "translation_z_m");
        Quat q{trajectoryNumber(fields[4], "quaternion_x"),
               trajectoryNumber(fields[5], "quaternion_y"),
               trajectoryNumber(fields[6], "quaternion_z"),
               trajectoryNumber(fields[7], "quaternion_w")};
        if (fields[8] == "1") {
            pose.active = 1;
        } else if (fields[8] == "0") {
            pose.active = 0;
        } else {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE, "grip trajectory active must be 0 or 1");
        }
        if (pose.timeSeconds < 0.0 ||
            (!result->poses.empty() &&
             !(pose.timeSeconds > result->poses.back().timeSeconds))) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                 "grip trajectory times must be nonnegative and strictly increasing");
        }
        if (std::abs(std::sqrt(quatDot(q, q)) - 1.0) > 1.0e-3) {
            fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                 "grip trajectory quaternion must be unit length within 1e-3");
        }
        q = quatNormalized(q);
        if (!result->poses.empty()) {
            const double* o = result->poses.back().orientation;
            const Quat previous{o[0], o[1], o[2], o[3]};
            if (quatDot(previous, q) < 0.0) q = {-q.x, -q.y, -q.z, -q.w};
        }
        if (!previousActive && pose.active) {
            if (!reactivation) {
                fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                     "grip trajectory v1 cannot reactivate after release");
            }
            if (generation == std::numeric_limits<uint32_t>::max()) {
                fail(NUMI_CLOTH_HIP_ERROR_PARSE,
                     "grip trajectory has too many attachment generations");
            }
            ++generation;
        }
        pose.orientation[0] = q.x;
        pose.orientation[1] = q.y;
        pose.orientation[2] = q.z;
        pose.orientation[3] = q.w;
        pose.attachmentGeneration = generation;
        previousActive = pose.active != 0;
        result->poses.push_back(pose);
    }
    if (!haveSchema || !haveHeader || result->poses.size() < 2) {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE,
             "grip trajectory requires schema, header, and at least two poses");
    }
    const NumiClothGripPose& first = result->poses.front();
    const double offset = std::sqrt(first.translation[0] * first.translation[0] +
                                    first.translation[1] * first.translation[1] +
                                    first.translation[2] * first.translation[2]);
    const double rotation =
        2.0 * std::acos(std::min(1.0, std::max(0.0, std::abs(first.orientation[3]))));
    if (std::abs(first.timeSeconds) > 1.0e-12 || offset > 1.0e-9 || rotation > 1.0e-9 ||
        !first.active) {
        fail(NUMI_CLOTH_HIP_ERROR_PARSE,
             "grip trajectory must start at t=0 with zero translation, identity "
             "orientation, and an active seam grip");
    }
}

NumiClothGripPose samplePose(const NumiClothGripTrajectory& trajectory, double t) {
    const std::vector<NumiClothGripPose>& poses = trajectory.poses;
    if (!std::isfinite(t) || t < 0.0 || poses.empty() ||
        t > poses.back().timeSeconds + 1.0e-9) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT,
             "grip trajectory sample is outside its duration");
    }
    if (t <= poses.front().timeSeconds) return poses.front();
    if (t >= poses.back().timeSeconds) return poses.back();
    // First pose with time >= t; t lies in (previous, that].
    std::size_t upper = 0;
    while (upper < poses.size() && poses[upper].timeSeconds < t) ++upper;
    if (upper == poses.size()) return poses.back();
    const NumiClothGripPose& b = poses[upper];
    const NumiClothGripPose& a = poses[upper - 1];
    const double f = (t - a.timeSeconds) / (b.timeSeconds - a.timeSeconds);
    NumiClothGripPose out{};
    out.timeSeconds = t;
    for (int k = 0; k < 3; ++k) {
        out.translation[k] = a.translation[k] + f * (b.translation[k] - a.translation[k]);
    }
    const Quat q = quatSlerp({a.orientation[0], a.orientation[1], a.orientation[2], a.orientation[3]},
                             {b.orientation[0], b.orientation[1], b.orientation[2], b.orientation[3]}, f);
    out.orientation[0] = q.x;
    out.orientation[1] = q.y;
    out.orientation[2] = q.z;
    out.orientation[3] = q.w;
    out.active = f >= 1.0 ? b.active : a.active;
    out.attachmentGeneration = f >= 1.0 ? b.attachmentGeneration : a.attachmentGeneration;
    return out;
}

}  // namespace
}  // namespace numi_hip

using namespace numi_hip;

extern "C" {

const char* numi_cloth_hip_last_error(void) { return gLastError.c_str(); }

uint64_t numi_cloth_fnv1a64(const void* data, size_t bytes, uint64_t seed) {
    return fnv1a(data, bytes, seed == 0 ? 1469598103934665603ull : seed);
}

int numi_cloth_material_defaults(NumiClothMaterialArtifact* out) {
    return guarded([&] {
        if (out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null output");
        setMaterialDefaults(out);
    });
}

int numi_cloth_material_load(const char* path, NumiClothMaterialArtifact* out) {
    return guarded([&] {
        if (path == nullptr || out == nullptr) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        }
        loadMaterial(path, out);
    });
}

int numi_cloth_grip_trajectory_load(const char* path, NumiClothGripTrajectory** out) {
    return guarded([&] {
        if (path == nullptr || out == nullptr) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        }
        *out = nullptr;
        auto* trajectory = new NumiClothGripTrajectory();
        try {
            loadTrajectory(path, trajectory);
        } catch (...) {
            delete trajectory;
            throw;
        }
        *out = trajectory;
    });
}

int numi_cloth_grip_trajectory_info(const NumiClothGripTrajectory* trajectory,
                                    NumiClothGripTrajectoryInfo* out) {
    return guarded([&] {
        if (trajectory == nullptr || out == nullptr) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        }
        std::memset(out, 0, sizeof(*out));
        out->poseCount = static_cast<uint32_t>(trajectory->poses.size());
        out->selectNearestCuffPatch = trajectory->nearestPatch ? 1u : 0u;
        double rotation = 0.0;
        uint32_t generations = 0;
        for (const NumiClothGripPose& pose : trajectory->poses) {
            rotation = std::max(
                rotation,
                2.0 * std::acos(std::min(1.0, std::max(0.0, std::abs(pose.orientation[3])))));
            generations = std::max(generations, pose.attachmentGeneration);
        }
        out->attachmentGenerations = generations;
        out->maximumRotation = rotation;
        out->durationSeconds = trajectory->poses.back().timeSeconds;
        copyText(out->schema, sizeof(out->schema), trajectory->schema);
        copyText(out->contentFingerprint, sizeof(out->contentFingerprint),
                 trajectory->fingerprint);
    });
}

int numi_cloth_grip_trajectory_pose(const NumiClothGripTrajectory* trajectory,
                                    uint32_t index, NumiClothGripPose* out) {
    return guarded([&] {
        if (trajectory == nullptr || out == nullptr || index >= trajectory->poses.size()) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "pose index out of range");
        }
        *out = trajectory->poses[index];
    });
}

int numi_cloth_grip_trajectory_sample(const NumiClothGripTrajectory* trajectory,
                                      double timeSeconds, NumiClothGripPose* out) {
    return guarded([&] {
        if (trajectory == nullptr || out == nullptr) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        }
        *out = samplePose(*trajectory, timeSeconds);
    });
}

void numi_cloth_grip_trajectory_destroy(NumiClothGripTrajectory* trajectory) {
    delete trajectory;
}

}  // extern "C"
