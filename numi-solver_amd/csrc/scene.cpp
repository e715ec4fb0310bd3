// Scene builders: topology, materials and greedy colour batches for the
// woven sheet (a regular warp/weft grid) and for the reference's produce bag.
//
// Colouring rule (reference tools/cloth_bag_metal.mm:627-676): constraints are
// visited in construction order; each takes the lowest colour absent from the
// 64-bit colour masks of its participating knots; the list is then stably
// sorted by colour and every colour becomes one batch {first, count, colour, 0}.
// The rule is kept identical because the colourings are compared bit for bit.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"
#include "numi_cloth_scene.h"
#include "scene_internal.hpp"


namespace numi_hip {
namespace {

struct D3 {
    double x, y, z;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double len3(D3 a) { return std::sqrt(dot3(a, a)); }
D3 asD3(const mr_float4& v) { return {v.x, v.y, v.z}; }

mr_float4 f4(float x, float y, float z, float w = 0.0f) { return {x, y, z, w}; }
mr_uint4 u4(uint32_t x, uint32_t y, uint32_t z = 0, uint32_t w = 0) { return {x, y, z, w}; }

double unitDouble(uint64_t* state) {
    return static_cast<double>(numi_cloth_splitmix64(state) >> 11) * 0x1.0p-53;
}

NumiClothBagGPUParticle makeParticle(D3 p, float mass, bool pinned) {
    NumiClothBagGPUParticle particle{};
    const float x = static_cast<float>(p.x);
    const float y = static_cast<float>(p.y);
    const float z = static_cast<float>(p.z);
    particle.positionAndInverseMass = f4(x, y, z, pinned ? 0.0f : 1.0f / mass);
    particle.previousAndMass = f4(x, y, z, mass);
    particle.velocity = f4(0.0f, 0.0f, 0.0f, 0.0f);
    return particle;
}

// Greedy first-free colouring in visiting order; returns per-item colours.
template <std::size_t N>
std::vector<uint32_t> greedyColours(const std::vector<std::array<uint32_t, N>>& items,
                                    std::size_t nodeCount, uint32_t* colourCount,
                                    const char* what) {
    std::vector<uint64_t> masks(nodeCount, 0u);
    std::vector<uint32_t> colours(items.size());
    uint32_t used = 0;
    for (std::size_t i = 0; i < items.size(); ++i) {
        uint64_t taken = 0;
        for (uint32_t node : items[i]) taken |= masks[node];
        if (~taken == 0u) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string(what) + " coloring exceeds 64 colors");
        const uint32_t colour = static_cast<uint32_t>(__builtin_ctzll(~taken));
        colours[i] = colour;
        used = std::max(used, colour + 1u);
        for (uint32_t node : items[i]) masks[node] |= uint64_t{1} << colour;
    }
    *colourCount = used;
    return colours;
}

// Stable order by colour plus one batch per colour.
std::vector<uint32_t> colourOrder(const std::vector<uint32_t>& colours, uint32_t colourCount,
                                  std::vector<NumiClothBagGPUBatch>* batches) {
    std::vector<uint32_t> order(colours.size());
    for (std::size_t i = 0; i < order.size(); ++i) order[i] = static_cast<uint32_t>(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return colours[a] < colours[b]; });
    batches->clear();
    uint32_t cursor = 0;
    for (uint32_t colour = 0; colour < colourCount; ++colour) {
        const uint32_t first = cursor;
        while (cursor < order.size() && colours[order[cursor]] == colour) ++cursor;
        batches->push_back({u4(first, cursor - first, colour, 0u)});
    }
    return order;
}

struct EdgeItem {
    uint32_t a, b, kind;
    float compliance;
};
struct BendItem {
    uint32_t a, m, c;
    float compliance;
};
struct KnotItem {
    uint32_t p[4];
};

void emitDistances(NumiClothHipScene& s, const std::vector<EdgeItem>& edges, float strainLimit) {
    std::vector<std::array<uint32_t, 2>> items;
    items.reserve(edges.size());
    for (const EdgeItem& e : edges) items.push_back({e.a, e.b});
    uint32_t colourCount = 0;
    const std::vector<uint32_t> colours =
        greedyColours(items, s.particles.size(), &colourCount, "cloth edge");
    const std::vector<uint32_t> order = colourOrder(colours, colourCount, &s.distanceBatches);
    s.distances.clear();
    s.distances.reserve(edges.size());
    for (uint32_t index : order) {
        const EdgeItem& e = edges[index];
        const double rest = len3(asD3(s.particles[e.b].positionAndInverseMass) -
                                 asD3(s.particles[e.a].positionAndInverseMass));
        NumiClothBagGPUDistance d{};
        d.particlesAndColor = u4(e.a, e.b, colours[index], e.kind);
        d.material = f4(static_cast<float>(rest), e.compliance, 0.0f, strainLimit);
        s.distances.push_back(d);
    }
}

void emitBends(NumiClothHipScene& s, const std::vector<BendItem>& bends) {
    std::vector<std::array<uint32_t, 2>> items;
    items.reserve(bends.size());
    for (const BendItem& b : bends) items.push_back({b.a, b.c});
    uint32_t colourCount = 0;
    const std::vector<uint32_t> colours =
        greedyColours(items, s.particles.size(), &colourCount, "cloth bend");
    const std::vector<uint32_t> order = colourOrder(colours, colourCount, &s.bendBatches);
    s.bends.clear();
    s.bends.reserve(bends.size());
    for (uint32_t index : order) {
        const BendItem& b = bends[index];
        const D3 pa = asD3(s.particles[b.a].positionAndInverseMass);
        const D3 pm = asD3(s.particles[b.m].positionAndInverseMass);
        const D3 pc = asD3(s.particles[b.c].positionAndInverseMass);
        NumiClothBagGPUBend bend{};
        bend.particlesAndColor = u4(b.a, b.m, b.c, colours[index]);
        bend.material = f4(static_cast<float>(len3(pc - pa)),
                           static_cast<float>(len3(pm - pa) + len3(pc - pm)), b.compliance, 0.0f);
        s.bends.push_back(bend);
    }
}

void emitKnots(NumiClothHipScene& s, const std::vector<KnotItem>& knots, float compliance) {
    std::vector<std::array<uint32_t, 4>> items;
    items.reserve(knots.size());
    for (const KnotItem& k : knots) items.push_back({k.p[0], k.p[1], k.p[2], k.p[3]});
    uint32_t colourCount = 0;
    const std::vector<uint32_t> colours =
        greedyColours(items, s.particles.size(), &colourCount, "cloth knot");
    const std::vector<uint32_t> order = colourOrder(colours, colourCount, &s.knotBatches);
    s.knots.clear();
    s.knots.reserve(knots.size());
    for (uint32_t index : order) {
        const KnotItem& k = knots[index];
        const D3 warp = asD3(s.particles[k.p[1]].positionAndInverseMass) -
                        asD3(s.particles[k.p[0]].positionAndInverseMass);
        const D3 weft = asD3(s.particles[k.p[3]].positionAndInverseMass) -
                        asD3(s.particles[k.p[2]].positionAndInverseMass);
        NumiClothBagGPUKnot knot{};
        knot.particles = u4(k.p[0], k.p[1], k.p[2], k.p[3]);
        knot.control = u4(colours[index], 0u, 0u, 0u);
        knot.material = f4(static_cast<float>(dot3(warp * (1.0 / len3(warp)), weft * (1.0 / len3(weft)))),
                           compliance, 0.0f, 0.0f);
        s.knots.push_back(knot);
    }
}

// ---------------------------------------------------------------- sheet ----

void buildSheet(const NumiClothHipSheetSpec& spec, NumiClothHipScene& s) {
    if (spec.width < 2 || spec.height < 2 || !(spec.spacing > 0.0f) ||
        !(spec.arealDensity > 0.0f) || !(spec.timestep > 0.0f)) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "sheet spec needs width,height >= 2 and positive spacing, density, timestep");
    }
    const uint32_t W = spec.width;
    const uint32_t H = spec.height;
    const uint64_t count = static_cast<uint64_t>(W) * H;
    if (count > (1ull << 28)) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "sheet too large");
    uint64_t rng = spec.seed;
    const double h = spec.spacing;
    const float mass = static_cast<float>(static_cast<double>(spec.arealDensity) * h * h);
    s.particles.reserve(count);
    for (uint32_t y = 0; y < H; ++y) {
        for (uint32_t x = 0; x < W; ++x) {
            const double jitter = spec.jitter * (2.0 * unitDouble(&rng) - 1.0);
            const D3 p{(x - 0.5 * (W - 1)) * h, (y - 0.5 * (H - 1)) * h, spec.elevation + jitter};
            s.particles.push_back(makeParticle(p, mass, spec.pinTopRow && y + 1 == H));
        }
    }
    auto at = [W](uint32_t x, uint32_t y) { return y * W + x; };
    std::vector<EdgeItem> edges;
    edges.reserve(2 * count);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x + 1 < W; ++x) edges.push_back({at(x, y), at(x + 1, y), 1u, spec.stretchCompliance});
    for (uint32_t y = 0; y + 1 < H; ++y)
        for (uint32_t x = 0; x < W; ++x) edges.push_back({at(x, y), at(x, y + 1), 0u, spec.stretchCompliance});
    emitDistances(s, edges, spec.strainLimit);

    std::vector<BendItem> bends;
    bends.reserve(2 * count);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 1; x + 1 < W; ++x) bends.push_back({at(x - 1, y), at(x, y), at(x + 1, y), spec.bendCompliance});
    for (uint32_t y = 1; y + 1 < H; ++y)
        for (uint32_t x = 0; x < W; ++x) bends.push_back({at(x, y - 1), at(x, y), at(x, y + 1), spec.bendCompliance});
    emitBends(s, bends);

    if (spec.crossingKnots) {
        std::vector<KnotItem> knots;
        for (uint32_t y = 1; y + 1 < H; ++y)
            for (uint32_t x = 1; x + 1 < W; ++x)
                knots.push_back({{at(x, y - 1), at(x, y + 1), at(x - 1, y), at(x + 1, y)}});
        emitKnots(s, knots, spec.knotCompliance);
    }

    NumiClothBagGPUConfig& c = s.config;
    c.control = u4(NUMI_CLOTH_BAG_GPU_ABI_VERSION, static_cast<uint32_t>(s.particles.size()),
                   static_cast<uint32_t>(s.distances.size()), 0u);
    c.constraintCounts = u4(static_cast<uint32_t>(s.knots.size()), static_cast<uint32_t>(s.bends.size()),
                            spec.ground ? 1u : 0u, 0u);
    c.contactCounts = u4(0u, 0u, 0u, 0u);
    c.gravityAndTimestep = f4(0.0f, 0.0f, spec.gravityZ, spec.timestep);
    c.gripTargetAndActive = f4(0.0f, 0.0f, 0.0f, 0.0f);
    c.gripOrientation = f4(0.0f, 0.0f, 0.0f, 1.0f);
    c.gripControl = u4(0u, 0u, 0u, 0u);
    c.gripMaterial = f4(0.0f, 0.0f, 0.0f, 0.0f);
    c.clothMaterial = f4(spec.yarnRadius, spec.cellSize, spec.groundFriction, spec.selfFriction);
    c.fruitMaterial = f4(0.0f, 0.0f, 0.0f, 0.0f);
    c.airVelocityAndDensity = f4(0.0f, 0.0f, 0.0f, 0.0f);
    c.aerodynamicCoefficients = f4(0.0f, 0.0f, 0.0f, 0.0f);
    c.mouthControl = u4(0u, 0u, 0u, 0u);
    c.mouthMaterial = f4(0.0f, 0.0f, 0.0f, 0.0f);

    if (spec.sphereMode == 1) {
        s.spheres.push_back(f4(spec.fixedSphere[0], spec.fixedSphere[1], spec.fixedSphere[2], spec.fixedSphere[3]));
    } else if (spec.sphereMode == 2) {
        const double halfX = 0.5 * (W - 1) * h;
        const double halfY = 0.5 * (H - 1) * h;
        for (uint32_t k = 0; k < spec.sphereCount; ++k) {
            const double cx = -halfX + 2.0 * halfX * unitDouble(&rng);
            const double cy = -halfY + 2.0 * halfY * unitDouble(&rng);
            const double r = spec.sphereRadiusMin +
                             (spec.sphereRadiusMax - spec.sphereRadiusMin) * unitDouble(&rng);
            s.spheres.push_back(f4(static_cast<float>(cx), static_cast<float>(cy),
                                   static_cast<float>(r), static_cast<float>(r)));
        }
    }
    s.publish();
    s.desc.extension.selfContact = spec.selfContact;
    // Broad-phase margin: a quarter yarn radius of travel within a substep.
    s.desc.extension.selfContactMargin = 0.25f * spec.yarnRadius;
}

// ---------------------------------------------------------------- strip ----

constexpr double kPiScene = 3.14159265358979323846;

void buildStrip(const NumiClothHipStripSpec& spec, NumiClothHipScene& s) {
    if (spec.width < 2 || spec.height < 2 || !(spec.spacing > 0.0f) || !(spec.arealDensity > 0.0f) ||
        !(spec.timestep > 0.0f) || !(spec.damping >= 0.0f)) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT,
             "strip spec needs width,height >= 2, positive spacing, density, timestep and damping >= 0");
    }
    const uint32_t W = spec.width;
    const uint32_t H = spec.height;
    const uint64_t count = static_cast<uint64_t>(W) * H;
    if (count > (1ull << 26)) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "strip too large");
    uint64_t rng = spec.seed;
    const double h = spec.spacing;
    const float mass = static_cast<float>(static_cast<double>(spec.arealDensity) * h * h);
    s.particles.reserve(count);
    for (uint32_t j = 0; j < H; ++j) {
        for (uint32_t i = 0; i < W; ++i) {
            const double jitter = spec.jitter * (2.0 * unitDouble(&rng) - 1.0);
            const D3 p{spec.origin[0] + i * h, spec.origin[1] + jitter, spec.origin[2] - j * h};
            s.particles.push_back(makeParticle(p, mass, j == 0));
        }
    }
    auto at = [W](uint32_t i, uint32_t j) { return j * W + i; };
    std::vector<EdgeItem> edges;
    edges.reserve(2 * count);
    for (uint32_t j = 0; j < H; ++j)
        for (uint32_t i = 0; i + 1 < W; ++i) edges.push_back({at(i, j), at(i + 1, j), 1u, spec.stretchCompliance});
    for (uint32_t j = 0; j + 1 < H; ++j)
        for (uint32_t i = 0; i < W; ++i) edges.push_back({at(i, j), at(i, j + 1), 0u, spec.stretchCompliance});
    emitDistances(s, edges, spec.strainLimit);
    std::vector<BendItem> bends;
    bends.reserve(2 * count);
    for (uint32_t j = 0; j < H; ++j)
        for (uint32_t i = 1; i + 1 < W; ++i) bends.push_back({at(i - 1, j), at(i, j), at(i + 1, j), spec.bendCompliance});
    for (uint32_t j = 1; j + 1 < H; ++j)
        for (uint32_t i = 0; i < W; ++i) bends.push_back({at(i, j - 1), at(i, j), at(i, j + 1), spec.bendCompliance});
    emitBends(s, bends);

    NumiClothBagGPUConfig& c = s.config;
    c.control = u4(NUMI_CLOTH_BAG_GPU_ABI_VERSION, static_cast<uint32_t>(s.particles.size()),
                   static_cast<uint32_t>(s.distances.size()), 0u);
    c.constraintCounts = u4(0u, static_cast<uint32_t>(s.bends.size()), 0u, 0u);
    c.gravityAndTimestep = f4(0.0f, 0.0f, spec.gravityZ, spec.timestep);
    c.gripOrientation = f4(0.0f, 0.0f, 0.0f, 1.0f);
    c.clothMaterial = f4(spec.yarnRadius, 0.0f, 0.0f, 0.0f);
    s.publish();
    s.desc.extension.driveAmplitude = spec.driveAmplitude;
    s.desc.extension.driveOmega = static_cast<float>(2.0 * kPiScene * spec.driveFrequency);
    s.desc.extension.drivePhase = spec.drivePhase;
    s.desc.extension.damping = spec.damping;
}

// ---------------------------------------------------------------- merge ----

template <typename Record, typename ColourOf, typename Shift>
void mergeFamily(const std::vector<const NumiClothHipSceneDesc*>& specimens, const std::vector<uint32_t>& offsets,
                 const Record* (*records)(const NumiClothHipSceneDesc&), uint32_t (*countOf)(const NumiClothHipSceneDesc&),
                 const NumiClothBagGPUBatch* (*batchesOf)(const NumiClothHipSceneDesc&),
                 uint32_t (*batchCountOf)(const NumiClothHipSceneDesc&), ColourOf colourOf, Shift shift,
                 std::vector<Record>& merged, std::vector<NumiClothBagGPUBatch>& mergedBatches,
                 std::vector<uint32_t>& start, std::vector<uint32_t>& index, const char* what) {
    const uint32_t n = static_cast<uint32_t>(specimens.size());
    start.assign(n + 1, 0);
    for (uint32_t s = 0; s < n; ++s) start[s + 1] = start[s] + countOf(*specimens[s]);
    index.assign(start[n], ~0u);
    uint32_t colours = 0;
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t batches = batchCountOf(*specimens[s]);
        const NumiClothBagGPUBatch* b = batchesOf(*specimens[s]);
        if (batches != 0 && b == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, std::string("null batches: ") + what);
        for (uint32_t k = 0; k < batches; ++k) colours = std::max(colours, b[k].control.z + 1);
    }
    merged.clear();
    mergedBatches.clear();
    for (uint32_t colour = 0; colour < colours; ++colour) {
        const uint32_t first = static_cast<uint32_t>(merged.size());
        for (uint32_t s = 0; s < n; ++s) {
            const NumiClothHipSceneDesc& d = *specimens[s];
            const Record* r = records(d);
            const NumiClothBagGPUBatch* b = batchesOf(d);
            for (uint32_t k = 0; k < batchCountOf(d); ++k) {
                if (b[k].control.z != colour) continue;
                if (static_cast<uint64_t>(b[k].control.x) + b[k].control.y > countOf(d)) {
                    fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string("batch out of range (FAILURE_RANGE): ") + what);
                }
                for (uint32_t j = b[k].control.x; j < b[k].control.x + b[k].control.y; ++j) {
                    if (colourOf(r[j]) != colour) {
                        fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY,
                             std::string("record colour differs from its batch (FAILURE_BATCH): ") + what);
                    }
                    if (index[start[s] + j] != ~0u) {
                        fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string("record in two batches: ") + what);
                    }
                    index[start[s] + j] = static_cast<uint32_t>(merged.size());
                    Record copy = r[j];
                    shift(copy, offsets[s]);
                    merged.push_back(copy);
                }
            }
        }
        const uint32_t count = static_cast<uint32_t>(merged.size()) - first;
        if (count != 0) mergedBatches.push_back({u4(first, count, colour, 0u)});
    }
    for (uint32_t k = 0; k < index.size(); ++k) {
        if (index[k] == ~0u) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string("record in no batch: ") + what);
    }
}

void mergeScenes(const std::vector<const NumiClothHipSceneDesc*>& specimens, NumiClothHipScene& m) {
    if (specimens.empty()) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "no specimens");
    const NumiClothBagGPUConfig* c0 = specimens[0]->config;
    for (const NumiClothHipSceneDesc* d : specimens) {
        if (d == nullptr || d->config == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null specimen");
        const NumiClothBagGPUConfig& c = *d->config;
        if (d->gripCount || d->fruitCount || d->fruitPairCount || d->yarnContactCount || d->mouthRimCount ||
            d->cuffCount || d->extension.staticSphereCount || d->extension.selfContact ||
            d->extension.pointSelfCollision || d->extension.sweptGround || d->extension.certificatePasses) {
            fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED,
                 "merged specimens may not have grips, fruits, mouth, cuff, static spheres, self-contact, point "
                 "self-collision, the swept ground pass or certificate passes");
        }
        if (c.airVelocityAndDensity.w > 0.0f) {
            fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "merged specimens may not have air loads (their reduction is global)");
        }
        if (std::memcmp(&c.gravityAndTimestep, &c0->gravityAndTimestep, sizeof(mr_float4)) != 0 ||
            std::memcmp(&c.clothMaterial, &c0->clothMaterial, sizeof(mr_float4)) != 0 ||
            c.constraintCounts.z != c0->constraintCounts.z) {
            fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED,
                 "merged specimens must share timestep, gravity, cloth material and the ground flag");
        }
        if (d->particleCount != 0 && d->particles == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null particles");
    }
    const uint32_t n = static_cast<uint32_t>(specimens.size());
    std::vector<uint32_t> offsets(n + 1, 0);
    for (uint32_t s = 0; s < n; ++s) {
        const uint64_t next = static_cast<uint64_t>(offsets[s]) + specimens[s]->particleCount;
        if (next >= (1ull << 31)) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "merged scene too large");
        offsets[s + 1] = static_cast<uint32_t>(next);
    }
    m.particles.clear();
    m.motion.clear();
    m.mergeStart[0] = offsets;
    m.mergeIndex[0].resize(offsets[n]);
    for (uint32_t s = 0; s < n; ++s) {
        const NumiClothHipSceneDesc& d = *specimens[s];
        const NumiClothHipSceneExtension& e = d.extension;
        for (uint32_t k = 0; k < d.particleCount; ++k) {
            m.mergeIndex[0][offsets[s] + k] = offsets[s] + k;
            m.particles.push_back(d.particles[k]);
            m.motion.push_back(e.particleMotion != nullptr ? e.particleMotion[k]
                                                           : f4(e.driveAmplitude, e.driveOmega, e.drivePhase, e.damping));
        }
    }
    auto checkIndex = [&](uint32_t index, uint32_t s) {
        if (index >= specimens[s]->particleCount) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "particle index out of range");
    };
    mergeFamily<NumiClothBagGPUDistance>(
        specimens, offsets, [](const NumiClothHipSceneDesc& d) { return d.distances; },
        [](const NumiClothHipSceneDesc& d) { return d.distanceCount; },
        [](const NumiClothHipSceneDesc& d) { return d.distanceBatches; },
        [](const NumiClothHipSceneDesc& d) { return d.distanceBatchCount; },
        [](const NumiClothBagGPUDistance& r) { return r.particlesAndColor.z; },
        [](NumiClothBagGPUDistance& r, uint32_t o) {
            r.particlesAndColor.x += o;
            r.particlesAndColor.y += o;
        },
        m.distances, m.distanceBatches, m.mergeStart[1], m.mergeIndex[1], "distance");
    mergeFamily<NumiClothBagGPUBend>(
        specimens, offsets, [](const NumiClothHipSceneDesc& d) { return d.bends; },
        [](const NumiClothHipSceneDesc& d) { return d.bendCount; },
        [](const NumiClothHipSceneDesc& d) { return d.bendBatches; },
        [](const NumiClothHipSceneDesc& d) { return d.bendBatchCount; },
        [](const NumiClothBagGPUBend& r) { return r.particlesAndColor.w; },
        [](NumiClothBagGPUBend& r, uint32_t o) {
            r.particlesAndColor.x += o;
            r.particlesAndColor.y += o;
            r.particlesAndColor.z += o;
        },
        m.bends, m.bendBatches, m.mergeStart[2], m.mergeIndex[2], "bend");
    mergeFamily<NumiClothBagGPUKnot>(
        specimens, offsets, [](const NumiClothHipSceneDesc& d) { return d.knots; },
        [](const NumiClothHipSceneDesc& d) { return d.knotCount; },
        [](const NumiClothHipSceneDesc& d) { return d.knotBatches; },
        [](const NumiClothHipSceneDesc& d) { return d.knotBatchCount; },
        [](const NumiClothBagGPUKnot& r) { return r.control.x; },
        [](NumiClothBagGPUKnot& r, uint32_t o) {
            r.particles.x += o;
            r.particles.y += o;
            r.particles.z += o;
            r.particles.w += o;
        },
        m.knots, m.knotBatches, m.mergeStart[3], m.mergeIndex[3], "knot");
    // Participant indices must stay inside their own specimen.
    for (uint32_t s = 0; s < n; ++s) {
        const NumiClothHipSceneDesc& d = *specimens[s];
        for (uint32_t k = 0; k < d.distanceCount; ++k) {
            checkIndex(d.distances[k].particlesAndColor.x, s);
            checkIndex(d.distances[k].particlesAndColor.y, s);
        }
        for (uint32_t k = 0; k < d.bendCount; ++k) {
            checkIndex(d.bends[k].particlesAndColor.x, s);
            checkIndex(d.bends[k].particlesAndColor.y, s);
            checkIndex(d.bends[k].particlesAndColor.z, s);
        }
        for (uint32_t k = 0; k < d.knotCount; ++k) {
            checkIndex(d.knots[k].particles.x, s);
            checkIndex(d.knots[k].particles.y, s);
            checkIndex(d.knots[k].particles.z, s);
            checkIndex(d.knots[k].particles.w, s);
        }
    }
    m.config = *c0;
    m.config.control = u4(NUMI_CLOTH_BAG_GPU_ABI_VERSION, offsets[n], static_cast<uint32_t>(m.distances.size()), 0u);
    m.config.constraintCounts = u4(static_cast<uint32_t>(m.knots.size()), static_cast<uint32_t>(m.bends.size()),
                                   c0->constraintCounts.z, 0u);
    m.config.contactCounts = u4(0u, 0u, 0u, 0u);
    m.publish();
}

// ------------------------------------------------------------------ bag ----
// Geometry of the reference bag (reference tools/cloth_bag_metal.mm:33-38,
// :327-440): 48 knots around, 28 wall levels, bottom 13 x 13 grid whose
// boundary maps onto ring 0 of level 0.

constexpr uint32_t kAround = 48;
constexpr uint32_t kLevels = 28;
constexpr uint32_t kGrid = 13;
constexpr uint32_t kInterior = kGrid - 2;
constexpr double kPi = 3.14159265358979323846;

double smooth01(double v) {
    const double x = std::min(1.0, std::max(0.0, v));
    return x * x * (3.0 - 2.0 * x);
}

uint32_t wallIndex(uint32_t level, uint32_t ring) { return level * kAround + ring % kAround; }

// Authored wall knot: radius and height profiles with wrinkle harmonics.
D3 wallPosition(uint32_t level, uint32_t ring) {
    const double v = static_cast<double>(level) / (kLevels - 1);
    const double body = smooth01(v / 0.55);
    const double fold = std::min(1.0, std::max(0.0, (v - 0.72) / 0.28));
    const double skirt = 1.0 - smooth01(v / 0.18);
    const double a = 2.0 * kPi * ring / kAround + 0.08 * v;
    const double baseRadius = 0.266 + 0.030 * body - 0.116 * fold + 0.045 * skirt;
    const double wrinkle = 1.0 + 0.035 * std::sin(5.0 * a + 5.0 * kPi * v) +
                           0.018 * std::sin(9.0 * a - 3.0 * kPi * v);
    const double rimLoose = fold * (0.007 * std::sin(3.0 * a + 0.35) + 0.004 * std::sin(7.0 * a - 0.80));
    const double skirtLoose = skirt * (0.020 * std::sin(3.0 * a - 0.40) + 0.012 * std::sin(7.0 * a + 0.70));
    const double rimSag = fold * (0.016 * std::sin(a + 0.55) + 0.008 * std::sin(3.0 * a - 0.30) +
                                  0.004 * std::sin(6.0 * a + 0.90));
    const double radius = baseRadius * wrinkle + rimLoose + skirtLoose;
    const double bodySag = body * (0.018 * std::sin(2.0 * a + 0.30) + 0.009 * std::sin(5.0 * a - 0.70));
    const double skirtSag = skirt * (0.012 * std::sin(4.0 * a + 0.20) + 0.006 * std::sin(9.0 * a - 0.50));
    const double bodyHeight = 0.012 + 0.155 * std::min(v / 0.72, 1.0) + bodySag;
    const double foldHeight = fold < 0.35 ? 0.167 + 0.038 * (fold / 0.35) : 0.205 - 0.025 * ((fold - 0.35) / 0.65);
    return {radius * std::cos(a), radius * std::sin(a),
            (v < 0.72 ? bodyHeight : foldHeight) + rimSag + skirtSag + 0.009};
}

// Square-to-disk ("concentric") map of a bottom grid node.
void diskCoordinate(uint32_t row, uint32_t column, double* x, double* y) {
    const double half = 0.5 * (kGrid - 1);
    const double u = (column - half) / half;
    const double v = (row - half) / half;
    if (std::abs(u) < 1.0e-12 && std::abs(v) < 1.0e-12) {
        *x = 0.0;
        *y = 0.0;
        return;
    }
    double r, angle;
    if (std::abs(u) >= std::abs(v)) {
        r = u;
        angle = (kPi / 4.0) * (v / u);
    } else {
        r = v;
        angle = kPi / 2.0 - (kPi / 4.0) * (u / v);
    }
    *x = r * std::cos(angle);
    *y = r * std::sin(angle);
}

bool onGridBoundary(uint32_t row, uint32_t column) {
    return row == 0 || column == 0 || row + 1 == kGrid || column + 1 == kGrid;
}

double positiveAngle(double x, double y) {
    double angle = std::atan2(y, x);
    if (angle < 0.0) angle += 2.0 * kPi;
    return angle;
}

uint32_t gridIndex(uint32_t row, uint32_t column) {
    if (onGridBoundary(row, column)) {
        double x, y;
        diskCoordinate(row, column, &x, &y);
        const double ring = positiveAngle(x, y) * kAround / (2.0 * kPi);
        return wallIndex(0, static_cast<uint32_t>(std::llround(ring)) % kAround);
    }
    return kAround * kLevels + (row - 1) * kInterior + column - 1;
}

D3 bottomPosition(uint32_t row, uint32_t column) {
    double x, y;
    diskCoordinate(row, column, &x, &y);
    const double radial = std::hypot(x, y);
    if (radial < 1.0e-12) return {0.0, 0.0, 0.021};
    const double ringCoordinate = positiveAngle(x, y) * kAround / (2.0 * kPi);
    const uint32_t first = static_cast<uint32_t>(std::floor(ringCoordinate)) % kAround;
    const uint32_t second = (first + 1) % kAround;
    const double f = ringCoordinate - std::floor(ringCoordinate);
    const D3 outer = wallPosition(0, first) * (1.0 - f) + wallPosition(0, second) * f;
    return {outer.x * radial, outer.y * radial, 0.021 + radial * (outer.z - 0.021)};
}

void buildBag(const NumiClothHipBagSpec& spec, NumiClothHipScene& s) {
    const NumiClothMaterialValues& m = spec.material;
    const float ordinaryMass = static_cast<float>(m.ordinaryNodeMassKg);
    const float hemMass = static_cast<float>(m.hemNodeMassKg);
    const float yarnRadius = static_cast<float>(m.yarnRadiusM);
    const float axialBody = spec.stretchComplianceOverride > 0.0f
                                ? spec.stretchComplianceOverride
                                : static_cast<float>(m.axialBodyComplianceMPerN);
    const float axialCuff = spec.stretchComplianceOverride > 0.0f
                                ? spec.stretchComplianceOverride
                                : static_cast<float>(m.axialCuffComplianceMPerN);
    const float bendBody = spec.bendComplianceOverride > 0.0f
                               ? spec.bendComplianceOverride
                               : static_cast<float>(m.bendBodyComplianceMPerN);
    const float bendCuff = spec.bendComplianceOverride > 0.0f
                               ? spec.bendComplianceOverride
                               : static_cast<float>(m.bendCuffComplianceMPerN);
    const float strainLimit = spec.strainLimitOverride > 0.0f ? spec.strainLimitOverride : 0.285f;
    const float knotCompliance = static_cast<float>(m.knotCompliance);
    const float gripCompliance = static_cast<float>(m.gripComplianceMPerN);

    for (uint32_t level = 0; level < kLevels; ++level)
        for (uint32_t ring = 0; ring < kAround; ++ring)
            s.particles.push_back(makeParticle(wallPosition(level, ring),
                                               level + 2 >= kLevels ? hemMass : ordinaryMass, false));
    for (uint32_t row = 1; row + 1 < kGrid; ++row)
        for (uint32_t column = 1; column + 1 < kGrid; ++column)
            s.particles.push_back(makeParticle(bottomPosition(row, column), ordinaryMass, false));

    std::vector<EdgeItem> edges;
    for (uint32_t level = 0; level < kLevels; ++level)
        for (uint32_t ring = 0; ring < kAround; ++ring)
            edges.push_back({wallIndex(level, ring), wallIndex(level, ring + 1), 1u,
                             level + 2 >= kLevels ? axialCuff : axialBody});
    for (uint32_t level = 0; level + 1 < kLevels; ++level)
        for (uint32_t ring = 0; ring < kAround; ++ring)
            edges.push_back({wallIndex(level, ring), wallIndex(level + 1, ring), 0u, axialBody});
    for (uint32_t row = 0; row < kGrid; ++row)
        for (uint32_t column = 0; column + 1 < kGrid; ++column)
            if (!(onGridBoundary(row, column) && onGridBoundary(row, column + 1)))
                edges.push_back({gridIndex(row, column), gridIndex(row, column + 1), 2u, axialBody});
    for (uint32_t column = 0; column < kGrid; ++column)
        for (uint32_t row = 0; row + 1 < kGrid; ++row)
            if (!(onGridBoundary(row, column) && onGridBoundary(row + 1, column)))
                edges.push_back({gridIndex(row, column), gridIndex(row + 1, column), 2u, axialBody});
    emitDistances(s, edges, strainLimit);

    std::vector<KnotItem> knots;
    for (uint32_t level = 1; level + 1 < kLevels; ++level)
        for (uint32_t ring = 0; ring < kAround; ++ring)
            knots.push_back({{wallIndex(level - 1, ring), wallIndex(level + 1, ring),
                              wallIndex(level, ring + kAround - 1), wallIndex(level, ring + 1)}});
    for (uint32_t row = 1; row + 1 < kGrid; ++row)
        for (uint32_t column = 1; column + 1 < kGrid; ++column)
            knots.push_back({{gridIndex(row - 1, column), gridIndex(row + 1, column),
                              gridIndex(row, column - 1), gridIndex(row, column + 1)}});
    emitKnots(s, knots, knotCompliance);

    std::vector<BendItem> bends;
    for (uint32_t level = 0; level < kLevels; ++level)
        for (uint32_t ring = 0; ring < kAround; ++ring)
            bends.push_back({wallIndex(level, ring + kAround - 1), wallIndex(level, ring),
                             wallIndex(level, ring + 1), level + 2 >= kLevels ? bendCuff : bendBody});
    for (uint32_t level = 1; level + 1 < kLevels; ++level)
        for (uint32_t ring = 0; ring < kAround; ++ring)
            bends.push_back({wallIndex(level - 1, ring), wallIndex(level, ring), wallIndex(level + 1, ring), bendBody});
    for (uint32_t row = 1; row + 1 < kGrid; ++row) {
        for (uint32_t column = 1; column + 1 < kGrid; ++column) {
            bends.push_back({gridIndex(row, column - 1), gridIndex(row, column), gridIndex(row, column + 1), bendBody});
            bends.push_back({gridIndex(column - 1, row), gridIndex(column, row), gridIndex(column + 1, row), bendBody});
        }
    }
    emitBends(s, bends);

    const D3 base = wallPosition(kLevels - 1, 0);
    NumiClothBagGPUConfig& c = s.config;
    const uint32_t fruitCount = 12;
    c.control = u4(NUMI_CLOTH_BAG_GPU_ABI_VERSION, static_cast<uint32_t>(s.particles.size()),
                   static_cast<uint32_t>(s.distances.size()), 10u);
    c.constraintCounts = u4(static_cast<uint32_t>(s.knots.size()), static_cast<uint32_t>(s.bends.size()), 0u,
                            fruitCount);
    c.contactCounts = u4(fruitCount * (fruitCount - 1) / 2,
                         fruitCount * static_cast<uint32_t>(s.distances.size()), 0u, 0u);
    c.gravityAndTimestep = f4(0.0f, 0.0f, -9.81f, 1.0f / 5760.0f);
    c.gripTargetAndActive = f4(static_cast<float>(base.x + 0.002), static_cast<float>(base.y - 0.001),
                               static_cast<float>(base.z + 0.006), 1.0f);
    c.gripOrientation = f4(0.0f, 0.0f, 0.0f, 1.0f);
    c.gripControl = u4(1u, 0u, kAround, 2u);
    c.gripMaterial = f4(0.12f, 0.0f, 0.0f, 0.0f);
    float longest = 0.0f;
    for (const NumiClothBagGPUDistance& d : s.distances)
        longest = std::max(longest, d.material.x * (1.0f + d.material.w));
    c.clothMaterial = f4(yarnRadius, std::max(0.1f, longest + 0.008001f),
                         static_cast<float>(m.clothGroundFriction), static_cast<float>(m.clothSelfFriction));
    c.fruitMaterial = f4(static_cast<float>(m.fruitPairFriction), static_cast<float>(m.fruitGroundFriction),
                         static_cast<float>(m.fruitRollingResistance), static_cast<float>(m.fruitYarnFriction));
    c.airVelocityAndDensity = f4(0.0f, 0.0f, 0.0f, 1.225f);
    c.aerodynamicCoefficients = f4(static_cast<float>(m.yarnCrossflowDrag), static_cast<float>(m.yarnSkinFriction),
                                   static_cast<float>(m.fruitDrag), static_cast<float>(m.fruitRotationalDrag));
    c.mouthControl = u4(kAround, gridIndex(kGrid / 2, kGrid / 2), 0u, 0u);
    c.mouthMaterial = f4(0.025f, 0.0f, 0.0f, 0.0f);
    for (uint32_t ring = 0; ring < kAround; ++ring) s.mouthRim.push_back(wallIndex(kLevels - 1, ring));
    for (uint32_t level = kLevels - 2; level < kLevels; ++level)
        for (uint32_t ring = 0; ring < kAround; ++ring) s.cuff.push_back(wallIndex(level, ring));

    for (uint32_t level = kLevels - 2; level < kLevels; ++level) {
        for (int offset = -2; offset <= 2; ++offset) {
            const uint32_t ring = static_cast<uint32_t>((static_cast<int>(kAround) + offset) % static_cast<int>(kAround));
            const uint32_t particle = wallIndex(level, ring);
            const D3 rest = asD3(s.particles[particle].positionAndInverseMass);
            NumiClothBagGPUGrip grip{};
            grip.particle = u4(particle, 1u, 0u, 0u);
            grip.targetOffsetAndCompliance = f4(static_cast<float>(rest.x - base.x), static_cast<float>(rest.y - base.y),
                                                static_cast<float>(rest.z - base.z), gripCompliance);
            grip.lambda = f4(0.0f, 0.0f, 0.0f, 0.0f);
            s.grips.push_back(grip);
        }
    }

    // Fruit centres, radii, masses and appearances of the reference bag
    // (reference tools/cloth_bag_metal.mm:988-1013).
    static const double centres[12][3] = {
        {0.190, 0.000, 0.088}, {0.118, 0.149, 0.083}, {-0.042, 0.185, 0.093}, {-0.171, 0.082, 0.078},
        {-0.171, -0.082, 0.086}, {-0.042, -0.185, 0.090}, {0.118, -0.149, 0.081}, {0.000, 0.000, 0.084},
        {0.105, 0.000, 0.210}, {0.000, 0.105, 0.210}, {-0.105, 0.000, 0.215}, {0.000, -0.105, 0.210}};
    static const float radii[12] = {0.070f, 0.065f, 0.075f, 0.060f, 0.068f, 0.072f,
                                    0.063f, 0.066f, 0.072f, 0.064f, 0.070f, 0.067f};
    static const float masses[12] = {0.21f, 0.17f, 0.26f, 0.14f, 0.19f, 0.23f,
                                     0.16f, 0.18f, 0.23f, 0.17f, 0.21f, 0.18f};
    static const uint32_t looks[12] = {0, 1, 2, 1, 3, 0, 2, 3, 0, 1, 2, 3};
    uint64_t rng = spec.seed;
    for (uint32_t i = 0; i < fruitCount; ++i) {
        const D3 p{centres[i][0], centres[i][1], centres[i][2] + 0.009};
        float radius = radii[i];
        if (spec.fruitRadiusMin > 0.0f && spec.fruitRadiusMax >= spec.fruitRadiusMin) {
            radius = static_cast<float>(spec.fruitRadiusMin +
                                        (spec.fruitRadiusMax - spec.fruitRadiusMin) * unitDouble(&rng));
        }
        NumiClothBagGPUFruit fruit{};
        fruit.positionAndInverseMass = f4(static_cast<float>(p.x), static_cast<float>(p.y), static_cast<float>(p.z),
                                          1.0f / masses[i]);
        fruit.previousAndRadius = f4(static_cast<float>(p.x), static_cast<float>(p.y), static_cast<float>(p.z), radius);
        fruit.orientation = f4(0.0f, 0.0f, 0.0f, 1.0f);
        fruit.identity = u4(looks[i], 0u, 0u, 0u);
        s.fruits.push_back(fruit);
    }
    std::vector<std::array<uint32_t, 2>> pairItems;
    std::vector<uint32_t> stable;
    for (uint32_t a = 0, k = 0; a < fruitCount; ++a)
        for (uint32_t b = a + 1; b < fruitCount; ++b, ++k) {
            pairItems.push_back({a, b});
            stable.push_back(k);
        }
    uint32_t pairColours = 0;
    const std::vector<uint32_t> colours = greedyColours(pairItems, fruitCount, &pairColours, "fruit pair");
    const std::vector<uint32_t> order = colourOrder(colours, pairColours, &s.fruitPairBatches);
    for (uint32_t index : order) {
        NumiClothBagGPUFruitPair pair{};
        pair.fruitsAndColor = u4(pairItems[index][0], pairItems[index][1], colours[index], stable[index]);
        s.fruitPairs.push_back(pair);
    }
    for (uint32_t f = 0; f < fruitCount; ++f) {
        for (uint32_t segment = 0; segment < s.distances.size(); ++segment) {
            NumiClothBagGPUYarnContact contact{};
            contact.identity = u4(f, segment, s.distances[segment].particlesAndColor.x,
                                  s.distances[segment].particlesAndColor.y);
            s.yarnContacts.push_back(contact);
        }
    }
    s.publish();
    s.desc.extension.selfContact = spec.selfContact;
    s.desc.extension.selfContactMargin = 0.25f * yarnRadius;
}

}  // namespace

void bagAuthoredPosition(uint32_t level, uint32_t ring, double out[3]) {
    const D3 p = wallPosition(level, ring);
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
}

uint32_t bagWallIndex(uint32_t level, uint32_t ring) { return wallIndex(level, ring); }

uint32_t bagGridIndex(uint32_t row, uint32_t column) { return gridIndex(row, column); }

static_assert(kBagAround == kAround && kBagLevels == kLevels && kBagGrid == kGrid, "bag geometry");

}  // namespace numi_hip

using namespace numi_hip;

extern "C" {

uint64_t numi_cloth_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

int numi_cloth_scene_sheet(const NumiClothHipSheetSpec* spec, NumiClothHipScene** out) {
    return guarded([&] {
        if (spec == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        auto scene = std::make_unique<NumiClothHipScene>();
        buildSheet(*spec, *scene);
        *out = scene.release();
    });
}

int numi_cloth_scene_bag(const NumiClothHipBagSpec* spec, NumiClothHipScene** out) {
    return guarded([&] {
        if (spec == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        auto scene = std::make_unique<NumiClothHipScene>();
        buildBag(*spec, *scene);
        *out = scene.release();
    });
}

int numi_cloth_scene_strip(const NumiClothHipStripSpec* spec, NumiClothHipScene** out) {
    return guarded([&] {
        if (spec == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        auto scene = std::make_unique<NumiClothHipScene>();
        buildStrip(*spec, *scene);
        *out = scene.release();
    });
}

int numi_cloth_scene_merge(const NumiClothHipSceneDesc* const* specimens, uint32_t count, NumiClothHipScene** out) {
    return guarded([&] {
        if (specimens == nullptr || out == nullptr || count == 0) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument or no specimens");
        }
        *out = nullptr;
        auto scene = std::make_unique<NumiClothHipScene>();
        mergeScenes(std::vector<const NumiClothHipSceneDesc*>(specimens, specimens + count), *scene);
        *out = scene.release();
    });
}

int numi_cloth_scene_merge_indices(const NumiClothHipScene* merged, uint32_t family, uint32_t specimen,
                                   const uint32_t** indices, uint32_t* count) {
    return guarded([&] {
        if (merged == nullptr || indices == nullptr || count == nullptr || family > 3) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument or unknown family");
        }
        const std::vector<uint32_t>& start = merged->mergeStart[family];
        if (start.empty()) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "scene was not merged");
        if (specimen + 1 >= start.size()) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "specimen out of range");
        *indices = merged->mergeIndex[family].data() + start[specimen];
        *count = start[specimen + 1] - start[specimen];
    });
}

const NumiClothHipSceneDesc* numi_cloth_scene_desc(const NumiClothHipScene* scene) {
    return scene == nullptr ? nullptr : &scene->desc;
}

void numi_cloth_scene_destroy(NumiClothHipScene* scene) { delete scene; }

}  // extern "C"
