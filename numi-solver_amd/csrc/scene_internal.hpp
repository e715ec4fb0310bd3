#pragma once

// Owning storage behind a NumiClothHipScene handle (scene builders, merged
// specimen scenes). Host code only.

#include <cstdint>
#include <vector>

#include "numi_cloth_scene.h"

struct NumiClothHipScene {
    NumiClothBagGPUConfig config{};
    std::vector<NumiClothBagGPUParticle> particles;
    std::vector<NumiClothBagGPUDistance> distances;
    std::vector<NumiClothBagGPUBatch> distanceBatches;
    std::vector<NumiClothBagGPUGrip> grips;
    std::vector<NumiClothBagGPUKnot> knots;
    std::vector<NumiClothBagGPUBatch> knotBatches;
    std::vector<NumiClothBagGPUBend> bends;
    std::vector<NumiClothBagGPUBatch> bendBatches;
    std::vector<NumiClothBagGPUFruit> fruits;
    std::vector<NumiClothBagGPUFruitPair> fruitPairs;
    std::vector<NumiClothBagGPUBatch> fruitPairBatches;
    std::vector<NumiClothBagGPUYarnContact> yarnContacts;
    std::vector<uint32_t> mouthRim;
    std::vector<uint32_t> cuff;
    std::vector<mr_float4> spheres;
    std::vector<mr_float4> motion;  // per-particle drive and damping, or empty
    // Merged scenes: per family, mergeIndex[mergeStart[s] + k] is the merged
    // index of record k of specimen s (mergeStart has specimens + 1 entries).
    std::vector<uint32_t> mergeStart[4];
    std::vector<uint32_t> mergeIndex[4];
    NumiClothHipSceneDesc desc{};

    // Points the descriptor at the arrays (after they changed); the scalar
    // extension fields the builder set (self-contact, margin, drive,
    // damping, extra passes) are kept.
    void publish() {
        const NumiClothHipSceneExtension extension = desc.extension;
        desc = NumiClothHipSceneDesc{};
        desc.extension = extension;
        desc.config = &config;
        desc.particles = particles.data();
        desc.distances = distances.data();
        desc.distanceBatches = distanceBatches.data();
        desc.grips = grips.data();
        desc.knots = knots.data();
        desc.knotBatches = knotBatches.data();
        desc.bends = bends.data();
        desc.bendBatches = bendBatches.data();
        desc.fruits = fruits.data();
        desc.fruitPairs = fruitPairs.data();
        desc.fruitPairBatches = fruitPairBatches.data();
        desc.yarnContacts = yarnContacts.data();
        desc.mouthRimParticles = mouthRim.data();
        desc.cuffParticles = cuff.data();
        desc.particleCount = static_cast<uint32_t>(particles.size());
        desc.distanceCount = static_cast<uint32_t>(distances.size());
        desc.distanceBatchCount = static_cast<uint32_t>(distanceBatches.size());
        desc.gripCount = static_cast<uint32_t>(grips.size());
        desc.knotCount = static_cast<uint32_t>(knots.size());
        desc.knotBatchCount = static_cast<uint32_t>(knotBatches.size());
        desc.bendCount = static_cast<uint32_t>(bends.size());
        desc.bendBatchCount = static_cast<uint32_t>(bendBatches.size());
        desc.fruitCount = static_cast<uint32_t>(fruits.size());
        desc.fruitPairCount = static_cast<uint32_t>(fruitPairs.size());
        desc.fruitPairBatchCount = static_cast<uint32_t>(fruitPairBatches.size());
        desc.yarnContactCount = static_cast<uint32_t>(yarnContacts.size());
        desc.mouthRimCount = static_cast<uint32_t>(mouthRim.size());
        desc.cuffCount = static_cast<uint32_t>(cuff.size());
        desc.extension.staticSpheres = spheres.data();
        desc.extension.staticSphereCount = static_cast<uint32_t>(spheres.size());
        desc.extension.particleMotion = motion.empty() ? nullptr : motion.data();
    }
};

namespace numi_hip {

// Produce-bag geometry shared by the scene builder and the frame driver
// (reference tools/cloth_bag_metal.mm:33-38, :302-440): 48 knots around,
// 28 wall levels, a 13 x 13 bottom grid whose boundary is ring 0 of level 0.
constexpr uint32_t kBagAround = 48;
constexpr uint32_t kBagLevels = 28;
constexpr uint32_t kBagGrid = 13;
constexpr uint32_t kBagKnots = kBagAround * kBagLevels + (kBagGrid - 2) * (kBagGrid - 2);

// Authored (rest) position of wall knot (level, ring), metres.
void bagAuthoredPosition(uint32_t level, uint32_t ring, double out[3]);
// Knot index of wall knot (level, ring) and of bottom-grid node (row, column).
uint32_t bagWallIndex(uint32_t level, uint32_t ring);
uint32_t bagGridIndex(uint32_t row, uint32_t column);

}  // namespace numi_hip
