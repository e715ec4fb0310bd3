#pragma once

// Host planner for LDS-resident tile programs.
//
// A tile program is a fixed sequence of colour passes (distance, knot, bend,
// strain-limit colours and the per-knot ground/sphere pass) that one launch
// runs entirely in LDS: every workgroup loads one tile of owned knots plus a
// halo, runs all passes with a workgroup barrier between colours, and writes
// back only its owned knots and the multipliers of the constraints it owns.
//
// Exactness. Within a colour no two constraints share a knot, so the
// per-colour global schedule is Gauss-Seidel over colours. A tile reproduces
// it bit for bit for a knot as long as every constraint that feeds that knot,
// through the chain of passes, had all its inputs loaded and exact. The
// planner simulates that chain on the host: every loaded knot starts exact; a
// constraint with all participants loaded and exact keeps them exact; a
// constraint with an input missing or inexact makes its loaded participants
// inexact. The halo grows (breadth-first over the program's constraint graph)
// until every owned knot, and every constraint owned by the tile (owner = the
// tile of its first participant), is exact at the end. Halo results are
// discarded, so the written state equals the per-colour schedule's.
//
// Tiles are consecutive chunks of the knots ordered by connected component
// and then by the Morton code of their positions at create, so they are
// spatially compact for any mesh and never mix independent specimens more
// than a chunk boundary requires.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace numi_hip {

enum TilePassKind : uint32_t {
    kTileDistance = 0,
    kTileKnot = 1,
    kTileBend = 2,
    kTileStrain = 3,
    kTileCollide = 4
};

struct TilePass {
    uint32_t kind;
    uint32_t first;  // colour batch range in the family's array
    uint32_t count;
};

// Participants of the constraint families (global knot indices).
struct TileTopology {
    uint32_t particles = 0;
    const uint32_t* distancePairs = nullptr;  // 2 per distance
    const uint32_t* bendPairs = nullptr;      // 2 per bend (first, third)
    const uint32_t* knotQuads = nullptr;      // 4 per knot
    const float* positions = nullptr;         // 4 per knot (xyz w) for ordering
    // Constants packed into the items (the device arrays' exact values).
    const float* distanceRest = nullptr;
    const float* distanceCompliance = nullptr;
    const float* distanceMaxLength = nullptr;
    const float* bendRest = nullptr;
    const float* bendCompliance = nullptr;
};

struct TilePlan {
    uint32_t tiles = 0;
    uint32_t ownedPerTile = 0;
    uint32_t maxLoaded = 0;
    uint32_t halo = 0;
    std::vector<uint32_t> loadStart;    // tiles + 1
    std::vector<uint32_t> loadIds;      // owned first, then halo
    std::vector<uint32_t> ownedCount;   // tiles
    std::vector<uint32_t> itemStart;    // tiles * passes + 1
    // One 16-byte item per computed constraint, x = local a | local b << 16:
    //   distance y = constraint | owner << 31, z = rest, w = compliance (bits)
    //   bend     y = constraint | owner << 31, z = rest chord, w = compliance
    //   strain   y = constraint, z = rest * (1 + limit), w = 0
    //   knot     y = local c | local d << 16, z = constraint | owner << 31, w = 0
    std::vector<uint32_t> items;        // 4 per item
    uint64_t computed = 0;              // items over all tiles
    uint64_t unique = 0;                // constraint-passes in the program
};

inline uint64_t spreadBits21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

inline uint32_t participantsOf(const TileTopology& t, uint32_t kind, uint32_t c, uint32_t* out) {
    if (kind == kTileDistance || kind == kTileStrain) {
        out[0] = t.distancePairs[2 * c];
        out[1] = t.distancePairs[2 * c + 1];
        return 2;
    }
    if (kind == kTileBend) {
        out[0] = t.bendPairs[2 * c];
        out[1] = t.bendPairs[2 * c + 1];
        return 2;
    }
    for (int k = 0; k < 4; ++k) out[k] = t.knotQuads[4 * c + k];
    return 4;
}

// Drops the items of one tile whose results reach neither an owned knot nor
// an owned multiplier. Walking the passes backwards from the end, where the
// owned knots are needed: an item of pass k is kept when it touches a knot
// needed after pass k, and then all its participants are needed before
// pass k (within a colour no two items share a knot, so the order inside a
// pass does not matter). Every kept item reads exactly the values it read
// in the full program, so the owned knots and multipliers are the same
// bytes; only halo work whose results were discarded goes. Owner items touch
// their owned first participant, so they are always kept.
inline void pruneUnneededItems(const std::vector<TilePass>& program, uint32_t owned, uint32_t loaded,
                               std::vector<uint32_t>& items, std::vector<uint32_t>& passItemStart) {
    const uint32_t passes = static_cast<uint32_t>(program.size());
    const uint32_t total = static_cast<uint32_t>(items.size() / 4);
    std::vector<uint8_t> needed(loaded, 0), keep(total, 0);
    for (uint32_t k = 0; k < owned; ++k) needed[k] = 1;
    for (uint32_t k = passes; k-- > 0;) {
        const uint32_t first = passItemStart[k];
        const uint32_t last = k + 1 < passes ? passItemStart[k + 1] : total;
        const bool quad = program[k].kind == kTileKnot;
        for (uint32_t j = first; j < last; ++j) {
            const uint32_t* it = &items[4 * j];
            const uint32_t l[4] = {it[0] & 0xffffu, it[0] >> 16, quad ? (it[1] & 0xffffu) : 0u, quad ? it[1] >> 16 : 0u};
            const uint32_t n = quad ? 4u : 2u;
            bool any = false;
            for (uint32_t m = 0; m < n; ++m) any = any || needed[l[m]] != 0;
            if (!any) continue;
            keep[j] = 1;
        }
        for (uint32_t j = first; j < last; ++j) {
            if (!keep[j]) continue;
            const uint32_t* it = &items[4 * j];
            needed[it[0] & 0xffffu] = 1;
            needed[it[0] >> 16] = 1;
            if (quad) {
                needed[it[1] & 0xffffu] = 1;
                needed[it[1] >> 16] = 1;
            }
        }
    }
    // Within a pass (one colour: no two items share a knot, so their order
    // does not change any result) the kept items go in ascending LDS slot of
    // their first knot, so the lanes of a wave read neighbouring LDS slots
    // rather than the scattered slots of constraint order.
    std::vector<uint32_t> kept;
    kept.reserve(items.size());
    std::vector<uint32_t> starts(passes), order;
    for (uint32_t k = 0; k < passes; ++k) {
        starts[k] = static_cast<uint32_t>(kept.size() / 4);
        const uint32_t first = passItemStart[k];
        const uint32_t last = k + 1 < passes ? passItemStart[k + 1] : total;
        order.clear();
        for (uint32_t j = first; j < last; ++j)
            if (keep[j]) order.push_back(j);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            return (items[4 * x] & 0xffffu) < (items[4 * y] & 0xffffu);
        });
        for (uint32_t j : order) kept.insert(kept.end(), items.begin() + 4 * j, items.begin() + 4 * j + 4);
    }
    items.swap(kept);
    passItemStart.swap(starts);
}

// Builds a plan for `program`; returns false when some tile would need more
// than `capacity` loaded knots or more than `maxHalo` hops of halo.
inline bool planTiles(const TileTopology& topo, const std::vector<TilePass>& program, uint32_t owned,
                      uint32_t capacity, uint32_t maxHalo, TilePlan* plan) {
    const uint32_t P = topo.particles;
    *plan = TilePlan{};
    if (P == 0 || owned == 0) return false;
    // Morton order of the create-time positions.
    float lo[3] = {topo.positions[0], topo.positions[1], topo.positions[2]};
    float hi[3] = {lo[0], lo[1], lo[2]};
    for (uint32_t i = 0; i < P; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], topo.positions[4 * i + a]);
            hi[a] = std::max(hi[a], topo.positions[4 * i + a]);
        }
    // One scale for all axes (the largest extent), so that a thin axis such
    // as the jitter of a flat sheet contributes only its true share of bits
    // and consecutive codes form compact patches.
    double span = 0.0;
    for (int a = 0; a < 3; ++a) span = std::max(span, static_cast<double>(hi[a]) - lo[a]);
    // Connected components of the program's constraint graph (union-find,
    // each labelled by its lowest knot): knots are ordered by component
    // first, so independent specimens merged into one scene tile apart even
    // where they overlap in space.
    std::vector<uint32_t> parent(P);
    for (uint32_t i = 0; i < P; ++i) parent[i] = i;
    auto root = [&](uint32_t i) {
        while (parent[i] != i) {
            parent[i] = parent[parent[i]];
            i = parent[i];
        }
        return i;
    };
    for (const TilePass& pass : program) {
        if (pass.kind == kTileCollide) continue;
        for (uint32_t c = pass.first; c < pass.first + pass.count; ++c) {
            uint32_t ids[4];
            const uint32_t n = participantsOf(topo, pass.kind, c, ids);
            for (uint32_t m = 1; m < n; ++m) {
                const uint32_t a = root(ids[0]), b = root(ids[m]);
                if (a != b) parent[std::max(a, b)] = std::min(a, b);
            }
        }
    }
    struct Keyed {
        uint32_t component;
        uint64_t code;
        uint32_t knot;
        bool operator<(const Keyed& o) const {
            if (component != o.component) return component < o.component;
            if (code != o.code) return code < o.code;
            return knot < o.knot;
        }
    };
    std::vector<Keyed> keyed(P);
    for (uint32_t i = 0; i < P; ++i) {
        uint64_t code = 0;
        for (int a = 0; a < 3; ++a) {
            const double u = span > 0.0 ? (static_cast<double>(topo.positions[4 * i + a]) - lo[a]) / span : 0.0;
            code |= spreadBits21(static_cast<uint64_t>(u * 2097151.0)) << a;
        }
        keyed[i] = {root(i), code, i};
    }
    std::sort(keyed.begin(), keyed.end());
    const uint32_t tiles = (P + owned - 1) / owned;
    std::vector<uint32_t> tileOf(P);
    for (uint32_t k = 0; k < P; ++k) tileOf[keyed[k].knot] = k / owned;

    // Per pass: the constraint touching each knot (colours are disjoint).
    const uint32_t passes = static_cast<uint32_t>(program.size());
    std::vector<std::vector<uint32_t>> touching(passes);
    uint64_t unique = 0;
    std::vector<std::vector<uint32_t>> neighbours(P);
    for (uint32_t k = 0; k < passes; ++k) {
        const TilePass& pass = program[k];
        if (pass.kind == kTileCollide) continue;
        unique += pass.count;
        touching[k].assign(P, ~0u);
        for (uint32_t c = pass.first; c < pass.first + pass.count; ++c) {
            uint32_t ids[4];
            const uint32_t n = participantsOf(topo, pass.kind, c, ids);
            for (uint32_t m = 0; m < n; ++m) {
                touching[k][ids[m]] = c;
                for (uint32_t o = 0; o < n; ++o)
                    if (o != m) neighbours[ids[m]].push_back(ids[o]);
            }
        }
    }
    for (auto& list : neighbours) {
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
    }

    std::vector<int32_t> localOf(P, -1);
    std::vector<uint32_t> stamp(P, ~0u);
    plan->tiles = tiles;
    plan->ownedPerTile = owned;
    plan->unique = unique;
    plan->loadStart.push_back(0);
    plan->itemStart.push_back(0);
    std::vector<uint32_t> loaded;
    std::vector<uint8_t> exact;
    std::vector<uint32_t> tileItems;
    // Tiles of one mesh need about the same halo: each search starts at the
    // previous tile's (any sufficient halo is exact; the smallest only saves
    // redundant work).
    uint32_t startHalo = 0;
    for (uint32_t t = 0; t < tiles; ++t) {
        const uint32_t begin = t * owned;
        const uint32_t end = std::min(P, begin + owned);
        bool done = false;
        for (uint32_t halo = startHalo; halo <= maxHalo && !done; ++halo) {
            // Owned knots first (Morton order), then the halo by BFS depth.
            loaded.clear();
            for (uint32_t k = begin; k < end; ++k) loaded.push_back(keyed[k].knot);
            for (uint32_t p : loaded) localOf[p] = -1;
            for (uint32_t k = 0; k < loaded.size(); ++k) {
                localOf[loaded[k]] = static_cast<int32_t>(k);
                stamp[loaded[k]] = t * (maxHalo + 1) + halo;
            }
            std::size_t frontier = 0;
            for (uint32_t depth = 0; depth < halo; ++depth) {
                const std::size_t levelEnd = loaded.size();
                for (std::size_t k = frontier; k < levelEnd; ++k) {
                    for (uint32_t q : neighbours[loaded[k]]) {
                        if (stamp[q] == t * (maxHalo + 1) + halo) continue;
                        stamp[q] = t * (maxHalo + 1) + halo;
                        localOf[q] = static_cast<int32_t>(loaded.size());
                        loaded.push_back(q);
                    }
                }
                frontier = levelEnd;
            }
            if (loaded.size() > capacity || loaded.size() > 65535) {
                for (uint32_t p : loaded) stamp[p] = ~0u;
                return false;
            }
            auto isLoaded = [&](uint32_t p) { return stamp[p] == t * (maxHalo + 1) + halo; };
            exact.assign(loaded.size(), 1);
            tileItems.clear();
            std::vector<uint32_t> passItemStart;
            bool ownersExact = true;
            for (uint32_t k = 0; k < passes && ownersExact; ++k) {
                passItemStart.push_back(static_cast<uint32_t>(tileItems.size() / 4));
                const TilePass& pass = program[k];
                if (pass.kind == kTileCollide) continue;
                // Unique constraints of this pass touching the loaded set, in
                // ascending constraint order.
                std::vector<uint32_t> cs;
                for (uint32_t p : loaded) {
                    const uint32_t c = touching[k][p];
                    if (c != ~0u) cs.push_back(c);
                }
                std::sort(cs.begin(), cs.end());
                cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
                for (uint32_t c : cs) {
                    uint32_t ids[4];
                    const uint32_t n = participantsOf(topo, pass.kind, c, ids);
                    bool all = true, ok = true;
                    for (uint32_t m = 0; m < n; ++m) {
                        if (!isLoaded(ids[m])) {
                            all = false;
                        } else if (!exact[localOf[ids[m]]]) {
                            ok = false;
                        }
                    }
                    const bool owner = tileOf[ids[0]] == t && pass.kind != kTileStrain;
                    if (!all) {
                        for (uint32_t m = 0; m < n; ++m)
                            if (isLoaded(ids[m])) exact[localOf[ids[m]]] = 0;
                        if (tileOf[ids[0]] == t) ownersExact = false;
                        continue;
                    }
                    if (!ok) {
                        for (uint32_t m = 0; m < n; ++m) exact[localOf[ids[m]]] = 0;
                        if (tileOf[ids[0]] == t) ownersExact = false;
                    }
                    const uint32_t l[4] = {static_cast<uint32_t>(localOf[ids[0]]),
                                           static_cast<uint32_t>(localOf[ids[1]]),
                                           n > 2 ? static_cast<uint32_t>(localOf[ids[2]]) : 0u,
                                           n > 3 ? static_cast<uint32_t>(localOf[ids[3]]) : 0u};
                    const uint32_t tagged = c | (owner ? 0x80000000u : 0u);
                    auto bits = [](float v) {
                        uint32_t u;
                        std::memcpy(&u, &v, sizeof(u));
                        return u;
                    };
                    tileItems.push_back(l[0] | (l[1] << 16));
                    if (pass.kind == kTileKnot) {
                        tileItems.push_back(l[2] | (l[3] << 16));
                        tileItems.push_back(tagged);
                        tileItems.push_back(0u);
                    } else if (pass.kind == kTileDistance) {
                        tileItems.push_back(tagged);
                        tileItems.push_back(bits(topo.distanceRest[c]));
                        tileItems.push_back(bits(topo.distanceCompliance[c]));
                    } else if (pass.kind == kTileBend) {
                        tileItems.push_back(tagged);
                        tileItems.push_back(bits(topo.bendRest[c]));
                        tileItems.push_back(bits(topo.bendCompliance[c]));
                    } else {
                        tileItems.push_back(c);
                        tileItems.push_back(bits(topo.distanceMaxLength[c]));
                        tileItems.push_back(0u);
                    }
                }
            }
            bool ownedExact = ownersExact;
            for (uint32_t k = 0; k < end - begin && ownedExact; ++k) ownedExact = exact[k] != 0;
            if (!ownedExact) {
                for (uint32_t p : loaded) stamp[p] = ~0u;
                continue;
            }
            while (passItemStart.size() < passes) passItemStart.push_back(static_cast<uint32_t>(tileItems.size() / 4));
            pruneUnneededItems(program, end - begin, static_cast<uint32_t>(loaded.size()), tileItems, passItemStart);
            const uint32_t base = static_cast<uint32_t>(plan->items.size() / 4);
            for (uint32_t k = 1; k < passes; ++k) plan->itemStart.push_back(base + passItemStart[k]);
            plan->items.insert(plan->items.end(), tileItems.begin(), tileItems.end());
            plan->itemStart.push_back(static_cast<uint32_t>(plan->items.size() / 4));
            plan->loadIds.insert(plan->loadIds.end(), loaded.begin(), loaded.end());
            plan->loadStart.push_back(static_cast<uint32_t>(plan->loadIds.size()));
            plan->ownedCount.push_back(end - begin);
            plan->maxLoaded = std::max<uint32_t>(plan->maxLoaded, static_cast<uint32_t>(loaded.size()));
            plan->halo = std::max(plan->halo, halo);
            startHalo = halo;
            plan->computed += tileItems.size() / 4;
            for (uint32_t p : loaded) stamp[p] = ~0u;
            done = true;
        }
        if (!done) return false;
    }
    return true;
}

}  // namespace numi_hip
