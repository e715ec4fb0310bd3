"""Scene builders and colourings (CPU only).

Counts of the produce bag are the reference's published ones
(reference docs/METAL_CLOTH.md:308-311: 1,465 particles, 2,904 distances,
1,369 knots, 2,834 bends, 10 grips, 12 fruits, 66 fruit pairs, colour
counts 5/6/5/15). Colourings are checked structurally: batches tile the
records in order, no two records of a batch share a particle, and the
colouring is first-fit (every record of colour c touches a record of each
lower colour, which a greedy lowest-free-colour pass always produces).
"""

import numpy as np


FAMILIES = {
    "distances": ("particlesAndColor", [0, 1], 2, "distanceBatches"),
    "bends": ("particlesAndColor", [0, 2], 3, "bendBatches"),
    "knots": (None, None, None, "knotBatches"),
    "fruitPairs": ("fruitsAndColor", [0, 1], 2, "fruitPairBatches"),
}


def _family(scene, name):
    records = getattr(scene, name)
    field, members, lane, batches = FAMILIES[name]
    if name == "knots":
        ids = records["particles"].astype(np.int64)
        colours = records["control"][:, 0].astype(np.int64)
    else:
        ids = records[field][:, members].astype(np.int64)
        colours = records[field][:, lane].astype(np.int64)
    return ids, colours, getattr(scene, batches)["control"]


def check_colouring(scene):
    for name in FAMILIES:
        ids, colours, batches = _family(scene, name)
        if len(ids) == 0:
            continue
        # Batches tile the records in order, one colour each.
        assert batches[0, 0] == 0
        assert np.all(batches[1:, 0] == batches[:-1, 0] + batches[:-1, 1])
        assert batches[-1, 0] + batches[-1, 1] == len(ids)
        for first, count, colour, _ in batches:
            assert np.all(colours[first:first + count] == colour), name
            members = ids[first:first + count].ravel()
            assert len(np.unique(members)) == len(members), f"{name} colour {colour} shares a particle"
        # First-fit: a record of colour c meets every colour below c.
        touching = {}
        for k, row in enumerate(ids):
            for p in row:
                touching.setdefault(int(p), []).append(k)
        for k, row in enumerate(ids):
            seen = set()
            for p in row:
                seen.update(int(colours[j]) for j in touching[int(p)] if j != k)
            assert set(range(colours[k])) <= seen, f"{name} record {k} is not first-fit"


def test_bag_counts_match_reference(cloth):
    bag = cloth.Scene.bag()
    assert len(bag.particles) == 1465
    assert len(bag.distances) == 2904
    assert len(bag.knots) == 1369
    assert len(bag.bends) == 2834
    assert len(bag.grips) == 10
    assert len(bag.fruits) == 12
    assert len(bag.fruitPairs) == 66
    assert len(bag.yarnContacts) == 12 * 2904
    assert len(bag.mouthRimParticles) == 48
    # Published colour counts 5 / 6 / 5 / 15.
    assert len(bag.distanceBatches) == 5
    assert len(bag.knotBatches) == 6
    assert len(bag.bendBatches) == 5
    assert len(bag.fruitPairBatches) == 15
    check_colouring(bag)
    c = bag.config[0]
    assert tuple(c["control"]) == (13, 1465, 2904, 10)
    # The reference's bag state leaves the ground flag off
    # (reference tools/cloth_bag_metal.mm:901-903).
    assert tuple(c["constraintCounts"]) == (1369, 2834, 0, 12)
    assert tuple(c["contactCounts"][:2]) == (66, 12 * 2904)


def test_bag_is_deterministic(cloth):
    a, b = cloth.Scene.bag(seed=5), cloth.Scene.bag(seed=5)
    for name, _, _ in cloth.Scene.ARRAYS:
        assert getattr(a, name).tobytes() == getattr(b, name).tobytes(), name


def test_sheet_counts_and_colouring(cloth):
    W, H = 9, 7
    s = cloth.Scene.sheet(width=W, height=H, crossingKnots=1)
    assert len(s.particles) == W * H
    assert len(s.distances) == (W - 1) * H + W * (H - 1)
    assert len(s.bends) == (W - 2) * H + W * (H - 2)
    assert len(s.knots) == (W - 2) * (H - 2)
    check_colouring(s)
    # A grid's distances need 4 colours under first-fit in row order, its
    # chord bends 4 as well.
    assert len(s.distanceBatches) == 4
    assert len(s.bendBatches) == 4


def test_sheet_rest_lengths_and_masses(cloth):
    s = cloth.Scene.sheet(width=5, height=4, spacing=0.02, arealDensity=0.3)
    rest = s.distances["material"][:, 0]
    assert np.allclose(rest, 0.02, rtol=1e-6)
    mass = s.particles["previousAndMass"][:, 3]
    assert np.allclose(mass, 0.3 * 0.02 * 0.02, rtol=1e-6)
    chord = s.bends["material"][:, 0]
    assert np.allclose(chord, 0.04, rtol=1e-6)
