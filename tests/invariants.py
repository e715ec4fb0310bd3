"""The reference's host-side qualification metrics, restated in FP64 numpy
over a read-back state (reference tools/cloth_bag_metal.mm:5176-5400):
strain-limit violation, ground penetration, yarn/yarn self penetration,
minimum cloth height, fruit clearance, grounded knot and fruit counts, the
escape mask, shape change and kinetic energy. Test infrastructure only.
"""

import numpy as np


def _xyz(records, field="positionAndInverseMass"):
    return records[field][:, :3].astype(np.float64)


def strain_violation(state, scene):
    """max(|x_j - x_i| - rest (1 + limit)) over distances (:5176-5191)."""
    x = _xyz(state["particles"])
    ij = scene.distances["particlesAndColor"][:, :2]
    rest = scene.distances["material"][:, 0].astype(np.float64)
    limit = scene.distances["material"][:, 3].astype(np.float64)
    length = np.linalg.norm(x[ij[:, 1]] - x[ij[:, 0]], axis=1)
    return max(0.0, float(np.max(length - rest * (1.0 + limit)))) if len(ij) else 0.0


def ground_penetration(state, radius):
    """max(r - z) over knots and (radius - z) over fruits (:5193-5214)."""
    z = state["particles"]["positionAndInverseMass"][:, 2].astype(np.float64)
    worst = float(np.max(radius - z)) if len(z) else 0.0
    fruits = state["fruits"]
    if len(fruits):
        worst = max(worst, float(np.max(fruits["previousAndRadius"][:, 3].astype(np.float64)
                                        - fruits["positionAndInverseMass"][:, 2].astype(np.float64))))
    return max(0.0, worst)


def _segment_distances(a0, a1, b0, b1):
    """Closest distances of segment pairs (rows), clamped textbook solve."""
    da, db, r = a1 - a0, b1 - b0, a0 - b0
    a = np.einsum("ij,ij->i", da, da)
    e = np.einsum("ij,ij->i", db, db)
    f = np.einsum("ij,ij->i", db, r)
    c = np.einsum("ij,ij->i", da, r)
    b = np.einsum("ij,ij->i", da, db)
    den = a * e - b * b
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(den > 1e-20, np.clip((b * f - c * e) / den, 0.0, 1.0), 0.0)
        t = np.where(e > 1e-20, (b * s + f) / e, 0.0)
        low, high = t < 0.0, t > 1.0
        s = np.where(low, np.clip(-c / np.where(a > 0, a, 1.0), 0.0, 1.0), s)
        s = np.where(high, np.clip((b - c) / np.where(a > 0, a, 1.0), 0.0, 1.0), s)
    t = np.clip(t, 0.0, 1.0)
    return np.linalg.norm((a0 + da * s[:, None]) - (b0 + db * t[:, None]), axis=1)


def self_penetration(state, scene, radius, oracle):
    """max(2r - segment distance) over every non-local segment pair whose
    boxes come within 2r (:5216-5276). The pairs are the oracle's FP32 broad
    phase on the state, with the previous positions set to the current ones
    and the boxes grown by r + margin, a superset of the pairs within 2r."""
    particles = state["particles"].copy()
    particles["previousAndMass"][:, :3] = particles["positionAndInverseMass"][:, :3]
    config = scene.config.copy()
    if not config["clothMaterial"][0, 1] > 0:
        config["clothMaterial"][0, 1] = 4.0 * radius
    pairs = oracle.self_candidates_fp32(scene, particles, config)
    if len(pairs) == 0:
        return 0.0, 0
    x = _xyz(particles)
    ij = scene.distances["particlesAndColor"][:, :2].astype(np.int64)
    s, t = ij[pairs[:, 0].astype(np.int64)], ij[pairs[:, 1].astype(np.int64)]
    d = _segment_distances(x[s[:, 0]], x[s[:, 1]], x[t[:, 0]], x[t[:, 1]])
    return max(0.0, float(np.max(2.0 * radius - d))), len(pairs)


def minimum_cloth_height(state):
    return float(np.min(state["particles"]["positionAndInverseMass"][:, 2].astype(np.float64)))


def minimum_fruit_clearance(state):
    f = state["fruits"]
    return float(np.min(f["positionAndInverseMass"][:, 2].astype(np.float64) - f["previousAndRadius"][:, 3]))


def grounded_cloth_count(state, radius):
    z = state["particles"]["positionAndInverseMass"][:, 2].astype(np.float64)
    return int(np.sum(np.abs(z - radius) <= 2e-6))


def grounded_fruit_count(state):
    f = state["fruits"]
    clearance = (f["positionAndInverseMass"][:, 2] - f["previousAndRadius"][:, 3]).astype(np.float64)
    return int(np.sum(np.abs(clearance) <= 2e-6))


def escape_mask(state, grounded):
    """Fruits outside the scenario's box (:5331-5350)."""
    mask = 0
    for k, x in enumerate(_xyz(state["fruits"])):
        radial = float(np.hypot(x[0], x[1]))
        outside = (x[2] > 0.90 or x[2] < 0.0 or radial > 0.75) if grounded else \
            (x[2] > 5.0 or x[2] < -5.0 or radial > 5.0)
        if not np.all(np.isfinite(x)) or outside:
            mask |= 1 << k
    return mask


def shape_change(initial_particles, state):
    """Largest change of the distance between sampled knots (every 8th,
    :5352-5385)."""
    a = _xyz(initial_particles)[::8]
    b = _xyz(state["particles"])[::8]
    worst = 0.0
    for k in range(len(a) - 1):
        da = np.linalg.norm(a[k + 1:] - a[k], axis=1)
        db = np.linalg.norm(b[k + 1:] - b[k], axis=1)
        worst = max(worst, float(np.max(np.abs(db - da))))
    return worst


def released_and_grounded(state):
    """Released mask and how many released fruits rest on the ground
    (|z - radius| <= 2e-6), the pickup gate (:6214-6229)."""
    mask = int(state["releaseStatus"]["masks"][0, 1])
    f = state["fruits"]
    grounded = 0
    for k in range(len(f)):
        if mask & (1 << k) and abs(float(f["positionAndInverseMass"][k, 2]) - float(f["previousAndRadius"][k, 3])) \
                <= 2e-6:
            grounded += 1
    return mask, grounded
