"""Tiny hand-authored scenes with closed-form answers.

The first three restate the reference's focused probe states
(reference tools/cloth_bag_metal.mm:1276-1352, :1411-1440); the rest are
textbook cases (free fall, hanging XPBD spring, sphere push-out, grip target,
crossing-angle constraint).
"""

import numpy as np


def config(cloth, particles=0, distances=0, grips=0, knots=0, bends=0, ground=0, dt=0.01, gravity=0.0,
           radius=0.004):
    c = np.zeros(1, cloth.CONFIG)
    c["control"] = (13, particles, distances, grips)
    c["constraintCounts"] = (knots, bends, ground, 0)
    c["gravityAndTimestep"] = (0.0, 0.0, gravity, dt)
    c["gripOrientation"] = (0.0, 0.0, 0.0, 1.0)
    # yarn radius, self cell, cloth/ground friction, cloth/self friction
    # (material defaults 0.45 / 0.34, reference include/numi/cloth_material.h:28-29)
    c["clothMaterial"] = (radius, 0.0, 0.45, 0.34)
    return c


def particles(cloth, positions, inverse_masses, velocities=None):
    p = np.zeros(len(positions), cloth.PARTICLE)
    for i, (x, w) in enumerate(zip(positions, inverse_masses)):
        p[i]["positionAndInverseMass"] = (*x, w)
        p[i]["previousAndMass"] = (*x, (1.0 / w) if w > 0 else 0.0)
        if velocities is not None:
            p[i]["velocity"] = (*velocities[i], 0.0)
    return p


def distance(cloth, i, j, rest, compliance=0.0, colour=0, limit=0.285):
    d = np.zeros(1, cloth.DISTANCE)
    d["particlesAndColor"] = (i, j, colour, 0)
    d["material"] = (rest, compliance, 0.0, limit)
    return d


def batch(cloth, first, count, colour):
    b = np.zeros(1, cloth.BATCH)
    b["control"] = (first, count, colour, 0)
    return b


def strain_probe(cloth):
    """One stretched (1.5 vs rest 1, limit 0.285) and one compressed pair."""
    p = particles(cloth, [(0, 0, 0), (1.5, 0, 0), (3.0, 0, 0), (3.6, 0, 0)], [1.0, 2.0, 1.0, 2.0])
    d = np.concatenate([distance(cloth, 0, 1, 1.0), distance(cloth, 2, 3, 1.0)])
    return cloth.Scene(config(cloth, particles=4, distances=2), particles=p, distances=d,
                       distanceBatches=batch(cloth, 0, 2, 0))


def ground_bend_probe(cloth):
    """Chord bend with its first end resting on the ground (z = yarn radius)."""
    p = particles(cloth, [(0, 0, 0.004), (0.5, 0, 0.054), (1.0, 0, 0.104)], [1.0, 1.0, 1.0])
    b = np.zeros(1, cloth.BEND)
    b["particlesAndColor"] = (0, 1, 2, 0)
    b["material"] = (1.2, 1.2, 0.0, 0.0)
    return cloth.Scene(config(cloth, particles=3, bends=1, ground=1), particles=p, bends=b,
                       bendBatches=batch(cloth, 0, 1, 0))


def ground_contact_probe(cloth):
    """A knot on the ground sliding at 1 m/s and a unit fruit sunk half its
    radius, sliding at 3 m/s and rolling at 2 rad/s (reference
    tools/cloth_bag_metal.mm:1411-1440)."""
    p = particles(cloth, [(0, 0, 0)], [1.0], velocities=[(1.0, 0, 0)])
    c = bag_config(cloth, fruits=1, particles=1, ground=1)
    return cloth.Scene(c, particles=p, fruits=fruit(cloth, (0, 0, 0.5), 2.0, 1.0, velocity=(3, 0, 0),
                                                    angular=(2, 0, 0)))


def ground_contact_cloth_probe(cloth):
    """The knot of ground_contact_probe alone."""
    p = particles(cloth, [(0, 0, 0)], [1.0], velocities=[(1.0, 0, 0)])
    return cloth.Scene(config(cloth, particles=1, ground=1), particles=p)


def free_fall(cloth, dt=1.0e-3, gravity=-9.81):
    p = particles(cloth, [(0.1, -0.2, 2.0)], [4.0], velocities=[(0.5, 0.0, 1.0)])
    return cloth.Scene(config(cloth, particles=1, dt=dt, gravity=gravity), particles=p)


def hanging_spring(cloth, mass=0.01, compliance=1.0e-3, damping=40.0, dt=1.0e-3):
    """A knot hanging from a fixed knot by one compliant distance."""
    p = particles(cloth, [(0, 0, 1.0), (0, 0, 0.9)], [0.0, 1.0 / mass])
    d = distance(cloth, 0, 1, 0.1, compliance=compliance, limit=10.0)
    return cloth.Scene(config(cloth, particles=2, distances=1, dt=dt, gravity=-9.81), particles=p, distances=d,
                       distanceBatches=batch(cloth, 0, 1, 0), damping=damping)


def sphere_push(cloth, radius=0.5):
    """Two knots inside a static sphere and one outside it."""
    p = particles(cloth, [(0.1, 0.0, 0.3), (0.0, 0.0, -0.2), (0.0, 2.0, 0.0)], [1.0, 1.0, 1.0])
    return cloth.Scene(config(cloth, particles=3), particles=p,
                       staticSpheres=np.array([[0.0, 0.0, 0.0, radius]], np.float32))


def grip_probe(cloth, compliance=0.0):
    """One knot held by a grip towards handle + q * offset."""
    p = particles(cloth, [(0.3, 0.1, 0.2)], [2.0])
    g = np.zeros(1, cloth.GRIP)
    g["particle"] = (0, 1, 0, 0)
    g["targetOffsetAndCompliance"] = (0.1, 0.0, 0.0, compliance)
    c = config(cloth, particles=1, grips=1)
    c["gripTargetAndActive"] = (0.2, 0.1, 0.2, 1.0)
    # 90 degrees about z: the offset (0.1, 0, 0) maps to (0, 0.1, 0).
    c["gripOrientation"] = (0.0, 0.0, np.sin(np.pi / 4), np.cos(np.pi / 4))
    c["gripControl"] = (1, 0, 0, 0)
    c["gripMaterial"] = (0.5, 0, 0, 0)
    return cloth.Scene(c, particles=p, grips=g)


def knot_probe(cloth, rest_cosine=0.0):
    """Warp and weft chords crossing at 60 degrees, pulled to `rest_cosine`."""
    p = particles(cloth, [(-1, 0, 0), (1, 0, 0), (-0.5, -np.sqrt(3) / 2, 0), (0.5, np.sqrt(3) / 2, 0)],
                  [1.0, 1.0, 1.0, 1.0])
    k = np.zeros(1, cloth.KNOT)
    k["particles"] = (0, 1, 2, 3)
    k["material"] = (rest_cosine, 0.0, 0.0, 0.0)
    return cloth.Scene(config(cloth, particles=4, knots=1), particles=p, knots=k,
                       knotBatches=batch(cloth, 0, 1, 0))


# ------------------------------------------------------------ bag probes --
# Fruit/yarn/air/mouth probes of the reference (tools/cloth_bag_metal.mm:
# 1354-1410, :1442-1495, :1556-1703) with the material defaults of
# include/numi/cloth_material.h:30-38.

FRUIT_MATERIAL = (0.30, 0.42, 0.015, 0.36)    # pair, ground, rolling, yarn friction
AERODYNAMICS = (1.10, 0.010, 0.47, 0.010)     # crossflow, skin, fruit drag, rotational
AIR_DENSITY = 1.225


def bag_config(cloth, fruits=0, pairs=0, yarn=0, **fields):
    c = config(cloth, **fields)
    c["constraintCounts"][0, 3] = fruits
    c["contactCounts"] = (pairs, yarn, 0, 0)
    c["fruitMaterial"] = FRUIT_MATERIAL
    return c


def fruit(cloth, position, inverse_mass, radius, velocity=(0, 0, 0), angular=(0, 0, 0)):
    f = np.zeros(1, cloth.FRUIT)
    f["positionAndInverseMass"] = (*position, inverse_mass)
    f["previousAndRadius"] = (*position, radius)
    f["velocityAndGroundImpulse"] = (*velocity, 0.0)
    f["angularVelocity"] = (*angular, 0.0)
    f["orientation"] = (0.0, 0.0, 0.0, 1.0)
    return f


def fruit_pair_probe(cloth):
    """Two unit fruits 1.5 m apart (overlap 0.5 m) sliding past each other."""
    f = np.concatenate([fruit(cloth, (0, 0, 2), 1.0, 1.0, velocity=(0, 1, 0)),
                        fruit(cloth, (1.5, 0, 2), 2.0, 1.0, velocity=(0, -2, 0))])
    pair = np.zeros(1, cloth.FRUIT_PAIR)
    pair["fruitsAndColor"] = (0, 1, 0, 0)
    return cloth.Scene(bag_config(cloth, fruits=2, pairs=1), fruits=f, fruitPairs=pair,
                       fruitPairBatches=batch(cloth, 0, 1, 0))


def yarn_ccd_probe(cloth):
    """A 2 cm fruit at 20 m/s tunnels through a fixed 20 cm segment in one
    0.01 s substep unless the swept test stops it."""
    p = particles(cloth, [(0, -0.1, 1.0), (0, 0.1, 1.0)], [0.0, 0.0])
    d = distance(cloth, 0, 1, 0.2)
    y = np.zeros(1, cloth.YARN_CONTACT)
    y["identity"] = (0, 0, 0, 1)
    c = bag_config(cloth, fruits=1, yarn=1, particles=2, distances=1)
    return cloth.Scene(c, particles=p, distances=d, distanceBatches=batch(cloth, 0, 1, 0),
                       fruits=fruit(cloth, (-0.1, 0, 1.0), 1.0, 0.02, velocity=(20, 2, 0)), yarnContacts=y)


def yarn_aero_probe(cloth, refined=False, co_moving=False):
    """A 2 m yarn moving at (3, 4, 0) m/s through still (or co-moving) air,
    as one segment or two."""
    if refined:
        p = particles(cloth, [(-1, 0, 2), (0, 0, 2), (1, 0, 2)], [2.0, 1.0, 2.0],
                      velocities=[(3, 4, 0)] * 3)
        d = np.concatenate([distance(cloth, 0, 1, 1.0), distance(cloth, 1, 2, 1.0, colour=1)])
        b = np.concatenate([batch(cloth, 0, 1, 0), batch(cloth, 1, 1, 1)])
    else:
        p = particles(cloth, [(-1, 0, 2), (1, 0, 2)], [1.0, 1.0], velocities=[(3, 4, 0)] * 2)
        d = distance(cloth, 0, 1, 2.0)
        b = batch(cloth, 0, 1, 0)
    c = bag_config(cloth, particles=len(p), distances=len(d))
    c["constraintCounts"] = (0, 0, 0, 0)
    c["fruitMaterial"] = (0, 0, 0, 0)
    c["airVelocityAndDensity"] = (3.0, 4.0, 0.0, AIR_DENSITY) if co_moving else (0.0, 0.0, 0.0, AIR_DENSITY)
    c["aerodynamicCoefficients"] = AERODYNAMICS
    return cloth.Scene(c, particles=p, distances=d, distanceBatches=b)


def fruit_aero_probe(cloth):
    """A 0.21 kg, 7 cm fruit flying and spinning through still air."""
    c = bag_config(cloth, fruits=1)
    c["airVelocityAndDensity"] = (0.0, 0.0, 0.0, AIR_DENSITY)
    c["aerodynamicCoefficients"] = AERODYNAMICS
    return cloth.Scene(c, fruits=fruit(cloth, (0, 0, 10), 1.0 / 0.21, 0.07, velocity=(5, -1, 2),
                                       angular=(3, 4, 2)))


MOUTH_RIM = 48


def mouth_probe(cloth, position, rotated=False, initial_candidates=0):
    """A fixed 48-knot unit rim (at z = 1, or in the plane x = 1 when
    `rotated`) above a bottom knot at the origin, and one 10 cm fruit."""
    angles = 2.0 * np.pi * np.arange(MOUTH_RIM) / MOUTH_RIM
    if rotated:
        rim = [(1.0, np.cos(a), np.sin(a)) for a in angles]
    else:
        rim = [(np.cos(a), np.sin(a), 1.0) for a in angles]
    p = particles(cloth, rim + [(0.0, 0.0, 0.0)], [0.0] * (MOUTH_RIM + 1))
    c = bag_config(cloth, fruits=1, particles=MOUTH_RIM + 1)
    c["mouthControl"] = (MOUTH_RIM, MOUTH_RIM, 0, 0)
    c["mouthMaterial"] = (0.025, 0.0, 0.0, 0.0)
    return cloth.Scene(c, particles=p, fruits=fruit(cloth, position, 1.0, 0.10),
                       mouthRimParticles=np.arange(MOUTH_RIM, dtype=np.uint32),
                       initialCandidateMask=initial_candidates)


# (position, rotated, initial candidate mask) of the reference's seven mouth
# cases (tools/cloth_bag_metal.mm:5811-5831) and their published released
# masks (docs/METAL_CLOTH.md:404-406).
MOUTH_CASES = {
    "inside": (((0.0, 0.0, 0.50), False, 0), 0),
    "released": (((0.0, 0.0, 1.130), False, 0), 1),
    "outside": (((1.50, 0.0, 1.20), False, 0), 0),
    "grazing": (((0.0, 0.0, 1.114), False, 0), 0),
    "edge_clearance": (((1.05, 0.0, 1.130), False, 0), 1),
    "edge_exit": (((1.30, 0.0, 1.02), False, 1), 1),
    "rotated": (((1.130, 0.0, 0.0), True, 0), 1),
}


def crossing_segments_probe(cloth, mu, slip=0.2, overlap=0.001, dt=1.0 / 960.0, radius=0.004):
    """Two free 2 cm segments crossing at right angles, B (along y) lying
    `overlap` deeper than contact on A (along x) and sliding along x at
    `slip`; no gravity, no ground; self-contact with cloth/self friction mu."""
    h = 2.0 * radius - overlap
    p = particles(cloth, [(-0.01, 0, 0), (0.01, 0, 0), (0, -0.01, h), (0, 0.01, h)], [1.0, 1.0, 1.0, 1.0],
                  velocities=[(0, 0, 0), (0, 0, 0), (slip, 0, 0), (slip, 0, 0)])
    d = np.concatenate([distance(cloth, 0, 1, 0.02), distance(cloth, 2, 3, 0.02)])
    c = config(cloth, particles=4, distances=2, dt=dt, radius=radius)
    c["clothMaterial"] = (radius, 0.02, 0.45, mu)
    return cloth.Scene(c, particles=p, distances=d, distanceBatches=batch(cloth, 0, 2, 0), selfContact=1,
                       selfContactMargin=0.25 * radius)


def self_ccd_probe(cloth):
    """The reference's yarn/yarn swept-contact probe (tools/cloth_bag_metal.mm:
    1497-1554): a fixed-length 2 m segment at rest on the x axis and a second
    2 m segment along y, 8 cm above it, moving at (2, 0, -16) m/s, so that in
    one 0.01 s substep it would pass 8 cm through the first. Unit inverse
    masses, no gravity, no ground, 0 iterations and 0 strain sweeps: only the
    swept self-contact pass acts. The hash cell is the reference's
    2 + 2r + 1e-6 m."""
    radius = 0.004
    p = particles(cloth, [(-1.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, -1.0, 0.08), (0.0, 1.0, 0.08)],
                  [1.0, 1.0, 1.0, 1.0],
                  velocities=[(0, 0, 0), (0, 0, 0), (2.0, 0.0, -16.0), (2.0, 0.0, -16.0)])
    d = np.concatenate([distance(cloth, 0, 1, 0.0, limit=0.0), distance(cloth, 2, 3, 0.0, limit=0.0)])
    c = config(cloth, particles=4, distances=2, dt=0.01, radius=radius)
    c["clothMaterial"] = (radius, np.float32(2.0) + np.float32(2.0) * np.float32(radius) + np.float32(1.0e-6),
                          0.45, 0.34)
    return cloth.Scene(c, particles=p, distances=d, distanceBatches=batch(cloth, 0, 2, 0), selfContact=1)
