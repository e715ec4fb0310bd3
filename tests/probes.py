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
    """A knot on the ground sliding at 1 m/s (cloth half of the reference's probe)."""
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
