"""Shared trajectory-loss problem for the adjoint tests: two hanging strips
of different size and material in one merged scene, tracked knots near the
free end, and a seeded target trajectory (each tracked knot's start
position plus a seeded sinusoidal offset)."""

import numpy as np

DT = 1.0 / 960.0


def build(cloth, substeps=200, seed=5):
    a = cloth.Scene.strip(width=6, height=10, seed=3, driveFrequency=2.0, damping=2.0, stretchCompliance=3e-5,
                          bendCompliance=1e-4, driveAmplitude=0.05, timestep=DT)
    b = cloth.Scene.strip(width=5, height=8, seed=4, driveFrequency=1.7, damping=1.5, origin=(0.5, 0.0, 1.0),
                          stretchCompliance=2e-5, bendCompliance=5e-5, driveAmplitude=0.05, timestep=DT)
    specimens = [a, b]
    merged, maps = cloth.Scene.merge(specimens)
    local = np.array([[20, 35, 59], [15, 33, 39]], np.uint32)
    tracked = np.concatenate([maps[k]["particles"][local[k]] for k in range(2)]).astype(np.uint32)
    rng = np.random.default_rng(seed)
    x0 = merged.particles["positionAndInverseMass"][tracked, :3].astype(np.float64)
    amplitude = rng.uniform(-0.02, 0.02, (len(tracked), 3))
    frequency = rng.uniform(0.5, 2.0, (len(tracked), 1))
    t = (np.arange(1, substeps + 1) * np.float32(DT))[:, None, None]
    targets = (x0[None] + amplitude[None] * np.sin(2 * np.pi * frequency[None] * t)).astype(np.float32)
    return dict(specimens=specimens, merged=merged, maps=maps, local=local, tracked=tracked,
                targets=targets.astype(np.float64), configs=merged.configs(substeps), iterations=8, sweeps=2)


def build_strained(cloth, substeps=48, iterations=1, seed=3):
    """A problem in which the strain limits act: a soft strip pulled 50 %
    beyond rest (limit 10 %) stepped with one solver iteration per substep,
    so the distance constraints cannot restore it and the strain passes do,
    beside an ordinary strip."""
    a = cloth.Scene.strip(width=4, height=10, driveAmplitude=0.03, driveFrequency=1.5, jitter=1e-4,
                          stretchCompliance=1e-5, bendCompliance=1e-5, damping=1.5, timestep=DT)
    x = a.particles["positionAndInverseMass"]
    top = x[0, 2]
    x[:, 2] = top + (x[:, 2] - top) * np.float32(1.5)
    a.particles["previousAndMass"][:, :3] = x[:, :3]
    a.refresh()
    b = cloth.Scene.strip(width=5, height=8, seed=4, stretchCompliance=2e-5, driveAmplitude=0.05, driveFrequency=2.0,
                          damping=1.0, origin=(0.5, 0.0, 1.0), timestep=DT)
    specimens = [a, b]
    merged, maps = cloth.Scene.merge(specimens)
    local = np.array([[13, 22, 37], [20, 30, 39]], np.uint32)
    tracked = np.concatenate([maps[k]["particles"][local[k]] for k in range(2)]).astype(np.uint32)
    rng = np.random.default_rng(seed)
    x0 = merged.particles["positionAndInverseMass"][tracked, :3].astype(np.float64)
    targets = (x0[None] + rng.uniform(-0.02, 0.02, (substeps, len(tracked), 3))).astype(np.float32)
    return dict(specimens=specimens, merged=merged, maps=maps, local=local, tracked=tracked,
                targets=targets.astype(np.float64), configs=merged.configs(substeps), iterations=iterations, sweeps=2)
