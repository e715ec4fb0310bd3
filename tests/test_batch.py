"""Strip specimens and specimen merging (CPU only).

A merged scene must step every specimen exactly as that specimen alone:
the oracle run of the merged scene is compared with the oracle runs of the
specimens one by one, bit for bit in both precisions.
"""

import numpy as np
import pytest

from test_scene import check_colouring


def _strips(cloth, count=5, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        out.append(cloth.Scene.strip(width=int(rng.integers(3, 9)), height=int(rng.integers(4, 12)), seed=k + 1,
                                     stretchCompliance=float(10 ** rng.uniform(-8, -5)),
                                     bendCompliance=float(10 ** rng.uniform(-8, -5)),
                                     arealDensity=float(rng.uniform(0.1, 0.4)),
                                     driveFrequency=float(rng.uniform(0.5, 2.0)), damping=float(k % 2),
                                     origin=(0.2 * k, 0.0, 1.0)))
    return out


def test_strip_topology(cloth):
    s = cloth.Scene.strip(width=6, height=10)
    W, H = 6, 10
    assert len(s.particles) == W * H
    assert len(s.distances) == (W - 1) * H + W * (H - 1)
    assert len(s.bends) == (W - 2) * H + W * (H - 2)
    w = s.particles["positionAndInverseMass"][:, 3]
    assert np.all(w[:W] == 0) and np.all(w[W:] > 0)
    x = s.particles["positionAndInverseMass"][:, :3]
    # Row j hangs j spacings below the pinned top row.
    assert np.allclose(x[W * 3 + 2, [0, 2]], [0.02, 1.0 - 0.03], atol=1e-6)
    check_colouring(s)
    mass = s.particles["previousAndMass"][W:, 3]
    assert np.allclose(mass, 0.2 * 0.01 * 0.01, rtol=1e-6)


def test_merge_layout(cloth):
    specimens = _strips(cloth)
    merged, maps = cloth.Scene.merge(specimens)
    assert len(merged.particles) == sum(len(s.particles) for s in specimens)
    assert len(merged.distances) == sum(len(s.distances) for s in specimens)
    check_colouring(merged)
    offset = 0
    for s, m in zip(specimens, maps):
        n = len(s.particles)
        assert np.array_equal(m["particles"], np.arange(offset, offset + n))
        d = merged.distances[m["distances"]]
        assert np.array_equal(d["particlesAndColor"][:, :2] - offset, s.distances["particlesAndColor"][:, :2])
        assert np.array_equal(d["material"], s.distances["material"])
        b = merged.bends[m["bends"]]
        assert np.array_equal(b["particlesAndColor"][:, :3] - offset, s.bends["particlesAndColor"][:, :3])
        motion = merged.particleMotion[offset:offset + n]
        assert np.all(motion[:, 0] == np.float32(0.01))
        assert np.all(motion[:, 3] == np.float32(s.damping))
        offset += n
    # One batch per colour, colours ascending.
    colours = merged.distanceBatches["control"][:, 2]
    assert np.array_equal(colours, np.arange(len(colours)))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_merged_scene_steps_each_specimen_as_alone(cloth, oracle, precision):
    specimens = _strips(cloth)
    merged, maps = cloth.Scene.merge(specimens)
    substeps, iterations, sweeps = 24, 6, 2
    together = oracle.run(merged, merged.configs(substeps), iterations, sweeps, precision=precision)
    for s, m in zip(specimens, maps):
        alone = oracle.run(s, s.configs(substeps), iterations, sweeps, precision=precision)
        assert np.array_equal(together["positions"][m["particles"]], alone["positions"])
        assert np.array_equal(together["velocities"][m["particles"]], alone["velocities"])
        assert np.array_equal(together["distanceLambda"][m["distances"]], alone["distanceLambda"])
        assert np.array_equal(together["bendLambda"][m["bends"]], alone["bendLambda"])


def test_driven_top_row_follows_sinusoid(cloth, oracle):
    s = cloth.Scene.strip(width=4, height=5, driveAmplitude=0.02, driveFrequency=1.25, drivePhase=0.3)
    dt = float(np.float32(1.0 / 2880.0))
    n = 37
    out = oracle.run(s, s.configs(n), 4, 1)
    t = n * dt
    omega, phase = float(np.float32(2 * np.pi * 1.25)), float(np.float32(0.3))
    # The drive starts where the row was created: A (sin(omega t + phase) - sin(phase)).
    expected = 1.0 + float(np.float32(0.02 * (np.sin(omega * t + phase) - np.sin(phase))))
    assert out["positions"][:4, 2] == pytest.approx([expected] * 4, abs=1e-7)
    start = oracle.run(s, s.configs(1), 4, 1)
    assert start["positions"][:4, 2] == pytest.approx([1.0] * 4, abs=2e-4)


def test_merge_rejects_coupled_specimens(cloth):
    import probes
    with pytest.raises(cloth.ClothError, match="unsupported"):
        cloth.Scene.merge([cloth.Scene.strip(width=3, height=4), probes.fruit_pair_probe(cloth)])
    with pytest.raises(cloth.ClothError, match="share timestep"):
        cloth.Scene.merge([cloth.Scene.strip(width=3, height=4), cloth.Scene.strip(width=3, height=4, timestep=1e-3)])
