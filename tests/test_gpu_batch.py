"""Batched specimens on the HIP path (needs a GPU).

A batch of strips of uneven sizes, stepped together, must leave every
specimen in exactly the state the plain one-specimen solver reaches (bit for
bit), and the merged scene must meet the frame gate against the oracle. The
persistent launch (Batch.step_resident: one workgroup per specimen, state
resident in LDS) must end where the merged tile path ends, bit for bit.
"""

import numpy as np
import pytest

from test_batch import _strips
from test_gpu_core import _check_frame, _positions

pytestmark = pytest.mark.gpu


def _alone(cloth, scene, substeps, iterations, sweeps):
    with cloth.Solver(scene) as solver:
        solver.step(scene.configs(substeps), iterations, sweeps)
        return solver.read()


def test_batch_equals_each_specimen_alone(cloth):
    specimens = _strips(cloth, count=9, seed=3)
    substeps, iterations, sweeps = 48, 16, 3
    with cloth.Batch(specimens) as batch:
        batch.step(specimens[0].configs(substeps), iterations, sweeps)
        states = [batch.read(k) for k in range(len(specimens))]
    for k, (scene, state) in enumerate(zip(specimens, states)):
        alone = _alone(cloth, scene, substeps, iterations, sweeps)
        assert state["failure"] == 0 and alone["failure"] == 0
        for name in ("particles", "distances", "bends"):
            assert state[name].tobytes() == alone[name].tobytes(), f"specimen {k} {name}"


def test_batch_split_steps_equal_one_step(cloth):
    specimens = _strips(cloth, count=4, seed=5)
    configs = specimens[0].configs(30)
    with cloth.Batch(specimens) as whole:
        whole.step(configs, 8, 2)
        a = [whole.read(k)["particles"] for k in range(4)]
    with cloth.Batch(specimens) as split:
        split.step(configs[:11], 8, 2)
        split.step(configs[11:], 8, 2)
        b = [split.read(k)["particles"] for k in range(4)]
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()


def test_batch_frame_parity_with_oracle(cloth, oracle):
    specimens = _strips(cloth, count=6, seed=11)
    merged, maps = cloth.Scene.merge(specimens)
    substeps, iterations, sweeps = 48, 16, 3
    with cloth.Batch(specimens) as batch:
        batch.step(merged.configs(substeps), iterations, sweeps)
        parts = [batch.read(k)["particles"] for k in range(len(specimens))]
    state = {"particles": np.concatenate(parts)}
    _check_frame(oracle, merged, state, merged.configs(substeps), iterations, sweeps, "strip batch")


def test_large_batch_uses_tile_programs(cloth):
    specimens = _strips(cloth, count=64, seed=13)
    with cloth.Batch(specimens) as batch:
        info = batch.tile_info()
        batch.step(specimens[0].configs(8), 8, 3)
        assert batch.read(0)["failure"] == 0
    print(f"batch tiles: {info.tiles} owned {info.ownedPerTile} workgroup {info.workgroupSize}")
    assert info.enabled == 1 and info.tiles >= 1


def _calibration_strips(cloth, count, seed):
    """Strips drawn from the batched calibration distribution (width 8-64,
    height 64-256 knots, log-uniform compliances, driven top row), kept to
    those whose state fits one workgroup's LDS (16 B per knot, 4 B per
    distance and bend, at most 5120 knots)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        width, height = int(rng.integers(8, 65)), int(rng.integers(64, 257))
        scene = cloth.Scene.strip(width=width, height=height, spacing=0.01, arealDensity=float(rng.uniform(0.1, 0.4)),
                                  stretchCompliance=float(10.0 ** rng.uniform(-8, -5)),
                                  bendCompliance=float(10.0 ** rng.uniform(-8, -5)), strainLimit=0.10,
                                  timestep=1.0 / 2880.0, driveAmplitude=0.01,
                                  driveFrequency=float(rng.uniform(0.5, 2.0)),
                                  drivePhase=float(rng.uniform(0.0, 2.0 * np.pi)),
                                  origin=(-0.005 * (width - 1), 0.0, 0.005 * (height - 1)), jitter=1e-4,
                                  seed=seed + len(out))
        lds = 16 * len(scene.particles) + 4 * (len(scene.distances) + len(scene.bends)) + 1024
        if len(scene.particles) <= 5120 and lds <= 160 * 1024:
            out.append(scene)
    return out


def test_resident_step_equals_batch_step(cloth):
    specimens = _calibration_strips(cloth, 12, seed=31)
    configs = specimens[0].configs(48)
    iterations, sweeps = 32, 3
    with cloth.Batch(specimens) as merged:
        merged.step(configs, iterations, sweeps)
        merged.step(configs, iterations, sweeps)
        expected = [merged.read(k) for k in range(len(specimens))]
    with cloth.Batch(specimens) as resident:
        resident.step_resident(configs, iterations, sweeps)
        resident.step_resident(configs[:17], iterations, sweeps)   # a call may end mid-frame
        resident.step_resident(configs[17:], iterations, sweeps)
        got = [resident.read(k) for k in range(len(specimens))]
    for k, (a, b) in enumerate(zip(expected, got)):
        assert a["failure"] == 0 and b["failure"] == 0
        for name in ("particles", "distances", "bends"):
            assert a[name].tobytes() == b[name].tobytes(), f"specimen {k} {name}"


def test_resident_and_merged_steps_interleave(cloth):
    specimens = _calibration_strips(cloth, 5, seed=37)
    configs = specimens[0].configs(24)
    with cloth.Batch(specimens) as merged:
        for _ in range(3):
            merged.step(configs, 16, 3)
        expected = [merged.read(k)["particles"] for k in range(len(specimens))]
    with cloth.Batch(specimens) as mixed:
        mixed.step(configs, 16, 3)
        mixed.step_resident(configs, 16, 3)
        mixed.step(configs, 16, 3)
        got = [mixed.read(k)["particles"] for k in range(len(specimens))]
    for a, b in zip(expected, got):
        assert a.tobytes() == b.tobytes()


def test_resident_step_refuses_specimens_beyond_lds(cloth):
    big = cloth.Scene.strip(width=64, height=100)   # 6400 knots
    with cloth.Batch([cloth.Scene.strip(width=8, height=64), big]) as batch:
        with pytest.raises(cloth.ClothError, match="5120"):
            batch.step_resident(big.configs(4), 8, 3)
        batch.step(big.configs(4), 8, 3)              # the merged path still runs
        assert batch.read(1)["failure"] == 0
