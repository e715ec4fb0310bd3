"""Batched specimens on the HIP path (needs a GPU).

A batch of strips of uneven sizes, stepped together, must leave every
specimen in exactly the state the plain one-specimen solver reaches (bit for
bit), and the merged scene must meet the frame gate against the oracle.
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
