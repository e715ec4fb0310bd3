"""Trajectory replays on the GPU (needs a GPU): the frame loop steps the
produce bag frame after frame with the reference's grip scenarios, hashes
every frame, measures the handle lag and writes renderer OBJs
(include/numi_cloth_replay.h; reference runTrajectoryReplay,
tools/cloth_bag_metal.mm:5050-5174).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITERATIONS, SWEEPS = 32, 3


def test_pickup_replay_is_deterministic_and_hashes_the_state(cloth):
    c = cloth.cloth
    scene = cloth.Scene.bag_replay(c.REPLAY_PICKUP, selfContact=1)
    runs = []
    for _ in range(2):
        with cloth.Solver(scene) as solver:
            result = solver.replay(c.REPLAY_PICKUP, 3, ITERATIONS, SWEEPS)
            state = solver.read()
        runs.append((result, state))
    (first, state), (second, _) = runs
    print(f"pickup hashes {[hex(h) for h in first['hashes']]}, handle lag {first['maximumHandleLag']:.3e} m")
    assert first["failureFree"] and first["framesRun"] == 3
    assert first["hashes"] == second["hashes"]
    assert len(set(first["hashes"])) == 3
    # the last frame's hash is the hash of the state read back afterwards
    assert c.frame_hash(scene, state) == first["hashes"][-1]
    assert 0.0 < first["maximumHandleLag"] < 0.05


def test_replay_frames_equal_one_step_over_their_configs(cloth, oracle):
    """Two replay frames give the same bytes as one step over the 96 configs
    the frame driver generates, and agree with the FP32 twin and the FP64
    oracle run over those configs."""
    c = cloth.cloth
    scene = cloth.Scene.bag_replay(c.REPLAY_PICKUP)
    configs = np.concatenate([c.replay_frame_configs(scene.config, c.REPLAY_PICKUP, f) for f in range(2)])
    with cloth.Solver(scene) as solver:
        solver.replay(c.REPLAY_PICKUP, 2, ITERATIONS, SWEEPS)
        replayed = solver.read()
    with cloth.Solver(scene) as solver:
        solver.step(configs, ITERATIONS, SWEEPS)
        stepped = solver.read()
    assert replayed["particles"].tobytes() == stepped["particles"].tobytes()
    assert replayed["fruits"].tobytes() == stepped["fruits"].tobytes()
    gpu = replayed["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)
    twin = oracle.run(scene, configs, ITERATIONS, SWEEPS, precision="fp32")
    fp64 = oracle.run(scene, configs, ITERATIONS, SWEEPS)
    x = scene.particles["positionAndInverseMass"][:, :3]
    extent = float(np.max(x.max(axis=0) - x.min(axis=0)))
    e_twin = np.max(np.abs(gpu - twin["positions"]))
    e_fp64 = np.max(np.abs(gpu - fp64["positions"]))
    e_ref = np.max(np.abs(twin["positions"] - fp64["positions"]))
    print(f"pickup 2 frames: |gpu-twin| {e_twin:.3e} |gpu-fp64| {e_fp64:.3e} |twin-fp64| {e_ref:.3e}")
    assert e_twin <= 1e-6 * extent
    assert e_fp64 <= max(1e-5 * extent, 1.5 * e_ref)


@pytest.mark.parametrize("scenario", [0, 1])
def test_grounded_and_spin_replays(cloth, scenario):
    c = cloth.cloth
    scene = cloth.Scene.bag_replay(scenario)
    with cloth.Solver(scene) as solver:
        result = solver.replay(scenario, 2, ITERATIONS, SWEEPS)
        state = solver.read()
    assert result["failureFree"]
    z = state["particles"]["positionAndInverseMass"][:, 2]
    if scenario == c.REPLAY_SPIN:
        assert z.min() > 0.4              # airborne, no ground
        assert 0.0 < result["maximumHandleLag"] < 0.05
    else:
        assert result["maximumHandleLag"] == 0.0   # grip inactive
        assert z.min() >= 0.0035          # resting on the ground (yarn radius 4 mm, FP32 slack)


def test_recorded_replay_and_obj_dumps(cloth, golden, tmp_path):
    c = cloth.cloth
    trajectory = cloth.GripTrajectory(golden / "trajectories" / "synthetic-regrab.csv")
    scene = cloth.Scene.bag_replay(c.REPLAY_RECORDED, trajectory)
    # The fixture lasts one frame (1/120 s); sampling past its end (beyond the
    # 1e-9 s tolerance) is an error, as in the reference.
    prefix = str(tmp_path / "recorded")
    with cloth.Solver(scene) as solver:
        result = solver.replay(c.REPLAY_RECORDED, 1, ITERATIONS, SWEEPS, trajectory=trajectory, dump_every=1,
                               dump_prefix=prefix)
        state = solver.read()
    with cloth.Solver(scene) as solver, pytest.raises(cloth.ClothError):
        solver.replay(c.REPLAY_RECORDED, 2, ITERATIONS, SWEEPS, trajectory=trajectory)
    assert result["failureFree"] and result["framesRun"] == 1
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["recorded-0.obj", "recorded-1.obj"]
    vertices = np.array([[float(v) for v in line.split()[1:]]
                         for line in open(tmp_path / "recorded-1.obj") if line.startswith("v ")], np.float32)
    assert np.array_equal(vertices, state["particles"]["positionAndInverseMass"][:, :3])
