"""The pinned probes on the GPU (needs a GPU): the reference's self-CCD probe,
a present yarn/yarn contact and the nearest-cuff-patch re-grab, each
stepped through the C ABI and compared with the FP32 twin (bytes), the
reference's published values and the FP64 oracle. The CPU pins of the same
functions are in test_oracle_self_ccd.py and test_oracle_regrab.py.
"""

import numpy as np
import pytest

import probes
from test_oracle_regrab import GRIP_CAPTURE, _check_patch, _replay
from test_oracle_self_ccd import RADIUS, _crossing, _probe_metrics

pytestmark = pytest.mark.gpu


def _run(cloth, scene, configs, iterations, sweeps):
    with cloth.Solver(scene) as solver:
        solver.step(configs, iterations, sweeps)
        return solver.read()


def _f64(state, field="positionAndInverseMass"):
    return state["particles"][field][:, :3].astype(np.float64)


def test_self_ccd_probe_on_the_gpu(cloth, oracle):
    scene = probes.self_ccd_probe(cloth)
    state = _run(cloth, scene, scene.configs(1), 0, 0)
    twin = oracle.run(scene, scene.configs(1), 0, 0, precision="fp32")
    fp64 = oracle.run(scene, scene.configs(1), 0, 0)
    assert state["failure"] == 0
    assert np.array_equal(_f64(state), twin["positions"])
    assert np.array_equal(_f64(state, "velocity"), twin["velocities"])
    counters = state["selfStatus"]["counters"][0]
    friction = state["frictionStatus"]["metrics"][0]
    # Published (docs/METAL_CLOTH.md:388-395): present 0, swept 1,
    # max_correction=0.087999925017, cone_ratio=0.668487787247, one
    # friction contact (the gate frictionStatus.metrics.w == oracle count).
    assert (int(counters[0]), int(counters[1])) == (0, 1)
    assert float(np.uint32(counters[2]).view(np.float32)) == pytest.approx(0.087999925017, abs=1e-12)
    assert float(np.uint32(friction[0]).view(np.float32)) == pytest.approx(0.668487787247, abs=1e-12)
    assert int(friction[3]) == 1
    gpu = {"positions": _f64(state), "velocities": _f64(state, "velocity")}
    m = _probe_metrics(scene, gpu, fp64["selfImpulses"])
    assert m["final_moving_height"] == pytest.approx(-0.036001820117, abs=1e-12)
    assert m["final_separation"] == pytest.approx(0.000004277196, abs=1e-12)
    assert m["slip_after"] == pytest.approx(0.000050152179, abs=1e-12)
    assert m["energy_after"] == pytest.approx(155.924161600600, abs=1e-9)
    # The reference's gates (tools/cloth_bag_metal.mm:7761-7775).
    assert np.max(np.abs(gpu["positions"] - fp64["positions"])) <= 2e-6
    assert np.max(np.abs(gpu["velocities"] - fp64["velocities"])) <= 1e-3
    assert m["momentum_error"] <= 1e-5


@pytest.mark.parametrize("inverse_masses,offset_x", [([1.0, 1.0, 1.0, 1.0], 0.0), ([1.0, 3.0, 2.0, 0.5], 0.004)])
def test_present_contact_on_the_gpu(cloth, oracle, inverse_masses, offset_x):
    scene = _crossing(cloth, inverse_masses, offset_x)
    state = _run(cloth, scene, scene.configs(1), 1, 0)
    twin = oracle.run(scene, scene.configs(1), 1, 0, precision="fp32")
    fp64 = oracle.run(scene, scene.configs(1), 1, 0)
    assert np.array_equal(_f64(state), twin["positions"])
    x = _f64(state)
    x0 = scene.particles["positionAndInverseMass"][:, :3].astype(np.float64)
    wa = (x0[2, 0] - x0[0, 0]) / (x0[1, 0] - x0[0, 0])
    gap = np.linalg.norm(0.5 * (x[2] + x[3]) - ((1 - wa) * x[0] + wa * x[1]))
    assert gap == pytest.approx(2 * RADIUS, abs=2e-9)
    assert np.max(np.abs(x - fp64["positions"])) <= 2e-9
    assert int(state["selfStatus"]["counters"][0, 0]) == 1


def test_patch_regrab_on_the_gpu(cloth, oracle, golden):
    scene, configs = _replay(cloth, golden, "synthetic-patch-regrab.csv")
    first_regrab = int(np.argmax(configs["gripControl"][:, 0] == 2)) + 1
    with cloth.Solver(scene) as solver:
        solver.step(configs[:first_regrab], 32, 3)
        mid = solver.read()
        solver.step(configs[first_regrab:], 32, 3)
        end = solver.read()
    assert mid["failure"] == 0 and end["failure"] == 0
    for state, substeps, centre, generation in [(mid, first_regrab, 12, 2), (end, len(configs), 24, 3)]:
        records = state["grips"]["particle"]
        _check_patch(scene, records, centre, generation)
        twin = oracle.run(scene, configs[:substeps], 32, 3, precision="fp32")
        assert np.array_equal(records, twin["gripRecords"])
    assert float(np.max(end["grips"]["lambda"][:, 3])) <= 2e-6
    # Positions: the GPU is the FP32 twin's arithmetic; the stiff bag
    # (axial compliance 1e-8) amplifies FP32 rounding over a 48-substep frame
    # to ~1.6e-4 m between the exact FP32 restatement and FP64, so the frame
    # gate is the twin's own distance from FP64 (one substep of the bag meets
    # the reference's 2e-5 m gate: test_gpu_bag.py).
    twin = oracle.run(scene, configs, 32, 3, precision="fp32")
    fp64 = oracle.run(scene, configs, 32, 3)
    x = scene.particles["positionAndInverseMass"][:, :3]
    extent = float(np.max(x.max(axis=0) - x.min(axis=0)))
    e_twin = float(np.max(np.abs(_f64(end) - twin["positions"])))
    e_fp64 = float(np.max(np.abs(_f64(end) - fp64["positions"])))
    e_ref = float(np.max(np.abs(twin["positions"] - fp64["positions"])))
    print(f"patch re-grab frame: |gpu-twin| {e_twin:.3e} |gpu-fp64| {e_fp64:.3e} |twin-fp64| {e_ref:.3e} m")
    assert e_twin <= 1e-6 * extent
    assert e_fp64 <= max(1e-5 * extent, 1.5 * e_ref)


def test_patch_regrab_replay_matches_the_reference_metrics(cloth, golden):
    """The frame driver's recorded replay (the reference's
    --grip-trajectory ... --recorded-steps 1 run): two replays with equal
    hashes, generation 3 on every grip, patch centre 24, exact topology."""
    c = cloth.cloth
    trajectory = cloth.GripTrajectory(golden / "trajectories" / "synthetic-patch-regrab.csv")
    scene = cloth.Scene.bag_replay(c.REPLAY_RECORDED, trajectory)
    runs = []
    for _ in range(2):
        with cloth.Solver(scene) as solver:
            result = solver.replay(c.REPLAY_RECORDED, 1, 32, 3, trajectory=trajectory)
            runs.append((result, solver.read()))
    (first, state), (second, _) = runs
    assert first["failureFree"] and first["hashes"] == second["hashes"]
    _check_patch(scene, state["grips"]["particle"], 24, 3)


def test_distant_patch_regrab_is_refused_on_the_gpu(cloth, golden):
    scene, configs = _replay(cloth, golden, "invalid-distant-patch-regrab.csv")
    state = _run(cloth, scene, configs, 32, 3)
    assert state["failure"] & GRIP_CAPTURE
