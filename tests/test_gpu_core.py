"""HIP path against the FP64 oracle on the same inputs (needs a GPU).

Tolerances: probes are O(1 m) scenes run for one substep, so FP32 rounding
bounds the difference at a few ulp (1e-6 m absolute). Sheets use the gate the
parity contract states: positions within 1e-5 of the cloth extent after one
frame. Determinism and split-versus-persistent execution are bit-exact.
"""

import numpy as np
import pytest

import probes

pytestmark = pytest.mark.gpu


def _positions(state):
    return state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)


def _run(cloth, scene, substeps, iterations, sweeps):
    with cloth.Solver(scene) as solver:
        solver.step(scene.configs(substeps), iterations, sweeps)
        return solver.read()


@pytest.mark.parametrize("name, iterations, sweeps", [
    ("strain_probe", 0, 1), ("ground_bend_probe", 1, 0), ("ground_contact_probe", 1, 0),
    ("sphere_push", 1, 0), ("grip_probe", 1, 0), ("knot_probe", 60, 0),
])
def test_probe_parity(cloth, oracle, name, iterations, sweeps):
    scene = getattr(probes, name)(cloth)
    state = _run(cloth, scene, 1, iterations, sweeps)
    reference = oracle.run(scene, scene.configs(1), iterations, sweeps)
    assert state["failure"] == 0
    print(f"{name}: max |dx| = {np.max(np.abs(_positions(state) - reference['positions'])):.3e}")
    assert np.max(np.abs(_positions(state) - reference["positions"])) < 1e-6
    velocity = state["particles"]["velocity"][:, :3].astype(np.float64)
    # Velocities are position differences over dt = 0.01: 1e-6 m -> 1e-4 m/s.
    assert np.max(np.abs(velocity - reference["velocities"])) < 1e-4


def test_ground_contact_friction_matches_published_probe(cloth):
    state = _run(cloth, probes.ground_contact_probe(cloth), 1, 1, 0)
    p = state["particles"][0]
    # Published FP32 values (reference docs/METAL_CLOTH.md:372-375):
    # cloth_height=0.004000000190, cloth_slip_after=0.819999992847.
    assert p["positionAndInverseMass"][2] == np.float32(0.004)
    assert p["velocity"][0] == pytest.approx(0.819999992847, abs=2e-7)
    counters = state["frictionStatus"]["counters"][0]
    assert counters[2] == 1


def test_strain_probe_final_violation(cloth):
    state = _run(cloth, probes.strain_probe(cloth), 1, 0, 1)
    x = _positions(state)
    limit = 1.0 * (1.0 + float(np.float32(0.285)))
    # Published strain_probe_final_violation=0.000000007451 (FP32).
    assert np.linalg.norm(x[1] - x[0]) - limit < 2e-7
    assert np.linalg.norm(x[3] - x[2]) == pytest.approx(0.6, abs=1e-7)


def test_hanging_spring_parity(cloth, oracle):
    scene = probes.hanging_spring(cloth)
    state = _run(cloth, scene, 600, 1, 0)
    reference = oracle.run(scene, scene.configs(600), 1, 0)
    assert np.max(np.abs(_positions(state) - reference["positions"])) < 1e-6


def _sheet(cloth, **overrides):
    fields = dict(width=48, height=40, elevation=0.12, jitter=1.0e-4, seed=11, sphereMode=1,
                  fixedSphere=(0.05, -0.03, 0.0, 0.1))
    fields.update(overrides)
    return cloth.Scene.sheet(**fields)


def _check_frame(oracle, scene, state, configs, iterations, sweeps, label):
    """GPU vs the FP32 twin (same schedule, same FP32 operation order) and vs
    FP64. The FP64 gate is 1e-5 of the cloth extent; a stiff pinned sheet
    amplifies FP32 rounding beyond it, and then the GPU must be as close to
    FP64 as the exact FP32 restatement is."""
    gpu = _positions(state)
    twin = oracle.run(scene, configs, iterations, sweeps, precision="fp32")
    fp64 = oracle.run(scene, configs, iterations, sweeps)
    x = scene.particles["positionAndInverseMass"][:, :3]
    extent = float(np.max(x.max(axis=0) - x.min(axis=0)))
    e_twin = np.max(np.abs(gpu - twin["positions"]))
    e_fp64 = np.max(np.abs(gpu - fp64["positions"]))
    e_twin_fp64 = np.max(np.abs(twin["positions"] - fp64["positions"]))
    print(f"{label}: |gpu-twin| {e_twin:.3e}  |gpu-fp64| {e_fp64:.3e}  |twin-fp64| {e_twin_fp64:.3e}  "
          f"gate {1e-5 * extent:.3e}")
    assert e_twin <= 1e-6 * extent
    assert e_fp64 <= max(1e-5 * extent, 1.5 * e_twin_fp64)
    return fp64, e_fp64, extent


@pytest.mark.parametrize("overrides", [
    {},
    {"crossingKnots": 1},
    {"pinTopRow": 1, "ground": 0, "sphereMode": 0},
])
def test_sheet_frame_parity(cloth, oracle, overrides):
    scene = _sheet(cloth, **overrides)
    substeps, iterations, sweeps = 32, 16, 3
    configs = scene.configs(substeps)
    state = _run(cloth, scene, substeps, iterations, sweeps)
    assert state["failure"] == 0
    reference, error, extent = _check_frame(oracle, scene, state, configs, iterations, sweeps, f"sheet {overrides}")
    if overrides.get("ground", 1):
        # Draped on ground and sphere (the benchmark's class): the plain gate.
        assert error <= 1.0e-5 * extent
    lam = state["distances"]["material"][:, 2].astype(np.float64)
    # The reference's multiplier gate for its GPU-vs-FP64 check is 2e-8
    # (reference tools/cloth_bag_metal.mm:7656-7661).
    assert np.max(np.abs(lam - reference["distanceLambda"])) <= 2e-8
    bend = state["bends"]["material"][:, 3].astype(np.float64)
    assert np.max(np.abs(bend - reference["bendLambda"])) <= 2e-8


def test_runs_are_bit_identical(cloth):
    scene = _sheet(cloth)
    first = _run(cloth, scene, 16, 16, 3)
    second = _run(cloth, scene, 16, 16, 3)
    assert cloth.state_hash(first) == cloth.state_hash(second)
    assert first["particles"].tobytes() == second["particles"].tobytes()


def test_split_steps_equal_one_step(cloth):
    scene = _sheet(cloth)
    configs = scene.configs(24)
    with cloth.Solver(scene) as whole:
        whole.step(configs, 8, 2)
        a = whole.read()
    with cloth.Solver(scene) as split:
        split.step(configs[:10], 8, 2)
        split.step(configs[10:13], 8, 2)
        split.step(configs[13:], 8, 2)
        b = split.read()
    assert a["particles"].tobytes() == b["particles"].tobytes()
    assert a["distances"].tobytes() == b["distances"].tobytes()


def test_kinematic_drive_parity(cloth, oracle):
    scene = cloth.Scene.sheet(width=16, height=24, elevation=0.5, pinTopRow=1, ground=0, seed=5)
    scene.driveAmplitude, scene.driveOmega, scene.drivePhase = 0.02, 2 * np.pi * 1.5, 0.3
    scene.damping = 0.5
    scene.refresh()
    state = _run(cloth, scene, 48, 8, 2)
    _check_frame(oracle, scene, state, scene.configs(48), 8, 2, "driven strip")


def test_diagnostics_equal_host_reduction(cloth):
    scene = _sheet(cloth)
    with cloth.Solver(scene) as solver:
        solver.step(scene.configs(8), 4, 0)
        state = solver.read()
        diag = solver.diagnostics()
    x = state["particles"]["positionAndInverseMass"]
    d = state["distances"]
    i, j = d["particlesAndColor"][:, 0], d["particlesAndColor"][:, 1]
    delta = x[j, :3] - x[i, :3]
    length = np.sqrt((delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1]) + delta[:, 2] * delta[:, 2])
    limit = d["material"][:, 0] * (np.float32(1.0) + d["material"][:, 3])
    expected = max(np.float32(0.0), np.max(length - limit))
    # Same FP32 operations in the same order: the reduction is exact.
    assert np.float32(diag.maximumStrainViolation) == expected
    assert np.float32(diag.maximumGroundPenetration) == max(np.float32(0.0), np.max(np.float32(0.004) - x[:, 2]))


def test_nonfinite_config_sets_failure_and_skips_substep(cloth):
    scene = _sheet(cloth)
    configs = scene.configs(2)
    configs["gravityAndTimestep"][1, 2] = np.nan
    with cloth.Solver(scene) as solver:
        solver.step(configs[:1], 4, 0)
        before = solver.read()
        solver.step(configs[1:], 4, 0)
        after = solver.read()
    assert before["failure"] == 0
    assert after["failure"] & cloth.cloth.FAILURE_NONFINITE
    assert before["particles"].tobytes() == after["particles"].tobytes()


def test_grip_capture_beyond_radius_is_flagged(cloth):
    scene = probes.grip_probe(cloth)
    configs = scene.configs(2)
    # Re-grab (generation 2) with the handle 1 m away from the knot
    # (reference tools/cloth_bag_metal.mm:5958-5969 checks the same rejection).
    configs["gripControl"][1, 0] = 2
    configs["gripTargetAndActive"][1, :3] = (1.3, 0.1, 0.2)
    with cloth.Solver(scene) as solver:
        solver.step(configs, 1, 0)
        state = solver.read()
    assert state["failure"] & cloth.cloth.FAILURE_GRIP_CAPTURE


def test_topology_change_is_rejected(cloth):
    scene = _sheet(cloth)
    configs = scene.configs(1)
    configs["control"][0, 1] += 1
    with cloth.Solver(scene) as solver:
        with pytest.raises(cloth.ClothError) as error:
            solver.step(configs, 1, 0)
    assert error.value.status == 2
