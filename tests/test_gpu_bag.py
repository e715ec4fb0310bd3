"""Produce-bag families on the HIP path (needs a GPU): fruit pairs, fruit
ground contact and friction, fruit/yarn swept contact, air loads, mouth
release, and one frame of the 1,465-knot bag with 12 fruits.

Each probe is compared with the FP32 twin of the oracle (same schedule and
FP32 operation order; 1e-6 of the probe's scale) and with the FP32 values
the reference publishes for the same probe (reference
docs/METAL_CLOTH.md:366-406). The bag frame uses the sheet frame gate.
"""

import numpy as np
import pytest

import probes
from test_gpu_core import _check_frame, _positions

pytestmark = pytest.mark.gpu


def _fruits(state):
    f = state["fruits"]
    return np.concatenate([f[name].astype(np.float64) for name in
                           ("positionAndInverseMass", "previousAndRadius", "velocityAndGroundImpulse",
                            "angularVelocity", "orientation")], axis=1)


def _yarn(state):
    y = state["yarnContacts"]
    return np.concatenate([y[name].astype(np.float64) for name in
                           ("currentNormalAndSeparation", "sweptNormalAndAdvance", "weightsAndTime", "control",
                            "fruitNormalAndImpulse", "segmentImpulse")], axis=1)


def _bits(word):
    return float(np.array([word], np.uint32).view(np.float32)[0])


def _run(cloth, scene, substeps, iterations, sweeps):
    with cloth.Solver(scene) as solver:
        solver.step(scene.configs(substeps), iterations, sweeps)
        return solver.read()


def _compare(label, gpu, twin, tolerance):
    error = float(np.max(np.abs(gpu - twin))) if gpu.size else 0.0
    print(f"{label}: max |gpu - twin| = {error:.3e}")
    assert error <= tolerance


def test_fruit_pair_probe(cloth, oracle):
    scene = probes.fruit_pair_probe(cloth)
    state = _run(cloth, scene, 1, 1, 0)
    twin = oracle.run(scene, scene.configs(1), 1, 0, precision="fp32")
    assert state["failure"] == 0
    fruits = _fruits(state)
    _compare("fruit pair fruits", fruits, twin["fruits"], 4e-6)
    _compare("fruit pair contact", state["fruitPairs"]["contact"].astype(np.float64), twin["fruitPairs"], 4e-6)
    separation = np.linalg.norm(fruits[1, :3] - fruits[0, :3])
    assert separation == pytest.approx(1.999999962672, abs=1e-7)
    assert state["fruitPairs"]["contact"][0, 3] == pytest.approx(16.656669616699, abs=4e-6)
    counters = state["frictionStatus"]["counters"][0]
    assert counters[0] == 1
    assert _bits(state["frictionStatus"]["metrics"][0, 0]) == pytest.approx(0.057165719569, abs=1e-6)


def test_fruit_ground_probe(cloth, oracle):
    scene = probes.ground_contact_probe(cloth)
    state = _run(cloth, scene, 1, 1, 0)
    twin = oracle.run(scene, scene.configs(1), 1, 0, precision="fp32")
    assert state["failure"] == 0
    fruits = _fruits(state)
    _compare("fruit ground", fruits, twin["fruits"], 1e-6)
    assert fruits[0, 2] == 1.0
    assert fruits[0, 11] == pytest.approx(25.0, abs=1e-5)
    counters = state["frictionStatus"]["counters"][0]
    metrics = state["frictionStatus"]["metrics"][0]
    assert tuple(counters[2:4]) == (1, 1) and metrics[2] == 1
    assert _bits(metrics[0]) == pytest.approx(1.0, abs=1e-6)
    assert _bits(metrics[1]) == pytest.approx(1.0, abs=1e-6)
    omega = fruits[0, 12:15]
    # Published rolling_speed_after=0.342739083985.
    assert np.hypot(omega[0], omega[1]) == pytest.approx(0.342739083985, abs=1e-6)


def test_yarn_ccd_probe(cloth, oracle):
    scene = probes.yarn_ccd_probe(cloth)
    state = _run(cloth, scene, 1, 0, 0)
    twin = oracle.run(scene, scene.configs(1), 0, 0, precision="fp32")
    assert state["failure"] == 0
    yarn = _yarn(state)
    _compare("yarn ccd record", yarn, twin["yarnContacts"], 2e-6)
    _compare("yarn ccd fruit", _fruits(state), twin["fruits"], 2e-6)
    # Published impact_time=0.380000025034 removed_advance=0.123999990523
    # normal_impulse=12.399999618530 final_fruit_x=-0.023999996483.
    assert yarn[0, 22] == pytest.approx(0.380000025034, abs=1e-7)
    assert yarn[0, 23] == pytest.approx(0.123999990523, abs=1e-7)
    assert yarn[0, 19] == pytest.approx(12.399999618530, abs=2e-6)
    assert yarn[0, 15] == 1
    assert state["fruits"]["positionAndInverseMass"][0, 0] == pytest.approx(-0.023999996483, abs=1e-7)
    assert state["frictionStatus"]["counters"][0, 1] == 1


@pytest.mark.parametrize("refined, co_moving", [(False, False), (True, False), (False, True)])
def test_yarn_air_probe(cloth, oracle, refined, co_moving):
    scene = probes.yarn_aero_probe(cloth, refined, co_moving)
    state = _run(cloth, scene, 1, 0, 0)
    twin = oracle.run(scene, scene.configs(1), 0, 0, precision="fp32")
    assert state["failure"] == 0
    velocity = state["particles"]["velocity"][:, :3].astype(np.float64)
    _compare(f"yarn air refined={refined} co_moving={co_moving}", velocity, twin["velocities"], 4e-6)
    force = _bits(state["aerodynamicsStatus"]["metrics"][0, 0])
    assert force == pytest.approx(twin["maximumYarnAeroForce"], rel=1e-6, abs=1e-12)
    if not refined and not co_moving:
        assert force == pytest.approx(0.172502279282, abs=1e-7)


def test_fruit_air_probe(cloth, oracle):
    scene = probes.fruit_aero_probe(cloth)
    state = _run(cloth, scene, 1, 0, 0)
    twin = oracle.run(scene, scene.configs(1), 0, 0, precision="fp32")
    assert state["failure"] == 0
    _compare("fruit air", _fruits(state), twin["fruits"], 2e-6)
    metrics = state["aerodynamicsStatus"]["metrics"][0]
    assert _bits(metrics[1]) == pytest.approx(0.132944747806, abs=1e-7)
    assert _bits(metrics[2]) == pytest.approx(0.000002209812, abs=1e-11)


@pytest.mark.parametrize("case", sorted(probes.MOUTH_CASES))
def test_mouth_release_probe(cloth, oracle, case):
    (position, rotated, initial), released = probes.MOUTH_CASES[case]
    scene = probes.mouth_probe(cloth, position, rotated, initial)
    state = _run(cloth, scene, 1, 0, 0)
    twin = oracle.run(scene, scene.configs(1), 0, 0, precision="fp32")
    assert state["failure"] == 0
    masks = state["releaseStatus"]["masks"][0]
    assert masks[1] == released == twin["releasedMask"]
    assert masks[0] == twin["candidateMask"]
    # The reference's mouth gate: clearance within 2 um of FP64.
    fp64 = oracle.run(scene, scene.configs(1), 0, 0)
    assert abs(float(state["mouthClearance"][0]) - fp64["mouthClearance"][0]) <= 2.0e-6


def _bag(cloth):
    return cloth.Scene.bag()


def test_bag_one_substep_meets_the_reference_gate(cloth, oracle):
    """One substep of the bag (32 iterations, 3 strain sweeps) against FP64
    under the reference's own FP32-vs-FP64 gates: positions 2e-5 m and
    multipliers 2e-8 (reference tools/cloth_bag_metal.mm:7656-7661; its
    published errors for this substep are 5.55e-7 m and 2.2e-11,
    docs/METAL_CLOTH.md:313-318). Over a whole 48-substep frame the stiff bag
    amplifies FP32 rounding beyond 1e-5 of its extent in the exact FP32
    restatement itself, which the frame tests below account for."""
    scene = _bag(cloth)
    state = _run(cloth, scene, 1, 32, 3)
    fp64 = oracle.run(scene, scene.configs(1), 32, 3)
    gpu = state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)
    error = float(np.max(np.abs(gpu - fp64["positions"])))
    lam = state["distances"]["material"][:, 2].astype(np.float64)
    lam_error = float(np.max(np.abs(lam - fp64["distanceLambda"])))
    print(f"bag one substep: max |x_gpu - x_fp64| {error:.3e} m, max |lambda error| {lam_error:.3e}")
    assert state["failure"] == 0
    assert error <= 2e-5
    assert lam_error <= 2e-8


def test_bag_frame_parity(cloth, oracle):
    """One 1/120 s frame (48 substeps) of the 1,465-knot bag holding 12 fruits."""
    scene = _bag(cloth)
    substeps, iterations, sweeps = 48, 32, 3
    configs = scene.configs(substeps)
    state = _run(cloth, scene, substeps, iterations, sweeps)
    assert state["failure"] == 0
    _check_frame(oracle, scene, state, configs, iterations, sweeps, "bag")
    twin = oracle.run(scene, configs, iterations, sweeps, precision="fp32")
    fp64 = oracle.run(scene, configs, iterations, sweeps)
    fruits = _fruits(state)
    x = scene.particles["positionAndInverseMass"][:, :3]
    extent = float(np.max(x.max(axis=0) - x.min(axis=0)))
    e_twin = np.max(np.abs(fruits[:, :3] - twin["fruits"][:, :3]))
    e_fp64 = np.max(np.abs(fruits[:, :3] - fp64["fruits"][:, :3]))
    e_ref = np.max(np.abs(twin["fruits"][:, :3] - fp64["fruits"][:, :3]))
    print(f"bag fruits: |gpu-twin| {e_twin:.3e} |gpu-fp64| {e_fp64:.3e} |twin-fp64| {e_ref:.3e}")
    assert e_twin <= 1e-6 * extent
    assert e_fp64 <= max(1e-5 * extent, 1.5 * e_ref)
    masks = state["releaseStatus"]["masks"][0]
    assert masks[0] == twin["candidateMask"] and masks[1] == twin["releasedMask"]


def test_bag_runs_are_bit_identical(cloth):
    scene = _bag(cloth)
    first = _run(cloth, scene, 24, 32, 3)
    second = _run(cloth, scene, 24, 32, 3)
    for name in ("particles", "fruits", "fruitPairs", "yarnContacts", "frictionStatus", "aerodynamicsStatus"):
        assert first[name].tobytes() == second[name].tobytes(), name


def test_bag_with_self_contact_frame_parity(cloth, oracle):
    """The bag with yarn/yarn self-contact (hash broad phase, normal response,
    impulse aggregation and self friction) over 12 substeps: knots and fruits
    as the FP32 twin computes them, the same self-contact and friction
    counts, and the FP64 gate."""
    scene = cloth.Scene.bag(selfContact=1)
    substeps, iterations, sweeps = 12, 32, 3
    configs = scene.configs(substeps)
    state = _run(cloth, scene, substeps, iterations, sweeps)
    assert state["failure"] == 0
    twin = oracle.run(scene, configs, iterations, sweeps, precision="fp32")
    fp64 = oracle.run(scene, configs, iterations, sweeps)
    gpu = state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)
    x = scene.particles["positionAndInverseMass"][:, :3]
    extent = float(np.max(x.max(axis=0) - x.min(axis=0)))
    e_twin = np.max(np.abs(gpu - twin["positions"]))
    e_fp64 = np.max(np.abs(gpu - fp64["positions"]))
    e_ref = np.max(np.abs(twin["positions"] - fp64["positions"]))
    counters = state["selfStatus"]["counters"][0]
    metrics = state["frictionStatus"]["metrics"][0]
    print(f"bag+self: |gpu-twin| {e_twin:.3e} |gpu-fp64| {e_fp64:.3e} |twin-fp64| {e_ref:.3e}; contacts gpu "
          f"{counters[:2]} twin {twin['presentContacts']}/{twin['sweptContacts']}; self friction gpu {metrics[3]} "
          f"twin {twin['selfFrictionContacts']}")
    assert twin["presentContacts"] + twin["sweptContacts"] > 0, "the bag made no yarn/yarn contact"
    assert e_twin <= 1e-6 * extent
    assert e_fp64 <= max(1e-5 * extent, 1.5 * e_ref)
    # Contact activation is a threshold; the reference's own gate on the self
    # friction count is a difference of at most 4 (reference
    # tools/cloth_bag_metal.mm:6769-6772, :7699).
    assert abs(int(counters[0]) - twin["presentContacts"]) <= max(4, twin["presentContacts"] // 100)
    assert abs(int(counters[1]) - twin["sweptContacts"]) <= 4
    assert abs(int(metrics[3]) - twin["selfFrictionContacts"]) <= 4


def _lifted_bag(cloth, frames):
    """The bag with self-contact and 12 fruits of 2-3 cm, its cuff grip
    lifted by 0.4 m * smoothstep(t / 2 s); 48 substeps per 1/120 s frame."""
    scene = cloth.Scene.bag(stretchComplianceOverride=1.0e-7, bendComplianceOverride=1.0e-4,
                            strainLimitOverride=0.10, fruitRadiusMin=0.02, fruitRadiusMax=0.03, seed=20261203,
                            selfContact=1)
    scene.config["constraintCounts"][0, 2] = 1
    substeps, dt = 48, 1.0 / (120.0 * 48)
    scene.config["gravityAndTimestep"][0, 3] = dt
    scene.refresh()
    base = scene.config[0]["gripTargetAndActive"].astype(np.float64)
    configs = scene.configs(substeps * frames)
    t = np.arange(1, substeps * frames + 1) * dt
    u = np.clip(t / 2.0, 0.0, 1.0)
    configs["gripTargetAndActive"][:, 2] = base[2] + 0.4 * u * u * (3.0 - 2.0 * u)
    configs["gravityAndTimestep"][:, 3] = dt
    return scene, configs


def test_yarn_kept_list_gives_the_full_listing_bytes(cloth):
    """Present-mode fruit/yarn solves that take their candidates from the
    kept list of visits near contact (valid while the knots and fruits have
    drifted less than its room) end in the bytes of solves that test every
    (fruit, segment) visit."""
    import os
    scene, configs = _lifted_bag(cloth, 12)  # the fruits reach the yarn within these frames
    out = {}
    for mode in ("1", "0"):
        saved = os.environ.get("NUMI_CLOTH_HIP_YARN_LIST")
        os.environ["NUMI_CLOTH_HIP_YARN_LIST"] = mode
        try:
            with cloth.Solver(scene) as solver:
                solver.step(configs, 32, 3)
                out[mode] = (solver.read(), solver.diagnostics())
        finally:
            if saved is None:
                os.environ.pop("NUMI_CLOTH_HIP_YARN_LIST", None)
            else:
                os.environ["NUMI_CLOTH_HIP_YARN_LIST"] = saved
    (kept, d), (full, d0) = out["1"], out["0"]
    print(f"kept list: {d.yarnListBuilds} builds, {d.yarnListSolves} solves from the list, most candidates "
          f"{d.yarnCandidatesMax}, reruns {d.yarnReruns}; without: reruns {d0.yarnReruns}")
    assert kept["failure"] == full["failure"] == 0
    assert d.yarnListBuilds >= 1 and d.yarnListSolves > 10 * d.yarnListBuilds
    assert d0.yarnListBuilds == 0 and d0.yarnListSolves == 0
    assert d.yarnCandidatesMax > 0
    for name in ("particles", "fruits", "fruitPairs", "yarnContacts", "frictionStatus", "selfStatus"):
        assert kept[name].tobytes() == full[name].tobytes(), name


def test_sparse_yarn_solve_equals_visit_by_visit_order(cloth):
    """Frames 21-22 of the reference's pickup replay, after the fruits have
    come to rest on the yarn (hundreds of candidate visits per solve), with
    the sparse fruit/yarn solve and with every solve run visit by visit in
    the reference's literal order: the same bytes. Both start from the state
    the sparse solve reaches after 20 frames."""
    import os
    c = cloth.cloth
    scene = cloth.Scene.bag_replay(c.REPLAY_PICKUP, selfContact=1)
    with cloth.Solver(scene) as solver:
        for f in range(20):
            solver.step(c.replay_frame_configs(scene.config, c.REPLAY_PICKUP, f), 32, 3)
        later = scene.continued(solver.read())
    out = {}
    for mode in ("0", "1"):
        saved = os.environ.get("NUMI_CLOTH_HIP_YARN_VISIT_BY_VISIT")
        os.environ["NUMI_CLOTH_HIP_YARN_VISIT_BY_VISIT"] = mode
        try:
            with cloth.Solver(later) as solver:
                for f in (20, 21):
                    solver.step(c.replay_frame_configs(scene.config, c.REPLAY_PICKUP, f), 32, 3)
                out[mode] = (solver.read(), solver.diagnostics())
        finally:
            if saved is None:
                os.environ.pop("NUMI_CLOTH_HIP_YARN_VISIT_BY_VISIT", None)
            else:
                os.environ["NUMI_CLOTH_HIP_YARN_VISIT_BY_VISIT"] = saved
    (sparse, d), (dense, d1) = out["0"], out["1"]
    print(f"sparse: most candidates {d.yarnCandidatesMax}, reruns {d.yarnReruns}, kept-list solves "
          f"{d.yarnListSolves}; visit by visit: reruns {d1.yarnReruns}")
    assert sparse["failure"] == dense["failure"] == 0
    assert d.yarnCandidatesMax > 128 and d.yarnReruns == 0 and d1.yarnReruns > 0
    for name in ("particles", "fruits", "fruitPairs", "yarnContacts", "frictionStatus", "selfStatus"):
        assert sparse[name].tobytes() == dense[name].tobytes(), name


def test_self_solve_in_the_yarn_workgroup_gives_the_same_bytes(cloth):
    """The bag runs each iteration's present self-contact solve in the
    fruit/yarn solve's workgroup, after its positions are written; the bytes
    equal those of the self-contact solve as a launch of its own."""
    import os
    scene, configs = _lifted_bag(cloth, 4)
    out = {}
    for mode in ("1", "0"):
        saved = os.environ.get("NUMI_CLOTH_HIP_SELF_IN_YARN")
        os.environ["NUMI_CLOTH_HIP_SELF_IN_YARN"] = mode
        try:
            with cloth.Solver(scene) as solver:
                solver.step(configs, 32, 3)
                out[mode] = solver.read()
        finally:
            if saved is None:
                os.environ.pop("NUMI_CLOTH_HIP_SELF_IN_YARN", None)
            else:
                os.environ["NUMI_CLOTH_HIP_SELF_IN_YARN"] = saved
    fused, own = out["1"], out["0"]
    counters = fused["selfStatus"]["counters"][0]
    print(f"self contacts (present, swept) {counters[:2]}")
    assert fused["failure"] == own["failure"] == 0
    assert counters[0] > 0
    for name in ("particles", "fruits", "fruitPairs", "yarnContacts", "frictionStatus", "selfStatus"):
        assert fused[name].tobytes() == own[name].tobytes(), name
