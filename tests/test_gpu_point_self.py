"""The CPU reference's extra passes on the GPU (needs a GPU): point
self-collision and the swept ground pass, against the oracle's FP32 twin
(same operation order, fixed-point sums: equal bytes) and the FP64 oracle,
plus the closed-form knot pair.
"""

import numpy as np
import pytest

from test_gpu_self import folded_sheet
from test_oracle_point_self import RF, free_knots

pytestmark = pytest.mark.gpu


def run_both(cloth, oracle, scene, substeps, iterations, sweeps):
    configs = scene.configs(substeps)
    with cloth.Solver(scene) as solver:
        solver.step(configs, iterations, sweeps)
        state = solver.read()
        diag = solver.diagnostics()
    twin = oracle.run(scene, configs, iterations, sweeps, precision="fp32")
    fp64 = oracle.run(scene, configs, iterations, sweeps)
    return state, diag, twin, fp64


def test_point_pair_closed_form(cloth, oracle):
    scene = free_knots(cloth, [(0.0, 0.0, 0.5), (0.003, 0.004, 0.5)], [1.0, 3.0], pointSelfCollision=1)
    state, diag, twin, _ = run_both(cloth, oracle, scene, 1, 1, 0)
    x = state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)
    assert diag.pointSelfContacts == 1
    assert np.linalg.norm(x[1] - x[0]) == pytest.approx(2 * RF, abs=1e-8)
    assert np.array_equal(x, twin["positions"])


@pytest.mark.parametrize("segments", [False, True])
def test_point_self_collision_matches_the_fp32_twin(cloth, oracle, segments):
    """Folded sheet with the layers 3 mm apart (2r = 8 mm): point pushes
    alone, and after the segment self-contact of each iteration."""
    scene = folded_sheet(cloth)
    scene.selfContact = 1 if segments else 0
    scene.pointSelfCollision = 1
    scene.refresh()
    state, diag, twin, fp64 = run_both(cloth, oracle, scene, 6, 8, 2)
    gpu = state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)
    x = scene.particles["positionAndInverseMass"][:, :3]
    extent = float(np.max(x.max(axis=0) - x.min(axis=0)))
    e_twin = np.max(np.abs(gpu - twin["positions"]))
    e_fp64 = np.max(np.abs(gpu - fp64["positions"]))
    e_ref = np.max(np.abs(twin["positions"] - fp64["positions"]))
    print(f"segments={segments}: |gpu-twin| {e_twin:.3e} |gpu-fp64| {e_fp64:.3e} |twin-fp64| {e_ref:.3e} "
          f"last-pass contacts gpu {diag.pointSelfContacts} twin {twin['pointSelfContacts']}")
    assert state["failure"] == 0
    assert e_twin == 0.0
    assert diag.pointSelfContacts == twin["pointSelfContacts"]
    assert e_fp64 <= max(1e-5 * extent, 1.5 * e_ref)
    z = gpu[:, 2]
    assert z.max() - z.min() > 2 * RF * 0.99  # the layers end at least 2r apart


def test_point_self_collision_first_pass_count(cloth, oracle):
    scene = folded_sheet(cloth)
    scene.selfContact = 0
    scene.pointSelfCollision = 1
    scene.refresh()
    _, diag, twin, _ = run_both(cloth, oracle, scene, 1, 1, 0)
    assert diag.pointSelfContacts == twin["pointSelfContacts"] > 0


def test_swept_ground_matches_the_fp32_twin(cloth, oracle):
    """A sheet thrown at the ground at 3 m/s from just above contact: the
    swept pass catches the knots that cross in one advance."""
    scene = cloth.Scene.sheet(width=24, height=20, elevation=0.0045, jitter=1e-4, seed=9, ground=1)
    scene.particles["velocity"][:, 2] = -3.0
    scene.sweptGround = 1
    scene.refresh()
    state, _, twin, fp64 = run_both(cloth, oracle, scene, 3, 8, 2)
    gpu = state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)
    assert state["failure"] == 0
    assert np.array_equal(gpu, twin["positions"])
    assert np.max(np.abs(gpu - fp64["positions"])) < 1e-6
    # without the pass the first substep differs (the knots sit below contact until the ground pass)
    plain = cloth.Scene.sheet(width=24, height=20, elevation=0.0045, jitter=1e-4, seed=9, ground=1)
    plain.particles["velocity"][:, 2] = -3.0
    plain.refresh()
    assert not np.array_equal(oracle.run(plain, plain.configs(1), 0, 0, precision="fp32")["positions"],
                              oracle.run(scene, scene.configs(1), 0, 0, precision="fp32")["positions"])
