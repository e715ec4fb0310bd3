"""The certificate loop on the GPU (needs a GPU): the number of rounds and the
state after them equal the FP32 twin's (the loop's stopping decisions are
taken on the device's residuals, so equal rounds mean equal residual
decisions), through the graph-replayed substeps, the direct launches and the
no-allocation entry point.
"""

import numpy as np
import pytest

import probes
from test_gpu_self import two_layer_impact
from test_oracle_certificate import _cap, _stretched_strip

pytestmark = pytest.mark.gpu


def _positions(state):
    return state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)


def _gpu(cloth, scene, substeps, iterations, sweeps, noalloc=False):
    with cloth.Solver(scene) as solver:
        step = solver.step_noalloc if noalloc else solver.step
        step(scene.configs(substeps), iterations, sweeps)
        state = solver.read()
        passes = solver.diagnostics().certificatePasses
    return state, passes


@pytest.mark.parametrize("substeps", [1, 3])
def test_stretched_strip_rounds_equal_the_twin(cloth, oracle, substeps):
    scene = _stretched_strip(cloth, 1.5, 4, 128)
    state, passes = _gpu(cloth, scene, substeps, 0, 1)
    twin = oracle.run(scene, scene.configs(substeps), 0, 1, precision="fp32")
    print(f"stretched strip, {substeps} substeps: {passes} rounds (twin {twin['certificatePasses']})")
    assert state["failure"] == 0
    # One substep needs several rounds; after three, the schedule's own
    # strain sweeps have done most of the work.
    assert passes == twin["certificatePasses"] >= (2 if substeps == 1 else 1)
    assert np.array_equal(_positions(state), twin["positions"])


def test_unfixable_contact_runs_to_the_bound(cloth, oracle):
    scene = probes.crossing_segments_probe(cloth, 0.3, slip=0.0, overlap=0.001)
    for field in ("positionAndInverseMass", "previousAndMass"):
        scene.particles[field][:, 3] = 0.0
    _cap(scene, 5)
    state, passes = _gpu(cloth, scene, 2, 2, 1)
    assert passes == 5 == oracle.run(scene, scene.configs(2), 2, 1, precision="fp32")["certificatePasses"]
    assert np.array_equal(_positions(state), scene.particles["positionAndInverseMass"][:, :3].astype(np.float64))


def test_self_contact_impact_with_certificates_equals_the_twin(cloth, oracle):
    scene = _cap(two_layer_impact(cloth), 128)
    state, passes = _gpu(cloth, scene, 8, 16, 3)
    twin = oracle.run(scene, scene.configs(8), 16, 3, precision="fp32")
    print(f"two-layer impact: {passes} rounds (twin {twin['certificatePasses']})")
    assert passes == twin["certificatePasses"] >= 1
    assert np.array_equal(_positions(state), twin["positions"])


def test_bag_frame_with_certificates_equals_the_twin(cloth, oracle):
    scene = _cap(cloth.Scene.bag(selfContact=1), 128)
    state, passes = _gpu(cloth, scene, 48, 32, 3)
    twin = oracle.run(scene, scene.configs(48), 32, 3, precision="fp32")
    assert passes == twin["certificatePasses"] == 1
    assert np.array_equal(_positions(state), twin["positions"])
    assert np.array_equal(state["fruits"]["positionAndInverseMass"][:, :3].astype(np.float64),
                          twin["fruits"][:, :3])


@pytest.mark.parametrize("substeps", [1, 3])
def test_noalloc_step_runs_the_same_rounds(cloth, substeps):
    # The no-allocation entry point runs the reference's 3 strain sweeps.
    scene = _stretched_strip(cloth, 1.5, 6, 128)
    a, passes_a = _gpu(cloth, scene, substeps, 0, 3)
    b, passes_b = _gpu(cloth, scene, substeps, 0, 3, noalloc=True)
    assert passes_a == passes_b >= (2 if substeps == 1 else 1)
    assert a["particles"].tobytes() == b["particles"].tobytes()
