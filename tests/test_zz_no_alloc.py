"""numi_cloth_hip_step_noalloc on the GPU (needs a GPU): the step entry point
that allocates nothing gives the same bytes as numi_cloth_hip_step on two
calls in a row, the second with fewer substeps, and refuses calls beyond
its staging limit with a clear error.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("particles", "distances", "bends", "grips", "fruits", "fruitPairs", "yarnContacts", "selfStatus",
         "frictionStatus", "aerodynamicsStatus", "releaseStatus", "mouthClearance")


def _sheet(cloth):
    """Tile programs, sphere windows (6 spheres) and hash self-contact."""
    scene = cloth.Scene.sheet(width=48, height=40, spacing=0.01, elevation=0.06, jitter=1.0e-4, seed=21,
                              selfContact=1, sphereMode=2, sphereCount=6, sphereRadiusMin=0.02, sphereRadiusMax=0.05,
                              cellSize=0.01, timestep=1.0 / 960.0)
    scene.particles["velocity"][:, 2] = -1.0
    scene.refresh()
    return scene


def _bag(cloth):
    return cloth.Scene.bag(selfContact=1)


@pytest.mark.parametrize("build, first, second", [(_sheet, 24, 9), (_bag, 6, 3)])
def test_noalloc_step_equals_step(cloth, build, first, second):
    scene = build(cloth)
    configs = scene.configs(first + second)
    with cloth.Solver(scene) as plain, cloth.Solver(scene) as noalloc:
        for call in (configs[:first], configs[first:]):
            plain.step(call, 32, 3)
            noalloc.step_noalloc(call, 32, 3)
            a, b = plain.read(), noalloc.read()
            assert a["failure"] == b["failure"] == 0
            for name in STATE:
                assert np.asarray(a[name]).tobytes() == np.asarray(b[name]).tobytes(), name


def test_noalloc_step_refuses_beyond_its_limits(cloth):
    scene = _sheet(cloth)
    with cloth.Solver(scene) as solver:
        with pytest.raises(cloth.ClothError) as error:
            solver.step_noalloc(scene.configs(1025), 32, 3)
        assert "at most 1024 substeps" in error.value.message
        if solver.tile_info().enabled:
            with pytest.raises(cloth.ClothError) as error:
                solver.step_noalloc(scene.configs(2), 32, 2)
            assert "prepared for 3 sweeps" in error.value.message
        # After a refusal the solver steps on as before.
        solver.step_noalloc(scene.configs(2), 32, 3)
        assert solver.read()["failure"] == 0
