"""Substep launch modes (needs a GPU): graph replay and direct launches run
the same kernels in the same order, so the first step()'s probe, which
times both on the live scene and keeps the faster, changes no byte; the
NUMI_CLOTH_HIP_GRAPHS variable fixes the mode."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(cloth):
    return cloth.Scene.sheet(width=48, height=40, elevation=0.12, jitter=1e-4, seed=11, sphereMode=1,
                             fixedSphere=(0.05, -0.03, 0.0, 0.1), selfContact=1)


def _run(cloth, scene, env):
    saved = os.environ.get("NUMI_CLOTH_HIP_GRAPHS")
    if env is None:
        os.environ.pop("NUMI_CLOTH_HIP_GRAPHS", None)
    else:
        os.environ["NUMI_CLOTH_HIP_GRAPHS"] = env
    try:
        with cloth.Solver(scene) as solver:
            before = solver.launch_mode()
            for _ in range(2):
                solver.step(scene.configs(12), 8, 3)
            mode = solver.launch_mode()
            return solver.read(), (before.decided, mode.decided, mode.graphs, mode.probeGraphMilliseconds,
                                   mode.probeDirectMilliseconds)
    finally:
        if saved is None:
            os.environ.pop("NUMI_CLOTH_HIP_GRAPHS", None)
        else:
            os.environ["NUMI_CLOTH_HIP_GRAPHS"] = saved


def test_launch_modes_give_the_same_bytes(cloth):
    scene = _scene(cloth)
    probed, m = _run(cloth, scene, None)
    print("probe: graph %.3f ms, direct %.3f ms -> %s" % (m[3], m[4], "graph" if m[2] else "direct"))
    assert m[0] == 0 and m[1] == 1 and m[3] > 0.0 and m[4] > 0.0
    assert m[2] == (1 if m[3] <= m[4] else 0)
    direct, d = _run(cloth, scene, "0")
    graph, g = _run(cloth, scene, "1")
    assert d[:3] == (1, 1, 0) and g[:3] == (1, 1, 1)
    for key in ["particles", "distances", "bends", "selfStatus", "frictionStatus"]:
        assert probed[key].tobytes() == direct[key].tobytes() == graph[key].tobytes(), key
    assert probed["failure"] == 0


def test_short_calls_leave_the_mode_undecided(cloth):
    scene = _scene(cloth)
    with cloth.Solver(scene) as solver:
        solver.step(scene.configs(4), 8, 3)
        assert solver.launch_mode().decided == 0
        solver.step(scene.configs(5), 8, 3)
        assert solver.launch_mode().decided == 1
