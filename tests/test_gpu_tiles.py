"""LDS tile programs against the per-colour schedule (needs a GPU).

The tile path fuses whole iterations and strain-limit blocks into one launch
per program; the host planner proves each owned knot exact, so the state must
equal the per-colour launches' state bit for bit.
"""

import os

import numpy as np
import pytest

import probes
from test_gpu_self import two_layer_impact

pytestmark = pytest.mark.gpu


def _run(cloth, scene, substeps, iterations, sweeps, tiles, staged=None):
    """staged: None lets the autotuner choose; True / False force the staged
    or the streaming tile kernel (NUMI_CLOTH_HIP_TILE_STAGED, read at create)."""
    saved = {k: os.environ.get(k) for k in ("NUMI_CLOTH_HIP_TILES", "NUMI_CLOTH_HIP_TILE_STAGED")}
    os.environ["NUMI_CLOTH_HIP_TILES"] = "1" if tiles else "0"
    if staged is not None:
        os.environ["NUMI_CLOTH_HIP_TILE_STAGED"] = "1" if staged else "0"
    try:
        with cloth.Solver(scene) as solver:
            info = solver.tile_info()
            solver.step(scene.configs(substeps), iterations, sweeps)
            return solver.read(), info
    finally:
        for key, value in saved.items():
            if value is None:
                os.environ.pop(key, None)
            else:
                os.environ[key] = value


SCENES = {
    "draped": lambda c: c.Scene.sheet(width=48, height=40, elevation=0.12, jitter=1e-4, seed=11, sphereMode=1,
                                      fixedSphere=(0.05, -0.03, 0.0, 0.1)),
    "knotted": lambda c: c.Scene.sheet(width=40, height=36, elevation=0.1, jitter=1e-4, seed=12, crossingKnots=1),
    "hanging": lambda c: c.Scene.sheet(width=20, height=64, elevation=0.5, pinTopRow=1, ground=0, seed=13),
    "large": lambda c: c.Scene.sheet(width=160, height=150, elevation=0.05, jitter=1e-4, seed=14, sphereMode=2,
                                     sphereCount=8, sphereRadiusMin=0.05, sphereRadiusMax=0.2),
    "self_contact": two_layer_impact,
    "strain_probe": probes.strain_probe,
}


@pytest.mark.parametrize("staged", [False, True])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_tiles_equal_per_colour_schedule(cloth, name, staged):
    """Both tile kernels (whole program staged in LDS first, or items and
    multipliers streamed pass by pass) against the per-colour launches."""
    scene = SCENES[name](cloth)
    substeps, iterations, sweeps = (1, 0, 1) if name == "strain_probe" else (12, 9, 3)
    tiled, info = _run(cloth, scene, substeps, iterations, sweeps, True, staged)
    plain, off = _run(cloth, scene, substeps, iterations, sweeps, False)
    print(f"{name}: tiles={info.tiles} owned={info.ownedPerTile} loaded<= {info.maxLoaded} halo={info.halo} "
          f"block={info.workgroupSize} staged={info.staged} redundancy={info.redundancy:.2f}")
    assert info.enabled == 1 and off.enabled == 0
    if not staged:
        assert info.staged == 0
    for key in ["particles", "distances", "bends", "knots", "grips", "selfStatus", "frictionStatus"]:
        assert tiled[key].tobytes() == plain[key].tobytes(), key
    assert tiled["failure"] == plain["failure"] == 0


def test_large_sheet_uses_several_tiles_with_halo(cloth):
    scene = SCENES["large"](cloth)
    with cloth.Solver(scene) as solver:
        info = solver.tile_info()
        timing = solver.time_pass(cloth.cloth.PASS_TILE_ITERATION, 0, 20)
    assert info.enabled and info.tiles > 1 and info.halo > 0
    assert info.maxLoaded > info.ownedPerTile
    assert timing.averageMilliseconds > 0.0 and timing.algorithmicBytes > 0.0


def test_reconciliation_rounds_skip_unchanged_tiles(cloth):
    """A reconciliation round's strain program skips the tiles whose loaded
    knots the previous round left unchanged, and so does a certificate
    round's on a scene without self-contact; the state equals the
    per-colour schedule's, which runs every round in full."""
    scene = SCENES["large"](cloth)
    saved = os.environ.get("NUMI_CLOTH_HIP_TILES")
    os.environ["NUMI_CLOTH_HIP_TILES"] = "1"
    try:
        with cloth.Solver(scene) as solver:
            info = solver.tile_info()
            solver.step(scene.configs(24), 9, 3)
            tiled = solver.read()
            d = solver.diagnostics()
    finally:
        if saved is None:
            os.environ.pop("NUMI_CLOTH_HIP_TILES", None)
        else:
            os.environ["NUMI_CLOTH_HIP_TILES"] = saved
    plain, _ = _run(cloth, scene, 24, 9, 3, False)
    print(f"tiles={info.tiles}: repeated-round tiles skipped {d.tileRoundsSkipped}, run {d.tileRoundsRun}")
    assert info.tiles > 1 and not scene.selfContact
    # 24 substeps x (8 reconciliation + 9 certificate rounds), every tile
    # counted once per round.
    assert d.tileRoundsSkipped + d.tileRoundsRun == 24 * (8 + 9) * info.tiles
    assert d.tileRoundsSkipped > 0
    for key in ["particles", "distances", "bends", "selfStatus", "frictionStatus"]:
        assert tiled[key].tobytes() == plain[key].tobytes(), key
    assert tiled["failure"] == plain["failure"] == 0
