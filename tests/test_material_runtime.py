"""Runtime parity of a fitted material artifact (the reference's
cmake/CheckClothCalibrationRuntime.cmake, against this library).

The 19-parameter contract (tests/golden/cloth-parameters-v1.csv, the
reference's calibration/cloth-parameters-v1.csv) gets an observation CSV of
unit sensitivities written by numi_cloth_calibration_write; the reference's
fitter (oracle/_ref, built from its source when the reference checkout is
present) qualifies it and writes a numi.cloth.material.v1 artifact. This
library must load that artifact with the fitter's provenance hashes, build
the bag from it, and (with zero log deltas) step exactly as the authored
defaults do; a fitted change must reach the scene.
"""

import re
import subprocess

import numpy as np
import pytest

from test_calibration import FITTER

FIELDS = ["ordinaryNodeMassKg", "hemNodeMassKg", "yarnRadiusM", "axialBodyComplianceMPerN",
          "axialCuffComplianceMPerN", "knotCompliance", "bendBodyComplianceMPerN", "bendCuffComplianceMPerN",
          "gripComplianceMPerN", "clothGroundFriction", "clothSelfFriction", "fruitYarnFriction",
          "fruitGroundFriction", "fruitPairFriction", "fruitRollingResistance", "yarnCrossflowDrag",
          "yarnSkinFriction", "fruitDrag", "fruitRotationalDrag"]

needs_fitter = pytest.mark.skipif(not FITTER.exists(), reason="reference fitter not built (no reference checkout)")


def contract(golden):
    lines = (golden / "cloth-parameters-v1.csv").read_text().split()
    assert lines[0] == "name,base,minimum,maximum"
    return [line.split(",")[0] for line in lines[1:]]


def fitted_artifact(cloth, golden, tmp_path, deltas):
    """Fit unit-sensitivity rows whose measurements are the given log deltas
    (calibration and heldout row per parameter); returns (artifact path,
    fitter report)."""
    names = contract(golden)
    assert len(names) == 19
    rows = []
    for j, name in enumerate(names):
        dlog = np.zeros(len(names))
        dlog[j] = 1.0
        for split in ("cal", "heldout"):
            rows.append({"trial": f"{split}_{name}", "heldout": split == "heldout", "observable": f"unit_{name}",
                         "measured": float(deltas[j]), "sigma": 1.0, "baseline": 0.0, "dlog": dlog})
    observations, material = tmp_path / "full-contract-observations.csv", tmp_path / "full-contract-material.txt"
    cloth.cloth.write_calibration(observations, names, rows)
    result = subprocess.run([str(FITTER), "--parameters", str(golden / "cloth-parameters-v1.csv"), "--observations",
                             str(observations), "--material-output", str(material),
                             "--allow-no-heldout-improvement"], capture_output=True, text=True, timeout=60)
    assert result.returncode == 0 and "qualified=true" in result.stdout, result.stdout + result.stderr
    return material, result.stdout


@needs_fitter
def test_fitted_artifact_loads_with_its_provenance(cloth, golden, tmp_path):
    material, report = fitted_artifact(cloth, golden, tmp_path, np.zeros(19))
    artifact = cloth.cloth.load_material(material)
    defaults = cloth.cloth.material_defaults()
    assert artifact.loaded == 1
    for field in FIELDS:
        assert getattr(artifact.values, field) == getattr(defaults.values, field), field
    parameters_hash = re.search(r"parameters_hash=(0x[0-9a-f]+)", report).group(1)
    observations_hash = re.search(r"observations_hash=(0x[0-9a-f]+)", report).group(1)
    assert artifact.parametersHash.decode() == parameters_hash
    assert artifact.observationsHash.decode() == observations_hash


@needs_fitter
def test_fitted_change_reaches_the_scene(cloth, golden, tmp_path):
    names = contract(golden)
    deltas = np.zeros(19)
    deltas[names.index("cloth_ground_friction")] = 0.2
    material, _ = fitted_artifact(cloth, golden, tmp_path, deltas)
    artifact = cloth.cloth.load_material(material)
    expected = 0.45 * np.exp(0.2)  # the fit (default ridge) and 12 printed decimals: within 1e-9
    assert artifact.values.clothGroundFriction == pytest.approx(expected, abs=1e-9)
    scene = cloth.Scene.bag(material=artifact)
    assert scene.config["clothMaterial"][0, 2] == np.float32(artifact.values.clothGroundFriction)
    plain = cloth.Scene.bag()
    assert plain.config["clothMaterial"][0, 2] == np.float32(0.45)


@needs_fitter
@pytest.mark.gpu
def test_artifact_run_equals_authored_defaults(cloth, golden, tmp_path):
    """The artifact path (fitter -> artifact -> loader -> scene -> solver)
    steps the bag to the same bytes and frame hash as the authored defaults."""
    material, _ = fitted_artifact(cloth, golden, tmp_path, np.zeros(19))
    artifact = cloth.cloth.load_material(material)
    runs = []
    for m in (artifact, None):
        scene = cloth.Scene.bag(material=m)
        with cloth.Solver(scene) as solver:
            solver.step(scene.configs(12), 32, 3)
            state = solver.read()
        runs.append((state, cloth.cloth.frame_hash(scene, state)))
    (a, ha), (b, hb) = runs
    assert a["failure"] == 0
    assert a["particles"].tobytes() == b["particles"].tobytes()
    assert ha == hb
