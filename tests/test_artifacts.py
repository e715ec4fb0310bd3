"""--material and --grip-trajectory inputs, checked against the reference's own
fixtures (copied as data under tests/golden/) and their documented outcomes
(reference cmake/CheckClothMaterialParity.cmake, CheckGripTrajectory.cmake,
CMakeLists.txt:589-662)."""

import math

import pytest

FIELDS = {
    "ordinary_node_mass_kg": "ordinaryNodeMassKg", "hem_node_mass_kg": "hemNodeMassKg",
    "yarn_radius_m": "yarnRadiusM", "axial_body_compliance_m_per_n": "axialBodyComplianceMPerN",
    "axial_cuff_compliance_m_per_n": "axialCuffComplianceMPerN", "knot_compliance": "knotCompliance",
    "bend_body_compliance_m_per_n": "bendBodyComplianceMPerN",
    "bend_cuff_compliance_m_per_n": "bendCuffComplianceMPerN", "grip_compliance_m_per_n": "gripComplianceMPerN",
    "cloth_ground_friction": "clothGroundFriction", "cloth_self_friction": "clothSelfFriction",
    "fruit_yarn_friction": "fruitYarnFriction", "fruit_ground_friction": "fruitGroundFriction",
    "fruit_pair_friction": "fruitPairFriction", "fruit_rolling_resistance": "fruitRollingResistance",
    "yarn_crossflow_drag": "yarnCrossflowDrag", "yarn_skin_friction": "yarnSkinFriction",
    "fruit_drag": "fruitDrag", "fruit_rotational_drag": "fruitRotationalDrag",
}


def _pairs(path):
    out = {}
    for line in path.read_text().splitlines():
        line = line.strip()
        if line and not line.startswith("#"):
            key, value = line.split("=", 1)
            out[key.strip()] = value.strip()
    return out


def test_default_material_fixture_equals_authored_defaults(cloth, golden):
    loaded = cloth.load_material(golden / "default-material.txt")
    defaults = cloth.material_defaults()
    text = _pairs(golden / "default-material.txt")
    assert loaded.loaded == 1 and defaults.loaded == 0
    assert loaded.parametersHash == b"synthetic-defaults"
    assert defaults.parametersHash == b"authored-defaults" and defaults.observationsHash == b"none"
    for key, field in FIELDS.items():
        assert getattr(loaded.values, field) == float(text[key])
        # The authored default artifact must reproduce the defaults exactly.
        assert getattr(loaded.values, field) == getattr(defaults.values, field)


def test_invalid_material_fixture_is_rejected(cloth, golden):
    with pytest.raises(cloth.ClothError) as error:
        cloth.load_material(golden / "invalid-material.txt")
    assert error.value.message == "cloth material artifact is missing: fruit_rotational_drag"
    assert error.value.status == 6


@pytest.mark.parametrize("edit, message", [
    (lambda t: t.replace("schema=numi.cloth.material.v1", "schema=numi.cloth.material.v2"),
     "cloth material schema mismatch"),
    (lambda t: t + "unknown_field=1\n", "unknown cloth material field: unknown_field"),
    (lambda t: t + "fruit_drag=0.5\n", "cloth material fields must be nonempty and unique"),
    (lambda t: t.replace("fruit_drag=0.47", "fruit_drag=9.0"),
     "cloth material parameter outside contract bounds: fruit_drag"),
    (lambda t: t.replace("fruit_drag=0.47", "fruit_drag=0.47x"), "non-finite cloth material value for fruit_drag"),
    (lambda t: t.replace("fruit_drag=0.47", "fruit_drag=abc"), "invalid cloth material value for fruit_drag"),
    (lambda t: t.replace("parameters_hash=synthetic-defaults\n", ""), "cloth material provenance hashes are required"),
    (lambda t: t.replace("fruit_drag=0.47", "fruit_drag 0.47"), "cloth material line must contain '='"),
])
def test_material_errors(cloth, golden, tmp_path, edit, message):
    path = tmp_path / "material.txt"
    path.write_text(edit((golden / "default-material.txt").read_text()))
    with pytest.raises(cloth.ClothError) as error:
        cloth.load_material(path)
    assert error.value.message == message


def _fnv1a64(data: bytes) -> str:
    # FNV-1a with the offset basis the reference uses everywhere
    # (1469598103934665603, reference include/numi/grip_trajectory.h:179),
    # which is not the textbook 14695981039346656037.
    value = 1469598103934665603
    for byte in data:
        value = ((value ^ byte) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return hex(value)


@pytest.mark.parametrize("name, schema, generations, nearest", [
    ("synthetic-identity.csv", "numi.grip.trajectory.v1", 1, 0),
    ("synthetic-rotation.csv", "numi.grip.trajectory.v1", 1, 0),
    ("synthetic-regrab.csv", "numi.grip.trajectory.v2", 3, 0),
    ("synthetic-patch-regrab.csv", "numi.grip.trajectory.v3", 3, 1),
    ("invalid-distant-regrab.csv", "numi.grip.trajectory.v2", 2, 0),
])
def test_trajectory_fixtures_load(cloth, golden, name, schema, generations, nearest):
    path = golden / "trajectories" / name
    info = cloth.GripTrajectory(path).info()
    assert info.schema.decode() == schema
    assert info.attachmentGenerations == generations
    assert info.selectNearestCuffPatch == nearest
    assert info.contentFingerprint.decode() == _fnv1a64(path.read_bytes())


def test_v1_reactivation_is_rejected(cloth, golden):
    with pytest.raises(cloth.ClothError) as error:
        cloth.GripTrajectory(golden / "trajectories" / "invalid-reactivation.csv")
    assert error.value.message == "grip trajectory v1 cannot reactivate after release"


def test_trajectory_sampling_closed_forms(cloth, golden):
    identity = cloth.GripTrajectory(golden / "trajectories" / "synthetic-identity.csv")
    half = identity.sample(0.004166666666667 / 2)
    assert half.translation[2] == pytest.approx(0.0005, abs=1e-15)
    rotation = cloth.GripTrajectory(golden / "trajectories" / "synthetic-rotation.csv")
    # Maximum rotation 2 acos(w) of the last pose: 2 degrees, the 0.0349 rad the
    # reference's comparator checks.
    assert rotation.info().maximumRotation == pytest.approx(math.radians(2.0), abs=2e-9)
    # Slerp halfway between identity and a 1 degree turn about y is 0.5 degrees.
    pose = rotation.sample(0.004166666666667 / 2)
    angle = 2.0 * math.acos(pose.orientation[3])
    assert angle == pytest.approx(math.radians(0.5), abs=2e-9)
    assert pose.orientation[0] == 0.0 and pose.orientation[2] == 0.0
    with pytest.raises(cloth.ClothError):
        rotation.sample(1.0)


def test_regrab_generations_step_at_reactivation(cloth, golden):
    trajectory = cloth.GripTrajectory(golden / "trajectories" / "synthetic-patch-regrab.csv")
    actives = [trajectory.pose(i).active for i in range(trajectory.info().poseCount)]
    generations = [trajectory.pose(i).attachmentGeneration for i in range(trajectory.info().poseCount)]
    assert actives == [1, 1, 0, 1, 1, 0, 1, 1]
    assert generations == [1, 1, 1, 2, 2, 2, 3, 3]
