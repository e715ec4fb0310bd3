"""Frame driver of the trajectory replays on the CPU: pose generators,
per-frame configs, the scenario's initial state, the frame hash and the
renderer OBJ (include/numi_cloth_replay.h; reference runTrajectoryReplay,
tools/cloth_bag_metal.mm:5050-5174).

Pose pins use points where the smoothsteps of the reference generators are
0, 1/2 or 1 and the geometry of the spin path (a circle), not a retyped
formula; the hash is checked against FNV-1a restated over the raw bytes; the
OBJ is parsed back.
"""

import math

import numpy as np
import pytest

CUFF_RING0 = 27 * 48  # wall knot (level 27, ring 0): the grip base point


def base_point(cloth):
    scene = cloth.Scene.bag()
    return scene.particles["positionAndInverseMass"][CUFF_RING0, :3].astype(np.float64)


def test_base_point_is_the_authored_cuff_knot(cloth):
    c = cloth.cloth
    pose = c.replay_pose(c.REPLAY_GROUNDED, 0.0)
    assert np.allclose(list(pose.target), base_point(cloth), atol=1e-7)
    assert pose.active == 0 and pose.attachmentGeneration == 1
    assert list(pose.orientation) == [0.0, 0.0, 0.0, 1.0]


def test_pickup_pose_at_smoothstep_nodes(cloth):
    c = cloth.cloth
    b = np.array(list(c.replay_pose(c.REPLAY_GROUNDED, 0.0).target))
    at = lambda t: np.array(list(c.replay_pose(c.REPLAY_PICKUP, t).target)) - b
    assert np.allclose(at(0.0), [0, 0, 0], atol=1e-15)
    # lift smoothstep(0.4 / 0.8) = 1/2 -> 1.25 / 2
    assert np.allclose(at(0.4), [0, 0, 0.625], atol=1e-15)
    # lift done, snap smoothstep((1.115 - 1.0) / 0.23) = 1/2, no recovery yet
    assert np.allclose(at(1.115), [-0.05, 0.04, 1.25 - 0.425], atol=1e-12)
    # every phase done: lift 1.25, snap -0.85 (and sin(pi) sideways), recovery +0.75
    assert np.allclose(at(2.5), [-0.10, 0.0, 1.15], atol=1e-12)
    assert c.replay_pose(c.REPLAY_PICKUP, 0.3).active == 1


def test_spin_pose_runs_on_a_circle(cloth):
    c = cloth.cloth
    b = np.array(list(c.replay_pose(c.REPLAY_GROUNDED, 0.0).target))
    centre = b + np.array([-0.28, 0.0, 0.52])
    angles, heights = [], []
    for t in np.linspace(0.0, 3.0, 301):
        p = np.array(list(c.replay_pose(c.REPLAY_SPIN, float(t)).target))
        assert abs(math.hypot(p[0] - centre[0], p[1] - centre[1]) - 0.28) < 1e-12
        angles.append(math.atan2(p[1] - centre[1], p[0] - centre[0]))
        heights.append(p[2] - centre[2])
    assert np.allclose(list(c.replay_pose(c.REPLAY_SPIN, 0.0).target), centre + [0.28, 0.0, 0.0], atol=1e-15)
    a = np.unwrap(angles)  # the angle swept from t = 0 (steps of at most 0.048 rad)
    assert np.allclose(heights, 0.035 * np.sin(0.5 * a), atol=1e-12)
    # after the ramp (time constant 0.18 s) the angular speed is 4.8 rad/s
    # (the ramp's tail, 4.8 * 0.18 / 0.5 * exp(-2.5 / 0.18) = 1.5e-6, is below the gate)
    rate = (a[-1] - a[-51]) / (3.0 - 2.5)
    assert abs(rate - 4.8) < 1e-5


def test_recorded_pose_follows_the_trajectory(cloth, golden):
    c = cloth.cloth
    trajectory = cloth.GripTrajectory(golden / "trajectories" / "synthetic-patch-regrab.csv")
    b = np.array(list(c.replay_pose(c.REPLAY_GROUNDED, 0.0).target))
    for t in [0.0, 0.0015, 0.0031, 0.0045]:
        sample = trajectory.sample(t)
        pose = c.replay_pose(c.REPLAY_RECORDED, t, trajectory)
        assert np.allclose(list(pose.target), b + np.array(list(sample.translation)), atol=1e-15)
        assert list(pose.orientation) == list(sample.orientation)
        assert pose.active == sample.active and pose.attachmentGeneration == sample.attachmentGeneration
    with pytest.raises(cloth.ClothError):
        c.replay_pose(c.REPLAY_RECORDED, 0.0)


GRIP_FIELDS = ["gripTargetAndActive", "gripOrientation", "gripControl"]


@pytest.mark.parametrize("scenario", [0, 1, 2])
def test_frame_configs(cloth, scenario):
    c = cloth.cloth
    scene = cloth.Scene.bag_replay(scenario)
    frame = 3
    configs = c.replay_frame_configs(scene.config, scenario, frame)
    assert len(configs) == c.REPLAY_SUBSTEPS_PER_FRAME
    dt = np.float32(1.0) / np.float32(5760.0)
    for s, cfg in enumerate(configs):
        pose = c.replay_pose(scenario, (frame * 48 + s + 1) * float(dt))
        assert np.array_equal(cfg["gripTargetAndActive"],
                              np.array([*pose.target, 1.0 if pose.active else 0.0], np.float32))
        assert np.array_equal(cfg["gripOrientation"], np.array(list(pose.orientation), np.float32))
        assert cfg["gripControl"][0] == pose.attachmentGeneration
        rest = [name for name in c.CONFIG.names if name not in GRIP_FIELDS]
        for name in rest:
            assert np.array_equal(cfg[name], scene.config[0][name]), name


def test_prepare_keeps_the_scene_extension(cloth):
    """Preparing a replay re-points the scene's arrays without dropping the
    builder's extension fields (self-contact and its margin)."""
    c = cloth.cloth
    plain = cloth.Scene.bag(selfContact=1)
    for scenario in (c.REPLAY_GROUNDED, c.REPLAY_SPIN, c.REPLAY_PICKUP):
        scene = cloth.Scene.bag_replay(scenario, selfContact=1)
        assert scene.selfContact == 1
        assert scene.selfContactMargin == plain.selfContactMargin > 0.0


def test_prepare_sets_ground_lift_and_initial_pose(cloth, golden):
    c = cloth.cloth
    plain = cloth.Scene.bag()
    spin = cloth.Scene.bag_replay(c.REPLAY_SPIN)
    lift = np.float32(0.52)
    assert np.array_equal(spin.particles["positionAndInverseMass"][:, 2],
                          plain.particles["positionAndInverseMass"][:, 2] + lift)
    assert np.array_equal(spin.particles["previousAndMass"][:, 2], plain.particles["previousAndMass"][:, 2] + lift)
    assert np.array_equal(spin.fruits["positionAndInverseMass"][:, 2], plain.fruits["positionAndInverseMass"][:, 2] + lift)
    assert spin.config["constraintCounts"][0, 2] == 0
    pickup = cloth.Scene.bag_replay(c.REPLAY_PICKUP)
    assert pickup.config["constraintCounts"][0, 2] == 1
    pose = c.replay_pose(c.REPLAY_PICKUP, 0.0)
    assert np.array_equal(pickup.config["gripTargetAndActive"][0], np.array([*pose.target, 1.0], np.float32))
    trajectory = cloth.GripTrajectory(golden / "trajectories" / "synthetic-patch-regrab.csv")
    recorded = cloth.Scene.bag_replay(c.REPLAY_RECORDED, trajectory)
    assert recorded.config["gripControl"][0, 1] == 1  # v3: nearest cuff patch
    with pytest.raises(cloth.ClothError):
        c.replay_pose(7, 0.0)


def _fnv(data: bytes, value: int) -> int:
    for byte in data:
        value = ((value ^ byte) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return value


def test_frame_hash_is_fnv_over_the_state_in_order(cloth):
    """The library's frame hash equals FNV-1a (the reference's offset basis)
    restated here over the raw record bytes in hashGPUResult's order."""
    scene = cloth.Scene.sheet(width=6, height=5, selfContact=0, seed=4)
    rng = np.random.default_rng(5)
    state = {}
    for name in ["particles", "distances", "grips", "knots", "bends", "fruits", "fruitPairs", "yarnContacts"]:
        array = getattr(scene, name).copy()
        raw = array.view(np.uint8)
        raw[...] = rng.integers(0, 256, raw.shape, dtype=np.uint8)
        state[name] = array
    for name, dtype in [("selfStatus", cloth.cloth.SELF_STATUS), ("frictionStatus", cloth.cloth.FRICTION_STATUS),
                        ("aerodynamicsStatus", cloth.cloth.AERO_STATUS), ("releaseStatus", cloth.cloth.RELEASE_STATUS)]:
        array = np.zeros(1, dtype)
        array.view(np.uint8)[...] = rng.integers(0, 256, array.nbytes, dtype=np.uint8)
        state[name] = array
    state["mouthClearance"] = np.zeros(len(scene.fruits), np.float32)
    state["failure"] = 0x12345678
    value = 1469598103934665603
    for name in ["particles", "distances", "grips", "knots", "bends", "fruits", "fruitPairs", "yarnContacts",
                 "selfStatus", "frictionStatus", "aerodynamicsStatus", "releaseStatus", "mouthClearance"]:
        value = _fnv(state[name].tobytes(), value)
    value = ((value ^ 0x12345678) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    assert cloth.cloth.frame_hash(scene, state) == value


def _parse_obj(path):
    vertices, faces, grip, balls, header = [], [], None, [], []
    for line in open(path):
        if line.startswith("v "):
            vertices.append([float(v) for v in line.split()[1:]])
        elif line.startswith("f "):
            faces.append([int(v) for v in line.split()[1:]])
        elif line.startswith("# grip center"):
            grip = line.split()
        elif line.startswith("# ball"):
            balls.append(line.split())
        elif line.startswith("#"):
            header.append(line)
    return np.array(vertices), np.array(faces), grip, balls, header


def test_obj_round_trip(cloth, tmp_path):
    scene = cloth.Scene.bag_replay(cloth.cloth.REPLAY_PICKUP)
    path = tmp_path / "bag.obj"
    cloth.cloth.write_obj(path, scene.particles, scene.fruits, scene.grips, scene.config)
    vertices, faces, grip, balls, header = _parse_obj(path)
    # 9 significant digits round-trip float32 exactly
    assert np.array_equal(vertices.astype(np.float32), scene.particles["positionAndInverseMass"][:, :3])
    assert header[1].split() == ["#", "vertices", "1465", "render_triangles", str(2 * 48 * 27 + 2 * 12 * 12)]
    assert len(faces) == 2 * 48 * 27 + 2 * 12 * 12
    assert faces.min() == 1 and faces.max() <= 1465
    assert all(len(set(f)) == 3 for f in faces)
    # every wall quad (level, ring) is split into two triangles that share its diagonal
    wall = {tuple(sorted(f)) for f in faces[: 2 * 48 * 27]}
    for level in range(27):
        for ring in range(48):
            a, b = level * 48 + ring + 1, level * 48 + (ring + 1) % 48 + 1
            c, d = a + 48, b + 48
            assert tuple(sorted((a, b, c))) in wall and tuple(sorted((b, d, c))) in wall
    # the surface is closed except at the mouth: every edge has one or two faces
    edges = {}
    for f in faces:
        for i in range(3):
            e = tuple(sorted((f[i], f[(i + 1) % 3])))
            edges[e] = edges.get(e, 0) + 1
    boundary = [e for e, n in edges.items() if n == 1]
    assert all(n <= 2 for n in edges.values())
    mouth = set(range(27 * 48 + 1, 28 * 48 + 1))
    assert all(set(e) <= mouth for e in boundary) and len(boundary) == 48
    target = scene.config["gripTargetAndActive"][0]
    assert np.allclose([float(v) for v in grip[3:6]], target[:3], atol=0) and grip[7] == "1"
    assert len(balls) == len(scene.fruits)
    for k, ball in enumerate(balls):
        assert int(ball[2]) == k
        assert np.array_equal(np.array([float(v) for v in ball[4:7]], np.float32),
                              scene.fruits["positionAndInverseMass"][k, :3])
        assert np.float32(float(ball[8])) == scene.fruits["previousAndRadius"][k, 3]
    with pytest.raises(cloth.ClothError):
        sheet = cloth.Scene.sheet(width=6, height=5)
        cloth.cloth.write_obj(tmp_path / "sheet.obj", sheet.particles, sheet.fruits, sheet.grips, sheet.config)
