"""Pins of the FP64 oracle itself (CPU only).

Each check is a number the reference publishes for its probes
(reference docs/METAL_CLOTH.md:355-375) or a closed form of the mechanics,
chosen so that a dropped term, a wrong sign or a wrong index fails it.
"""

import math

import numpy as np
import pytest

import probes


def test_strain_probe_matches_published_values(cloth, oracle):
    scene = probes.strain_probe(cloth)
    x0 = scene.particles["positionAndInverseMass"][:, :3].astype(np.float64)
    limit = 1.0 * (1.0 + float(np.float32(0.285)))
    # Published strain_probe_initial_violation=0.215000003576.
    assert np.linalg.norm(x0[1] - x0[0]) - limit == pytest.approx(0.215000003576, abs=1e-12)
    out = oracle.run(scene, scene.configs(1), iterations=0, strain_sweeps=1)
    x = out["positions"]
    assert np.linalg.norm(x[1] - x[0]) == pytest.approx(limit, abs=1e-14)
    # Compressed pair untouched (published compression_length_change=0).
    assert np.linalg.norm(x[3] - x[2]) == np.linalg.norm(x0[3] - x0[2])
    # Centre of mass of the corrected pair preserved (masses 1 and 0.5).
    com0 = (x0[0] * 1.0 + x0[1] * 0.5) / 1.5
    com = (x[0] * 1.0 + x[1] * 0.5) / 1.5
    assert np.max(np.abs(com - com0)) < 1e-15


def test_ground_bend_probe_matches_published_values(cloth, oracle):
    scene = probes.ground_bend_probe(cloth)
    out = oracle.run(scene, scene.configs(1), iterations=1, strain_sweeps=0)
    x = out["positions"]
    # Published supported_height=0.004000000190 (FP32 0.004) and
    # free_endpoint_rise=0.009750500321 (FP32 run of the same probe).
    assert x[0, 2] == pytest.approx(float(np.float32(0.004)), abs=1e-15)
    rise = x[2, 2] - float(np.float32(0.104))
    # The published value is an FP32 result; one FP32 ulp at z = 0.11 is 7.5e-9.
    assert rise == pytest.approx(0.009750500321, abs=7.5e-9)
    # The grounded end keeps its height, so only the free end rises.
    assert x[0, 2] == float(np.float32(0.004))


def test_ground_contact_probe_matches_published_values(cloth, oracle):
    scene = probes.ground_contact_probe(cloth)
    out = oracle.run(scene, scene.configs(1), iterations=1, strain_sweeps=0)
    # Published cloth_height=0.004000000190, cloth_slip_after=0.819999992847:
    # normal impulse 0.4 (0.004 m over 0.01 s, unit mass), Coulomb limit
    # 0.45 * 0.4 = 0.18 removed from the 1 m/s slip.
    assert out["positions"][0, 2] == pytest.approx(0.004, abs=1e-9)
    assert out["groundImpulse"][0] == pytest.approx(float(np.float32(0.004)) / 0.01, rel=1e-6)
    assert out["velocities"][0, 0] == pytest.approx(0.819999992847, abs=1e-7)
    assert out["frictionContacts"] == 1 and out["maximumConeRatio"] == pytest.approx(1.0)


def test_free_fall_is_symplectic_euler(cloth, oracle):
    scene = probes.free_fall(cloth)
    n = 37
    out = oracle.run(scene, scene.configs(n), iterations=4, strain_sweeps=0)
    dt = float(np.float32(1.0e-3))
    g = float(np.float32(-9.81))
    x0 = np.array([0.1, -0.2, 2.0], np.float32).astype(np.float64)
    v0 = np.array([0.5, 0.0, 1.0])
    expected = x0 + n * dt * v0 + np.array([0.0, 0.0, g * dt * dt * n * (n + 1) / 2])
    assert np.max(np.abs(out["positions"][0] - expected)) < 1e-12
    # finalize differences positions of magnitude 2 over dt = 1e-3: ~1e-12 rounding.
    assert out["velocities"][0, 2] == pytest.approx(v0[2] + n * g * dt, abs=1e-10)


def test_hanging_spring_settles_at_hooke_extension(cloth, oracle):
    mass, compliance = 0.01, 1.0e-3
    scene = probes.hanging_spring(cloth, mass=mass, compliance=compliance)
    out = oracle.run(scene, scene.configs(3000), iterations=1, strain_sweeps=0)
    extension = (1.0 - out["positions"][1, 2]) - float(np.float32(0.1))
    g = -float(np.float32(-9.81))
    m = 1.0 / float(np.float32(1.0 / mass))
    # XPBD's fixed point is the compliance law: extension = c m g.
    assert extension == pytest.approx(float(np.float32(compliance)) * m * g, rel=1e-9)


def test_sphere_push_out(cloth, oracle):
    scene = probes.sphere_push(cloth)
    out = oracle.run(scene, scene.configs(1), iterations=1, strain_sweeps=0)
    x = out["positions"]
    reach = 0.5 + float(np.float32(0.004))
    x0 = scene.particles["positionAndInverseMass"][:, :3].astype(np.float64)
    assert np.linalg.norm(x[0]) == pytest.approx(reach, abs=1e-15)
    assert np.allclose(x[0] / np.linalg.norm(x[0]), x0[0] / np.linalg.norm(x0[0]), atol=1e-15)
    assert x[1] == pytest.approx([0.0, 0.0, -reach], abs=1e-15)
    assert np.array_equal(x[2], x0[2])


def test_rigid_grip_reaches_rotated_target(cloth, oracle):
    scene = probes.grip_probe(cloth)
    out = oracle.run(scene, scene.configs(1), iterations=1, strain_sweeps=0)
    # Handle (0.2, 0.1, 0.2), offset (0.1, 0, 0) turned 90 degrees about z.
    assert out["positions"][0] == pytest.approx([0.2, 0.2, 0.2], abs=1e-8)


def test_crossing_knot_converges_to_rest_cosine(cloth, oracle):
    scene = probes.knot_probe(cloth, rest_cosine=0.0)
    out = oracle.run(scene, scene.configs(1), iterations=60, strain_sweeps=0)
    x = out["positions"]
    warp = (x[1] - x[0]) / np.linalg.norm(x[1] - x[0])
    weft = (x[3] - x[2]) / np.linalg.norm(x[3] - x[2])
    assert abs(float(np.dot(warp, weft))) < 1e-9
    # Gradients sum to zero, so with equal masses the centroid is fixed.
    x0 = scene.particles["positionAndInverseMass"][:, :3].astype(np.float64)
    assert np.max(np.abs(x.mean(axis=0) - x0.mean(axis=0))) < 1e-15


def test_sheet_is_deterministic_and_stays_on_ground(cloth, oracle):
    scene = cloth.Scene.sheet(width=12, height=10, elevation=0.02, jitter=1e-4, seed=3)
    a = oracle.run(scene, scene.configs(20), 8, 2)
    b = oracle.run(scene, scene.configs(20), 8, 2)
    assert np.array_equal(a["positions"], b["positions"])
    assert a["positions"][:, 2].min() >= float(np.float32(0.004)) - 1e-15
