"""Pins of the oracle's produce-bag families (CPU only): fruit pairs, fruit
ground contact with Coulomb and rolling friction, fruit/yarn swept contact,
yarn and fruit air loads, and mouth release.

Two independent anchors per probe: the FP32 values the reference publishes
for the same probe states (reference docs/METAL_CLOTH.md:366-406, probe
states tools/cloth_bag_metal.mm:1379-1495, :1556-1703), which the FP32 twin
must reproduce to FP32 rounding, and closed forms of the mechanics, which the
FP64 oracle must meet to FP64 rounding.
"""

import math

import numpy as np
import pytest

import probes

def _tangential(v, n):
    return float(np.linalg.norm(v - n * np.dot(v, n)))


def _fruit(out, i=0):
    r = out["fruits"][i]
    return dict(x=r[0:3], w=r[3], radius=r[7], v=r[8:11], ground=r[11], omega=r[12:15], q=r[16:20])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fruit_pair_probe(cloth, oracle, precision):
    scene = probes.fruit_pair_probe(cloth)
    out = oracle.run(scene, scene.configs(1), 1, 0, precision=precision)
    a, b = _fruit(out, 0), _fruit(out, 1)
    separation = np.linalg.norm(b["x"] - a["x"])
    impulse = out["fruitPairs"][0, 3]
    change = np.max(np.abs(np.concatenate([a["q"], b["q"]]) - np.array([0, 0, 0, 1] * 2)))
    if precision == "fp32":
        # Published separation=1.999999962672 normal_impulse=16.656669616699
        # orientation_change=0.007141246926 cone_ratio=0.057165719569.
        assert separation == pytest.approx(1.999999962672, abs=2e-9)
        assert impulse == pytest.approx(16.656669616699, abs=2e-6)
        assert change == pytest.approx(0.007141246926, abs=2e-9)
        assert out["maximumConeRatio"] == pytest.approx(0.057165719569, abs=2e-9)
    else:
        # Fruits advance by v dt to (0, 0.01, 2) and (1.5, -0.02, 2); the pair
        # projection restores the 2 m contact distance with multiplier
        # (2 - |d|) / (w0 + w1), and the impulse is that over dt.
        dt = float(np.float32(0.01))
        d = math.hypot(1.5, 3.0 * dt)
        assert separation == pytest.approx(2.0, abs=1e-14)
        assert impulse == pytest.approx((2.0 - d) / 3.0 / dt, rel=1e-12)
    assert out["pairFrictionContacts"] == 1
    # Momentum is conserved by the contact and the friction impulse (masses 1, 0.5).
    x0 = scene.fruits["positionAndInverseMass"][:, :3].astype(np.float64)
    dt = float(np.float32(0.01))
    masses = np.array([1.0, 0.5])
    before = sum(m * (f["x"] - x) / dt for m, f, x in zip(masses, [a, b], x0))
    after = sum(m * f["v"] for m, f in zip(masses, [a, b]))
    assert np.max(np.abs(after - before)) < (1e-6 if precision == "fp32" else 1e-10)
    # Tangential slip at the contact falls (published slip 2.9994 -> 1.7e-7).
    n = (b["x"] - a["x"]) / separation
    slip_after = _tangential(b["v"] + np.cross(b["omega"], -n) - a["v"] - np.cross(a["omega"], n), n)
    assert slip_after < 1e-5
    for f in (a, b):
        assert np.linalg.norm(f["q"]) == pytest.approx(1.0, abs=1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fruit_ground_probe(cloth, oracle, precision):
    scene = probes.ground_contact_probe(cloth)
    out = oracle.run(scene, scene.configs(1), 1, 0, precision=precision)
    f = _fruit(out)
    offset = np.array([0.0, 0.0, -1.0])
    up = np.array([0.0, 0.0, 1.0])
    slip = _tangential(f["v"] + np.cross(f["omega"], offset), up)
    rolling = math.hypot(f["omega"][0], f["omega"][1])
    # Fruit of radius 1 sunk 0.5 m: lifted to z = 1 with normal impulse
    # 0.5 / (w dt) = 25 (published fruit_height=1 fruit_normal_impulse=25).
    assert f["x"][2] == 1.0
    assert f["ground"] == pytest.approx(25.0, rel=1e-7)
    assert out["fruitGroundContacts"] == 1 and out["rollingContacts"] == 1
    assert out["maximumConeRatio"] == pytest.approx(1.0, abs=1e-7)
    assert out["maximumRollingRatio"] == pytest.approx(1.0, abs=1e-7)
    # Published fruit_slip_after=1.875000083742 rolling_speed_after=0.342739083985.
    assert slip == pytest.approx(1.875000083742, abs=5e-7)
    assert rolling == pytest.approx(0.342739083985, abs=5e-7)
    if precision == "fp64":
        # Sliding friction stops the contact point (slip sqrt(13) over the
        # effective inverse mass 2 + 5 = 7 is below the 0.42 * 25 cap); the
        # rolling cap 0.015 * 1 * 25 = 0.375 then removes 0.375 * 5 rad/s of
        # the x spin, so the contact point slips at r * that spin change.
        assert slip == pytest.approx(1.875, abs=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_yarn_ccd_probe(cloth, oracle, precision):
    scene = probes.yarn_ccd_probe(cloth)
    out = oracle.run(scene, scene.configs(1), 0, 0, precision=precision)
    y = out["yarnContacts"][0]
    f = _fruit(out)
    # The fruit (r 2 cm) moves x: -0.1 -> 0.1 in one substep; it meets the
    # 4 mm yarn at x = -0.024, i.e. t = 0.076 / 0.2 = 0.38, and the swept
    # response removes the remaining 0.124 m of advance, impulse 0.124 / dt.
    # Published impact_time=0.380000025034 removed_advance=0.123999990523
    # normal_impulse=12.399999618530 final_fruit_x=-0.023999996483.
    # The conservative-advancement search stops within 1e-9 m of contact, so
    # the impact time is exact to (1e-9 m) / (0.2 m per substep) plus steps.
    tolerance = 1e-7 if precision == "fp32" else 1e-8
    assert y[22] == pytest.approx(0.38, abs=5e-8 if precision == "fp64" else tolerance)
    assert y[23] == pytest.approx(0.124, abs=tolerance)
    assert y[15] == 1  # one swept response
    assert y[19] == pytest.approx(12.4, abs=1e-6 if precision == "fp32" else 1e-6)
    assert f["x"][0] == pytest.approx(-0.024, abs=tolerance)
    assert abs(y[3]) < 1e-8
    if precision == "fp32":
        assert y[19] == pytest.approx(12.399999618530, abs=1e-9)
        assert f["x"][0] == pytest.approx(-0.023999996483, abs=1e-12)
        # Published cone_ratio=0.128008306026 (yarn friction of the contact).
        assert out["maximumConeRatio"] == pytest.approx(0.128008306026, abs=1e-7)
    assert out["yarnFrictionContacts"] == 1


def _yarn_energy(scene, out, air):
    m = scene.particles["previousAndMass"][:, 3].astype(np.float64)
    rel = out["velocities"] - air
    return 0.5 * float(np.sum(m * np.sum(rel * rel, axis=1)))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_yarn_air_probe(cloth, oracle, precision):
    scene = probes.yarn_aero_probe(cloth)
    out = oracle.run(scene, scene.configs(1), 0, 0, precision=precision)
    # Crossflow (0, 4, 0) and axial (3, 0, 0) drag of a 2 m yarn of radius 4 mm.
    rho, cd, cf, r, length = 1.225, 1.10, 0.010, 0.004, 2.0
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    cross = 0.5 * f32(rho) * f32(cd) * 2 * f32(r) * length * 4.0 * 4.0
    skin = 0.5 * f32(rho) * f32(cf) * 2 * math.pi * f32(r) * length * 3.0 * 3.0
    assert out["maximumYarnAeroForce"] == pytest.approx(math.hypot(cross, skin), rel=1e-6)
    energy = _yarn_energy(scene, out, np.zeros(3))
    # Published force=0.172502279282 energy 25 -> 24.993015079104.
    assert out["maximumYarnAeroForce"] == pytest.approx(0.172502279282, abs=2e-8)
    assert energy == pytest.approx(24.993015079104, abs=1e-5)
    assert energy < 25.0
    # Subdivision invariance: the centre velocity of the two-segment yarn
    # equals the one-segment yarn's (published refinement_error=4.8e-7).
    refined = probes.yarn_aero_probe(cloth, refined=True)
    out2 = oracle.run(refined, refined.configs(1), 0, 0, precision=precision)
    m1 = scene.particles["previousAndMass"][:, 3].astype(np.float64)
    m2 = refined.particles["previousAndMass"][:, 3].astype(np.float64)
    c1 = (m1[:, None] * out["velocities"]).sum(0) / m1.sum()
    c2 = (m2[:, None] * out2["velocities"]).sum(0) / m2.sum()
    assert np.linalg.norm(c1 - c2) < (1e-6 if precision == "fp32" else 1e-12)
    # A yarn moving with the air feels no drag (published delta 2.9e-6 in FP32).
    still = probes.yarn_aero_probe(cloth, co_moving=True)
    out3 = oracle.run(still, still.configs(1), 0, 0, precision=precision)
    assert out3["maximumYarnAeroForce"] == 0.0
    assert np.max(np.abs(out3["velocities"] - [3.0, 4.0, 0.0])) < (4e-6 if precision == "fp32" else 1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fruit_air_probe(cloth, oracle, precision):
    scene = probes.fruit_aero_probe(cloth)
    out = oracle.run(scene, scene.configs(1), 0, 0, precision=precision)
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    rho, cd, cr, r = f32(1.225), f32(0.47), f32(0.010), f32(0.07)
    force = 0.5 * rho * cd * math.pi * r * r * 30.0          # |v|^2 = 25 + 1 + 4
    torque = 3.0 * math.pi ** 2 / 8.0 * rho * cr * r ** 5 * 29.0   # |w|^2 = 9 + 16 + 4
    assert out["maximumFruitAeroForce"] == pytest.approx(force, rel=1e-6)
    assert out["maximumFruitAeroTorque"] == pytest.approx(torque, rel=1e-6)
    # Published force=0.132944747806 torque=0.000002209812.
    assert out["maximumFruitAeroForce"] == pytest.approx(0.132944747806, abs=1e-8)
    assert out["maximumFruitAeroTorque"] == pytest.approx(0.000002209812, abs=1e-12)
    # Exact quadratic decay: |v| -> |v| / (1 + w k |v| dt) along the same direction.
    f = _fruit(out)
    w = f32(1.0 / 0.21)
    k = 0.5 * rho * cd * math.pi * r * r
    speed = math.sqrt(30.0)
    expected = np.array([5.0, -1.0, 2.0]) / (1.0 + w * k * speed * f32(0.01))
    # FP32 finalizes v from positions near z = 10 (ulp 9.5e-7 m over dt =
    # 0.01 s); the reference publishes the same FP32-vs-FP64 velocity error,
    # fruit_aerodynamics_velocity_error=0.000029432718.
    error = np.max(np.abs(f["v"] - expected))
    if precision == "fp32":
        assert error == pytest.approx(0.000029432718, abs=1e-9)
    else:
        assert error < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", sorted(probes.MOUTH_CASES))
def test_mouth_release_probe(cloth, oracle, precision, case):
    (position, rotated, initial), released = probes.MOUTH_CASES[case]
    scene = probes.mouth_probe(cloth, position, rotated, initial)
    out = oracle.run(scene, scene.configs(1), 0, 0, precision=precision)
    assert out["releasedMask"] == released
    if case in ("released", "rotated"):
        # Through the cap: axial clearance 1.13 - 1 - (0.1 + 0.004) = 0.026
        # (published clearance=rotated_clearance=0.025999993086).
        assert out["mouthClearance"][0] == pytest.approx(0.026, abs=1e-7)
        if precision == "fp32":
            assert out["mouthClearance"][0] == pytest.approx(0.025999993086, abs=2e-9)
