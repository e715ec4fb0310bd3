"""The CPU reference's extra passes in the oracle (CPU only): the swept
ground pass after the advance and point self-collision.

Pins are closed forms on hand-made scenes: a knot crossing the ground in one
advance lands at contact height; a single knot pair ends exactly 2r apart
with its mass-weighted centre fixed (the reference's correction for one
contact, where the snapshot mean changes nothing); pairs within two
distance hops are skipped.
"""

import numpy as np
import pytest

import probes

R = 0.004
RF = float(np.float32(R))  # the config holds the yarn radius in FP32


def free_knots(cloth, positions, inverse_masses, velocities=None, distances=(), ground=0, dt=1.0e-3, **flags):
    p = probes.particles(cloth, positions, inverse_masses, velocities)
    records = [probes.distance(cloth, i, j, rest, compliance=1.0e-9, colour=k)
               for k, (i, j, rest) in enumerate(distances)]
    c = probes.config(cloth, particles=len(p), distances=len(records), ground=ground, dt=dt, radius=R)
    arrays = dict(particles=p)
    if records:
        arrays["distances"] = np.concatenate(records)
        arrays["distanceBatches"] = np.concatenate([probes.batch(cloth, k, 1, k) for k in range(len(records))])
    return cloth.Scene(c, **arrays, **flags)


def test_swept_ground_puts_a_crossing_knot_at_contact_height(cloth, oracle):
    scene = free_knots(cloth, [(0.0, 0.0, 0.005), (0.1, 0.0, 0.003)], [1.0, 1.0],
                       velocities=[(0.0, 0.0, -2.0), (0.0, 0.0, -2.0)], ground=1, sweptGround=1)
    # no iterations and no strain sweeps: no ground pass runs after the advance
    out = oracle.run(scene, scene.configs(1), 0, 0)
    z = out["positions"][:, 2]
    assert z[0] == pytest.approx(RF, abs=1e-15)     # 0.005 -> 0.003 crossed contact height: back at r
    assert z[1] == pytest.approx(0.001, abs=1e-9)   # started below contact height: not swept (FP32 input)
    scene.sweptGround = 0
    scene.refresh()
    plain = oracle.run(scene, scene.configs(1), 0, 0)
    assert plain["positions"][0, 2] == pytest.approx(0.003, abs=1e-9)


@pytest.mark.parametrize("w0, w1", [(1.0, 1.0), (1.0, 3.0)])
def test_point_pair_ends_two_radii_apart(cloth, oracle, w0, w1):
    x0, x1 = np.array([0.0, 0.0, 0.5]), np.array([0.003, 0.004, 0.5])  # 5 mm apart, 2r = 8 mm
    scene = free_knots(cloth, [tuple(x0), tuple(x1)], [w0, w1], pointSelfCollision=1)
    x0, x1 = (scene.particles["positionAndInverseMass"][k, :3].astype(np.float64) for k in range(2))
    out = oracle.run(scene, scene.configs(1), 1, 0)
    a, b = out["positions"]
    assert out["pointSelfContacts"] == 1
    assert np.linalg.norm(b - a) == pytest.approx(2 * RF, abs=1e-15)
    m0, m1 = 1 / w0, 1 / w1
    np.testing.assert_allclose(m0 * a + m1 * b, m0 * x0 + m1 * x1, atol=1e-15)
    # the push is along the pair's axis
    np.testing.assert_allclose(np.cross(b - a, x1 - x0), 0.0, atol=1e-15)


def test_pairs_within_two_hops_are_skipped(cloth, oracle):
    # chain 0-1-2 with 5 mm links: all three pairs are within two hops;
    # knot 3 (no links) sits 5.7 mm from knot 0 (and 9.8 mm from knot 1) and is pushed
    positions = [(0.0, 0.0, 0.5), (0.005, 0.0, 0.5), (0.010, 0.0, 0.5), (-0.004, 0.004, 0.5)]
    scene = free_knots(cloth, positions, [1.0] * 4, distances=[(0, 1, 0.005), (1, 2, 0.005)],
                       pointSelfCollision=1)
    out = oracle.run(scene, scene.configs(1), 1, 0)
    assert out["pointSelfContacts"] == 1
    assert np.linalg.norm(out["positions"][3] - out["positions"][0]) > 0.0057


def test_fp32_twin_matches_fp64_on_a_folded_sheet(cloth, oracle):
    from test_gpu_self import folded_sheet
    scene = folded_sheet(cloth, width=12, height=16)
    scene.selfContact = 0
    scene.pointSelfCollision = 1
    scene.refresh()
    # the count is the last pass's: one iteration still finds the layers in contact
    first = oracle.run(scene, scene.configs(1), 1, 0)
    assert first["pointSelfContacts"] > 0
    z = first["positions"][:, 2]
    assert z.max() - z.min() > 0.003  # the layers were pushed apart
    fp64 = oracle.run(scene, scene.configs(4), 4, 1)
    twin = oracle.run(scene, scene.configs(4), 4, 1, precision="fp32")
    assert np.max(np.abs(fp64["positions"] - twin["positions"])) < 1e-6
