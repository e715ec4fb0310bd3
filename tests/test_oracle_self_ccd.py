"""Pins of the oracle's yarn/yarn narrow phase (CPU only).

Swept contact: the reference's self-CCD probe (tests/probes.py
self_ccd_probe; state tools/cloth_bag_metal.mm:1497-1554, run with 0
iterations and 0 strain sweeps at :5740-5751, metrics :7321-7445). Its
published FP32 values (docs/METAL_CLOTH.md:388-395) must be reproduced by
the FP32 twin to the printed 12 digits; the FP64 oracle must meet the
closed forms of the same mechanics.

Present contact: two free segments crossing at right angles with their
capsules overlapping. One present pass must leave the segments exactly 2r
apart along the common normal, keep the mass-weighted centre fixed, and
split the correction by inverse mass and barycentric weight (reference
solve_self_contact, src/metal/ClothBag.metal:2125-2448).

For one candidate pair the hash redesign's Jacobi mean equals the
reference's sequential update, so both pins apply to it directly.
"""

import numpy as np
import pytest

import probes

DT = float(np.float32(0.01))
RADIUS = float(np.float32(0.004))
SELF_FRICTION = float(np.float32(0.34))


def _segment_distance(a0, a1, b0, b1):
    """Closest distance of two segments (textbook clamped solve, in FP64)."""
    da, db, r = a1 - a0, b1 - b0, a0 - b0
    a, e, f = da @ da, db @ db, db @ r
    c, b = da @ r, da @ db
    den = a * e - b * b
    s = min(max((b * f - c * e) / den, 0.0), 1.0) if den > 0 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, min(max(-c / a, 0.0), 1.0)
    elif t > 1:
        t, s = 1.0, min(max((b - c) / a, 0.0), 1.0)
    return float(np.linalg.norm((a0 + da * s) - (b0 + db * t)))


def _probe_metrics(scene, out, impulses):
    """The reference's probe metrics (tools/cloth_bag_metal.mm:7321-7445):
    velocities before friction are (x_final - x_initial) / dt, the contact
    velocity uses the FP64 oracle's impulse normal and endpoint weights."""
    x = out["positions"]
    v = out["velocities"]
    x0 = scene.particles["positionAndInverseMass"][:, :3].astype(np.float64)
    before = (x - x0) / DT
    row = impulses[0]
    n = row[3:6] / np.linalg.norm(row[3:6])
    first = row[7:9] / row[7:9].sum()
    second = row[9:11] / row[9:11].sum()

    def slip(vel):
        rel = vel[0] * first[0] + vel[1] * first[1] - vel[2] * second[0] - vel[3] * second[1]
        return float(np.linalg.norm(rel - n * (rel @ n)))

    return {
        "final_separation": _segment_distance(x[0], x[1], x[2], x[3]) - 0.008,
        "final_moving_height": 0.5 * (x[2, 2] + x[3, 2]),
        "slip_before": slip(before),
        "slip_after": slip(v),
        "energy_before": 0.5 * float(np.sum(before * before)),
        "energy_after": 0.5 * float(np.sum(v * v)),
        "momentum_error": float(np.linalg.norm(v.sum(axis=0) - before.sum(axis=0))),
    }


@pytest.fixture(scope="module")
def ccd(cloth, oracle):
    scene = probes.self_ccd_probe(cloth)
    fp64 = oracle.run(scene, scene.configs(1), 0, 0)
    fp32 = oracle.run(scene, scene.configs(1), 0, 0, precision="fp32")
    return scene, fp64, fp32


def test_self_ccd_probe_fp32_reproduces_the_published_values(ccd):
    scene, fp64, fp32 = ccd
    assert fp32["failure"] == 0
    assert fp32["presentContacts"] == 0 and fp32["sweptContacts"] == 1
    assert fp32["selfFrictionContacts"] == 1
    m = _probe_metrics(scene, fp32, fp64["selfImpulses"])
    print({k: f"{val:.12f}" for k, val in m.items()}, f"cone {fp32['maximumConeRatio']:.12f}")
    published = {  # docs/METAL_CLOTH.md:389-395
        "final_separation": 0.000004277196, "final_moving_height": -0.036001820117,
        "slip_before": 2.000063489379, "slip_after": 0.000050152179,
        "energy_before": 157.924313130071, "energy_after": 155.924161600600,
        "momentum_error": 0.000000478928,
    }
    for key, value in published.items():
        assert m[key] == pytest.approx(value, abs=1e-12 * max(1.0, abs(value)) + 1e-12), key
    assert fp32["maximumConeRatio"] == pytest.approx(0.668487787247, abs=1e-12)


def test_self_ccd_probe_fp64_closed_forms(ccd):
    """In FP64 the moving segment (z from 0.08 to -0.08, x from 0 to 0.02)
    first comes within 2r of the fixed one at t* = (0.08 - 2r) / 0.16; the
    conservative advancement stops within its 1e-9 m tolerance of that
    time. The remaining advance below the impact plane is
    (1 - t*) * 0.16 = 0.088 m along +z (the normal points from the fixed
    segment to the moving one), shared by the four unit masses in
    proportion to their barycentric weights: with contact at x = 0.02 t*
    on the fixed segment (weight (1 + 0.02 t*) / 2) and at the moving
    segment's midpoint, lambda = 0.088 / (ua^2 + wa^2 + 1/2). Friction
    then removes the 2 m/s contact slip (the cone mu * lambda / dt is not
    reached), so slip_after = 0 and the cone ratio is (2 / den) / (mu J)."""
    scene, fp64, _ = ccd
    x, v = fp64["positions"], fp64["velocities"]
    assert fp64["presentContacts"] == 0 and fp64["sweptContacts"] == 1
    t_star = (0.08 - 2 * RADIUS) / 0.16
    wa = (1.0 + 0.02 * t_star) / 2.0
    den = (1 - wa) ** 2 + wa ** 2 + 0.5
    lam = (0.16 * (1.0 - t_star)) / den
    tol = 2e-9  # the advancement's 1e-9 m gap tolerance, seen through lambda
    assert x[2, 2] == pytest.approx(-0.08 + 0.5 * lam, abs=tol)
    assert x[3, 2] == pytest.approx(-0.08 + 0.5 * lam, abs=tol)
    assert x[0, 2] == pytest.approx(-(1 - wa) * lam, abs=tol)
    assert x[1, 2] == pytest.approx(-wa * lam, abs=tol)
    assert np.allclose(x[2:, 0], 0.02, atol=1e-15) and np.allclose(x[:2, 0], [-1.0, 1.0], atol=1e-15)
    # The unequal endpoint shares tilt the fixed segment, so the pair ends a
    # few micrometres beyond 2r (the reference's gate is 5e-6, :7766); the
    # separation is that of the closed-form endpoints.
    closed = np.array([[-1.0, 0.0, -(1 - wa) * lam], [1.0, 0.0, -wa * lam],
                       [0.02, -1.0, -0.08 + 0.5 * lam], [0.02, 1.0, -0.08 + 0.5 * lam]])
    separation = _segment_distance(x[0], x[1], x[2], x[3])
    assert separation == pytest.approx(_segment_distance(*closed), abs=tol)
    assert 0.0 < separation - 2 * RADIUS <= 5e-6
    row = fp64["selfImpulses"][0]
    assert (row[0], row[1], row[2]) == (0, 1, 1)
    assert row[6] == pytest.approx(lam / DT, rel=1e-7)
    assert np.allclose(row[3:6] / row[6], [0.0, 0.0, -1.0], atol=1e-12)
    m = _probe_metrics(scene, fp64, fp64["selfImpulses"])
    assert m["slip_before"] == pytest.approx(2.0, abs=1e-9)
    assert m["slip_after"] < 1e-12
    assert m["momentum_error"] < 1e-12
    assert m["energy_after"] < m["energy_before"]
    fw = row[7:9] / row[7:9].sum()
    sw = row[9:11] / row[9:11].sum()
    friction_den = fw @ fw + sw @ sw
    assert fp64["maximumConeRatio"] == pytest.approx((2.0 / friction_den) / (SELF_FRICTION * row[6]), rel=1e-9)


def _crossing(cloth, inverse_masses, offset_x, overlap=0.0015):
    """Segment A along x from -1 cm to 1 cm; segment B along y, crossing A at
    x = offset_x, `overlap` closer than 2r above it. No motion, no gravity."""
    h = 2.0 * RADIUS - overlap
    p = probes.particles(cloth, [(-0.01, 0, 0), (0.01, 0, 0), (offset_x, -0.01, h), (offset_x, 0.01, h)],
                         inverse_masses)
    d = np.concatenate([probes.distance(cloth, 0, 1, 0.02), probes.distance(cloth, 2, 3, 0.02)])
    c = probes.config(cloth, particles=4, distances=2, dt=1.0 / 960.0, radius=RADIUS)
    c["clothMaterial"] = (RADIUS, 0.02, 0.45, 0.0)
    return cloth.Scene(c, particles=p, distances=d, distanceBatches=probes.batch(cloth, 0, 2, 0), selfContact=1)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("inverse_masses,offset_x", [([1.0, 1.0, 1.0, 1.0], 0.0),
                                                     ([1.0, 3.0, 2.0, 0.5], 0.004),
                                                     ([0.0, 2.0, 1.0, 1.0], -0.006)])
def test_present_contact_ends_two_radii_apart(cloth, oracle, precision, inverse_masses, offset_x):
    """One iteration, no strain sweeps: the present pass is the only pass
    that moves a knot (the push is along z, normal to both segments, and
    the distances start at rest length). Each knot moves along z by
    w_i * share_i * lambda with lambda = overlap / sum(w_i share_i^2), so
    the two contact points (A at weight wa, B at its midpoint) end exactly
    2r apart and sum m_i dx_i = 0 over free knots. Unequal shares tilt a
    segment, which can only bring other points closer: the segment distance
    is at most 2r, and equals it for the symmetric case."""
    scene = _crossing(cloth, inverse_masses, offset_x)
    x0 = scene.particles["positionAndInverseMass"][:, :3].astype(np.float64)
    out = oracle.run(scene, scene.configs(1), 1, 0, precision=precision)
    x = out["positions"]
    tol = 1e-15 if precision == "fp64" else 2e-9
    assert out["presentContacts"] == 1 and out["sweptContacts"] == 0
    wa = (x0[2, 0] - x0[0, 0]) / (x0[1, 0] - x0[0, 0])  # from the float32 start positions
    contact_a = (1 - wa) * x[0] + wa * x[1]
    contact_b = 0.5 * (x[2] + x[3])
    assert np.linalg.norm(contact_b - contact_a) == pytest.approx(2 * RADIUS, abs=tol)
    separation = _segment_distance(x[0], x[1], x[2], x[3])
    assert separation <= 2 * RADIUS + tol
    if inverse_masses == [1.0, 1.0, 1.0, 1.0]:
        assert separation == pytest.approx(2 * RADIUS, abs=tol)
    dx = x - x0
    assert np.max(np.abs(dx[:, :2])) <= tol
    w = np.asarray(inverse_masses, np.float64)
    shares = np.array([1 - wa, wa, 0.5, 0.5])
    overlap = 2 * RADIUS - (x0[2, 2] - x0[0, 2])  # from the float32 start positions
    lam = overlap / float(np.sum(w * shares * shares))
    expected = np.array([-1.0, -1.0, 1.0, 1.0]) * w * shares * lam
    assert np.allclose(dx[:, 2], expected, rtol=0, atol=tol + 1e-12 * lam)
    free = w > 0
    if np.all(free):
        assert abs(float(np.sum(dx[:, 2] / w))) <= (1e-15 if precision == "fp64" else 1e-8) * float(np.sum(1 / w))
