"""Self-contact impulse aggregation and friction in the oracle (CPU).

Two free segments crossing at right angles, one sliding over the other
(tests/probes.py crossing_segments_probe). Each knot has one contact, so
the Jacobi mean equals the reference's ordered pass (reference
src/metal/ClothBag.metal:2580-2731) and closed forms hold:
  * momentum: the normal response and the friction impulse act in equal
    and opposite pairs on unit masses, so sum(v_x) stays 2 * slip (the
    float32 slip the particle records hold);
  * limited branch (small mu): the tangential impulse is mu * J, where J is
    the substep's summed normal impulse lambda / dt; the normal push moves
    A's two knots by -lambda * n in total, so J = -(z_A0 + z_A1) / dt and
    friction adds mu * J to A's summed v_x (over the same run with mu = 0,
    where A only gets the tiny x motion of its own tilt);
  * unlimited branch (large mu): the impulse removes the weighted relative
    slip, so A and B end with the same mean x velocity up to the squared
    offset of the contact point from A's midpoint (about 1e-4 here);
  * mu = 0: no friction contact.
"""

import numpy as np
import pytest

from probes import crossing_segments_probe

SLIP = 0.2
DT = 1.0 / 960.0


def run(cloth, oracle, mu, precision="fp64"):
    scene = crossing_segments_probe(cloth, mu, slip=SLIP, dt=DT)
    return oracle.run(scene, scene.configs(1), 4, 0, precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_momentum_is_conserved(cloth, oracle, precision):
    out = run(cloth, oracle, 0.34, precision)
    assert out["selfFrictionContacts"] == 1
    assert out["presentContacts"] >= 1
    momentum = 2 * float(np.float32(SLIP))
    assert abs(out["velocities"][:, 0].sum() - momentum) < (1e-12 if precision == "fp64" else 1e-6)


def test_limited_impulse_is_mu_times_normal_impulse(cloth, oracle):
    mu = 0.01
    out = run(cloth, oracle, mu)
    x = out["positions"]
    normal_impulse = -(x[0, 2] + x[1, 2]) / DT
    assert normal_impulse > 0.5  # about overlap / dt = 0.96
    free = run(cloth, oracle, 0.0)
    assert np.array_equal(free["positions"], out["positions"])  # friction acts on velocities only
    transferred = out["velocities"][:2, 0].sum() - free["velocities"][:2, 0].sum()
    print(f"J = {normal_impulse:.9f}, sum vA.x = {transferred:.9e}, mu J = {mu * normal_impulse:.9e}")
    # J from the z motion assumes a vertical normal; the contact point's
    # 0.2 mm shift tilts it by O(1e-4), hence 1e-6 rather than roundoff.
    assert abs(transferred - mu * normal_impulse) <= 1e-6 * mu * normal_impulse
    assert abs(out["maximumConeRatio"] - 1.0) < 1e-12


def test_unlimited_impulse_removes_the_slip(cloth, oracle):
    out = run(cloth, oracle, 100.0)
    v = out["velocities"][:, 0]
    mean_a, mean_b = v[:2].mean(), v[2:].mean()
    print(f"mean vA.x {mean_a:.6e}, mean vB.x {mean_b:.6e}")
    assert abs(mean_a - SLIP / 2) < 1e-3 * SLIP
    assert abs(mean_b - mean_a) < 1e-3 * SLIP
    assert out["maximumConeRatio"] < 1.0


def test_no_friction_without_coefficient(cloth, oracle):
    out = run(cloth, oracle, 0.0)
    assert out["presentContacts"] >= 1
    assert out["selfFrictionContacts"] == 0
    assert out["maximumConeRatio"] == 0.0
