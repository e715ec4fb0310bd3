"""The data-dependent certificate loop of the independent CPU reference
(tools/cloth_bag.cpp:5143-5185) in the oracle (CPU only).

At the last substep of a call with extension.certificatePasses = N the
schedule's certificate round (self-contact, strain sweeps, final contacts)
repeats until one round leaves every residual below the reference's
thresholds (strain-limit violation, fruit/yarn and self penetration < 1e-6,
ground < 1e-9), at most N rounds. The residual measures are pinned by closed
forms on hand-made states; the loop by its stopping rule, its bound, and a
settled frame that one round leaves unchanged.
"""

import numpy as np
import pytest

import probes

F32 = np.float32


def _cap(scene, passes):
    scene.certificatePasses = passes
    scene.refresh()
    return scene


def _stretched_strip(cloth, stretch, height, passes):
    """A 3-wide strip pinned at its top row, every row pulled down so each
    vertical link is `stretch` times its rest length (beyond the 10 % limit)."""
    s = cloth.Scene.strip(width=3, height=height, driveAmplitude=0.0, jitter=0.0)
    x = s.particles["positionAndInverseMass"]
    top = x[0, 2]
    x[:, 2] = top + (x[:, 2] - top) * F32(stretch)
    s.particles["previousAndMass"][:, :3] = x[:, :3]
    return _cap(s, passes)


def _state_scene(cloth, scene, out):
    """The scene with its knots moved to an oracle run's final positions."""
    arrays = {name: getattr(scene, name).copy() for name, _, _ in scene.ARRAYS}
    arrays["particles"]["positionAndInverseMass"][:, :3] = out["positions"]
    arrays["particles"]["previousAndMass"][:, :3] = out["positions"]
    return cloth.Scene(scene.config, staticSpheres=scene.staticSpheres, selfContact=scene.selfContact,
                       selfContactMargin=scene.selfContactMargin, **arrays)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_strain_residual_closed_form(cloth, oracle, precision):
    # Knot 1 at 1.5 from knot 0 (rest 1, limit 0.285): violation 0.215;
    # the compressed pair adds nothing.
    scene = probes.strain_probe(cloth)
    r = oracle.certificate_residuals(scene, precision=precision)
    if precision == "fp64":
        expected = 1.5 - 1.0 * (1.0 + float(F32(0.285)))
    else:
        expected = float(F32(1.5) - F32(1.0) * (F32(1.0) + F32(0.285)))
    assert r[0] == pytest.approx(expected, abs=1e-15)
    assert r[1] == r[2] == r[3] == 0.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ground_residual_closed_form(cloth, oracle, precision):
    # A knot at z = 0 under a 4 mm yarn radius; a unit fruit sunk half its radius.
    knot = oracle.certificate_residuals(probes.ground_contact_cloth_probe(cloth), precision=precision)
    assert knot[1] == float(F32(0.004))
    both = oracle.certificate_residuals(probes.ground_contact_probe(cloth), precision=precision)
    assert both[1] == 0.5
    # Without the ground the same states have no ground residual.
    scene = probes.ground_contact_cloth_probe(cloth)
    scene.config["constraintCounts"][0, 2] = 0
    assert oracle.certificate_residuals(scene, precision=precision)[1] == 0.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fruit_yarn_residual_closed_form(cloth, oracle, precision):
    # The fruit of the yarn CCD probe (radius 2 cm) 1 cm from the middle of
    # its 20 cm segment: penetration 2 cm + 4 mm - 1 cm; 10 cm away: none.
    scene = probes.yarn_ccd_probe(cloth)
    scene.fruits["positionAndInverseMass"][0, :3] = (0.01, 0.0, 1.0)
    scene.refresh()
    r = oracle.certificate_residuals(scene, precision=precision)
    if precision == "fp64":
        expected = float(F32(0.02)) + float(F32(0.004)) - float(F32(0.01))
    else:
        expected = float(F32(0.02) + F32(0.004) - F32(0.01))
    assert r[2] == pytest.approx(expected, abs=1e-15)
    far = probes.yarn_ccd_probe(cloth)
    assert oracle.certificate_residuals(far, precision=precision)[2] == 0.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_self_residual_closed_form(cloth, oracle, precision):
    # Crossing segments 2r - 1 mm apart: self penetration 1 mm; 1 mm beyond
    # contact: none.
    scene = probes.crossing_segments_probe(cloth, 0.3, slip=0.0, overlap=0.001)
    h = F32(2.0 * 0.004 - 0.001)
    r = oracle.certificate_residuals(scene, precision=precision)
    expected = 2.0 * float(F32(0.004)) - float(h) if precision == "fp64" else float(F32(2.0) * F32(0.004) - h)
    assert r[3] == pytest.approx(expected, abs=1e-15)
    clear = probes.crossing_segments_probe(cloth, 0.3, slip=0.0, overlap=-0.001)
    assert oracle.certificate_residuals(clear, precision=precision)[3] == 0.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_loop_stops_at_the_first_certified_round(cloth, oracle, precision):
    scene = _stretched_strip(cloth, 1.5, 4, 128)
    out = oracle.run(scene, scene.configs(1), 0, 1, precision=precision)
    k = out["certificatePasses"]
    assert 2 <= k < 128
    # The same run capped at k rounds is the same run; capped at k - 1 it
    # stops uncertified.
    exact = oracle.run(_cap(scene, k), scene.configs(1), 0, 1, precision=precision)
    assert exact["certificatePasses"] == k
    assert np.array_equal(exact["positions"], out["positions"])
    short = oracle.run(_cap(scene, k - 1), scene.configs(1), 0, 1, precision=precision)
    assert short["certificatePasses"] == k - 1
    assert not np.array_equal(short["positions"], out["positions"])
    # The stopping rule on the FP64 states, restated with numpy: strain-limit
    # violation max(len - rest (1 + limit)) below 1e-6 after k rounds only.
    if precision == "fp64":
        ij = scene.distances["particlesAndColor"][:, :2].astype(np.int64)
        m = scene.distances["material"].astype(np.float64)
        for run, below in ((out, True), (short, False)):
            x = run["positions"]
            violation = np.max(np.linalg.norm(x[ij[:, 1]] - x[ij[:, 0]], axis=1) - m[:, 0] * (1.0 + m[:, 3]))
            assert (max(violation, 0.0) < 1e-6) == below
        assert oracle.certificate_residuals(_state_scene(cloth, scene, short))[0] >= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_loop_runs_to_its_bound(cloth, oracle, precision):
    # Pinned crossing segments 1 mm into contact: nothing can move them, so
    # no round is certified and the loop stops at its bound.
    scene = probes.crossing_segments_probe(cloth, 0.3, slip=0.0, overlap=0.001)
    for field in ("positionAndInverseMass", "previousAndMass"):
        scene.particles[field][:, 3] = 0.0
    for bound in (1, 5):
        _cap(scene, bound)
        out = oracle.run(scene, scene.configs(2), 2, 1, precision=precision)
        assert out["certificatePasses"] == bound
    # Without strain sweeps the schedule has no certificate rounds.
    assert oracle.run(scene, scene.configs(2), 2, 0, precision=precision)["certificatePasses"] == 0


def test_settled_bag_frame_certifies_in_one_round(cloth, oracle):
    """The bag's first frame ends settled: one certificate round finds every
    residual below threshold and changes nothing."""
    plain = cloth.Scene.bag(selfContact=1)
    certified = _cap(cloth.Scene.bag(selfContact=1), 128)
    configs = plain.configs(48)
    a = oracle.run(plain, configs, 32, 3, precision="fp32")
    b = oracle.run(certified, configs, 32, 3, precision="fp32")
    assert a["certificatePasses"] == 0 and b["certificatePasses"] == 1
    for key in ("positions", "velocities", "fruits", "distanceLambda"):
        assert np.array_equal(a[key], b[key]), key
