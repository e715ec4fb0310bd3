"""The FP64 adjoint oracle (CPU only): its forward is the oracle's, and its
gradients equal central finite differences of the FP64 loss.

Finite differences of an FP64 trajectory are limited by rounding below a
step of about 1e-5 relative for the weaker parameters (the loss change
falls under the trajectory's rounding noise), so the full set is checked at
h = 1e-3 (truncation error ~h^2) and the strongest parameter, the first
specimen's damping (loss change far above the rounding noise), at the
h = 1e-6 the calibration contract names.
"""

import numpy as np
import pytest

import adjoint_problem


@pytest.fixture(scope="module")
def problem(cloth, oracle):
    p = adjoint_problem.build(cloth)
    p["out"] = oracle.adjoint(p["merged"], p["configs"], p["iterations"], p["sweeps"], p["tracked"], p["targets"])
    return p


def _loss(oracle, p, **overrides):
    return oracle.adjoint(p["merged"], p["configs"], p["iterations"], p["sweeps"], p["tracked"], p["targets"],
                          gradients=False, **overrides)["loss"]


def _fd(oracle, p, h, **make):
    """Central difference of L in the log of one parameter group."""
    plus = {k: f(1 + h) for k, f in make.items()}
    minus = {k: f(1 - h) for k, f in make.items()}
    return (_loss(oracle, p, **plus) - _loss(oracle, p, **minus)) / (2 * h)


def _scale(n, idx):
    def make(factor):
        s = np.ones(n)
        s[idx] = factor
        return s
    return make


def _density(n, idx):
    # density rho(1 + h) scales the inverse masses by 1 / (1 + h)
    def make(factor):
        s = np.ones(n)
        s[idx] = 1.0 / factor
        return s
    return make


def _damping(merged, idx):
    base = merged.particleMotion[:, 3].astype(np.float64)
    def make(factor):
        d = base.copy()
        d[idx] = base[idx] * factor
        return d
    return make


def test_adjoint_forward_is_the_oracle(cloth, oracle, problem):
    p = problem
    reference = oracle.run(p["merged"], p["configs"], p["iterations"], p["sweeps"])
    assert np.array_equal(p["out"]["positions"], reference["positions"])


def _analytic(p):
    m, out = p["merged"], p["out"]
    w = m.particles["positionAndInverseMass"][:, 3].astype(np.float64)
    damping = m.particleMotion[:, 3].astype(np.float64)
    values = {}
    for k in range(2):
        idx = p["maps"][k]["particles"]
        values[f"density{k}"] = -np.sum(out["inverseMass"][idx] * w[idx])
        values[f"damping{k}"] = np.sum(out["damping"][idx] * damping[idx])
    return values


def test_gradients_match_finite_differences(cloth, oracle, problem):
    p = problem
    m, out = p["merged"], p["out"]
    D, B, P = len(m.distances), len(m.bends), len(m.particles)
    h = 1e-3
    gd = out["distanceCompliance"] * m.distances["material"][:, 1]
    gb = out["bendCompliance"] * m.bends["material"][:, 2]
    checks = []
    for k in np.argsort(-np.abs(gd))[:3]:
        checks.append((gd[k], _fd(oracle, p, h, distance_scale=_scale(D, [k]))))
    for k in np.argsort(-np.abs(gb))[:3]:
        checks.append((gb[k], _fd(oracle, p, h, bend_scale=_scale(B, [k]))))
    analytic = _analytic(p)
    for k in range(2):
        idx = p["maps"][k]["particles"]
        checks.append((analytic[f"density{k}"], _fd(oracle, p, h, inverse_mass_scale=_density(P, idx))))
        checks.append((analytic[f"damping{k}"], _fd(oracle, p, h, damping=_damping(m, idx))))
    for value, fd in checks:
        assert value == pytest.approx(fd, rel=1e-3)


def test_gradients_match_finite_differences_at_1e6(cloth, oracle, problem):
    p = problem
    analytic = _analytic(p)
    idx0 = p["maps"][0]["particles"]
    assert analytic["damping0"] == pytest.approx(_fd(oracle, p, 1e-6, damping=_damping(p["merged"], idx0)), rel=1e-3)


def test_mass_compliance_scaling_identity(cloth, oracle, problem):
    """Scaling every inverse mass and compliance of a specimen by s leaves
    its trajectory unchanged (each projection is homogeneous of degree 0 in
    (w, alpha) with lambda scaled by 1/s), so
    sum_i w_i dL/dw_i = -sum_k c_k dL/dc_k over the specimen (Euler)."""
    p = problem
    m, out = p["merged"], p["out"]
    w = m.particles["positionAndInverseMass"][:, 3].astype(np.float64)
    cd = m.distances["material"][:, 1].astype(np.float64)
    cb = m.bends["material"][:, 2].astype(np.float64)
    for k in range(2):
        idx, di, bi = p["maps"][k]["particles"], p["maps"][k]["distances"], p["maps"][k]["bends"]
        mass = np.sum(out["inverseMass"][idx] * w[idx])
        compliance = np.sum(out["distanceCompliance"][di] * cd[di]) + np.sum(out["bendCompliance"][bi] * cb[bi])
        scale = np.sum(np.abs(out["inverseMass"][idx] * w[idx]))
        assert abs(mass + compliance) <= 1e-12 * scale


def test_unsupported_scene_is_rejected(cloth, oracle):
    scene = cloth.Scene.sheet(width=4, height=4)  # ground contact
    with pytest.raises(RuntimeError, match="status 5"):
        oracle.adjoint(scene, scene.configs(2), 2, 1, np.array([0], np.uint32), np.zeros((2, 1, 3)))
