"""The FP64 adjoint oracle (CPU only): its forward is the oracle's, and its
gradients equal central finite differences of the loss.

Every gradient family (distance compliance, bend compliance, density,
damping) of both specimens is checked at a step of h = 1e-6 relative, along
three directions each: every entry of the family scaled together, a seeded
positive random mix of the entries (a dropped or misindexed entry changes
the sum) and the family's largest single entry.

At h = 1e-6 the loss change of the weaker families (about 1e-13 of the
loss) lies below the rounding noise of the FP64 trajectory itself (tens of
ulp per tracked position after 200 substeps): FP64 differences there scatter
by 1e-2 to 1e-1 relative whatever the summation. Those differences are taken
of the same algorithm run in x87 extended precision (64-bit significand,
2^11 finer), formed term by term from the two tracked trajectories; the
analytic gradients stay those of the FP64 adjoint. The families whose loss
change stands far above the FP64 noise are also checked with FP64
differences at h = 1e-6, and the full set with FP64 differences at h = 1e-3.
"""

import math

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


def _family(p, k, family):
    """(override name, entry count, entries of specimen k, analytic gradient
    in the log of each entry) of one gradient family."""
    m, out = p["merged"], p["out"]
    maps = p["maps"][k]
    if family == "distance":
        return "distance_scale", len(m.distances), maps["distances"], out["distanceCompliance"] * m.distances["material"][:, 1]
    if family == "bend":
        return "bend_scale", len(m.bends), maps["bends"], out["bendCompliance"] * m.bends["material"][:, 2]
    w = m.particles["positionAndInverseMass"][:, 3].astype(np.float64)
    if family == "density":
        return "inverse_mass_scale", len(m.particles), maps["particles"], -out["inverseMass"] * w
    damping = m.particleMotion[:, 3].astype(np.float64)
    return "damping", len(m.particles), maps["particles"], out["damping"] * damping


def _direction(kind, n, idx, gradient, seed):
    u = np.zeros(n)
    if kind == "every":
        u[idx] = 1.0
    elif kind == "mix":
        u[idx] = np.random.default_rng(seed).uniform(0.5, 1.5, len(idx))
    else:
        u[idx[np.argmax(np.abs(gradient[idx]))]] = 1.0
    return u


def _override(p, name, u, step):
    """Parameters scaled by (1 + step u): compliances and damping directly,
    density through the inverse masses."""
    factor = 1.0 + step * u
    if name == "inverse_mass_scale":
        return {name: 1.0 / factor}
    if name == "damping":
        return {name: p["merged"].particleMotion[:, 3].astype(np.float64) * factor}
    return {name: factor}


def _trajectory(oracle, p, precision, overrides):
    out = oracle.adjoint(p["merged"], p["configs"], p["iterations"], p["sweeps"], p["tracked"], p["targets"],
                         gradients=False, precision=precision, **overrides)
    return out["trajectory"], out["trajectoryLow"]


def _fd_terms(oracle, p, h, precision, name, u):
    """Central difference (L(1 + h u) - L(1 - h u)) / 2h with the loss
    difference formed term by term, sum (x+ - x-) . (x+ + x- - 2 target),
    from the tracked trajectories (high and low parts)."""
    plus = _trajectory(oracle, p, precision, _override(p, name, u, h))
    minus = _trajectory(oracle, p, precision, _override(p, name, u, -h))
    difference = (plus[0] - minus[0]) + (plus[1] - minus[1])
    total = (plus[0] + minus[0] - 2.0 * p["targets"]) + (plus[1] + minus[1])
    return math.fsum((difference * total).ravel()) / (2.0 * h)


@pytest.mark.parametrize("specimen", [0, 1])
@pytest.mark.parametrize("family", ["distance", "bend", "density", "damping"])
def test_every_gradient_family_at_1e6(cloth, oracle, problem, specimen, family):
    p = problem
    name, n, idx, gradient = _family(p, specimen, family)
    for seed, kind in enumerate(("every", "mix", "largest")):
        u = _direction(kind, n, idx, gradient, seed=17 + 3 * specimen + seed)
        analytic = float(np.sum(u * gradient))
        fd = _fd_terms(oracle, p, 1e-6, "extended", name, u)
        assert analytic == pytest.approx(fd, rel=1e-3), (family, specimen, kind, analytic, fd)


@pytest.mark.parametrize("specimen,family,kind", [(0, "damping", "every"), (0, "density", "largest"),
                                                  (1, "density", "largest"), (1, "damping", "largest")])
def test_strong_families_at_1e6_in_fp64(cloth, oracle, problem, specimen, family, kind):
    p = problem
    name, n, idx, gradient = _family(p, specimen, family)
    u = _direction(kind, n, idx, gradient, seed=0)
    assert float(np.sum(u * gradient)) == pytest.approx(_fd_terms(oracle, p, 1e-6, "fp64", name, u), rel=1e-3)


def test_extended_trajectory_is_the_fp64_trajectory(cloth, oracle, problem):
    """The extended-precision run is the same algorithm: its tracked
    trajectory differs from the FP64 one by rounding only."""
    p = problem
    high, low = _trajectory(oracle, p, "extended", {})
    reference, zero = _trajectory(oracle, p, "fp64", {})
    assert not np.any(zero)
    assert np.max(np.abs((high - reference) + low)) <= 1e-12
    loss = oracle.adjoint(p["merged"], p["configs"], p["iterations"], p["sweeps"], p["tracked"], p["targets"],
                          gradients=False)["loss"]
    assert math.fsum(((reference - p["targets"]) ** 2).ravel()) == pytest.approx(loss, rel=1e-13)


@pytest.mark.parametrize("specimen", [0, 1])
def test_gradients_when_strain_limits_act(cloth, oracle, specimen):
    """Every family of both specimens of a problem whose strain limits act
    (a strip pulled 50 % beyond rest, one iteration per substep), against
    extended-precision central differences at h = 1e-6."""
    p = adjoint_problem.build_strained(cloth)
    p["out"] = oracle.adjoint(p["merged"], p["configs"], p["iterations"], p["sweeps"], p["tracked"], p["targets"])
    for family in ("distance", "bend", "density", "damping"):
        name, n, idx, gradient = _family(p, specimen, family)
        u = _direction("every", n, idx, gradient, seed=0)
        analytic = float(np.sum(u * gradient))
        fd = _fd_terms(oracle, p, 1e-6, "extended", name, u)
        assert analytic == pytest.approx(fd, rel=1e-4), (family, specimen, analytic, fd)


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
