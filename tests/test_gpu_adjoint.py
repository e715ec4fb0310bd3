"""Specimen adjoint on the HIP path (needs a GPU).

The forward half runs the specimen kernels (one workgroup per specimen, all
state in LDS); it must leave the batch exactly where the plain batch step
leaves it. The gradients (FP32, checkpoints every 16 substeps, in-segment
recomputation) are compared with the oracle's FP32 twin of the adjoint (same
operation order: equal up to the order of per-segment partial sums), and
with the FP64 adjoint (pinned by finite differences in
tests/test_oracle_adjoint.py) as closely as FP32 allows: no farther than the
FP32 twin is from FP64. Uniform mass scaling is a near-symmetry of XPBD
(exact with compliance scaled alongside), so the density derivative is a
relative sensitivity of ~1e-7, at FP32's resolution; the damping derivative
is resolved, and is also checked against central differences of the FP64
loss at h = 1e-6.
"""

import numpy as np
import pytest

import adjoint_problem
from test_oracle_adjoint import _analytic, _damping, _fd

pytestmark = pytest.mark.gpu


def _gpu(cloth, p, checkpoint=16):
    tracked_local = p["local"]
    targets = p["targets"].reshape(len(p["configs"]), 2, tracked_local.shape[1], 3)
    with cloth.Batch(p["specimens"]) as batch:
        batch.adjoint(p["configs"], p["iterations"], p["sweeps"], tracked_local, targets, checkpoint)
        grads = [batch.gradients(k) for k in range(2)]
        states = [batch.read(k) for k in range(2)]
    return grads, states


@pytest.fixture(scope="module")
def problem(cloth, oracle):
    p = adjoint_problem.build(cloth)
    args = (p["merged"], p["configs"], p["iterations"], p["sweeps"], p["tracked"], p["targets"])
    p["out"] = oracle.adjoint(*args)
    p["twin"] = oracle.adjoint(*args, precision="fp32")
    p["gpu"], p["states"] = _gpu(cloth, p)
    return p


def test_specimen_forward_equals_plain_batch_step(cloth, problem):
    p = problem
    with cloth.Batch(p["specimens"]) as batch:
        batch.step(p["configs"], p["iterations"], p["sweeps"])
        plain = [batch.read(k) for k in range(2)]
    for k in range(2):
        assert problem["states"][k]["failure"] == 0
        for name in ("particles", "distances", "bends"):
            assert p["states"][k][name].tobytes() == plain[k][name].tobytes(), f"specimen {k} {name}"


def test_loss_matches_oracle(problem):
    p = problem
    loss = sum(g["loss"] for g in p["gpu"])
    print(f"loss gpu {loss:.9e} twin {p['twin']['loss']:.9e} fp64 {p['out']['loss']:.9e}")
    assert loss == pytest.approx(p["twin"]["loss"], rel=1e-9)
    assert loss == pytest.approx(p["out"]["loss"], rel=1e-4)


def _gpu_full(p, name):
    """GPU per-record gradients scattered into merged order."""
    key = {"distanceCompliance": "distances", "bendCompliance": "bends", "inverseMass": "particles",
           "damping": "particles"}[name]
    out = np.zeros_like(p["out"][name])
    for k in range(2):
        out[p["maps"][k][key]] = p["gpu"][k][name]
    return out


NAMES = ["distanceCompliance", "bendCompliance", "inverseMass", "damping"]


@pytest.mark.parametrize("name", NAMES)
def test_gradients_match_fp32_twin(problem, name):
    p = problem
    gpu, twin = _gpu_full(p, name), p["twin"][name]
    error = np.linalg.norm(gpu - twin) / np.linalg.norm(twin)
    print(f"{name}: |gpu - twin| / |twin| = {error:.3e}")
    assert error < 1e-4


@pytest.mark.parametrize("name", NAMES)
def test_gradients_as_close_to_fp64_as_fp32_allows(problem, name):
    p = problem
    gpu, twin, ref = _gpu_full(p, name), p["twin"][name], p["out"][name]
    e_gpu = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    e_twin = np.linalg.norm(twin - ref) / np.linalg.norm(ref)
    print(f"{name}: |gpu - fp64| {e_gpu:.3e}  |twin - fp64| {e_twin:.3e}")
    assert e_gpu <= 1.5 * e_twin + 1e-4
    assert e_gpu < 0.1


def test_specimen_parameters(problem):
    p = problem
    m = p["merged"]
    w = m.particles["positionAndInverseMass"][:, 3].astype(np.float64)
    damping = m.particleMotion[:, 3].astype(np.float64)
    gw, gc = _gpu_full(p, "inverseMass"), _gpu_full(p, "damping")
    ref = _analytic(p)
    twin = _analytic(dict(p, out=p["twin"]))
    for k in range(2):
        idx = p["maps"][k]["particles"]
        density = -np.sum(gw[idx] * w[idx])
        damp = np.sum(gc[idx] * damping[idx])
        print(f"specimen {k}: density {density:.6e} twin {twin[f'density{k}']:.6e} fp64 {ref[f'density{k}']:.6e}; "
              f"damping {damp:.6e} twin {twin[f'damping{k}']:.6e} fp64 {ref[f'damping{k}']:.6e}")
        # Both sums cancel heavily (relative sensitivities ~1e-7 for density);
        # per-segment partial sums are added in a different order than the
        # twin's running sum, so the twin gate is 1e-2.
        assert density == pytest.approx(twin[f"density{k}"], rel=1e-2)
        assert damp == pytest.approx(twin[f"damping{k}"], rel=1e-2)
    # The first specimen's damping derivative is resolved in FP32.
    idx = p["maps"][0]["particles"]
    assert np.sum(gc[idx] * damping[idx]) == pytest.approx(ref["damping0"], rel=1e-3)


def test_damping_gradient_matches_finite_differences(oracle, problem):
    p = problem
    m = p["merged"]
    idx = p["maps"][0]["particles"]
    damping = m.particleMotion[:, 3].astype(np.float64)
    gpu = np.sum(_gpu_full(p, "damping")[idx] * damping[idx])
    fd = _fd(oracle, p, 1e-6, damping=_damping(m, idx))
    print(f"damping gradient gpu {gpu:.6e} central difference (h = 1e-6) {fd:.6e}")
    assert gpu == pytest.approx(fd, rel=1e-3)


def test_checkpoint_interval_does_not_change_gradients(cloth, problem):
    p = problem
    grads, _ = _gpu(cloth, p, checkpoint=7)
    # Same arithmetic; only the order of the per-segment partial sums of the
    # per-knot gradients changes.
    for k in range(2):
        for name in ("distanceCompliance", "bendCompliance", "inverseMass", "damping"):
            a, b = grads[k][name], p["gpu"][k][name]
            assert np.allclose(a, b, rtol=1e-4, atol=1e-6 * np.max(np.abs(b))), name


def test_adjoint_reports_the_strain_passes_it_tapes(cloth):
    """Strips pulled 20 % beyond rest (limit 10 %) and stepped without
    distance iterations, so the strain limit acts in the first substeps and
    some strain passes are taped, never more than the recomputations ran
    (substeps x 18 invocations x sweeps x colours x specimens). With eight
    iterations the stiff distance constraints pull the strips back within
    the limit before any strain pass."""
    from test_oracle_certificate import _stretched_strip
    specimens = [_stretched_strip(cloth, 1.2, 6, 0), cloth.Scene.strip(width=5, height=8, seed=4)]
    merged, _ = cloth.Scene.merge(specimens)
    colours = int(merged.distanceBatches["control"][:, 2].max()) + 1
    substeps, sweeps = 32, 2
    tracked = np.array([[10, 14], [20, 30]], np.uint32)
    targets = np.zeros((substeps, 2, 2, 3), np.float32)
    taped = {}
    for iterations in (0, 8):
        with cloth.Batch(specimens) as batch:
            batch.adjoint(specimens[0].configs(substeps), iterations, sweeps, tracked, targets, 16)
            stats = batch.adjoint_stats()
        print(f"{iterations} iterations: strain passes {stats.strainPasses}, taped {stats.strainPassesTaped}")
        assert stats.strainPasses == substeps * 18 * sweeps * colours * 2
        taped[iterations] = stats.strainPassesTaped
    assert 0 < taped[0] < substeps * 18 * sweeps * colours * 2
    assert taped[8] == 0


def test_gradients_when_strain_limits_act(cloth, oracle):
    """The strained problem (strain limits act in most substeps): GPU
    gradients equal the FP32 twin's and are as close to FP64 as the twin is;
    the adjoint reports the strain passes it taped."""
    p = adjoint_problem.build_strained(cloth)
    args = (p["merged"], p["configs"], p["iterations"], p["sweeps"], p["tracked"], p["targets"])
    ref = oracle.adjoint(*args)
    twin = oracle.adjoint(*args, precision="fp32")
    targets = p["targets"].reshape(len(p["configs"]), 2, p["local"].shape[1], 3)
    with cloth.Batch(p["specimens"]) as batch:
        batch.adjoint(p["configs"], p["iterations"], p["sweeps"], p["local"], targets, 16)
        gpu = [batch.gradients(k) for k in range(2)]
        stats = batch.adjoint_stats()
    print(f"strained problem: {stats.strainPassesTaped} of {stats.strainPasses} strain passes taped")
    assert stats.strainPassesTaped > 0
    q = dict(p, gpu=gpu, out=ref)
    for name in NAMES:
        g = _gpu_full(q, name)
        e_twin = np.linalg.norm(g - twin[name]) / np.linalg.norm(twin[name])
        e_gpu = np.linalg.norm(g - ref[name]) / np.linalg.norm(ref[name])
        e_ref = np.linalg.norm(twin[name] - ref[name]) / np.linalg.norm(ref[name])
        print(f"{name}: |gpu - twin| {e_twin:.3e}  |gpu - fp64| {e_gpu:.3e}  |twin - fp64| {e_ref:.3e}")
        assert e_twin < 1e-4
        assert e_gpu <= 1.5 * e_ref + 1e-4
