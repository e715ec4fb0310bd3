"""Calibration sensitivity columns from the HIP adjoint (needs a GPU).

One batch of three strips, one adjoint run per observable (linear
observable: a knot's height after the last substep): the library's dlog
columns (numi_cloth_calibration_dlog) against the oracle's FP32 twin and
FP64 adjoint of each strip; the observable itself against the FP64
forward; then the columns written as the fitter's CSV and fitted by the
reference's fitter (oracle/_ref, when built).
"""

import numpy as np
import pytest

import calibration_problem as cp
from test_calibration import FITTER, run_fitter

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def columns(cloth, oracle):
    scenes = cp.specimens(cloth)
    S, O = len(scenes), len(cp.OBSERVABLES)
    merged, _ = cloth.Scene.merge(scenes)
    configs = merged.configs(cp.SUBSTEPS)
    gpu = np.zeros((S, O, 4))
    gpu_baseline = np.zeros((S, O))
    failure = 0
    for o, name in enumerate(cp.OBSERVABLES):
        # The adjoint advances the batch's state, so every observable starts
        # from a fresh batch at the authored state.
        with cloth.Batch(scenes) as batch:
            tracked = np.array([[cp.observed_knots(s)[name]] for s in scenes], np.uint32)
            weights = np.broadcast_to(cp.weights()[:, None], (cp.SUBSTEPS, S, 1, 3))
            batch.adjoint(configs, cp.ITERATIONS, cp.SWEEPS, tracked, weights, linear=True)
            for s, scene in enumerate(scenes):
                c = cp.columns(scene)
                gpu[s, o], gpu_baseline[s, o] = batch.dlog(s, 4, distance=c["distance"], bend=c["bend"],
                                                          damping=c["damping"], mass=c["mass"])
            failure |= batch.read(0)["failure"]
    fp64 = np.zeros((S, O, 4))
    twin = np.zeros((S, O, 4))
    baseline = np.zeros((S, O))
    for s, scene in enumerate(scenes):
        for o, name in enumerate(cp.OBSERVABLES):
            out = cp.oracle_trial(oracle, scene, name)
            baseline[s, o] = out["loss"]
            fp64[s, o] = cp.dlog_from_gradients(scene, out)
            twin[s, o] = cp.dlog_from_gradients(scene, cp.oracle_trial(oracle, scene, name, precision="fp32"))
    return dict(gpu=gpu, gpu_baseline=gpu_baseline, fp64=fp64, twin=twin, baseline=baseline, failure=failure)


def test_observable_matches_the_fp64_forward(columns):
    assert columns["failure"] == 0
    error = np.max(np.abs(columns["gpu_baseline"] - columns["baseline"]))
    print(f"observable |gpu - fp64| {error:.3e} m")
    assert error <= 1e-5  # positions within 1e-5 of a 0.1 m strip


def test_dlog_columns_match_the_oracle(columns):
    """The GPU columns equal the FP32 twin's up to the order of partial sums
    (1e-3 relative; 1e-2 for the mass column, a sum over every knot of terms
    that cancel to about 1 % of their magnitude), and are no farther from
    FP64 than the twin is (plus the same margin)."""
    gpu, twin, fp64 = columns["gpu"], columns["twin"], columns["fp64"]
    for s in range(gpu.shape[0]):
        for o in range(gpu.shape[1]):
            for j in range(4):
                g, t, f = gpu[s, o, j], twin[s, o, j], fp64[s, o, j]
                print(f"strip {s} {cp.OBSERVABLES[o]} column {j}: gpu {g:.6e} twin {t:.6e} fp64 {f:.6e}")
    for s in range(gpu.shape[0]):
        for o in range(gpu.shape[1]):
            for j in range(4):
                g, t, f = gpu[s, o, j], twin[s, o, j], fp64[s, o, j]
                margin = 1e-2 if j == 3 else 1e-3
                assert abs(g - t) <= margin * abs(t) + 1e-12
                assert abs(g - f) <= 1.5 * abs(t - f) + margin * abs(f) + 1e-12


def test_mass_column_is_the_sum_of_compliance_columns(columns):
    gpu = columns["gpu"]
    for s in range(gpu.shape[0]):
        for o in range(gpu.shape[1]):
            axial, bend, _, mass = gpu[s, o]
            assert abs(mass - (axial + bend)) <= 1e-2 * (abs(axial) + abs(bend))


@pytest.mark.skipif(not FITTER.exists(), reason="reference fitter not built (no reference checkout)")
def test_reference_fitter_recovers_log_deltas_from_gpu_columns(cloth, tmp_path, columns):
    truth = np.array([-0.12, 0.18, 0.07])
    rows = cp.synthetic_rows(columns["gpu_baseline"], columns["gpu"][:, :, :3], truth, cp.SIGMA)
    code, report, result = run_fitter(tmp_path, cp.PARAMETERS, rows, cloth)
    print(result.stdout)
    assert code == 0, result.stderr
    assert report["qualified"] == "true"
    for name, value in zip(cp.PARAMETERS, truth):
        assert report["log_delta"][name] == pytest.approx(value, abs=1e-5)
