"""Calibration sensitivity columns on the CPU: the oracle's linear-observable
adjoint, its dlog columns against central differences (the reference
protocol's h = 0.05 and a tight h = 1e-3), the mass/compliance identity, the
observation CSV writer, and the reference's own fitter
(tools/cloth_calibration.cpp, built into oracle/_ref when the reference
checkout is present) reading that CSV and recovering known log deltas.
"""

import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import calibration_problem as cp

FITTER = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "numi-solver-cloth-calibration"


@pytest.fixture(scope="module")
def trials(cloth, oracle):
    """FP64 baselines [strip, observable] and dlog [strip, observable, 4]."""
    scenes = cp.specimens(cloth)
    baselines = np.zeros((len(scenes), len(cp.OBSERVABLES)))
    dlog = np.zeros((len(scenes), len(cp.OBSERVABLES), 4))
    for s, scene in enumerate(scenes):
        for o, name in enumerate(cp.OBSERVABLES):
            out = cp.oracle_trial(oracle, scene, name)
            baselines[s, o] = out["loss"]
            dlog[s, o] = cp.dlog_from_gradients(scene, out)
    return scenes, baselines, dlog


def test_linear_observable_is_the_weighted_final_position(cloth, oracle):
    scene = cp.specimens(cloth)[0]
    knots = cp.observed_knots(scene)
    forward = oracle.run(scene, scene.configs(cp.SUBSTEPS), cp.ITERATIONS, cp.SWEEPS)
    for name in cp.OBSERVABLES:
        out = cp.oracle_trial(oracle, scene, name, gradients=False)
        assert out["loss"] == forward["positions"][knots[name], 2]
        assert np.array_equal(out["positions"], forward["positions"])


def _observable(oracle, scene, name, column, factor):
    """The FP64 observable with every parameter of one column scaled by `factor`."""
    P, D, B = len(scene.particles), len(scene.distances), len(scene.bends)
    overrides = {}
    if column == 0:
        overrides["distance_scale"] = np.full(D, factor)
    elif column == 1:
        overrides["bend_scale"] = np.full(B, factor)
    elif column == 2:
        overrides["damping"] = np.full(P, np.float64(np.float32(scene.damping)) * factor)
    else:  # mass m -> m * factor, so w -> w / factor
        overrides["inverse_mass_scale"] = np.full(P, 1.0 / factor)
    return cp.oracle_trial(oracle, scene, name, gradients=False, **overrides)["loss"]


@pytest.mark.parametrize("strip", [0, 2])
def test_dlog_columns_match_central_differences(oracle, trials, strip):
    """dlog_j = dy/dx_j at x = 0 for theta_j -> theta_j e^{x_j}: central
    differences in x at h = 1e-3 agree to 1e-3 relative; the reference
    protocol's h = 0.05 agrees to its O(h^2) truncation (2e-2)."""
    scenes, _, dlog = trials
    scene = scenes[strip]
    for o, name in enumerate(cp.OBSERVABLES):
        for column in range(4):
            value = dlog[strip, o, column]
            fd = {}
            for h in (1e-3, 0.05):
                plus = _observable(oracle, scene, name, column, math.exp(h))
                minus = _observable(oracle, scene, name, column, math.exp(-h))
                fd[h] = (plus - minus) / (2 * h)
            print(f"strip {strip} {name} column {column}: adjoint {value:.9e} fd(1e-3) {fd[1e-3]:.9e} "
                  f"fd(0.05) {fd[0.05]:.9e}")
            assert value != 0.0
            assert value == pytest.approx(fd[1e-3], rel=1e-3)
            assert value == pytest.approx(fd[0.05], rel=2e-2)


def test_mass_column_is_the_sum_of_compliance_columns(trials):
    """Scaling every mass and compliance of a strip by the same factor leaves
    its trajectory unchanged, so dlog_mass = dlog_axial + dlog_bend."""
    _, _, dlog = trials
    for s in range(dlog.shape[0]):
        for o in range(dlog.shape[1]):
            axial, bend, _, mass = dlog[s, o]
            assert abs(mass - (axial + bend)) <= 1e-9 * (abs(axial) + abs(bend))


def _read_csv(path):
    lines = [line.rstrip("\n").split(",") for line in open(path)]
    return lines[0], lines[1:]


def test_writer_round_trip(cloth, tmp_path, trials):
    _, baselines, dlog = trials
    rows = cp.synthetic_rows(baselines, dlog[:, :, :3], np.array([0.1, -0.2, 0.15]), 1e-4)
    path = tmp_path / "observations.csv"
    cloth.cloth.write_calibration(path, cp.PARAMETERS, rows)
    header, body = _read_csv(path)
    assert header == ["trial_id", "split", "observable", "measured", "sigma", "baseline"] + \
        [f"dlog_{n}" for n in cp.PARAMETERS]
    assert len(body) == len(rows)
    for row, fields in zip(rows, body):
        assert fields[:3] == [row["trial"], "heldout" if row["heldout"] else "calibration", row["observable"]]
        values = [float(v) for v in fields[3:]]
        assert values[:3] == [row["measured"], row["sigma"], row["baseline"]]  # 17 digits: exact doubles
        assert values[3:] == list(row["dlog"])


@pytest.mark.parametrize("mutate, message", [
    (lambda names, rows: (names + [names[0]], rows), "duplicate"),
    (lambda names, rows: (["a,b"] + names[1:], rows), "plain"),
    (lambda names, rows: (names, [dict(rows[0], sigma=0.0)] + rows[1:]), "sigma"),
    (lambda names, rows: (names, [dict(rows[0], measured=float("nan"))] + rows[1:]), "finite"),
    (lambda names, rows: (names, [dict(rows[0], heldout=True)] + rows[1:]), "splits"),
    (lambda names, rows: (names, rows + [dict(rows[0])]), "duplicate trial"),
    (lambda names, rows: (names, [dict(rows[0], trial="#1")] + rows[1:]), "plain"),
])
def test_writer_rejects_what_the_fitter_rejects(cloth, tmp_path, trials, mutate, message):
    _, baselines, dlog = trials
    names = list(cp.PARAMETERS)
    rows = cp.synthetic_rows(baselines, dlog[:, :, :3], np.zeros(3), 1e-4)
    names, rows = mutate(names, rows)
    if len(names) != 3:
        rows = [dict(r, dlog=np.resize(r["dlog"], len(names))) for r in rows]
    path = tmp_path / "bad.csv"
    with pytest.raises(cloth.ClothError, match=message):
        cloth.cloth.write_calibration(path, names, rows)
    assert not path.exists()


def run_fitter(tmp_path, names, rows, cloth):
    parameters, observations = tmp_path / "parameters.csv", tmp_path / "observations.csv"
    cp.write_parameters(parameters, names)
    cloth.cloth.write_calibration(observations, names, rows)
    result = subprocess.run([str(FITTER), "--parameters", str(parameters), "--observations", str(observations)],
                            capture_output=True, text=True, timeout=60)
    return result.returncode, cp.parse_report(result.stdout), result


@pytest.mark.skipif(not FITTER.exists(), reason="reference fitter not built (no reference checkout)")
def test_reference_fitter_recovers_log_deltas_from_oracle_columns(cloth, tmp_path, trials):
    """Measurements made exactly linear in the columns: the reference's
    fitter must read the CSV, qualify the fit and return the log deltas that
    made them. With the mass column added the design is rank deficient
    (mass = axial + bend), which the fitter must refuse."""
    _, baselines, dlog = trials
    truth = np.array([0.1, -0.2, 0.15])
    rows = cp.synthetic_rows(baselines, dlog[:, :, :3], truth, cp.SIGMA)
    code, report, result = run_fitter(tmp_path, cp.PARAMETERS, rows, cloth)
    print(result.stdout)
    assert code == 0, result.stderr
    assert report["qualified"] == "true" and report["rank"] == "3"
    for name, value in zip(cp.PARAMETERS, truth):
        assert report["log_delta"][name] == pytest.approx(value, abs=1e-6)
    rows = cp.synthetic_rows(baselines, dlog, np.append(truth, 0.0), cp.SIGMA)
    code, report, result = run_fitter(tmp_path, cp.PARAMETERS + [cp.MASS], rows, cloth)
    assert code == 1 and report["qualified"] == "false" and report["rank"] == "3"
