"""Calibration trials for the sensitivity tests: three hanging strips of
different size, stiffness, drive and damping; two observables per strip, the
height after the last substep of the lowest knot ("tip_z") and of a knot
half way down ("mid_z"). Fitter columns: the strips' axial (distance) and
bend compliances and their damping coefficient. The knot-mass column is the
sum of the two compliance columns (scaling every mass and compliance of a
strip together leaves XPBD unchanged), so a fit that includes it is rank
deficient; the tests use that as an identity and as the fitter's rank check.
"""

import numpy as np

DT = 1.0 / 960.0
SUBSTEPS = 96
ITERATIONS, SWEEPS = 8, 2
PARAMETERS = ["axial_body_compliance_m_per_n", "bend_body_compliance_m_per_n", "strip_damping_per_s"]
OBSERVABLES = ["tip_z", "mid_z"]
MASS = "ordinary_node_mass_kg"
# Measurement sigma of the synthetic rows: the columns are 1e-9 .. 2e-6 m per
# unit log, so log deltas of 0.1 move the observables by a few sigma.
SIGMA = 1e-8


def specimens(cloth):
    return [
        cloth.Scene.strip(width=6, height=10, seed=3, driveFrequency=2.0, damping=2.0, stretchCompliance=3e-5,
                          bendCompliance=1e-4, driveAmplitude=0.05, timestep=DT),
        cloth.Scene.strip(width=5, height=8, seed=4, driveFrequency=1.7, damping=1.5, origin=(0.5, 0.0, 1.0),
                          stretchCompliance=2e-5, bendCompliance=5e-5, driveAmplitude=0.05, timestep=DT),
        cloth.Scene.strip(width=7, height=12, seed=6, driveFrequency=1.2, damping=3.0, origin=(1.0, 0.0, 1.0),
                          stretchCompliance=4e-5, bendCompliance=2e-4, driveAmplitude=0.04, timestep=DT),
    ]


def observed_knots(scene) -> dict:
    """Specimen-local knots of the two observables, from the authored positions."""
    x = scene.particles["positionAndInverseMass"][:, :3].astype(np.float64)
    free = scene.particles["positionAndInverseMass"][:, 3] > 0
    middle = 0.5 * (x[:, 0].min() + x[:, 0].max())
    z_low, z_high = x[free, 2].min(), x[:, 2].max()
    tip = np.where(free, np.abs(x[:, 2] - z_low) + np.abs(x[:, 0] - middle), np.inf)
    mid = np.where(free, np.abs(x[:, 2] - 0.5 * (z_low + z_high)) + np.abs(x[:, 0] - middle), np.inf)
    return {"tip_z": int(np.argmin(tip)), "mid_z": int(np.argmin(mid))}


def weights(substeps: int = SUBSTEPS) -> np.ndarray:
    """Linear-observable weights of one tracked knot: its z after the last substep."""
    w = np.zeros((substeps, 1, 3))
    w[-1, 0, 2] = 1.0
    return w


def columns(scene) -> dict:
    """Column of every parameter of a strip: distances -> 0, bends -> 1,
    damping of every knot -> 2; the mass column (3) is kept apart."""
    P, D, B = len(scene.particles), len(scene.distances), len(scene.bends)
    return {"distance": np.zeros(D, np.uint32), "bend": np.ones(B, np.uint32),
            "damping": np.full(P, 2, np.uint32), "mass": np.full(P, 3, np.uint32)}


def dlog_from_gradients(scene, g) -> np.ndarray:
    """[axial, bend, damping, mass] columns from per-record gradients of the
    observable: c dy/dc summed per compliance family, c dy/dc over the knots'
    damping coefficients, and m dy/dm = -w dy/dw over the free knots."""
    w = scene.particles["positionAndInverseMass"][:, 3].astype(np.float64)
    cd = scene.distances["material"][:, 1].astype(np.float64)
    cb = scene.bends["material"][:, 2].astype(np.float64)
    free = w > 0
    return np.array([
        np.sum(cd * np.asarray(g["distanceCompliance"], np.float64)),
        np.sum(cb * np.asarray(g["bendCompliance"], np.float64)),
        np.sum(np.float64(np.float32(scene.damping)) * np.asarray(g["damping"], np.float64)),
        -np.sum(w[free] * np.asarray(g["inverseMass"], np.float64)[free]),
    ])


def oracle_trial(oracle, scene, observable: str, **overrides) -> dict:
    """The FP64 observable of one strip and its gradients (oracle adjoint)."""
    knot = observed_knots(scene)[observable]
    return oracle.adjoint(scene, scene.configs(SUBSTEPS), ITERATIONS, SWEEPS, np.array([knot], np.uint32), weights(),
                          linear=True, **overrides)


def synthetic_rows(baselines: np.ndarray, dlog: np.ndarray, truth: np.ndarray, sigma: float) -> list[dict]:
    """Fitter rows whose measurements are exactly baseline + dlog . truth:
    strips 0 and 1 calibrate, strip 2 is held out."""
    rows = []
    for s in range(baselines.shape[0]):
        for o, name in enumerate(OBSERVABLES):
            rows.append({"trial": f"strip{s}", "heldout": s == 2, "observable": name,
                         "measured": float(baselines[s, o] + dlog[s, o] @ truth), "sigma": sigma,
                         "baseline": float(baselines[s, o]), "dlog": dlog[s, o]})
    return rows


def write_parameters(path, names) -> None:
    """Parameter CSV of the fitter: multiplicative factors around 1."""
    with open(path, "w") as f:
        f.write("name,base,minimum,maximum\n")
        for name in names:
            f.write(f"{name},1.0,0.5,2.0\n")


def parse_report(text: str) -> dict:
    """log_delta per parameter and the scalar fields of the fitter's report."""
    out = {"log_delta": {}}
    for line in text.splitlines():
        fields = dict(item.split("=", 1) for item in line.split() if "=" in item)
        if "parameter" in fields:
            out["log_delta"][fields["parameter"]] = float(fields["log_delta"])
        else:
            out.update(fields)
    return out
