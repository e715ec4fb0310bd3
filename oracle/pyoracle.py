"""ctypes front end of the CPU FP64 oracle (test infrastructure only).

Takes the same C scene descriptor as the product library (the SceneDesc
ctypes structure of numi-solver_amd/cloth.py) and returns double-precision
state after a run of substeps.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().parent / "libnumi_cloth_oracle.so"


class OracleOutput(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("previous", C.c_void_p), ("velocities", C.c_void_p),
                ("groundImpulse", C.c_void_p), ("distanceLambda", C.c_void_p), ("bendLambda", C.c_void_p),
                ("knotLambda", C.c_void_p), ("gripLambda", C.c_void_p), ("gripParticle", C.c_void_p),
                ("failure", C.c_uint32), ("frictionContacts", C.c_uint32), ("maximumConeRatio", C.c_double),
                ("presentContacts", C.c_uint32), ("sweptContacts", C.c_uint32), ("candidateCount", C.c_uint32),
                ("certificatePasses", C.c_uint32), ("fruits", C.c_void_p), ("fruitPairs", C.c_void_p),
                ("yarnContacts", C.c_void_p), ("mouthClearance", C.c_void_p), ("maximumYarnAeroForce", C.c_double),
                ("maximumFruitAeroForce", C.c_double), ("maximumFruitAeroTorque", C.c_double),
                ("candidateMask", C.c_uint32),
                ("releasedMask", C.c_uint32), ("pairFrictionContacts", C.c_uint32),
                ("yarnFrictionContacts", C.c_uint32), ("fruitGroundContacts", C.c_uint32),
                ("rollingContacts", C.c_uint32), ("maximumRollingRatio", C.c_double),
                ("selfFrictionContacts", C.c_uint32), ("pointSelfContacts", C.c_uint32),
                ("selfImpulses", C.c_void_p), ("selfImpulseCapacity", C.c_uint32),
                ("selfImpulseCount", C.c_uint32), ("gripRecords", C.c_void_p), ("gripCaptureError", C.c_void_p)]


_lib = None


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            from . import build as oracle_build
            oracle_build.build()
        lib = C.CDLL(str(LIBRARY_PATH))
        lib.numi_cloth_oracle_run.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.POINTER(OracleOutput)]
        lib.numi_cloth_oracle_run.restype = C.c_int
        lib.numi_cloth_oracle_run_fp32.argtypes = lib.numi_cloth_oracle_run.argtypes
        lib.numi_cloth_oracle_run_fp32.restype = C.c_int
        lib.numi_cloth_oracle_self_candidates_fp32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                                C.c_uint32]
        lib.numi_cloth_oracle_self_candidates_fp32.restype = C.c_int64
        _lib = lib
    return _lib


def run(scene, configs: np.ndarray, iterations: int, strain_sweeps: int, precision: str = "fp64") -> dict:
    """Run len(configs) substeps of `scene` (an object with .desc and the
    ABI arrays) in FP64, or with precision="fp32" as the FP32 twin that
    evaluates every expression in the device kernels' order."""
    configs = np.ascontiguousarray(configs)
    P = len(scene.particles)
    out = {
        "positions": np.zeros((P, 3)), "previous": np.zeros((P, 3)), "velocities": np.zeros((P, 3)),
        "groundImpulse": np.zeros(P), "distanceLambda": np.zeros(len(scene.distances)),
        "bendLambda": np.zeros(len(scene.bends)), "knotLambda": np.zeros(len(scene.knots)),
        "gripLambda": np.zeros((len(scene.grips), 3)), "gripParticle": np.zeros(len(scene.grips), np.uint32),
        "fruits": np.zeros((len(scene.fruits), 20)), "fruitPairs": np.zeros((len(scene.fruitPairs), 4)),
        "yarnContacts": np.zeros((len(scene.yarnContacts), 24)), "mouthClearance": np.zeros(len(scene.fruits)),
        "gripRecords": np.zeros((len(scene.grips), 4), np.uint32), "gripCaptureError": np.zeros(len(scene.grips)),
    }
    result = OracleOutput()
    for name, array in out.items():
        setattr(result, name, array.ctypes.data if array.size else None)
    impulses = np.zeros((max(1, min(len(scene.distances) * 16, 4096)), 11))
    result.selfImpulses = impulses.ctypes.data
    result.selfImpulseCapacity = len(impulses)
    entry = library().numi_cloth_oracle_run if precision == "fp64" else library().numi_cloth_oracle_run_fp32
    status = entry(C.addressof(scene.desc), configs.ctypes.data, len(configs), iterations, strain_sweeps,
                   C.byref(result))
    if status != 0:
        raise RuntimeError(f"oracle run failed with status {status}")
    out["failure"] = int(result.failure)
    out["frictionContacts"] = int(result.frictionContacts)
    out["maximumConeRatio"] = float(result.maximumConeRatio)
    out["presentContacts"] = int(result.presentContacts)
    out["sweptContacts"] = int(result.sweptContacts)
    out["candidateCount"] = int(result.candidateCount)
    out["certificatePasses"] = int(result.certificatePasses)
    for name in ["candidateMask", "releasedMask", "pairFrictionContacts", "yarnFrictionContacts",
                 "fruitGroundContacts", "rollingContacts", "selfFrictionContacts", "pointSelfContacts"]:
        out[name] = int(getattr(result, name))
    for name in ["maximumRollingRatio", "maximumYarnAeroForce", "maximumFruitAeroForce", "maximumFruitAeroTorque"]:
        out[name] = float(getattr(result, name))
    # Per candidate of the last substep: s, t, records, sum -n*impulse (3),
    # sum impulse, endpoint shares (4).
    out["selfImpulses"] = impulses[:min(int(result.selfImpulseCount), len(impulses))].copy()
    return out


def certificate_residuals(scene, config=None, precision: str = "fp64") -> np.ndarray:
    """The certificate loop's residuals of the scene's own state: strain-limit
    violation, ground penetration, fruit/yarn penetration, segment
    self-penetration (each max(0, ...); self candidates from a broad phase
    on that state)."""
    lib = library()
    if not hasattr(lib, "_residuals_ready"):
        lib.numi_cloth_oracle_certificate_residuals.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        lib.numi_cloth_oracle_certificate_residuals.restype = C.c_int
        lib._residuals_ready = True
    config = np.ascontiguousarray(scene.config if config is None else config)
    out = np.zeros(4)
    status = lib.numi_cloth_oracle_certificate_residuals(C.addressof(scene.desc), config.ctypes.data,
                                                         1 if precision == "fp32" else 0, out.ctypes.data)
    if status != 0:
        raise RuntimeError(f"oracle certificate residuals failed with status {status}")
    return out


def self_candidates_fp32(scene, particles: np.ndarray, config: np.ndarray) -> np.ndarray:
    """Sorted (s, t) candidate pairs of the broad phase on an FP32 snapshot."""
    particles = np.ascontiguousarray(particles)
    config = np.ascontiguousarray(config)
    total = library().numi_cloth_oracle_self_candidates_fp32(C.addressof(scene.desc), particles.ctypes.data,
                                                             config.ctypes.data, None, 0)
    pairs = np.zeros((max(total, 1), 2), np.uint32)
    library().numi_cloth_oracle_self_candidates_fp32(C.addressof(scene.desc), particles.ctypes.data,
                                                     config.ctypes.data, pairs.ctypes.data, total)
    return pairs[:total]


class AdjointProblem(C.Structure):
    _fields_ = [("tracked", C.c_void_p), ("trackedCount", C.c_uint32), ("observable", C.c_uint32),
                ("targets", C.c_void_p), ("distanceComplianceScale", C.c_void_p),
                ("bendComplianceScale", C.c_void_p), ("inverseMassScale", C.c_void_p), ("damping", C.c_void_p)]


class Gradients(C.Structure):
    _fields_ = [("loss", C.c_double), ("distanceCompliance", C.c_void_p), ("bendCompliance", C.c_void_p),
                ("inverseMass", C.c_void_p), ("damping", C.c_void_p), ("positions", C.c_void_p),
                ("trajectory", C.c_void_p), ("trajectoryLow", C.c_void_p)]


def _array_pointer(array):
    return None if array is None else array.ctypes.data


def adjoint(scene, configs: np.ndarray, iterations: int, strain_sweeps: int, tracked: np.ndarray,
            targets: np.ndarray, gradients: bool = True, distance_scale=None, bend_scale=None,
            inverse_mass_scale=None, damping=None, precision: str = "fp64", linear: bool = False) -> dict:
    """FP64 loss sum_s sum_k |x_k(s) - target_k(s)|^2 over the substeps
    (linear=True: the observable sum_s sum_k target_k(s) . x_k(s)) and,
    with gradients=True, its gradient with respect to distance and bend
    compliance, inverse mass and damping coefficient (per record / knot).
    The *_scale / damping arrays override parameters in double. The result
    also holds "trajectory": the tracked knots' positions after every
    substep (substeps x tracked x 3), from which differences of two losses
    can be formed term by term, and "trajectoryLow", the part of each
    position below FP64 (precision="extended": the same algorithm with x87
    80-bit arithmetic; zero in the other precisions)."""
    lib = library()
    if not hasattr(lib, "_adjoint_ready"):
        for name in ("numi_cloth_oracle_adjoint", "numi_cloth_oracle_adjoint_fp32",
                     "numi_cloth_oracle_adjoint_extended"):
            fn = getattr(lib, name)
            fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(AdjointProblem),
                           C.POINTER(Gradients)]
            fn.restype = C.c_int
        lib._adjoint_ready = True
    entries = {"fp64": lib.numi_cloth_oracle_adjoint, "fp32": lib.numi_cloth_oracle_adjoint_fp32,
               "extended": lib.numi_cloth_oracle_adjoint_extended}
    entry = entries[precision]
    configs = np.ascontiguousarray(configs)
    tracked = np.ascontiguousarray(tracked, np.uint32)
    targets = np.ascontiguousarray(targets, np.float64)
    assert targets.shape == (len(configs), len(tracked), 3)
    keep = [np.ascontiguousarray(a, np.float64) if a is not None else None
            for a in (distance_scale, bend_scale, inverse_mass_scale, damping)]
    problem = AdjointProblem(tracked.ctypes.data, len(tracked), 1 if linear else 0, targets.ctypes.data,
                             *[_array_pointer(a) for a in keep])
    P, D, B = len(scene.particles), len(scene.distances), len(scene.bends)
    out = {"distanceCompliance": np.zeros(D), "bendCompliance": np.zeros(B), "inverseMass": np.zeros(P),
           "damping": np.zeros(P), "positions": np.zeros((P, 3)),
           "trajectory": np.zeros((len(configs), len(tracked), 3)),
           "trajectoryLow": np.zeros((len(configs), len(tracked), 3))}
    g = Gradients(0.0, *(out[name].ctypes.data if gradients and out[name].size else None
                         for name in ("distanceCompliance", "bendCompliance", "inverseMass", "damping")),
                  out["positions"].ctypes.data, out["trajectory"].ctypes.data, out["trajectoryLow"].ctypes.data)
    status = entry(C.addressof(scene.desc), configs.ctypes.data, len(configs), iterations, strain_sweeps,
                   C.byref(problem), C.byref(g))
    if status != 0:
        raise RuntimeError(f"oracle adjoint failed with status {status}")
    out["loss"] = g.loss
    return out
