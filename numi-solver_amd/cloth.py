"""ctypes bindings of libnumi_cloth_hip.so.

The C ABI (include/numi_cloth_hip.h, include/numi_cloth_scene.h) is the
product boundary; this module only mirrors it for Python callers and tests.
ABI v13 records are exposed as numpy structured dtypes with the field names
of the reference header (reference include/numi/cloth_bag_gpu.h:19-190).

There is no CPU fallback: if the shared library is missing, loading fails.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PACKAGE = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE / "libnumi_cloth_hip.so"

ABI_VERSION = 13
FAILURE_ABI = 1
FAILURE_RANGE = 2
FAILURE_BATCH = 4
FAILURE_NONFINITE = 8
FAILURE_GRIP_CAPTURE = 16

STATUS_NAMES = {0: "ok", 1: "invalid argument", 2: "topology", 3: "device", 4: "allocation",
                5: "unsupported", 6: "parse", 7: "io"}

_u4 = ("<u4", (4,))
_f4 = ("<f4", (4,))


def _dtype(*fields):
    return np.dtype([(name, *kind) for name, kind in fields], align=True)


CONFIG = _dtype(("control", _u4), ("constraintCounts", _u4), ("contactCounts", _u4),
                ("gravityAndTimestep", _f4), ("gripTargetAndActive", _f4), ("gripOrientation", _f4),
                ("gripControl", _u4), ("gripMaterial", _f4), ("clothMaterial", _f4), ("fruitMaterial", _f4),
                ("airVelocityAndDensity", _f4), ("aerodynamicCoefficients", _f4), ("mouthControl", _u4),
                ("mouthMaterial", _f4))
PARTICLE = _dtype(("positionAndInverseMass", _f4), ("previousAndMass", _f4), ("velocity", _f4))
DISTANCE = _dtype(("particlesAndColor", _u4), ("material", _f4))
GRIP = _dtype(("particle", _u4), ("targetOffsetAndCompliance", _f4), ("lambda", _f4))
KNOT = _dtype(("particles", _u4), ("control", _u4), ("material", _f4))
BEND = _dtype(("particlesAndColor", _u4), ("material", _f4))
FRUIT = _dtype(("positionAndInverseMass", _f4), ("previousAndRadius", _f4), ("velocityAndGroundImpulse", _f4),
               ("angularVelocity", _f4), ("orientation", _f4), ("identity", _u4))
FRUIT_PAIR = _dtype(("fruitsAndColor", _u4), ("contact", _f4))
YARN_CONTACT = _dtype(("identity", _u4), ("currentNormalAndSeparation", _f4), ("sweptNormalAndAdvance", _f4),
                      ("weightsAndTime", _f4), ("control", _u4), ("fruitNormalAndImpulse", _f4),
                      ("segmentImpulse", _f4))
BATCH = _dtype(("control", _u4))
SELF_STATUS = _dtype(("counters", _u4))
FRICTION_STATUS = _dtype(("counters", _u4), ("metrics", _u4))
AERO_STATUS = _dtype(("metrics", _u4))
RELEASE_STATUS = _dtype(("masks", _u4))

RECORD_SIZES = {"config": 224, "particle": 48, "distance": 32, "grip": 48, "knot": 48, "bend": 32,
                "fruit": 96, "fruit_pair": 32, "yarn_contact": 112, "batch": 16}
for _name, _dt in [("config", CONFIG), ("particle", PARTICLE), ("distance", DISTANCE), ("grip", GRIP),
                   ("knot", KNOT), ("bend", BEND), ("fruit", FRUIT), ("fruit_pair", FRUIT_PAIR),
                   ("yarn_contact", YARN_CONTACT), ("batch", BATCH)]:
    assert _dt.itemsize == RECORD_SIZES[_name], _name


class Float4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class SceneExtension(C.Structure):
    _fields_ = [("staticSpheres", C.c_void_p), ("staticSphereCount", C.c_uint32), ("selfContact", C.c_uint32),
                ("selfContactMargin", C.c_float), ("selfCandidateCapacity", C.c_uint32),
                ("driveAmplitude", C.c_float), ("driveOmega", C.c_float), ("drivePhase", C.c_float),
                ("damping", C.c_float), ("particleMotion", C.c_void_p), ("sweptGround", C.c_uint32),
                ("pointSelfCollision", C.c_uint32), ("certificatePasses", C.c_uint32)]


_DESC_POINTERS = ["config", "particles", "distances", "distanceBatches", "grips", "knots", "knotBatches", "bends",
                  "bendBatches", "fruits", "fruitPairs", "fruitPairBatches", "yarnContacts", "mouthRimParticles",
                  "cuffParticles"]
_DESC_COUNTS = ["particleCount", "distanceCount", "distanceBatchCount", "gripCount", "knotCount",
                "knotBatchCount", "bendCount", "bendBatchCount", "fruitCount", "fruitPairCount",
                "fruitPairBatchCount", "yarnContactCount", "mouthRimCount", "cuffCount", "initialCandidateMask",
                "initialReleasedMask"]


class SceneDesc(C.Structure):
    _fields_ = ([(name, C.c_void_p) for name in _DESC_POINTERS] + [(name, C.c_uint32) for name in _DESC_COUNTS]
                + [("extension", SceneExtension)])


class State(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in
                ["particles", "distances", "grips", "knots", "bends", "fruits", "fruitPairs", "yarnContacts",
                 "selfStatus", "frictionStatus", "aerodynamicsStatus", "releaseStatus", "failure",
                 "mouthClearance"]]


class AdjointStats(C.Structure):
    _fields_ = [("strainPasses", C.c_uint64), ("strainPassesTaped", C.c_uint32), ("reserved", C.c_uint32)]


class LaunchMode(C.Structure):
    _fields_ = [("graphs", C.c_uint32), ("decided", C.c_uint32), ("probeGraphMilliseconds", C.c_float),
                ("probeDirectMilliseconds", C.c_float)]


class Diagnostics(C.Structure):
    _fields_ = [("maximumStrainViolation", C.c_float), ("maximumGroundPenetration", C.c_float),
                ("maximumSpherePenetration", C.c_float), ("maximumSelfPenetration", C.c_float),
                ("selfCandidateCount", C.c_uint32), ("selfCandidateOverflow", C.c_uint32),
                ("yarnReruns", C.c_uint32), ("yarnCandidatesMax", C.c_uint32),
                ("selfEntryCount", C.c_uint32), ("selfGridCells", C.c_uint32 * 3),
                ("pointSelfContacts", C.c_uint32), ("certificatePasses", C.c_uint32),
                ("tileRoundsSkipped", C.c_uint64), ("tileRoundsRun", C.c_uint64),
                ("yarnListBuilds", C.c_uint32), ("yarnListSolves", C.c_uint32)]


MATERIAL_FIELDS = ["ordinaryNodeMassKg", "hemNodeMassKg", "yarnRadiusM", "axialBodyComplianceMPerN",
                   "axialCuffComplianceMPerN", "knotCompliance", "bendBodyComplianceMPerN",
                   "bendCuffComplianceMPerN", "gripComplianceMPerN", "clothGroundFriction", "clothSelfFriction",
                   "fruitYarnFriction", "fruitGroundFriction", "fruitPairFriction", "fruitRollingResistance",
                   "yarnCrossflowDrag", "yarnSkinFriction", "fruitDrag", "fruitRotationalDrag"]


class SelfBroadPhase(C.Structure):
    _fields_ = [("pairs", C.c_void_p), ("cellIds", C.c_void_p), ("segments", C.c_void_p),
                ("pairCapacity", C.c_uint32), ("entryCapacity", C.c_uint32), ("pairCount", C.c_uint32),
                ("entryCount", C.c_uint32), ("origin", C.c_float * 3), ("cell", C.c_float),
                ("cells", C.c_uint32 * 3), ("valid", C.c_uint32)]


class PassTiming(C.Structure):
    _fields_ = [("averageMilliseconds", C.c_double), ("algorithmicBytes", C.c_double), ("launches", C.c_uint32),
                ("itemsPerLaunch", C.c_uint32)]


PASS_DISTANCE, PASS_BEND, PASS_STRAIN, PASS_COLLIDE, PASS_TILE_ITERATION = 0, 1, 2, 3, 4


class TileInfo(C.Structure):
    _fields_ = [("enabled", C.c_uint32), ("tiles", C.c_uint32), ("ownedPerTile", C.c_uint32),
                ("maxLoaded", C.c_uint32), ("halo", C.c_uint32), ("workgroupSize", C.c_uint32),
                ("passes", C.c_uint32), ("staged", C.c_uint32), ("redundancy", C.c_double)]


class MaterialValues(C.Structure):
    _fields_ = [(name, C.c_double) for name in MATERIAL_FIELDS]


class MaterialArtifact(C.Structure):
    _fields_ = [("values", MaterialValues), ("loaded", C.c_uint32), ("parametersHash", C.c_char * 128),
                ("observationsHash", C.c_char * 128)]


class GripPose(C.Structure):
    _fields_ = [("timeSeconds", C.c_double), ("translation", C.c_double * 3), ("orientation", C.c_double * 4),
                ("active", C.c_uint32), ("attachmentGeneration", C.c_uint32)]


class GripTrajectoryInfo(C.Structure):
    _fields_ = [("poseCount", C.c_uint32), ("selectNearestCuffPatch", C.c_uint32),
                ("attachmentGenerations", C.c_uint32), ("reserved", C.c_uint32), ("durationSeconds", C.c_double),
                ("maximumRotation", C.c_double), ("schema", C.c_char * 32), ("contentFingerprint", C.c_char * 32)]


class SheetSpec(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spacing", C.c_float), ("elevation", C.c_float),
                ("jitter", C.c_float), ("arealDensity", C.c_float), ("yarnRadius", C.c_float),
                ("stretchCompliance", C.c_float), ("bendCompliance", C.c_float), ("knotCompliance", C.c_float),
                ("strainLimit", C.c_float), ("cellSize", C.c_float), ("timestep", C.c_float),
                ("gravityZ", C.c_float), ("groundFriction", C.c_float), ("selfFriction", C.c_float),
                ("crossingKnots", C.c_uint32), ("ground", C.c_uint32), ("pinTopRow", C.c_uint32),
                ("selfContact", C.c_uint32), ("sphereMode", C.c_uint32), ("sphereCount", C.c_uint32),
                ("sphereRadiusMin", C.c_float), ("sphereRadiusMax", C.c_float), ("fixedSphere", C.c_float * 4),
                ("seed", C.c_uint64)]


class StripSpec(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spacing", C.c_float), ("arealDensity", C.c_float),
                ("yarnRadius", C.c_float), ("stretchCompliance", C.c_float), ("bendCompliance", C.c_float),
                ("strainLimit", C.c_float), ("timestep", C.c_float), ("gravityZ", C.c_float),
                ("driveAmplitude", C.c_float), ("driveFrequency", C.c_float), ("drivePhase", C.c_float),
                ("damping", C.c_float), ("origin", C.c_float * 3), ("jitter", C.c_float), ("seed", C.c_uint64)]


class AdjointProblem(C.Structure):
    _fields_ = [("configs", C.c_void_p), ("substeps", C.c_uint32), ("iterations", C.c_uint32),
                ("strainSweeps", C.c_uint32), ("checkpointInterval", C.c_uint32), ("tracked", C.c_void_p),
                ("trackedPerSpecimen", C.c_uint32), ("observable", C.c_uint32), ("targets", C.c_void_p)]


class CalibrationColumns(C.Structure):
    _fields_ = [("columnCount", C.c_uint32), ("reserved", C.c_uint32), ("distanceColumn", C.c_void_p),
                ("bendColumn", C.c_void_p), ("massColumn", C.c_void_p), ("dampingColumn", C.c_void_p)]


class CalibrationRow(C.Structure):
    _fields_ = [("trialId", C.c_char_p), ("heldout", C.c_uint32), ("reserved", C.c_uint32),
                ("observable", C.c_char_p), ("measured", C.c_double), ("sigma", C.c_double),
                ("baseline", C.c_double), ("dlog", C.c_void_p)]


NO_COLUMN = 0xFFFFFFFF


class SpecimenGradients(C.Structure):
    _fields_ = [("loss", C.c_void_p), ("distanceCompliance", C.c_void_p), ("bendCompliance", C.c_void_p),
                ("inverseMass", C.c_void_p), ("damping", C.c_void_p)]


class ReplayPose(C.Structure):
    _fields_ = [("target", C.c_double * 3), ("orientation", C.c_double * 4), ("active", C.c_uint32),
                ("attachmentGeneration", C.c_uint32)]


class ReplayResult(C.Structure):
    _fields_ = [("frameHashes", C.c_void_p), ("framesRun", C.c_uint32), ("failureFree", C.c_uint32),
                ("releasedMask", C.c_uint32), ("reserved", C.c_uint32), ("maximumHandleLag", C.c_double),
                ("stepSeconds", C.c_double)]


# Trajectory replay scenarios (include/numi_cloth_replay.h).
REPLAY_GROUNDED, REPLAY_SPIN, REPLAY_PICKUP, REPLAY_RECORDED = 0, 1, 2, 3
REPLAY_SUBSTEPS_PER_FRAME = 48
REPLAY_TIMESTEP = float(np.float32(1.0) / np.float32(5760.0))


class BagSpec(C.Structure):
    _fields_ = [("material", MaterialValues), ("stretchComplianceOverride", C.c_float),
                ("bendComplianceOverride", C.c_float), ("strainLimitOverride", C.c_float),
                ("fruitRadiusMin", C.c_float), ("fruitRadiusMax", C.c_float), ("selfContact", C.c_uint32),
                ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class ClothError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


_lib = None


def library() -> C.CDLL:
    """Load the in-tree shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY_PATH.exists():
        raise RuntimeError(f"{LIBRARY_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(str(LIBRARY_PATH))
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    signatures = {
        "numi_cloth_hip_abi_version": ([], u32),
        "numi_cloth_hip_last_error": ([], C.c_char_p),
        "numi_cloth_hip_create": ([vp, C.POINTER(vp)], i32),
        "numi_cloth_hip_step": ([vp, vp, u32, u32, u32], i32),
        "numi_cloth_hip_step_noalloc": ([vp, vp, u32, u32, u32], i32),
        "numi_cloth_hip_read": ([vp, C.POINTER(State)], i32),
        "numi_cloth_hip_diagnostics": ([vp, C.POINTER(Diagnostics)], i32),
        "numi_cloth_hip_launch_mode": ([vp, C.POINTER(LaunchMode)], i32),
        "numi_cloth_hip_batch_adjoint_stats": ([vp, C.POINTER(AdjointStats)], i32),
        "numi_cloth_hip_synchronize": ([vp], i32),
        "numi_cloth_hip_stream": ([vp], vp),
        "numi_cloth_hip_destroy": ([vp], i32),
        "numi_cloth_hip_time_pass": ([vp, u32, u32, u32, C.POINTER(PassTiming)], i32),
        "numi_cloth_hip_self_broad_phase": ([vp, C.POINTER(SelfBroadPhase)], i32),
        "numi_cloth_hip_tile_info": ([vp, C.POINTER(TileInfo)], i32),
        "numi_cloth_material_defaults": ([C.POINTER(MaterialArtifact)], i32),
        "numi_cloth_material_load": ([C.c_char_p, C.POINTER(MaterialArtifact)], i32),
        "numi_cloth_grip_trajectory_load": ([C.c_char_p, C.POINTER(vp)], i32),
        "numi_cloth_grip_trajectory_info": ([vp, C.POINTER(GripTrajectoryInfo)], i32),
        "numi_cloth_grip_trajectory_pose": ([vp, u32, C.POINTER(GripPose)], i32),
        "numi_cloth_grip_trajectory_sample": ([vp, C.c_double, C.POINTER(GripPose)], i32),
        "numi_cloth_grip_trajectory_destroy": ([vp], None),
        "numi_cloth_scene_sheet": ([C.POINTER(SheetSpec), C.POINTER(vp)], i32),
        "numi_cloth_scene_bag": ([C.POINTER(BagSpec), C.POINTER(vp)], i32),
        "numi_cloth_scene_strip": ([C.POINTER(StripSpec), C.POINTER(vp)], i32),
        "numi_cloth_scene_merge": ([C.POINTER(vp), u32, C.POINTER(vp)], i32),
        "numi_cloth_scene_merge_indices": ([vp, u32, u32, C.POINTER(C.POINTER(u32)), C.POINTER(u32)], i32),
        "numi_cloth_hip_batch_create": ([C.POINTER(vp), u32, C.POINTER(vp)], i32),
        "numi_cloth_hip_batch_step": ([vp, vp, u32, u32, u32], i32),
        "numi_cloth_hip_batch_step_resident": ([vp, vp, u32, u32, u32], i32),
        "numi_cloth_hip_batch_read": ([vp, u32, C.POINTER(State)], i32),
        "numi_cloth_hip_batch_solver": ([vp], vp),
        "numi_cloth_hip_batch_scene": ([vp], C.POINTER(SceneDesc)),
        "numi_cloth_hip_batch_destroy": ([vp], i32),
        "numi_cloth_hip_batch_adjoint": ([vp, C.POINTER(AdjointProblem)], i32),
        "numi_cloth_hip_batch_gradients": ([vp, u32, C.POINTER(SpecimenGradients)], i32),
        "numi_cloth_scene_desc": ([vp], C.POINTER(SceneDesc)),
        "numi_cloth_scene_destroy": ([vp], None),
        "numi_cloth_splitmix64": ([C.POINTER(C.c_uint64)], C.c_uint64),
        "numi_cloth_fnv1a64": ([vp, C.c_size_t, C.c_uint64], C.c_uint64),
        "numi_cloth_replay_pose": ([u32, C.c_double, vp, C.POINTER(ReplayPose)], i32),
        "numi_cloth_replay_prepare": ([vp, u32, vp], i32),
        "numi_cloth_replay_frame_configs": ([vp, u32, vp, u32, u32, vp], i32),
        "numi_cloth_state_hash": ([C.POINTER(SceneDesc), C.POINTER(State), C.POINTER(C.c_uint64)], i32),
        "numi_cloth_obj_write": ([C.c_char_p, vp, u32, vp, u32, vp, u32, vp], i32),
        "numi_cloth_replay_run": ([vp, C.POINTER(SceneDesc), u32, vp, u32, u32, u32, u32, C.c_char_p,
                                   C.POINTER(ReplayResult)], i32),
        "numi_cloth_calibration_dlog": ([vp, u32, C.POINTER(SceneDesc), C.POINTER(CalibrationColumns), vp, vp], i32),
        "numi_cloth_calibration_write": ([C.c_char_p, u32, vp, C.POINTER(CalibrationRow), u32], i32),
    }
    for name, (args, result) in signatures.items():
        function = getattr(lib, name)
        function.argtypes = args
        function.restype = result
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        raise ClothError(status, library().numi_cloth_hip_last_error().decode())


def _ptr(array) -> int | None:
    return None if array is None or len(array) == 0 else array.ctypes.data


# ---------------------------------------------------------------- artifacts --

def material_defaults() -> MaterialArtifact:
    artifact = MaterialArtifact()
    _check(library().numi_cloth_material_defaults(C.byref(artifact)))
    return artifact


def load_material(path) -> MaterialArtifact:
    artifact = MaterialArtifact()
    _check(library().numi_cloth_material_load(str(path).encode(), C.byref(artifact)))
    return artifact


class GripTrajectory:
    """numi.grip.trajectory v1/v2/v3 (reference include/numi/grip_trajectory.h:189-362)."""

    def __init__(self, path):
        handle = C.c_void_p()
        _check(library().numi_cloth_grip_trajectory_load(str(path).encode(), C.byref(handle)))
        self._handle = handle

    def __del__(self):
        if getattr(self, "_handle", None):
            library().numi_cloth_grip_trajectory_destroy(self._handle)
            self._handle = None

    def info(self) -> GripTrajectoryInfo:
        out = GripTrajectoryInfo()
        _check(library().numi_cloth_grip_trajectory_info(self._handle, C.byref(out)))
        return out

    def pose(self, index: int) -> GripPose:
        out = GripPose()
        _check(library().numi_cloth_grip_trajectory_pose(self._handle, index, C.byref(out)))
        return out

    def sample(self, time_seconds: float) -> GripPose:
        out = GripPose()
        _check(library().numi_cloth_grip_trajectory_sample(self._handle, time_seconds, C.byref(out)))
        return out


# -------------------------------------------------------------------- scene --

class Scene:
    """ABI v13 arrays plus the extension block, as numpy arrays.

    Built either by a C builder (sheet(), bag()) or directly from arrays
    (probe scenes). `desc` is the C descriptor pointing into these arrays.
    """

    ARRAYS = [("particles", PARTICLE, "particleCount"), ("distances", DISTANCE, "distanceCount"),
              ("distanceBatches", BATCH, "distanceBatchCount"), ("grips", GRIP, "gripCount"),
              ("knots", KNOT, "knotCount"), ("knotBatches", BATCH, "knotBatchCount"),
              ("bends", BEND, "bendCount"), ("bendBatches", BATCH, "bendBatchCount"),
              ("fruits", FRUIT, "fruitCount"), ("fruitPairs", FRUIT_PAIR, "fruitPairCount"),
              ("fruitPairBatches", BATCH, "fruitPairBatchCount"), ("yarnContacts", YARN_CONTACT, "yarnContactCount"),
              ("mouthRimParticles", np.dtype("<u4"), "mouthRimCount"), ("cuffParticles", np.dtype("<u4"), "cuffCount")]

    def __init__(self, config, **arrays):
        self.config = np.array(config, dtype=CONFIG).reshape(1)
        for name, dtype, _ in self.ARRAYS:
            value = arrays.pop(name, None)
            setattr(self, name, np.ascontiguousarray(np.zeros(0, dtype) if value is None else value, dtype=dtype))
        spheres = arrays.pop("staticSpheres", None)
        self.staticSpheres = np.ascontiguousarray(np.zeros((0, 4), np.float32) if spheres is None else spheres,
                                                  dtype=np.float32).reshape(-1, 4)
        self.selfContact = int(arrays.pop("selfContact", 0))
        self.selfContactMargin = float(arrays.pop("selfContactMargin", 0.0))
        self.selfCandidateCapacity = int(arrays.pop("selfCandidateCapacity", 0))
        self.driveAmplitude = float(arrays.pop("driveAmplitude", 0.0))
        self.driveOmega = float(arrays.pop("driveOmega", 0.0))
        self.drivePhase = float(arrays.pop("drivePhase", 0.0))
        self.damping = float(arrays.pop("damping", 0.0))
        motion = arrays.pop("particleMotion", None)
        self.particleMotion = None if motion is None else np.ascontiguousarray(motion, np.float32).reshape(-1, 4)
        self.sweptGround = int(arrays.pop("sweptGround", 0))
        self.pointSelfCollision = int(arrays.pop("pointSelfCollision", 0))
        self.certificatePasses = int(arrays.pop("certificatePasses", 0))
        self.initialCandidateMask = int(arrays.pop("initialCandidateMask", 0))
        self.initialReleasedMask = int(arrays.pop("initialReleasedMask", 0))
        if arrays:
            raise TypeError(f"unknown scene arrays: {sorted(arrays)}")
        self.desc = self._describe()

    def _describe(self) -> SceneDesc:
        desc = SceneDesc()
        desc.config = self.config.ctypes.data
        for name, _, count in self.ARRAYS:
            array = getattr(self, name)
            setattr(desc, name, _ptr(array))
            setattr(desc, count, len(array))
        desc.initialCandidateMask = self.initialCandidateMask
        desc.initialReleasedMask = self.initialReleasedMask
        e = desc.extension
        e.staticSpheres = _ptr(self.staticSpheres)
        e.staticSphereCount = len(self.staticSpheres)
        e.selfContact = self.selfContact
        e.selfContactMargin = self.selfContactMargin
        e.selfCandidateCapacity = self.selfCandidateCapacity
        e.driveAmplitude = self.driveAmplitude
        e.driveOmega = self.driveOmega
        e.drivePhase = self.drivePhase
        e.damping = self.damping
        e.particleMotion = _ptr(self.particleMotion)
        e.sweptGround = self.sweptGround
        e.pointSelfCollision = self.pointSelfCollision
        e.certificatePasses = self.certificatePasses
        return desc

    def refresh(self) -> None:
        """Re-point the descriptor after arrays or extension fields changed."""
        self.desc = self._describe()

    @classmethod
    def _from_builder(cls, handle, keep: bool = False):
        lib = library()
        d = lib.numi_cloth_scene_desc(handle).contents
        arrays = {}
        for name, dtype, count in cls.ARRAYS:
            n = getattr(d, count)
            pointer = getattr(d, name)
            if n == 0 or not pointer:
                arrays[name] = np.zeros(0, dtype)
            else:
                raw = (C.c_ubyte * (n * dtype.itemsize)).from_address(pointer)
                arrays[name] = np.frombuffer(raw, dtype=dtype, count=n).copy()
        config = np.frombuffer((C.c_ubyte * CONFIG.itemsize).from_address(d.config), dtype=CONFIG).copy()
        e = d.extension
        spheres = np.zeros((0, 4), np.float32)
        if e.staticSphereCount:
            raw = (C.c_float * (4 * e.staticSphereCount)).from_address(e.staticSpheres)
            spheres = np.frombuffer(raw, dtype=np.float32).reshape(-1, 4).copy()
        motion = None
        if e.particleMotion:
            raw = (C.c_float * (4 * d.particleCount)).from_address(e.particleMotion)
            motion = np.frombuffer(raw, dtype=np.float32).reshape(-1, 4).copy()
        scene = cls(config, staticSpheres=spheres, selfContact=e.selfContact, selfContactMargin=e.selfContactMargin,
                    selfCandidateCapacity=e.selfCandidateCapacity, driveAmplitude=e.driveAmplitude,
                    driveOmega=e.driveOmega, drivePhase=e.drivePhase, damping=e.damping, particleMotion=motion,
                    sweptGround=e.sweptGround, pointSelfCollision=e.pointSelfCollision,
                    certificatePasses=e.certificatePasses, initialCandidateMask=d.initialCandidateMask, initialReleasedMask=d.initialReleasedMask, **arrays)
        if keep:
            return scene, handle
        lib.numi_cloth_scene_destroy(handle)
        return scene

    @classmethod
    def sheet(cls, **fields) -> "Scene":
        spec = SheetSpec()
        defaults = dict(spacing=0.01, elevation=0.3, jitter=0.0, arealDensity=0.2, yarnRadius=0.004,
                        stretchCompliance=1.0e-7, bendCompliance=1.0e-4, knotCompliance=2.0e-6, strainLimit=0.10,
                        cellSize=0.01, timestep=1.0 / 1920.0, gravityZ=-9.81, groundFriction=0.45,
                        selfFriction=0.34, ground=1, seed=1)
        defaults.update(fields)
        for key, value in defaults.items():
            if key == "fixedSphere":
                spec.fixedSphere = (C.c_float * 4)(*value)
            else:
                setattr(spec, key, value)
        handle = C.c_void_p()
        _check(library().numi_cloth_scene_sheet(C.byref(spec), C.byref(handle)))
        return cls._from_builder(handle)

    @classmethod
    def bag(cls, material: MaterialArtifact | None = None, **fields) -> "Scene":
        spec = BagSpec()
        spec.material = (material or material_defaults()).values
        for key, value in fields.items():
            setattr(spec, key, value)
        handle = C.c_void_p()
        _check(library().numi_cloth_scene_bag(C.byref(spec), C.byref(handle)))
        return cls._from_builder(handle)

    @classmethod
    def bag_replay(cls, scenario: int, trajectory: "GripTrajectory | None" = None,
                   material: MaterialArtifact | None = None, **fields) -> "Scene":
        """The bag scene prepared as a trajectory replay's initial state
        (numi_cloth_replay_prepare: spin lift, ground flag, grip pose at t = 0)."""
        spec = BagSpec()
        spec.material = (material or material_defaults()).values
        for key, value in fields.items():
            setattr(spec, key, value)
        handle = C.c_void_p()
        lib = library()
        _check(lib.numi_cloth_scene_bag(C.byref(spec), C.byref(handle)))
        status = lib.numi_cloth_replay_prepare(handle, scenario, trajectory._handle if trajectory else None)
        if status != 0:
            lib.numi_cloth_scene_destroy(handle)
            _check(status)
        return cls._from_builder(handle)

    @classmethod
    def strip(cls, **fields) -> "Scene":
        """A hanging strip specimen (see NumiClothHipStripSpec)."""
        spec = StripSpec()
        defaults = dict(spacing=0.01, arealDensity=0.2, yarnRadius=0.002, stretchCompliance=1.0e-7,
                        bendCompliance=1.0e-5, strainLimit=0.10, timestep=1.0 / 2880.0, gravityZ=-9.81,
                        driveAmplitude=0.01, driveFrequency=1.0, drivePhase=0.0, damping=0.0, origin=(0.0, 0.0, 1.0),
                        jitter=1.0e-4, seed=1)
        defaults.update(fields)
        for key, value in defaults.items():
            if key == "origin":
                spec.origin = (C.c_float * 3)(*value)
            else:
                setattr(spec, key, value)
        handle = C.c_void_p()
        _check(library().numi_cloth_scene_strip(C.byref(spec), C.byref(handle)))
        return cls._from_builder(handle)

    @classmethod
    def merge(cls, specimens: list["Scene"]) -> tuple["Scene", list[dict]]:
        """The merged scene of independent specimens and, per specimen, the
        merged index of each of its particles / distances / bends / knots."""
        lib = library()
        pointers = (C.c_void_p * len(specimens))(*[C.addressof(s.desc) for s in specimens])
        handle = C.c_void_p()
        _check(lib.numi_cloth_scene_merge(pointers, len(specimens), C.byref(handle)))
        maps = []
        for k in range(len(specimens)):
            entry = {}
            for family, name in enumerate(["particles", "distances", "bends", "knots"]):
                indices = C.POINTER(C.c_uint32)()
                count = C.c_uint32()
                _check(lib.numi_cloth_scene_merge_indices(handle, family, k, C.byref(indices), C.byref(count)))
                entry[name] = np.ctypeslib.as_array(indices, (count.value,)).copy() if count.value else \
                    np.zeros(0, np.uint32)
            maps.append(entry)
        scene = cls._from_builder(handle)
        return scene, maps

    def configs(self, substeps: int) -> np.ndarray:
        """`substeps` copies of the initial configuration."""
        return np.repeat(self.config, substeps)

    def continued(self, state: dict) -> "Scene":
        """This scene with the knots, grips and fruits of a read-back state
        as its initial state, and the release masks it ended with (the
        reference's continuedInitialState, tools/cloth_bag_metal.mm:4764-4790):
        a solver created from it continues the run the state came from."""
        arrays = {name: getattr(self, name) for name, _, _ in self.ARRAYS}
        for name in ("particles", "grips", "fruits"):
            if len(arrays[name]):
                arrays[name] = np.ascontiguousarray(state[name]).copy()
        masks = state["releaseStatus"]["masks"][0]
        return Scene(self.config, staticSpheres=self.staticSpheres, selfContact=self.selfContact,
                     selfContactMargin=self.selfContactMargin, selfCandidateCapacity=self.selfCandidateCapacity,
                     driveAmplitude=self.driveAmplitude, driveOmega=self.driveOmega, drivePhase=self.drivePhase,
                     damping=self.damping, particleMotion=self.particleMotion, sweptGround=self.sweptGround,
                     pointSelfCollision=self.pointSelfCollision, certificatePasses=self.certificatePasses,
                     initialCandidateMask=int(masks[0]), initialReleasedMask=int(masks[1]), **arrays)


# ------------------------------------------------------------------- solver --

def _state_arrays(particles, distances, grips, knots, bends, fruits, fruit_pairs, yarn_contacts) -> dict:
    return {
        "particles": np.zeros(particles, PARTICLE),
        "distances": np.zeros(distances, DISTANCE),
        "grips": np.zeros(grips, GRIP),
        "knots": np.zeros(knots, KNOT),
        "bends": np.zeros(bends, BEND),
        "fruits": np.zeros(fruits, FRUIT),
        "fruitPairs": np.zeros(fruit_pairs, FRUIT_PAIR),
        "yarnContacts": np.zeros(yarn_contacts, YARN_CONTACT),
        "selfStatus": np.zeros(1, SELF_STATUS),
        "frictionStatus": np.zeros(1, FRICTION_STATUS),
        "aerodynamicsStatus": np.zeros(1, AERO_STATUS),
        "releaseStatus": np.zeros(1, RELEASE_STATUS),
        "failure": np.zeros(1, np.uint32),
        "mouthClearance": np.zeros(fruits, np.float32),
    }


class Batch:
    """Independent specimens stepped together by one device solver."""

    def __init__(self, specimens: list[Scene]):
        self.specimens = specimens
        pointers = (C.c_void_p * len(specimens))(*[C.addressof(s.desc) for s in specimens])
        handle = C.c_void_p()
        _check(library().numi_cloth_hip_batch_create(pointers, len(specimens), C.byref(handle)))
        self._handle = handle

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _check(library().numi_cloth_hip_batch_destroy(self._handle))
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def solver_handle(self) -> int:
        return library().numi_cloth_hip_batch_solver(self._handle)

    def step(self, configs: np.ndarray, iterations: int = 32, strain_sweeps: int = 3) -> None:
        configs = np.ascontiguousarray(configs, dtype=CONFIG)
        _check(library().numi_cloth_hip_batch_step(self._handle, configs.ctypes.data, len(configs), iterations,
                                                   strain_sweeps))

    def step_resident(self, configs: np.ndarray, iterations: int = 32, strain_sweeps: int = 3) -> None:
        """The same substeps in one persistent launch, one workgroup per
        specimen with its state resident in LDS (contact-free specimens that
        fit one workgroup's LDS); ends where step() ends, bit for bit."""
        configs = np.ascontiguousarray(configs, dtype=CONFIG)
        _check(library().numi_cloth_hip_batch_step_resident(self._handle, configs.ctypes.data, len(configs),
                                                            iterations, strain_sweeps))

    def synchronize(self) -> None:
        _check(library().numi_cloth_hip_synchronize(self.solver_handle))

    def tile_info(self) -> "TileInfo":
        info = TileInfo()
        _check(library().numi_cloth_hip_tile_info(self.solver_handle, C.byref(info)))
        return info

    def adjoint(self, configs: np.ndarray, iterations: int, strain_sweeps: int, tracked: np.ndarray,
                targets: np.ndarray, checkpoint_interval: int = 16, linear: bool = False) -> None:
        """Run len(configs) substeps with the specimen kernels and the reverse
        pass; `tracked` is specimens x T specimen-local knots, `targets`
        substeps x specimens x T x 3 (the weights of a linear observable when
        linear=True)."""
        configs = np.ascontiguousarray(configs, dtype=CONFIG)
        tracked = np.ascontiguousarray(tracked, np.uint32)
        targets = np.ascontiguousarray(targets, np.float32)
        assert tracked.shape[0] == len(self.specimens)
        assert targets.shape == (len(configs), len(self.specimens), tracked.shape[1], 3)
        problem = AdjointProblem(configs.ctypes.data, len(configs), iterations, strain_sweeps, checkpoint_interval,
                                 tracked.ctypes.data, tracked.shape[1], 1 if linear else 0, targets.ctypes.data)
        _check(library().numi_cloth_hip_batch_adjoint(self._handle, C.byref(problem)))

    def adjoint_stats(self) -> AdjointStats:
        """Strain-limit passes of the last adjoint call's recomputations, and
        how many of them acted (and were taped)."""
        out = AdjointStats()
        _check(library().numi_cloth_hip_batch_adjoint_stats(self._handle, C.byref(out)))
        return out

    def gradients(self, specimen: int) -> dict:
        s = self.specimens[specimen]
        out = {"loss": np.zeros(1), "distanceCompliance": np.zeros(len(s.distances), np.float32),
               "bendCompliance": np.zeros(len(s.bends), np.float32),
               "inverseMass": np.zeros(len(s.particles), np.float32),
               "damping": np.zeros(len(s.particles), np.float32)}
        g = SpecimenGradients(*(_ptr(out[k]) for k in ("loss", "distanceCompliance", "bendCompliance",
                                                        "inverseMass", "damping")))
        _check(library().numi_cloth_hip_batch_gradients(self._handle, specimen, C.byref(g)))
        out["loss"] = float(out["loss"][0])
        return out

    def dlog(self, specimen: int, column_count: int, distance=None, bend=None, mass=None,
             damping=None) -> tuple[np.ndarray, float]:
        """Calibration sensitivity columns of the last adjoint's observable for
        one specimen (numi_cloth_calibration_dlog): `distance`, `bend` (per
        record) and `mass`, `damping` (per knot) give each parameter's column
        or NO_COLUMN; None leaves the family out. Returns (dlog, baseline)."""
        keep = [None if a is None else np.ascontiguousarray(a, np.uint32) for a in (distance, bend, mass, damping)]
        columns = CalibrationColumns(column_count, 0, *[None if a is None else a.ctypes.data for a in keep])
        out = np.zeros(max(column_count, 1))
        baseline = C.c_double()
        _check(library().numi_cloth_calibration_dlog(self._handle, specimen, C.byref(self.specimens[specimen].desc),
                                                     C.byref(columns), out.ctypes.data, C.byref(baseline)))
        return out[:column_count], baseline.value

    def read(self, specimen: int) -> dict:
        s = self.specimens[specimen]
        out = _state_arrays(len(s.particles), len(s.distances), 0, len(s.knots), len(s.bends), 0, 0, 0)
        state = State()
        for name, array in out.items():
            setattr(state, name, _ptr(array))
        _check(library().numi_cloth_hip_batch_read(self._handle, specimen, C.byref(state)))
        out["failure"] = int(out["failure"][0])
        return out


class Solver:
    """Device-resident cloth state on the current HIP device."""

    def __init__(self, scene: Scene):
        self.scene = scene
        handle = C.c_void_p()
        _check(library().numi_cloth_hip_create(C.byref(scene.desc), C.byref(handle)))
        self._handle = handle

    @classmethod
    def borrow(cls, handle: int, scene: Scene) -> "Solver":
        """A non-owning view of a solver created elsewhere (a batch's)."""
        view = cls.__new__(cls)
        view.scene = scene
        view._handle = C.c_void_p(handle)
        view._borrowed = True
        return view

    def close(self) -> None:
        if getattr(self, "_handle", None):
            if not getattr(self, "_borrowed", False):
                _check(library().numi_cloth_hip_destroy(self._handle))
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def stream(self) -> int:
        return library().numi_cloth_hip_stream(self._handle)

    def step(self, configs: np.ndarray, iterations: int = 32, strain_sweeps: int = 3) -> None:
        configs = np.ascontiguousarray(configs, dtype=CONFIG)
        _check(library().numi_cloth_hip_step(self._handle, configs.ctypes.data, len(configs), iterations,
                                             strain_sweeps))

    def step_noalloc(self, configs: np.ndarray, iterations: int = 32, strain_sweeps: int = 3) -> None:
        """step() through numi_cloth_hip_step_noalloc: no device allocation,
        at most 1024 substeps per call."""
        configs = np.ascontiguousarray(configs, dtype=CONFIG)
        _check(library().numi_cloth_hip_step_noalloc(self._handle, configs.ctypes.data, len(configs), iterations,
                                                     strain_sweeps))

    def synchronize(self) -> None:
        _check(library().numi_cloth_hip_synchronize(self._handle))

    def read(self) -> dict:
        s = self.scene
        out = _state_arrays(len(s.particles), len(s.distances), len(s.grips), len(s.knots), len(s.bends),
                            len(s.fruits), len(s.fruitPairs), len(s.yarnContacts))
        state = State()
        for name, array in out.items():
            setattr(state, name, _ptr(array))
        _check(library().numi_cloth_hip_read(self._handle, C.byref(state)))
        out["failure"] = int(out["failure"][0])
        return out

    def replay(self, scenario: int, frames: int, iterations: int = 32, strain_sweeps: int = 3,
               trajectory: "GripTrajectory | None" = None, dump_every: int = 0,
               dump_prefix: str | None = None) -> dict:
        """Frames of a trajectory replay on this solver (numi_cloth_replay_run);
        the solver must have been created from Scene.bag_replay(scenario)."""
        hashes = np.zeros(max(frames, 1), np.uint64)
        out = ReplayResult()
        out.frameHashes = hashes.ctypes.data
        _check(library().numi_cloth_replay_run(self._handle, C.byref(self.scene.desc), scenario,
                                               trajectory._handle if trajectory else None, frames, iterations,
                                               strain_sweeps, dump_every,
                                               dump_prefix.encode() if dump_prefix else None, C.byref(out)))
        return {"hashes": [int(h) for h in hashes[:out.framesRun]], "framesRun": out.framesRun,
                "failureFree": bool(out.failureFree), "releasedMask": out.releasedMask,
                "maximumHandleLag": out.maximumHandleLag, "stepSeconds": out.stepSeconds}

    def time_pass(self, kind: int, colour: int = 0, repeats: int = 50) -> PassTiming:
        out = PassTiming()
        _check(library().numi_cloth_hip_time_pass(self._handle, kind, colour, repeats, C.byref(out)))
        return out

    def self_broad_phase(self) -> dict:
        """Run the self-contact broad phase on the live state; returns the
        sorted (cell id, segment) entries, the candidate pairs and the grid."""
        io = SelfBroadPhase()
        _check(library().numi_cloth_hip_self_broad_phase(self._handle, C.byref(io)))
        pairs = np.zeros((max(io.pairCount, 1), 2), np.uint32)
        cells = np.zeros(max(io.entryCount, 1), np.uint32)
        segments = np.zeros(max(io.entryCount, 1), np.uint32)
        io.pairs, io.cellIds, io.segments = pairs.ctypes.data, cells.ctypes.data, segments.ctypes.data
        io.pairCapacity, io.entryCapacity = io.pairCount, io.entryCount
        _check(library().numi_cloth_hip_self_broad_phase(self._handle, C.byref(io)))
        return {"pairs": pairs[:io.pairCount], "cellIds": cells[:io.entryCount],
                "segments": segments[:io.entryCount], "origin": np.array(io.origin[:], np.float32),
                "cell": np.float32(io.cell), "cells": tuple(io.cells), "valid": io.valid}

    def tile_info(self) -> TileInfo:
        out = TileInfo()
        _check(library().numi_cloth_hip_tile_info(self._handle, C.byref(out)))
        return out

    def diagnostics(self) -> Diagnostics:
        out = Diagnostics()
        _check(library().numi_cloth_hip_diagnostics(self._handle, C.byref(out)))
        return out

    def launch_mode(self) -> LaunchMode:
        out = LaunchMode()
        _check(library().numi_cloth_hip_launch_mode(self._handle, C.byref(out)))
        return out


def replay_pose(scenario: int, time_seconds: float, trajectory: "GripTrajectory | None" = None) -> ReplayPose:
    out = ReplayPose()
    _check(library().numi_cloth_replay_pose(scenario, time_seconds, trajectory._handle if trajectory else None,
                                            C.byref(out)))
    return out


def replay_frame_configs(config: np.ndarray, scenario: int, frame: int,
                         trajectory: "GripTrajectory | None" = None,
                         substeps: int = REPLAY_SUBSTEPS_PER_FRAME) -> np.ndarray:
    """The substep configs of one replay frame (numi_cloth_replay_frame_configs)."""
    base = np.ascontiguousarray(config, dtype=CONFIG).reshape(1)
    out = np.zeros(substeps, CONFIG)
    _check(library().numi_cloth_replay_frame_configs(base.ctypes.data, scenario,
                                                     trajectory._handle if trajectory else None, frame, substeps,
                                                     out.ctypes.data))
    return out


def frame_hash(scene: "Scene", state: dict) -> int:
    """FNV-1a 64 of a read state in the reference's frame-hash order
    (numi_cloth_state_hash)."""
    s = State()
    arrays = dict(state)
    arrays["failure"] = np.array([state["failure"]], np.uint32)
    for name in ["particles", "distances", "grips", "knots", "bends", "fruits", "fruitPairs", "yarnContacts",
                 "selfStatus", "frictionStatus", "aerodynamicsStatus", "releaseStatus", "failure",
                 "mouthClearance"]:
        setattr(s, name, _ptr(arrays[name]))
    out = C.c_uint64()
    _check(library().numi_cloth_state_hash(C.byref(scene.desc), C.byref(s), C.byref(out)))
    return int(out.value)


def write_obj(path, particles: np.ndarray, fruits: np.ndarray, grips: np.ndarray, config: np.ndarray) -> None:
    """Renderer OBJ of a bag state (numi_cloth_obj_write)."""
    base = np.ascontiguousarray(config, dtype=CONFIG).reshape(1)
    _check(library().numi_cloth_obj_write(str(path).encode(), _ptr(particles), len(particles), _ptr(fruits),
                                          len(fruits), _ptr(grips), len(grips), base.ctypes.data))


def write_calibration(path, names: list[str], rows: list[dict]) -> None:
    """Observation CSV of the reference's calibration fitter
    (numi_cloth_calibration_write). Each row: trial, heldout (bool),
    observable, measured, sigma, baseline, dlog (one value per name)."""
    encoded = [n.encode() for n in names]
    name_array = (C.c_char_p * max(len(names), 1))(*encoded)
    dlogs = [np.ascontiguousarray(r["dlog"], np.float64) for r in rows]
    for d in dlogs:
        assert d.shape == (len(names),)
    records = (CalibrationRow * max(len(rows), 1))()
    texts = []
    for k, r in enumerate(rows):
        trial, observable = r["trial"].encode(), r["observable"].encode()
        texts += [trial, observable]
        records[k] = CalibrationRow(trial, 1 if r.get("heldout") else 0, 0, observable, r["measured"], r["sigma"],
                                    r["baseline"], dlogs[k].ctypes.data)
    _check(library().numi_cloth_calibration_write(str(path).encode(), len(names), C.cast(name_array, C.c_void_p),
                                                  records, len(rows)))


def state_hash(arrays: dict) -> int:
    """FNV-1a 64 over the raw bytes of the state arrays, in a fixed order
    (reference tools/cloth_bag_metal.mm:4821-4860 hashes its raw buffers)."""
    lib = library()
    value = 1469598103934665603
    for name in ["particles", "distances", "grips", "knots", "bends", "fruits", "fruitPairs", "yarnContacts"]:
        array = arrays.get(name)
        if array is not None and len(array):
            value = lib.numi_cloth_fnv1a64(array.ctypes.data, array.nbytes, value)
    return int(value)
