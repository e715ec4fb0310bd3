#!/usr/bin/env python3
"""Throughput of the XPBD cloth substep loop on one MI355X.

A step is one frame of the workload (all of its substeps, each with the full
iteration / strain-limit / reconciliation schedule) on inputs already resident
in HBM. Prints one JSON line.

Workloads (synthetic, seeded):
  sheet256   256 x 256 woven sheet (65,536 knots), 1 cm spacing, z jitter
             +-0.1 mm, dropped from 0.3 m onto a fixed r = 0.25 m sphere and
             the ground, 32 substeps per 1/60 s frame (default).
  sheet1024  1024 x 1024 sheet (1,048,576 knots) over 64 static spheres of
             radius 2-10 cm at seeded positions, 16 substeps per 1/60 s frame.
Both use 32 XPBD iterations and 3 strain-limit sweeps per substep.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

import __graft_entry__  # noqa: E402

SEED = 20261203
PEAK_HBM_GBPS = 8000.0

WORKLOADS = {
    "sheet256": dict(width=256, height=256, substeps=32, frame=1.0 / 60.0, sphereMode=1,
                     fixedSphere=(0.0, 0.0, 0.0, 0.25)),
    "sheet1024": dict(width=1024, height=1024, substeps=16, frame=1.0 / 60.0, sphereMode=2, sphereCount=64,
                      sphereRadiusMin=0.02, sphereRadiusMax=0.10),
}
ITERATIONS = 32
STRAIN_SWEEPS = 3


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return hip


def build_scene(cloth, name: str):
    w = WORKLOADS[name]
    fields = {k: v for k, v in w.items() if k not in ("substeps", "frame")}
    scene = cloth.Scene.sheet(spacing=0.01, elevation=0.3, jitter=1.0e-4, seed=SEED, cellSize=0.01,
                              selfContact=1, timestep=w["frame"] / w["substeps"], **fields)
    return scene, scene.configs(w["substeps"])


def cpu_baseline(scene, configs, substeps: int) -> dict:
    """FP64 oracle, one core, on the first `substeps` substeps of the frame."""
    from oracle import pyoracle
    pyoracle.library()
    start = time.perf_counter()
    pyoracle.run(scene, configs[:substeps], ITERATIONS, STRAIN_SWEEPS)
    seconds = time.perf_counter() - start
    rate = len(scene.particles) * substeps / seconds
    return {"value": rate, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": f"{substeps} substep(s) of the same frame, FP64 oracle (g++ -O3 -fno-fast-math "
                      f"-ffp-contract=off), {seconds:.2f} s on {platform.processor() or platform.machine()}"}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--workload", choices=sorted(WORKLOADS), default="sheet256")
    parser.add_argument("--cpu-substeps", type=int, default=None,
                        help="substeps of the CPU FP64 baseline sample (default: sized to ~15 s)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    args = parser.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the cloth substep loop of this library runs on a single "
              "MI355X (multi-GPU sharding is a design note, not built); rerun with --gpus 1", file=sys.stderr)
        return 2

    cloth = __graft_entry__.load_package()
    hip = _hip()
    scene, configs = build_scene(cloth, args.workload)
    substeps = len(configs)
    particles = len(scene.particles)
    solver = cloth.Solver(scene)
    stream = ctypes.c_void_p(solver.stream)

    for _ in range(args.warmup):
        solver.step(configs, ITERATIONS, STRAIN_SWEEPS)
    solver.synchronize()
    assert hip.hipDeviceSynchronize() == 0

    start, stop = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(start))
    hip.hipEventCreate(ctypes.byref(stop))
    wall = time.perf_counter()
    hip.hipEventRecord(start, stream)
    for _ in range(args.steps):
        solver.step(configs, ITERATIONS, STRAIN_SWEEPS)
    hip.hipEventRecord(stop, stream)
    hip.hipEventSynchronize(stop)
    solver.synchronize()
    assert hip.hipDeviceSynchronize() == 0
    wall = time.perf_counter() - wall
    elapsed = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(elapsed), start, stop)
    seconds = elapsed.value / 1000.0
    state = solver.read()
    if state["failure"] != 0:
        print(f"device failure bits {state['failure']}", file=sys.stderr)
        return 1
    diagnostics = solver.diagnostics()
    self_status = state["selfStatus"]["counters"][0]

    # Dominant kernel, timed with HIP events on the solver's stream right after
    # the timed region (same state, same launch shape): the LDS tile program of
    # an even iteration when the solver runs tiles, else one distance colour.
    tiles = solver.tile_info()
    if tiles.enabled:
        timing = solver.time_pass(cloth.cloth.PASS_TILE_ITERATION, 0, repeats=50)
        kernel = "k_tile_program (even iteration: distance + bend colours fused in LDS)"
    else:
        timing = solver.time_pass(cloth.cloth.PASS_DISTANCE, 0, repeats=200)
        kernel = "k_distance (one colour batch)"
    achieved = timing.algorithmicBytes / (timing.averageMilliseconds * 1.0e-3) / 1.0e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None, "kernel": kernel,
                "bytes_per_launch": timing.algorithmicBytes,
                "avg_launch_us": round(timing.averageMilliseconds * 1000.0, 3),
                "items_per_launch": timing.itemsPerLaunch}
    tile_config = {"enabled": bool(tiles.enabled), "tiles": tiles.tiles, "owned_per_tile": tiles.ownedPerTile,
                   "max_loaded": tiles.maxLoaded, "halo_hops": tiles.halo, "workgroup": tiles.workgroupSize,
                   "passes_fused": tiles.passes, "redundancy": round(tiles.redundancy, 3)}
    solver.close()

    baseline = None
    if not args.no_cpu_baseline:
        cpu_substeps = args.cpu_substeps or max(1, min(substeps, int(32 * 65536 / particles)))
        baseline = cpu_baseline(scene, configs, cpu_substeps)

    value = particles * substeps * args.steps / seconds
    line = {
        "metric": "particle-substeps/s", "value": value, "unit": "particle-substeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": seconds * 1000.0 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded sheet, SplitMix64 seed %d)" % SEED,
        "config": {"workload": args.workload, "particles": particles, "distances": len(scene.distances),
                   "bends": len(scene.bends), "substeps_per_frame": substeps, "iterations": ITERATIONS,
                   "strain_sweeps": STRAIN_SWEEPS, "static_spheres": len(scene.staticSpheres),
                   "self_contact": bool(scene.selfContact), "frame_seconds": WORKLOADS[args.workload]["frame"],
                   "wall_seconds_timed": round(wall, 3), "tile_programs": tile_config},
        "roofline": roofline,
        "cpu_baseline": baseline,
        "diagnostics": {"max_strain_violation": diagnostics.maximumStrainViolation,
                        "max_ground_penetration": diagnostics.maximumGroundPenetration,
                        "max_sphere_penetration": diagnostics.maximumSpherePenetration,
                        "max_self_penetration": diagnostics.maximumSelfPenetration,
                        "self_candidates_last_substep": diagnostics.selfCandidateCount,
                        "self_candidate_overflow": bool(diagnostics.selfCandidateOverflow),
                        "self_contacts_last_step": {"present": int(self_status[0]), "swept": int(self_status[1])}},
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
