#!/usr/bin/env python3
"""Throughput of the XPBD cloth substep loop on one MI355X.

A step is one frame of the workload (all of its substeps, each with the full
iteration / strain-limit / reconciliation schedule) on inputs already resident
in HBM. Prints one JSON line.

Workloads (synthetic, seeded with SplitMix64 / numpy PCG64 from SEED):
  sheet1024  1024 x 1024 sheet (1,048,576 knots) dropped from 0.3 m over 64
             static spheres of radius 2-10 cm at seeded positions, self-contact
             with 2 cm cells (the hash cell only accelerates the candidate
             search; 2 cm cells enter a segment's box 3.4 times on average
             against 14 at 1 cm), 16 substeps per 1/60 s frame, 120 frames
             (default: the bandwidth-roofline workload). The timed frames
             start at frame 25 or after the warm-up, whichever is later, so
             the window is the draped sheet (knots on the spheres, the
             ground and each other), never the free fall; the window and its
             contact counts are reported.
  sheet256   256 x 256 woven sheet (65,536 knots), 1 cm spacing, z jitter
             +-0.1 mm, dropped from 0.3 m onto a fixed r = 0.25 m sphere and
             the ground, hash self-contact with 1 cm cells, 32 substeps per
             1/60 s frame, timed from frame 25 on as above.
  bag        the produce bag (1,465 knots, 12 fruits of radius 2-3 cm),
             stretch compliance 1e-7, bend 1e-4, strain limit 0.10, ground
             on, yarn/yarn self-contact and friction, cuff grip lifted 0.4 m
             over 2 s (smoothstep) then held, 48 substeps per 1/120 s frame.
  pickup     the reference's pickup replay of the same bag (default
             materials, self-contact): grip lifted 1.25 m over 0.8 s, snapped
             at 1.0 s, recovered at 1.45 s, 48 substeps of 1/5760 s per frame,
             the state read back and FNV-hashed after every frame as the
             reference's frame loop does; vs_baseline against the reference's
             Metal pickup time, which covers all 480 frames (vs_baseline is
             given only for a 480-frame window; shorter windows report the
             frames they cover, the later lift frames cost ~4x the first).
             The released mask and the invariant diagnostics of the last
             frame are reported.
  strips512  512 hanging strips, width U[8,64] x height U[64,256] knots,
             compliances log-U[1e-8,1e-5], density U[0.1,0.4] kg/m^2, top row
             driven by a 0.5-2 Hz vertical sinusoid, one batch (merged scene,
             one launch per pass, hipGraph substeps), 48 substeps per 1/60 s
             frame; reports specimen-frames/s and the speed-up over stepping
             the specimens one solver at a time.
  adjoint64  64 strips of 64 x 64 knots (compliances log-U[1e-8,1e-5],
             density U[0.1,0.4], damping U[0.5,2] 1/s), loss = L2 distance of
             32 tracked knots per strip to a seeded target trajectory over 200
             substeps of 1/960 s; gradients with respect to every distance and
             bend compliance, every knot's inverse mass (density) and damping,
             checkpoints every 16 substeps; reports gradients/s (one gradient
             = one specimen's full parameter gradient) and the cost relative
             to the plain forward of the same 200 substeps; every sample runs
             the window from the seeded state on a fresh batch.
All use 32 XPBD iterations and 3 strain-limit sweeps per substep.
"""

from __future__ import annotations

import argparse
import concurrent.futures
import ctypes
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

import __graft_entry__  # noqa: E402

SEED = 20261203
PEAK_HBM_GBPS = 8000.0
ITERATIONS = 32
STRAIN_SWEEPS = 3

SHEETS = {
    "sheet256": dict(width=256, height=256, substeps=32, frame=1.0 / 60.0, sphereMode=1,
                     fixedSphere=(0.0, 0.0, 0.0, 0.25), cellSize=0.01),
    "sheet1024": dict(width=1024, height=1024, substeps=16, frame=1.0 / 60.0, sphereMode=2, sphereCount=64,
                      sphereRadiusMin=0.02, sphereRadiusMax=0.10, cellSize=0.02),
}
# First timed frame of the sheet workloads (0-based): by then the sheet has
# fallen the 0.3 m onto the spheres and the ground.
SHEET_WINDOW_START = 24
WORKLOADS = sorted([*SHEETS, "bag", "pickup", "strips512", "adjoint64"])
# Reference Metal pickup replay: 480 frames x 48 substeps of the 1,465-knot
# bag in 7086.04 s of GPU command time on an Apple M4 Pro (reference
# docs/METAL_CLOTH.md:419, first_gpu_seconds).
PICKUP_REFERENCE = 1465 * 480 * 48 / 7086.04


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return hip


class Timer:
    """HIP events on one stream around a region."""

    def __init__(self, hip, stream: int):
        self.hip = hip
        self.stream = ctypes.c_void_p(stream)
        self.start, self.stop = ctypes.c_void_p(), ctypes.c_void_p()
        hip.hipEventCreate(ctypes.byref(self.start))
        hip.hipEventCreate(ctypes.byref(self.stop))

    def begin(self):
        self.wall = time.perf_counter()
        self.hip.hipEventRecord(self.start, self.stream)

    def end(self) -> float:
        self.hip.hipEventRecord(self.stop, self.stream)
        self.hip.hipEventSynchronize(self.stop)
        self.wall = time.perf_counter() - self.wall
        elapsed = ctypes.c_float()
        self.hip.hipEventElapsedTime(ctypes.byref(elapsed), self.start, self.stop)
        return elapsed.value / 1000.0


# ---------------------------------------------------------------- scenes --

def sheet_scene(cloth, name: str, cell_size: float | None = None):
    """The cell size of the self-contact hash only accelerates the candidate
    search (the candidate set is every overlapping non-local pair whatever the
    grid), so an override changes the time of a step, not its bytes."""
    w = dict(SHEETS[name])
    if cell_size is not None:
        w["cellSize"] = cell_size
    fields = {k: v for k, v in w.items() if k not in ("substeps", "frame")}
    scene = cloth.Scene.sheet(spacing=0.01, elevation=0.3, jitter=1.0e-4, seed=SEED, selfContact=1,
                              timestep=w["frame"] / w["substeps"], **fields)
    return scene, [scene.configs(w["substeps"])], w["frame"]


def bag_scene(cloth, frames: int):
    """Produce bag with the stated materials; frame f's 48 configs lift the
    cuff grip by 0.4 m * smoothstep(t / 2 s)."""
    scene = cloth.Scene.bag(stretchComplianceOverride=1.0e-7, bendComplianceOverride=1.0e-4,
                            strainLimitOverride=0.10, fruitRadiusMin=0.02, fruitRadiusMax=0.03, seed=SEED,
                            selfContact=1)
    scene.config["constraintCounts"][0, 2] = 1
    scene.refresh()
    substeps, frame = 48, 1.0 / 120.0
    base = scene.config[0]["gripTargetAndActive"].astype(np.float64)
    dt = frame / substeps
    out = []
    for f in range(frames):
        configs = scene.configs(substeps)
        t = (f * substeps + np.arange(1, substeps + 1)) * dt
        u = np.clip(t / 2.0, 0.0, 1.0)
        configs["gripTargetAndActive"][:, 2] = base[2] + 0.4 * u * u * (3.0 - 2.0 * u)
        configs["gravityAndTimestep"][:, 3] = dt
        out.append(configs)
    scene.config["gravityAndTimestep"][0, 3] = dt
    scene.refresh()
    return scene, out, frame


def strip_specimens(cloth, count: int = 512):
    """The batched calibration specimens. They share no knot and have no
    contact, so each hangs centred on the origin: every coordinate stays
    within 1.3 m of it, where an FP32 ulp is at most 1.2e-7 m (strips laid
    side by side would reach x ~ 160 m, an ulp of 1.5e-5 m, a 0.15 % error
    on the 1 cm spacing)."""
    rng = np.random.default_rng(SEED)
    specimens = []
    for k in range(count):
        width = int(rng.integers(8, 65))
        height = int(rng.integers(64, 257))
        specimens.append(cloth.Scene.strip(
            width=width, height=height, spacing=0.01, arealDensity=float(rng.uniform(0.1, 0.4)),
            stretchCompliance=float(10.0 ** rng.uniform(-8, -5)), bendCompliance=float(10.0 ** rng.uniform(-8, -5)),
            strainLimit=0.10, timestep=1.0 / (60.0 * 48.0), driveAmplitude=0.01,
            driveFrequency=float(rng.uniform(0.5, 2.0)), drivePhase=float(rng.uniform(0.0, 2.0 * np.pi)),
            origin=(-0.005 * (width - 1), 0.0, 0.005 * (height - 1)), jitter=1.0e-4, seed=SEED + k))
    return specimens


def adjoint_problem(cloth, specimens_count: int = 64, substeps: int = 200, tracked: int = 32):
    rng = np.random.default_rng(SEED + 1)
    dt = 1.0 / 960.0
    specimens = []
    for k in range(specimens_count):
        specimens.append(cloth.Scene.strip(
            width=64, height=64, spacing=0.01, arealDensity=float(rng.uniform(0.1, 0.4)),
            stretchCompliance=float(10.0 ** rng.uniform(-8, -5)), bendCompliance=float(10.0 ** rng.uniform(-8, -5)),
            strainLimit=0.10, timestep=dt, driveAmplitude=0.02, driveFrequency=float(rng.uniform(0.5, 2.0)),
            drivePhase=float(rng.uniform(0.0, 2.0 * np.pi)), damping=float(rng.uniform(0.5, 2.0)),
            origin=(-0.315, 0.0, 0.32), jitter=1.0e-4, seed=SEED + 1000 + k))
    local = np.stack([rng.choice(np.arange(64, 64 * 64), tracked, replace=False) for _ in specimens]).astype(np.uint32)
    x0 = np.stack([s.particles["positionAndInverseMass"][local[k], :3] for k, s in enumerate(specimens)])
    amplitude = rng.uniform(-0.02, 0.02, x0.shape)
    frequency = rng.uniform(0.5, 2.0, x0.shape[:2] + (1,))
    t = (np.arange(1, substeps + 1) * np.float32(dt))[:, None, None, None]
    targets = (x0[None] + amplitude[None] * np.sin(2 * np.pi * frequency[None] * t)).astype(np.float32)
    return specimens, local, targets, specimens[0].configs(substeps)


# ------------------------------------------------------------ baselines --

def host_cpu() -> dict:
    """The host the CPU baseline ran on: CPU model, hardware threads, and the
    threads this process may use."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "hardware_concurrency": os.cpu_count(), "usable_threads": usable}


def cpu_baseline(scene, configs, substeps: int, where: str = "") -> dict:
    """FP64 oracle, one core, on the first `substeps` substeps of the frame
    (from `scene`, the state the timed window starts from)."""
    from oracle import pyoracle
    pyoracle.library()
    start = time.perf_counter()
    pyoracle.run(scene, configs[:substeps], ITERATIONS, STRAIN_SWEEPS)
    seconds = time.perf_counter() - start
    rate = len(scene.particles) * substeps / seconds
    host = host_cpu()
    return {"value": rate, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": f"{substeps} substep(s) of the same workload{where}, FP64 oracle (g++ -O3 -fno-fast-math "
                      f"-ffp-contract=off), {seconds:.2f} s on 1 thread of {host['cpu_model']}", **host}


def cpu_baseline_strips(specimens, count: int) -> dict:
    """FP64 oracle, one specimen per thread (ctypes releases the GIL), one
    frame of each of the first `count` specimens."""
    from oracle import pyoracle
    pyoracle.library()
    threads = max(1, min(16, os.cpu_count() or 1))
    sample = specimens[:count]
    start = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as pool:
        list(pool.map(lambda s: pyoracle.run(s, s.configs(48), ITERATIONS, STRAIN_SWEEPS), sample))
    seconds = time.perf_counter() - start
    host = host_cpu()
    return {"value": len(sample) / seconds, "unit": "specimen-frames/s", "cores": threads, "kind": "port",
            "sample": f"1 frame (48 substeps) of {len(sample)} strips, FP64 oracle, {threads} threads, "
                      f"{seconds:.2f} s on {host['cpu_model']}", **host}


def cpu_baseline_adjoint(specimens, local, targets, configs, iterations: int = ITERATIONS) -> dict:
    """FP64 adjoint oracle, one specimen, all 200 substeps, one core."""
    from oracle import pyoracle
    pyoracle.library()
    scene = specimens[0]
    start = time.perf_counter()
    pyoracle.adjoint(scene, configs if len(configs) else scene.configs(1), iterations, STRAIN_SWEEPS,
                     local[0], targets[:, 0].astype(np.float64))
    seconds = time.perf_counter() - start
    host = host_cpu()
    return {"value": 1.0 / seconds, "unit": "gradients/s", "cores": 1, "kind": "port",
            "sample": f"one 64x64 strip, {len(configs)} substeps, FP64 adjoint oracle (tape recomputed per "
                      f"substep), {seconds:.2f} s on 1 thread of {host['cpu_model']}", **host}


# -------------------------------------------------------------- roofline --

def measured_traffic(workload: str):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (two counter
    runs, each scaled by a copy of known size); None when there is none."""
    path = REPO / "profiles" / f"traffic_{workload}.txt"
    if not path.exists():
        return None
    for text in path.read_text().splitlines():
        if text.startswith("traffic_bytes_per_launch:"):
            return int(text.split(":")[1])
    return None


def roofline(cloth, solver, workload: str) -> tuple[dict, dict]:
    """Dominant kernel, timed with HIP events on the solver's stream after the
    timed region (same state, same launch shape): the LDS tile program of an
    even iteration when the solver runs tiles, else one distance colour."""
    tiles = solver.tile_info()
    if tiles.enabled:
        timing = solver.time_pass(cloth.cloth.PASS_TILE_ITERATION, 0, repeats=50)
        kernel = "k_tile_program (even iteration: distance + bend colours fused in LDS)"
    else:
        timing = solver.time_pass(cloth.cloth.PASS_DISTANCE, 0, repeats=200)
        kernel = "k_distance (one colour batch)"
    achieved = timing.algorithmicBytes / (timing.averageMilliseconds * 1.0e-3) / 1.0e9
    line = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": measured_traffic(workload), "kernel": kernel,
            "bytes_per_launch": timing.algorithmicBytes,
            "avg_launch_us": round(timing.averageMilliseconds * 1000.0, 3),
            "items_per_launch": timing.itemsPerLaunch}
    tile_config = {"enabled": bool(tiles.enabled), "tiles": tiles.tiles, "owned_per_tile": tiles.ownedPerTile,
                   "max_loaded": tiles.maxLoaded, "halo_hops": tiles.halo, "workgroup": tiles.workgroupSize,
                   "passes_fused": tiles.passes, "staged": bool(tiles.staged),
                   "redundancy": round(tiles.redundancy, 3)}
    return line, tile_config


def launch_of(solver) -> dict:
    """How the solver launches substeps (graph replay or direct), as the
    first step()'s probe decided, with the probe's timings."""
    m = solver.launch_mode()
    return {"mode": "graph" if m.graphs else "direct", "probed": bool(m.decided) and m.probeGraphMilliseconds > 0,
            "probe_ms_two_substeps": {"graph": round(m.probeGraphMilliseconds, 3),
                                      "direct": round(m.probeDirectMilliseconds, 3)}}


def diagnostics_of(solver, state) -> dict:
    d = solver.diagnostics()
    self_status = state["selfStatus"]["counters"][0]
    return {"max_strain_violation": d.maximumStrainViolation, "max_ground_penetration": d.maximumGroundPenetration,
            "max_sphere_penetration": d.maximumSpherePenetration, "max_self_penetration": d.maximumSelfPenetration,
            "self_candidates_last_substep": d.selfCandidateCount, "self_candidate_overflow": bool(d.selfCandidateOverflow),
            "self_entries_last_substep": d.selfEntryCount, "self_grid_cells": list(d.selfGridCells),
            "self_contacts_last_step": {"present": int(self_status[0]), "swept": int(self_status[1])},
            "repeated_round_tiles_since_create": {"skipped": int(d.tileRoundsSkipped),
                                                        "run": int(d.tileRoundsRun)},
            "fruit_yarn_solves_since_create": {"reruns": int(d.yarnReruns), "most_candidates": int(d.yarnCandidatesMax),
                                               "list_builds": int(d.yarnListBuilds),
                                               "from_kept_list": int(d.yarnListSolves)}}


# ------------------------------------------------------------- workloads --

def contact_counts(scene, state) -> dict:
    """Knots resting on the ground and on a static sphere (within 1e-6 m of
    contact height) in a read-back state."""
    x = state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)
    r = float(scene.config["clothMaterial"][0, 0])
    ground = int(np.sum(np.abs(x[:, 2] - r) <= 1e-6)) if scene.config["constraintCounts"][0, 2] else 0
    sphere = 0
    for c in scene.staticSpheres.astype(np.float64):
        sphere += int(np.sum(np.abs(np.linalg.norm(x - c[:3], axis=1) - (c[3] + r)) <= 1e-6))
    return {"ground_knots": ground, "sphere_knots": sphere}


def run_plain(args, cloth, hip) -> dict:
    if args.workload == "bag":
        scene, frames, frame_seconds = bag_scene(cloth, args.warmup + args.steps)
        untimed = args.warmup
    else:
        scene, frames, frame_seconds = sheet_scene(cloth, args.workload, args.cell_size)
        untimed = max(args.warmup, SHEET_WINDOW_START)
    def frame(k):
        return frames[k % len(frames)]
    substeps = len(frames[0])
    particles = len(scene.particles)
    solver = cloth.Solver(scene)
    for k in range(untimed):
        solver.step(frame(k), ITERATIONS, STRAIN_SWEEPS)
    solver.synchronize()
    assert hip.hipDeviceSynchronize() == 0
    start_state = solver.read()
    start_contacts = contact_counts(scene, start_state)
    timer = Timer(hip, solver.stream)
    timer.begin()
    for k in range(args.steps):
        solver.step(frame(untimed + k), ITERATIONS, STRAIN_SWEEPS)
    seconds = timer.end()
    solver.synchronize()
    assert hip.hipDeviceSynchronize() == 0
    state = solver.read()
    if state["failure"] != 0:
        raise RuntimeError(f"device failure bits {state['failure']}")
    diagnostics = diagnostics_of(solver, state)
    diagnostics["window_start"] = start_contacts
    diagnostics["window_end"] = contact_counts(scene, state)
    line, tile_config = roofline(cloth, solver, args.workload)
    launch = launch_of(solver)
    released = int(state["releaseStatus"]["masks"][0, 1])
    solver.close()
    baseline = None
    if not args.no_cpu_baseline:
        budget = 32 * 65536 if args.workload != "bag" else 48 * 1465 * 4
        cpu_substeps = args.cpu_substeps or max(1, min(substeps, int(budget / particles)))
        baseline = cpu_baseline(scene.continued(start_state), frame(untimed), cpu_substeps,
                                f" from the state the timed window starts from (frame {untimed + 1})")
    value = particles * substeps * args.steps / seconds
    config = {"workload": args.workload, "particles": particles, "distances": len(scene.distances),
              "bends": len(scene.bends), "knots": len(scene.knots), "fruits": len(scene.fruits),
              "substeps_per_frame": substeps, "iterations": ITERATIONS, "strain_sweeps": STRAIN_SWEEPS,
              "static_spheres": len(scene.staticSpheres), "self_contact": bool(scene.selfContact),
              "self_cell_size": float(scene.config["clothMaterial"][0, 1]),
              "frame_seconds": frame_seconds, "wall_seconds_timed": round(timer.wall, 3),
              "timed_frames": f"{untimed + 1}-{untimed + args.steps} (1-based; {untimed} untimed before)",
              "tile_programs": tile_config, "launch": launch}
    if args.workload == "bag":
        config["released_mask"] = released
    return {"metric": "particle-substeps/s", "value": value, "unit": "particle-substeps/s",
            "ms_per_step": seconds * 1000.0 / args.steps, "dtype": "f32", "config": config, "roofline": line,
            "cpu_baseline": baseline, "diagnostics": diagnostics,
            "data": "synthetic (seeded scene, SplitMix64 seed %d)" % SEED}


def run_pickup(args, cloth, hip) -> dict:
    c = cloth.cloth
    scene = cloth.Scene.bag_replay(c.REPLAY_PICKUP, selfContact=1)
    particles = len(scene.particles)

    def replay(solver, frames, report):
        """The reference's frame loop: a frame's 48 substep configs, the step,
        the state read back and FNV-hashed (the library's frame driver does
        the same in numi_cloth_replay_run; this loop reports its progress)."""
        hashes = []
        for f in range(frames):
            solver.step(c.replay_frame_configs(scene.config, c.REPLAY_PICKUP, f), ITERATIONS, STRAIN_SWEEPS)
            state = solver.read()
            if state["failure"] != 0:
                raise RuntimeError(f"pickup replay raised failure bits {state['failure']} in frame {f + 1}")
            hashes.append(c.frame_hash(scene, state))
            if report and ((f + 1) % 24 == 0 or f + 1 == frames):
                print(f"pickup frame {f + 1}/{frames}: released mask {int(state['releaseStatus']['masks'][0, 1])}",
                      file=sys.stderr, flush=True)
        return hashes

    solver = cloth.Solver(scene)
    if args.warmup:
        replay(solver, args.warmup, False)
    # The timed frames start the replay again from t = 0 on a fresh solver.
    solver.close()
    solver = cloth.Solver(scene)
    solver.synchronize()
    assert hip.hipDeviceSynchronize() == 0
    timer = Timer(hip, solver.stream)
    timer.begin()
    hashes = replay(solver, args.steps, True)
    seconds = timer.end()
    state = solver.read()
    d = solver.diagnostics()
    # The reference's pickup qualification (tools/cloth_bag_metal.mm:6214-6240):
    # at least two fruits released, at least two of them resting on the
    # ground, strain and self violation <= 2e-6, ground penetration <= 1e-6.
    released = int(state["releaseStatus"]["masks"][0, 1])
    fruits = state["fruits"]
    grounded = sum(1 for k in range(len(fruits)) if released & (1 << k) and
                   abs(float(fruits["positionAndInverseMass"][k, 2]) - float(fruits["previousAndRadius"][k, 3])) <= 2e-6)
    qualification = {"released_mask": released, "released_count": bin(released).count("1"),
                     "grounded_released_count": grounded, "max_strain_violation": d.maximumStrainViolation,
                     "max_ground_penetration": d.maximumGroundPenetration,
                     "max_self_penetration_over_candidates": d.maximumSelfPenetration,
                     "reference_published": "480 frames: released_mask=1537 released_count=3 "
                                            "grounded_released_count=2, violations 0"}
    diagnostics = diagnostics_of(solver, state)
    line, tile_config = roofline(cloth, solver, args.workload)
    solver.close()
    baseline = None
    if not args.no_cpu_baseline:
        baseline = cpu_baseline(scene, c.replay_frame_configs(scene.config, c.REPLAY_PICKUP, 0), 48,
                                " (the replay's first frame)")
    value = particles * 48 * args.steps / seconds
    config = {"workload": "pickup", "particles": particles, "substeps_per_frame": 48, "iterations": ITERATIONS,
              "strain_sweeps": STRAIN_SWEEPS, "frames": args.steps, "self_contact": True,
              "frame_seconds": 48 * c.REPLAY_TIMESTEP, "wall_seconds_timed": round(timer.wall, 3),
              "last_frame_hashes": [hex(h) for h in hashes[-3:]],
              "qualification_at_last_frame": qualification, "tile_programs": tile_config,
              "reference": "Metal pickup replay, 480 frames x 48 substeps, 7086.04 s GPU command time on an Apple "
                           "M4 Pro -> %.1f particle-substeps/s" % PICKUP_REFERENCE}
    # The reference's figure covers all 480 frames (the later lift frames
    # cost ~4x the first ones), so only a 480-frame window compares with it.
    vs = value / PICKUP_REFERENCE if args.steps == 480 else None
    if vs is None:
        config["vs_baseline_note"] = ("null: the reference time covers frames 1-480; this window covers frames "
                                      f"1-{args.steps}")
    return {"metric": "particle-substeps/s", "value": value, "unit": "particle-substeps/s",
            "ms_per_step": seconds * 1000.0 / args.steps, "dtype": "f32", "config": config, "roofline": line,
            "cpu_baseline": baseline, "vs_baseline": vs, "diagnostics": diagnostics,
            "data": "reference pickup trajectory (deterministic), default bag materials"}


def run_strips(args, cloth, hip) -> dict:
    specimens = strip_specimens(cloth)
    configs = specimens[0].configs(48)
    particles = sum(len(s.particles) for s in specimens)
    batch = cloth.Batch(specimens)
    for _ in range(args.warmup):
        batch.step(configs, ITERATIONS, STRAIN_SWEEPS)
    batch.synchronize()
    assert hip.hipDeviceSynchronize() == 0
    solver_handle = batch.solver_handle
    stream = cloth.cloth.library().numi_cloth_hip_stream(solver_handle)
    timer = Timer(hip, stream)
    timer.begin()
    for _ in range(args.steps):
        batch.step(configs, ITERATIONS, STRAIN_SWEEPS)
    seconds = timer.end()
    batch.synchronize()
    failure = batch.read(0)["failure"]
    if failure != 0:
        raise RuntimeError(f"device failure bits {failure}")
    view = cloth.Solver.borrow(solver_handle, specimens[0])
    line, tile_config = roofline(cloth, view, args.workload)
    batch.close()
    batched = len(specimens) * args.steps / seconds

    # One specimen per solver: the plain single-specimen path, launched once
    # per specimen, on a seeded sample of the same specimens.
    sample = [specimens[k] for k in np.random.default_rng(SEED).choice(len(specimens), 16, replace=False)]
    total = 0.0
    for scene in sample:
        own = scene.configs(48)
        with cloth.Solver(scene) as solver:
            solver.step(own, ITERATIONS, STRAIN_SWEEPS)
            solver.synchronize()
            t = Timer(hip, solver.stream)
            t.begin()
            for _ in range(2):
                solver.step(own, ITERATIONS, STRAIN_SWEEPS)
            total += t.end()
    single = 2 * len(sample) / total
    resident = resident_comparison(cloth, hip, specimens, configs, max(args.steps, 1))
    split = split_job(cloth, specimens, configs, max(args.steps, 1))
    baseline = None if args.no_cpu_baseline else cpu_baseline_strips(specimens, 16)
    config = {"workload": "strips512", "specimens": len(specimens), "particles": particles,
              "substeps_per_frame": 48, "iterations": ITERATIONS, "strain_sweeps": STRAIN_SWEEPS,
              "frame_seconds": 1.0 / 60.0, "wall_seconds_timed": round(timer.wall, 3),
              "one_per_launch_specimen_frames_per_s": single, "speedup_vs_one_per_launch": batched / single,
              "one_per_launch_sample": "16 seeded specimens, 2 frames each, own solver each",
              "resident_subset": resident, "split_job": split,
              "split_over_merged": split["specimen_frames_per_s"] / batched,
              "particle_substeps_per_s": particles * 48 * args.steps / seconds, "tile_programs": tile_config}
    return {"metric": "batched specimen-frames/s", "value": batched, "unit": "specimen-frames/s",
            "ms_per_step": seconds * 1000.0 / args.steps, "dtype": "f32", "config": config, "roofline": line,
            "cpu_baseline": baseline, "data": "synthetic (seeded strips, PCG64 seed %d)" % SEED}


def fits_one_workgroup(scene) -> bool:
    """The specimen's state fits one workgroup's LDS (160 KiB on gfx950):
    16 B per knot, 4 B per distance and bend multiplier, at most 5120 knots."""
    lds = 16 * len(scene.particles) + 4 * (len(scene.distances) + len(scene.bends)) + 1024
    return len(scene.particles) <= 5120 and lds <= 160 * 1024


def resident_comparison(cloth, hip, specimens, configs, frames: int) -> dict:
    """The merged tile path (Batch.step) against one persistent launch per
    call (Batch.step_resident: a workgroup per specimen, its state resident
    in LDS for every substep), each on a fresh batch of the specimens that
    fit one workgroup, one untimed frame then `frames` timed frames; both
    must end in the same bytes."""
    subset = [s for s in specimens if fits_one_workgroup(s)]
    rates, ends = {}, {}
    for mode in ("tiles", "resident"):
        with cloth.Batch(subset) as batch:
            step = batch.step if mode == "tiles" else batch.step_resident
            step(configs, ITERATIONS, STRAIN_SWEEPS)
            batch.synchronize()
            stream = cloth.cloth.library().numi_cloth_hip_stream(batch.solver_handle)
            timer = Timer(hip, stream)
            timer.begin()
            for _ in range(frames):
                step(configs, ITERATIONS, STRAIN_SWEEPS)
            seconds = timer.end()
            batch.synchronize()
            ends[mode] = [batch.read(k)["particles"].tobytes() for k in range(len(subset))]
            rates[mode] = len(subset) * frames / seconds
            print(f"strips512 {mode}: {rates[mode]:.1f} specimen-frames/s over {len(subset)} specimens",
                  file=sys.stderr, flush=True)
    return {"specimens": len(subset), "particles": sum(len(s.particles) for s in subset),
            "largest": max(len(s.particles) for s in subset), "frames": frames,
            "merged_tiles_specimen_frames_per_s": rates["tiles"],
            "resident_specimen_frames_per_s": rates["resident"],
            "resident_over_tiles": rates["resident"] / rates["tiles"],
            "equal_bytes": ends["tiles"] == ends["resident"]}


def split_job(cloth, specimens, configs, frames: int) -> dict:
    """All specimens as two batches stepped side by side on their own
    streams: those that fit one workgroup by the persistent resident launch,
    the rest by the merged tile path; whole-job wall time from the host clock
    (both streams drained before and after)."""
    fit = [s for s in specimens if fits_one_workgroup(s)]
    rest = [s for s in specimens if not fits_one_workgroup(s)]
    with cloth.Batch(fit) as resident, cloth.Batch(rest) as merged:
        resident.step_resident(configs, ITERATIONS, STRAIN_SWEEPS)
        merged.step(configs, ITERATIONS, STRAIN_SWEEPS)
        resident.synchronize()
        merged.synchronize()
        start = time.perf_counter()
        for _ in range(frames):
            resident.step_resident(configs, ITERATIONS, STRAIN_SWEEPS)
            merged.step(configs, ITERATIONS, STRAIN_SWEEPS)
        resident.synchronize()
        merged.synchronize()
        seconds = time.perf_counter() - start
        if resident.read(0)["failure"] != 0 or merged.read(0)["failure"] != 0:
            raise RuntimeError("device failure in the split job")
    rate = len(specimens) * frames / seconds
    print(f"strips512 split: {rate:.1f} specimen-frames/s ({len(fit)} resident, {len(rest)} merged)",
          file=sys.stderr, flush=True)
    return {"resident_specimens": len(fit), "merged_specimens": len(rest), "frames": frames,
            "specimen_frames_per_s": rate, "ms_per_frame": seconds * 1000.0 / frames}


def run_adjoint(args, cloth, hip) -> dict:
    specimens, local, targets, configs = adjoint_problem(cloth)
    substeps = len(configs)
    particles = sum(len(s.particles) for s in specimens)
    start = args.adjoint_start
    iterations = ITERATIONS if args.adjoint_iterations is None else args.adjoint_iterations
    # One trajectory per path, as a calibration loop runs it: the untimed
    # substeps before `start`, `warm` untimed 200-substep calls (they capture
    # the forward's graph and size the adjoint's checkpoints and tape), then
    # `steps` timed calls over the next consecutive 200-substep windows. The
    # forward is the library's step path over the same windows.
    warm = max(1, args.warmup)
    forward = cloth.Batch(specimens)
    if start:
        forward.step(specimens[0].configs(start), iterations, STRAIN_SWEEPS)
    for _ in range(warm):
        forward.step(configs, iterations, STRAIN_SWEEPS)
    forward.synchronize()
    forward_timer = Timer(hip, cloth.cloth.library().numi_cloth_hip_stream(forward.solver_handle))
    forward_times = []
    for _ in range(args.steps):
        forward_timer.begin()
        forward.step(configs, iterations, STRAIN_SWEEPS)
        forward_times.append(forward_timer.end())
    view = cloth.Solver.borrow(forward.solver_handle, specimens[0])
    line, tile_config = roofline(cloth, view, args.workload)
    forward.close()
    forward_seconds = sum(forward_times) / len(forward_times)
    # Adjoint: forward with checkpoints, then the reverse; the library call
    # returns when the gradients are in device memory.
    batch = cloth.Batch(specimens)
    if start:
        batch.step(specimens[0].configs(start), iterations, STRAIN_SWEEPS)
    for _ in range(warm):
        batch.adjoint(configs, iterations, STRAIN_SWEEPS, local, targets, 16)
    batch.synchronize()
    timer = Timer(hip, cloth.cloth.library().numi_cloth_hip_stream(batch.solver_handle))
    adjoint_times = []
    for _ in range(args.steps):
        timer.begin()
        batch.adjoint(configs, iterations, STRAIN_SWEEPS, local, targets, 16)
        adjoint_times.append(timer.end())
    adjoint_seconds = sum(adjoint_times) / len(adjoint_times)
    grads = batch.gradients(0)
    stats = batch.adjoint_stats()
    failure = batch.read(0)["failure"]
    batch.close()
    if failure != 0:
        raise RuntimeError(f"device failure bits {failure}")
    config = {"workload": "adjoint64", "specimens": len(specimens), "particles": particles, "substeps": substeps,
              "checkpoint_interval": 16, "tracked_per_specimen": int(local.shape[1]), "iterations": iterations,
              "strain_sweeps": STRAIN_SWEEPS, "forward_seconds": forward_seconds,
              "adjoint_seconds": adjoint_seconds, "adjoint_over_forward": adjoint_seconds / forward_seconds,
              "window": f"substeps {start + warm * substeps}-{start + (warm + args.steps) * substeps} of the "
                        f"trajectory from the seeded state, {args.steps} consecutive calls of {substeps} substeps "
                        f"after {start + warm * substeps} untimed; the forward over the same substeps",
              "gradient_entries_per_specimen": int(grads["distanceCompliance"].size + grads["bendCompliance"].size
                                                   + grads["inverseMass"].size + grads["damping"].size),
              "loss_specimen0": grads["loss"], "tile_programs": tile_config,
              "strain_passes": int(stats.strainPasses), "strain_passes_acting": int(stats.strainPassesTaped),
              "strain_acting_fraction": stats.strainPassesTaped / max(1, stats.strainPasses)}
    return {"metric": "adjoint gradients/s", "value": len(specimens) / adjoint_seconds, "unit": "gradients/s",
            "ms_per_step": adjoint_seconds * 1000.0, "dtype": "f32", "config": config, "roofline": line,
            "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_adjoint(specimens, local, targets, configs, iterations),
            "data": "synthetic (seeded strips and targets, PCG64 seed %d)" % (SEED + 1)}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--workload", choices=WORKLOADS, default="sheet1024")
    parser.add_argument("--cpu-substeps", type=int, default=None,
                        help="substeps of the CPU FP64 baseline sample (default: sized to ~15 s)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--adjoint-start", type=int, default=0, metavar="SUBSTEPS",
                        help="adjoint64: substeps stepped (untimed) before the timed 200-substep window")
    parser.add_argument("--adjoint-iterations", type=int, default=None, metavar="N",
                        help="adjoint64: solver iterations per substep of both the forward and the adjoint "
                             "(default: the library's 32; few iterations leave the distance constraints "
                             "unconverged, so the strain limits act and their passes are taped)")
    parser.add_argument("--cell-size", type=float, default=None,
                        help="self-contact hash cell of the sheet workloads in metres (default: the workload's)")
    parser.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                        help="after the workload, copy a MIB-MiB float32 buffer three times with an elementwise "
                             "kernel (reads and writes MIB MiB each): the reference dispatch that scales the "
                             "FETCH_SIZE / WRITE_SIZE counters of a profiled run to bytes")
    args = parser.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the cloth substep loop of this library runs on a single "
              "MI355X (multi-GPU sharding is a design note, not built); rerun with --gpus 1", file=sys.stderr)
        return 2
    cloth = __graft_entry__.load_package()
    hip = _hip()
    if args.workload == "strips512":
        result = run_strips(args, cloth, hip)
    elif args.workload == "adjoint64":
        result = run_adjoint(args, cloth, hip)
    elif args.workload == "pickup":
        result = run_pickup(args, cloth, hip)
    else:
        result = run_plain(args, cloth, hip)
    line = {"metric": result["metric"], "value": result["value"], "unit": result["unit"], "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": result["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": result.get("vs_baseline"), "dtype": result["dtype"],
            "data": result["data"], "config": result["config"], "roofline": result["roofline"],
            "cpu_baseline": result["cpu_baseline"]}
    if "diagnostics" in result:
        line["diagnostics"] = result["diagnostics"]
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    print(json.dumps(line))
    return 0


def copy_calibration(mib: int) -> None:
    import torch
    n = mib * (1 << 20) // 4
    source = torch.ones(n, dtype=torch.float32, device="cuda")
    target = torch.empty_like(source)
    for _ in range(3):
        torch.mul(source, 1.0, out=target)
    torch.cuda.synchronize()


if __name__ == "__main__":
    sys.exit(main())
