"""Every benchmark workload at its own size on the GPU (needs a GPU), built
by the same functions bench.py times.

Two gates. Against the FP32 twin (the oracle's FP32 restatement in the
device's operation order) the GPU must agree bit for bit on every frame.
Against FP64 the parity gate, positions within 1e-5 of the cloth
extent after one frame, is asserted on the frames FP32 arithmetic can meet
it: a workload's first frame from its seeded state and the frame in which a
draping sheet first touches its spheres. Once a stiff sheet (axial
compliance 1e-7, 32 iterations) lies wrapped over a sphere, the exact FP32
restatement itself ends 1e-4 m from FP64 after one frame, while FP64 moves
by 1e-7 m under a one-ulp perturbation of its input: the gap is FP32
rounding of the algorithm, not conditioning, and those frames are gated on
the twin (bytes) alone, with the FP64 distance printed.

  sheet256   256 x 256 sheet (sphere, ground, 1 cm hash self-contact):
             frame 1 and frame 6 (first sphere contact) under the plain
             gate; the draped frame 31 bit for bit against the twin;
  sheet1024  the 1,048,576-knot sheet over 64 spheres: invariants after it
             has landed (strain, ground, sphere, self penetration), and the
             same bytes from a second solver, from per-colour launches
             instead of tile programs, and with sphere windows switched off;
             plus a 128 x 128 sheet over 16 spheres, where the windowed
             sphere pass meets the twin (bytes) and FP64 (plain gate, first
             sphere contact);
  strips512  16 seeded strips drawn from the batched workload's
             distribution (8-64 x 64-256 knots), batched: bytes of the twin,
             plain FP64 gate;
  adjoint64  two 64 x 64 strips at 32 iterations and 3 strain sweeps over
             the workload's 200 substeps: loss and gradients against the
             FP32 twin and FP64.
"""

import concurrent.futures
import os

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

ITERATIONS, SWEEPS = 32, 3


def _positions(state):
    return state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)


def _extent(scene):
    x = scene.particles["positionAndInverseMass"][:, :3]
    return float(np.max(x.max(axis=0) - x.min(axis=0)))


def _contacts(scene, state):
    """Knots on the ground and on a static sphere (within 1e-6 m of contact)."""
    x = _positions(state)
    r = float(scene.config["clothMaterial"][0, 0])
    ground = int(np.sum(np.abs(x[:, 2] - r) <= 1e-6))
    sphere = 0
    for s in scene.staticSpheres.astype(np.float64):
        sphere += int(np.sum(np.abs(np.linalg.norm(x - s[:3], axis=1) - (s[3] + r)) <= 1e-6))
    return ground, sphere


def _with_env(name, value, fn):
    saved = os.environ.get(name)
    os.environ[name] = value
    try:
        return fn()
    finally:
        if saved is None:
            os.environ.pop(name, None)
        else:
            os.environ[name] = saved


def _frames(cloth, scene, configs, frames):
    with cloth.Solver(scene) as solver:
        for _ in range(frames):
            solver.step(configs, ITERATIONS, SWEEPS)
        return solver.read(), solver.diagnostics()


def _oracle_frame(oracle, scene, configs):
    with concurrent.futures.ThreadPoolExecutor(2) as pool:
        twin = pool.submit(oracle.run, scene, configs, ITERATIONS, SWEEPS, "fp32")
        fp64 = pool.submit(oracle.run, scene, configs, ITERATIONS, SWEEPS)
        return twin.result(), fp64.result()


def _gaps(oracle, scene, configs, gpu_state):
    twin, fp64 = _oracle_frame(oracle, scene, configs)
    gpu = _positions(gpu_state)
    return (gpu, twin, fp64, float(np.max(np.abs(gpu - twin["positions"]))),
            float(np.max(np.abs(gpu - fp64["positions"]))))


def test_sheet256_frames_against_the_oracle(cloth, oracle):
    scene, frames, _ = bench.sheet_scene(cloth, "sheet256")
    configs = frames[0]
    keep = {1, 5, 6, 30, 31}
    states = {}
    with cloth.Solver(scene) as solver:
        for f in range(1, 32):
            solver.step(configs, ITERATIONS, SWEEPS)
            if f in keep:
                states[f] = solver.read()
        diag = solver.diagnostics()
    extent = _extent(scene)
    gate = 1e-5 * extent
    # Frame 1 (free fall) from the seeded state, frame 6 from the GPU's state
    # after frame 5: the sheet meets the sphere.
    for frame, start in [(1, scene), (6, scene.continued(states[5]))]:
        gpu, twin, fp64, e_twin, e_fp64 = _gaps(oracle, start, configs, states[frame])
        ground, sphere = _contacts(scene, states[frame])
        print(f"sheet256 frame {frame}: |gpu-twin| {e_twin:.3e} |gpu-fp64| {e_fp64:.3e} gate {gate:.3e}; "
              f"sphere knots {sphere}, ground knots {ground}")
        assert np.array_equal(gpu, twin["positions"])
        assert e_fp64 <= gate
        if frame == 6:
            assert sphere > 0
    # Frame 31, draped over the sphere and on the ground: the twin's bytes.
    start = scene.continued(states[30])
    with cloth.Solver(start) as solver:
        solver.step(configs, ITERATIONS, SWEEPS)
        again = solver.read()
    assert again["particles"].tobytes() == states[31]["particles"].tobytes()
    twin = oracle.run(start, configs, ITERATIONS, SWEEPS, precision="fp32")
    fp64 = oracle.run(start, configs, ITERATIONS, SWEEPS)
    gpu = _positions(states[31])
    ground, sphere = _contacts(scene, states[31])
    counters = states[31]["selfStatus"]["counters"][0]
    print(f"sheet256 frame 31: |gpu-twin| {np.max(np.abs(gpu - twin['positions'])):.3e} |twin-fp64| "
          f"{np.max(np.abs(twin['positions'] - fp64['positions'])):.3e}; ground knots {ground}, sphere knots "
          f"{sphere}, self candidates {diag.selfCandidateCount}, self contacts present/swept "
          f"{counters[0]}/{counters[1]} (twin {twin['presentContacts']}/{twin['sweptContacts']})")
    assert states[31]["failure"] == 0
    assert ground > 1000 and sphere > 100
    assert np.array_equal(gpu, twin["positions"])
    assert (int(counters[0]), int(counters[1])) == (twin["presentContacts"], twin["sweptContacts"])


def test_sheet1024_invariants_and_equal_bytes_on_every_path(cloth):
    scene, frames, _ = bench.sheet_scene(cloth, "sheet1024")
    configs = frames[0]
    landed = 20   # 0.33 s: the sheet lies on the spheres, its edges still falling
    state, diag = _frames(cloth, scene, configs, landed)
    assert state["failure"] == 0
    ground, sphere = _contacts(scene, state)
    print(f"sheet1024 after {landed} frames: strain {diag.maximumStrainViolation:.3e} ground "
          f"{diag.maximumGroundPenetration:.3e} sphere {diag.maximumSpherePenetration:.3e} self "
          f"{diag.maximumSelfPenetration:.3e}; ground knots {ground}, sphere knots {sphere}, self candidates "
          f"{diag.selfCandidateCount}, entries {diag.selfEntryCount}")
    assert sphere > 1000
    assert diag.maximumStrainViolation <= 1e-6
    assert diag.maximumGroundPenetration <= 1e-6
    assert diag.maximumSpherePenetration <= 1e-6
    assert diag.maximumSelfPenetration <= 2e-6
    assert not diag.selfCandidateOverflow
    reference = state["particles"].tobytes()
    assert _frames(cloth, scene, configs, landed)[0]["particles"].tobytes() == reference
    unwindowed = _with_env("NUMI_CLOTH_HIP_SPHERE_WINDOW", "0", lambda: _frames(cloth, scene, configs, landed))
    assert unwindowed[0]["particles"].tobytes() == reference
    per_colour = _with_env("NUMI_CLOTH_HIP_TILES", "0", lambda: _frames(cloth, scene, configs, landed))
    assert per_colour[0]["particles"].tobytes() == reference


def test_windowed_sphere_pass_meets_the_oracle(cloth, oracle):
    """128 x 128 sheet over 16 static spheres (the 1M sheet's sphere
    distribution on a smaller sheet). Frame 9, the first with knots on the
    spheres, and frame 13 (spheres and self-contact): windows forced on,
    forced off and the default give the same bytes, those of the FP32 twin
    (which tests every sphere); frame 9 meets the plain FP64 gate."""
    scene = cloth.Scene.sheet(width=128, height=128, spacing=0.01, elevation=0.3, jitter=1.0e-4, seed=bench.SEED,
                              selfContact=1, timestep=1.0 / 960.0, sphereMode=2, sphereCount=16,
                              sphereRadiusMin=0.02, sphereRadiusMax=0.10, cellSize=0.02)
    configs = scene.configs(16)
    states = {}
    with cloth.Solver(scene) as solver:
        for f in range(1, 13):
            solver.step(configs, ITERATIONS, SWEEPS)
            if f in (8, 12):
                states[f] = solver.read()
    extent = _extent(scene)
    for before in (8, 12):
        start = scene.continued(states[before])
        runs = {mode: _with_env("NUMI_CLOTH_HIP_SPHERE_WINDOW", mode, lambda: _frames(cloth, start, configs, 1)[0])
                for mode in ("0", "1")}
        default = _frames(cloth, start, configs, 1)[0]
        assert runs["0"]["particles"].tobytes() == default["particles"].tobytes()
        assert runs["1"]["particles"].tobytes() == default["particles"].tobytes()
        gpu, twin, fp64, e_twin, e_fp64 = _gaps(oracle, start, configs, default)
        ground, sphere = _contacts(scene, default)
        counters = default["selfStatus"]["counters"][0]
        print(f"16-sphere sheet frame {before + 1}: |gpu-twin| {e_twin:.3e} |gpu-fp64| {e_fp64:.3e} gate "
              f"{1e-5 * extent:.3e}; sphere knots {sphere}, self contacts {counters[0]}/{counters[1]}")
        assert np.array_equal(gpu, twin["positions"])
        assert sphere > 100
        if before == 8:
            assert e_fp64 <= 1e-5 * extent
        else:
            assert counters[0] > 0


def test_strips_from_the_batched_distribution(cloth, oracle):
    specimens = bench.strip_specimens(cloth)
    pick = sorted(np.random.default_rng(bench.SEED + 7).choice(len(specimens), 16, replace=False).tolist())
    sample = [specimens[k] for k in pick]
    sizes = [len(s.particles) for s in sample]
    configs = sample[0].configs(48)
    with cloth.Batch(sample) as batch:
        batch.step(configs, ITERATIONS, SWEEPS)
        states = [batch.read(k) for k in range(len(sample))]
    assert states[0]["failure"] == 0

    def reference(scene):
        own = scene.configs(48)
        return (oracle.run(scene, own, ITERATIONS, SWEEPS, precision="fp32"),
                oracle.run(scene, own, ITERATIONS, SWEEPS))

    with concurrent.futures.ThreadPoolExecutor(8) as pool:
        results = list(pool.map(reference, sample))
    worst = []
    for scene, state, (twin, fp64) in zip(sample, states, results):
        gpu = _positions(state)
        extent = _extent(scene)
        e_twin = float(np.max(np.abs(gpu - twin["positions"])))
        e_fp64 = float(np.max(np.abs(gpu - fp64["positions"])))
        e_ref = float(np.max(np.abs(twin["positions"] - fp64["positions"])))
        worst.append((e_fp64 / extent, e_twin / extent, e_ref / extent))
        assert np.array_equal(gpu, twin["positions"])
        assert e_fp64 <= 1e-5 * extent
    print(f"16 strips ({min(sizes)}-{max(sizes)} knots): worst |gpu-fp64|/extent {max(w[0] for w in worst):.3e}, "
          f"|gpu-twin|/extent {max(w[1] for w in worst):.3e}, |twin-fp64|/extent {max(w[2] for w in worst):.3e}")
    assert min(sizes) >= 8 * 64 and max(sizes) <= 64 * 256


def test_adjoint_on_64x64_strips(cloth, oracle):
    specimens, local, targets, configs = bench.adjoint_problem(cloth, specimens_count=2)
    with cloth.Batch(specimens) as batch:
        batch.adjoint(configs, ITERATIONS, SWEEPS, local, targets, 16)
        gpu = [batch.gradients(k) for k in range(2)]
        assert batch.read(0)["failure"] == 0

    def reference(args):
        k, precision = args
        return oracle.adjoint(specimens[k], configs, ITERATIONS, SWEEPS, local[k],
                              targets[:, k].astype(np.float64), precision=precision)

    jobs = [(k, p) for k in range(2) for p in ("fp32", "fp64")]
    with concurrent.futures.ThreadPoolExecutor(4) as pool:
        out = dict(zip(jobs, pool.map(reference, jobs)))
    for k in range(2):
        twin, fp64 = out[(k, "fp32")], out[(k, "fp64")]
        assert gpu[k]["loss"] == pytest.approx(twin["loss"], rel=1e-6)
        assert gpu[k]["loss"] == pytest.approx(fp64["loss"], rel=1e-4)
        # Gradients: the twin's to FP32 summation order; FP64's no farther
        # than the twin is (FP32 resolves the inverse-mass and damping
        # derivatives of 200 stiff substeps to a few per cent).
        for name in ("distanceCompliance", "bendCompliance", "inverseMass", "damping"):
            g, t, r = gpu[k][name].astype(np.float64), twin[name], fp64[name]
            e_twin = np.linalg.norm(g - t) / np.linalg.norm(t)
            e_fp64 = np.linalg.norm(g - r) / np.linalg.norm(r)
            e_ref = np.linalg.norm(t - r) / np.linalg.norm(r)
            print(f"64x64 strip {k} {name}: |gpu-twin| {e_twin:.3e} |gpu-fp64| {e_fp64:.3e} |twin-fp64| {e_ref:.3e}")
            assert e_twin < 1e-4
            assert e_fp64 <= 1.5 * e_ref + 1e-4
