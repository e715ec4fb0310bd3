"""Hash self-contact on the GPU (needs a GPU).

Integer outputs (sorted cell entries, candidate pairs) must equal, bit for
bit, the same quantities computed on the CPU from the same FP32 snapshot:
the cell entries by numpy below, the candidate set by the oracle's FP32
twin (its own CPU hash) and, for a small sheet, by brute force over all
pairs. Positions are compared with the FP32 twin and the FP64 oracle.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def folded_sheet(cloth, width=24, height=32, gap=0.003, spacing=0.01, gravity=0.0, margin=None):
    """A sheet folded in half along y with its layers `gap` apart (less than
    two yarn radii), so the layers start in self-contact."""
    scene = cloth.Scene.sheet(width=width, height=height, spacing=spacing, elevation=0.1, jitter=0.0,
                              ground=0, gravityZ=gravity, selfContact=1, seed=3)
    x = scene.particles["positionAndInverseMass"]
    fold = float(x[(height // 2) * width, 1])
    upper = x[:, 1] > fold + 1e-6
    x[upper, 1] = 2.0 * fold - x[upper, 1]
    x[upper, 2] = x[upper, 2] + gap
    scene.particles["previousAndMass"][:, :3] = x[:, :3]
    if margin is not None:
        scene.selfContactMargin = margin
    scene.refresh()
    return scene


def merge_scenes(cloth, a, b):
    """One scene holding both sheets; colour batches of equal colour are
    concatenated (the two particle sets are disjoint, so the union of two
    valid colourings is valid)."""
    offset = len(a.particles)
    arrays = {"particles": np.concatenate([a.particles, b.particles])}
    for name, field, index_lanes in [("distances", "particlesAndColor", 2), ("bends", "particlesAndColor", 3)]:
        colour_lane = 2 if name == "distances" else 3
        ra, rb = getattr(a, name).copy(), getattr(b, name).copy()
        rb[field][:, :index_lanes] += offset
        colours = max(ra[field][:, colour_lane].max(), rb[field][:, colour_lane].max()) + 1
        records, batches, first = [], [], 0
        for colour in range(colours):
            chunk = np.concatenate([ra[ra[field][:, colour_lane] == colour], rb[rb[field][:, colour_lane] == colour]])
            records.append(chunk)
            batch = np.zeros(1, cloth.BATCH)
            batch["control"] = (first, len(chunk), colour, 0)
            batches.append(batch)
            first += len(chunk)
        arrays[name] = np.concatenate(records)
        arrays[name[:-1] + "Batches"] = np.concatenate(batches)
    config = a.config.copy()
    config["control"][0, 1:3] = (len(arrays["particles"]), len(arrays["distances"]))
    config["constraintCounts"][0, 1] = len(arrays["bends"])
    return cloth.Scene(config, staticSpheres=a.staticSpheres, selfContact=a.selfContact,
                       selfContactMargin=a.selfContactMargin, **arrays)


def two_layer_impact(cloth):
    """A sheet resting on the ground and a second sheet 12 mm above it thrown
    down at 1 m/s: swept and present yarn/yarn contacts."""
    lower = cloth.Scene.sheet(width=30, height=24, elevation=0.004, ground=1, selfContact=1, seed=5)
    upper = cloth.Scene.sheet(width=26, height=20, elevation=0.016, ground=1, selfContact=1, seed=6)
    upper.particles["velocity"][:, 2] = -1.0
    upper.particles["positionAndInverseMass"][:, 0] += 0.003
    upper.particles["previousAndMass"][:, 0] += 0.003
    return merge_scenes(cloth, lower, upper)


def numpy_entries(scene, state):
    """(cell id, segment) entries of the broad phase, in FP32 as the device
    computes them, stably sorted by cell id."""
    f32 = np.float32
    pos = state["particles"]["positionAndInverseMass"][:, :3]
    prev = state["particles"]["previousAndMass"][:, :3]
    ij = scene.distances["particlesAndColor"][:, :2]
    pts = np.stack([pos[ij[:, 0]], pos[ij[:, 1]], prev[ij[:, 0]], prev[ij[:, 1]]])
    expand = f32(scene.config["clothMaterial"][0, 0]) + f32(scene.selfContactMargin)
    lo = pts.min(axis=0) - expand
    hi = pts.max(axis=0) + expand
    origin = lo.min(axis=0)
    cell = f32(scene.config["clothMaterial"][0, 1])
    c0 = np.floor((lo - origin) / cell).astype(np.int64)
    c1 = np.floor((hi - origin) / cell).astype(np.int64)
    n = np.floor((hi.max(axis=0) - origin) / cell).astype(np.int64) + 1
    keys, segs = [], []
    for s in range(len(ij)):
        for z in range(c0[s, 2], c1[s, 2] + 1):
            for y in range(c0[s, 1], c1[s, 1] + 1):
                for x in range(c0[s, 0], c1[s, 0] + 1):
                    keys.append((z * n[1] + y) * n[0] + x)
                    segs.append(s)
    keys = np.array(keys, np.int64)
    segs = np.array(segs, np.int64)
    order = np.argsort(keys, kind="stable")
    return keys[order].astype(np.uint32), segs[order].astype(np.uint32), origin, n


def brute_force_pairs(scene, state):
    f32 = np.float32
    pos = state["particles"]["positionAndInverseMass"][:, :3]
    prev = state["particles"]["previousAndMass"][:, :3]
    ij = scene.distances["particlesAndColor"][:, :2].astype(np.int64)
    pts = np.stack([pos[ij[:, 0]], pos[ij[:, 1]], prev[ij[:, 0]], prev[ij[:, 1]]])
    expand = f32(scene.config["clothMaterial"][0, 0]) + f32(scene.selfContactMargin)
    lo = pts.min(axis=0) - expand
    hi = pts.max(axis=0) + expand
    P = len(pos)
    adjacency = [set() for _ in range(P)]
    for a, b in ij:
        adjacency[a].add(b)
        adjacency[b].add(a)
    near = []
    for i in range(P):
        hood = {i} | adjacency[i]
        for a in adjacency[i]:
            hood |= adjacency[a]
        near.append(hood)
    overlap = ((lo[:, None, :] <= hi[None, :, :]) & (hi[:, None, :] >= lo[None, :, :])).all(axis=2)
    s_idx, t_idx = np.nonzero(np.triu(overlap, k=1))
    keep = [not (ij[t, 0] in near[ij[s, 0]] or ij[t, 1] in near[ij[s, 0]] or ij[t, 0] in near[ij[s, 1]]
                 or ij[t, 1] in near[ij[s, 1]]) for s, t in zip(s_idx, t_idx)]
    return np.stack([s_idx[keep], t_idx[keep]], axis=1).astype(np.uint32)


def _sorted(pairs):
    pairs = np.asarray(pairs, np.uint32).reshape(-1, 2)
    return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]


def use_sort(monkeypatch, sort):
    """The solver sorts cell entries by counting on grids that fit its
    per-cell counters (every grid here) and by radix otherwise;
    NUMI_CLOTH_HIP_SELF_SORT=radix, read when a solver is created, forces
    the radix sort."""
    if sort == "radix":
        monkeypatch.setenv("NUMI_CLOTH_HIP_SELF_SORT", "radix")
    else:
        monkeypatch.delenv("NUMI_CLOTH_HIP_SELF_SORT", raising=False)


@pytest.mark.parametrize("sort", ["counting", "radix"])
@pytest.mark.parametrize("make", [folded_sheet, two_layer_impact])
def test_cell_entries_and_candidates_are_bit_exact(cloth, oracle, make, sort, monkeypatch):
    use_sort(monkeypatch, sort)
    scene = make(cloth)
    # The folded layers separate within a few substeps; probe them early.
    substeps = 1 if make is folded_sheet else 6
    with cloth.Solver(scene) as solver:
        solver.step(scene.configs(substeps), 2, 1)
        state = solver.read()
        broad = solver.self_broad_phase()
    assert broad["valid"] == 1
    keys, segs, origin, n = numpy_entries(scene, state)
    assert np.array_equal(broad["origin"], origin.astype(np.float32))
    assert tuple(broad["cells"]) == tuple(int(v) for v in n)
    assert np.array_equal(broad["cellIds"], keys)
    assert np.array_equal(broad["segments"], segs)
    twin = oracle.self_candidates_fp32(scene, state["particles"], scene.config)
    gpu = _sorted(broad["pairs"])
    print(f"{make.__name__}: {len(keys)} entries, {len(gpu)} candidate pairs")
    assert len(gpu) > 0
    assert np.array_equal(gpu, _sorted(twin))
    assert np.all(broad["pairs"][:, 0] < broad["pairs"][:, 1])


@pytest.mark.parametrize("sort", ["counting", "radix"])
def test_candidates_equal_brute_force(cloth, sort, monkeypatch):
    use_sort(monkeypatch, sort)
    scene = folded_sheet(cloth, width=14, height=20)
    with cloth.Solver(scene) as solver:
        solver.step(scene.configs(3), 4, 0)
        state = solver.read()
        broad = solver.self_broad_phase()
    assert np.array_equal(_sorted(broad["pairs"]), _sorted(brute_force_pairs(scene, state)))


def test_counting_and_radix_sorts_step_identically(cloth, monkeypatch):
    """Both sorts give the same entry order, so whole substeps (broad phase,
    narrow phase, friction) give the same bytes."""
    scene = two_layer_impact(cloth)
    states = []
    for sort in ["counting", "radix"]:
        use_sort(monkeypatch, sort)
        with cloth.Solver(scene) as solver:
            solver.step(scene.configs(8), 8, 2)
            states.append(solver.read())
            assert solver.diagnostics().selfEntryCount > 0
    assert states[0]["particles"].tobytes() == states[1]["particles"].tobytes()
    assert states[0]["selfStatus"].tobytes() == states[1]["selfStatus"].tobytes()
    assert states[0]["frictionStatus"].tobytes() == states[1]["frictionStatus"].tobytes()


def test_cell_size_changes_no_byte_of_the_step(cloth):
    """The grid only accelerates the candidate search: cells of 1, 2 and 3
    segment lengths give the same candidate set and the same substep bytes
    (corrections and friction are summed per knot in fixed point, so the
    candidates' order does not matter), with fewer entries in larger cells."""
    scene = two_layer_impact(cloth)
    states, pairs, entries = [], [], []
    for cell in [0.01, 0.02, 0.03]:
        scene.config["clothMaterial"][0, 1] = cell
        scene.refresh()
        with cloth.Solver(scene) as solver:
            solver.step(scene.configs(8), 8, 2)
            states.append(solver.read())
            pairs.append(_sorted(solver.self_broad_phase()["pairs"]))
            entries.append(solver.diagnostics().selfEntryCount)
    print(f"entries per cell size: {entries}, candidates {len(pairs[0])}")
    assert len(pairs[0]) > 0 and entries[0] > entries[1] > entries[2]
    for k in (1, 2):
        assert np.array_equal(pairs[0], pairs[k])
        assert states[0]["failure"] == states[k]["failure"] == 0
        assert states[0]["particles"].tobytes() == states[k]["particles"].tobytes()
        assert states[0]["selfStatus"].tobytes() == states[k]["selfStatus"].tobytes()
        assert states[0]["frictionStatus"].tobytes() == states[k]["frictionStatus"].tobytes()


def test_folded_layers_are_pushed_apart(cloth):
    scene = folded_sheet(cloth)
    radius = float(scene.config["clothMaterial"][0, 0])
    with cloth.Solver(scene) as solver:
        solver.step(scene.configs(32), 32, 3)
        state = solver.read()
        diag = solver.diagnostics()
    counters = state["selfStatus"]["counters"][0]
    assert counters[0] > 0, "no present self-contacts were accepted"
    assert state["failure"] == 0
    x = state["particles"]["positionAndInverseMass"][:, :3]
    z = x[:, 2]
    # The two layers end at least ~two yarn radii apart.
    separation = z.max() - z.min()
    print(f"layer separation {separation:.5f} m, 2r = {2 * radius:.5f} m, "
          f"max self penetration {diag.maximumSelfPenetration:.2e} m")
    assert separation > 2 * radius * 0.9
    assert diag.maximumSelfPenetration < 0.05 * radius


@pytest.mark.parametrize("make", [folded_sheet, two_layer_impact])
def test_self_contact_frame_parity(cloth, oracle, make):
    scene = make(cloth)
    configs = scene.configs(16)
    with cloth.Solver(scene) as solver:
        solver.step(configs, 16, 3)
        state = solver.read()
    gpu = state["particles"]["positionAndInverseMass"][:, :3].astype(np.float64)
    twin = oracle.run(scene, configs, 16, 3, precision="fp32")
    fp64 = oracle.run(scene, configs, 16, 3)
    x = scene.particles["positionAndInverseMass"][:, :3]
    extent = float(np.max(x.max(axis=0) - x.min(axis=0)))
    e_twin = np.max(np.abs(gpu - twin["positions"]))
    e_fp64 = np.max(np.abs(gpu - fp64["positions"]))
    print(f"{make.__name__}: |gpu - fp32 twin| = {e_twin:.3e}, |gpu - fp64| = {e_fp64:.3e}, "
          f"gate {1e-5 * extent:.3e}; contacts gpu {state['selfStatus']['counters'][0][:2]} "
          f"twin {twin['presentContacts']}/{twin['sweptContacts']}")
    e_twin_fp64 = np.max(np.abs(twin["positions"] - fp64["positions"]))
    # Contact activation (distance < 2r) is a threshold: FP32 and FP64 runs
    # of the same schedule part where a pair sits on it. The GPU must be as
    # close to FP64 as the exact FP32 restatement is, and within the gate.
    assert e_twin <= 1e-6 * extent
    assert e_fp64 <= max(1e-5 * extent, 1.5 * e_twin_fp64)


def test_self_contact_runs_are_bit_identical(cloth):
    scene = two_layer_impact(cloth)
    results = []
    for _ in range(2):
        with cloth.Solver(scene) as solver:
            solver.step(scene.configs(12), 12, 3)
            results.append(solver.read())
    assert results[0]["particles"].tobytes() == results[1]["particles"].tobytes()
    assert results[0]["selfStatus"].tobytes() == results[1]["selfStatus"].tobytes()


@pytest.mark.parametrize("mu", [0.01, 0.34, 100.0])
def test_self_friction_matches_fp32_twin(cloth, oracle, mu):
    """Impulse aggregation and self friction (crossing segments, one slides
    over the other): positions and velocities as the FP32 twin computes
    them, the same friction contact count and cone ratio."""
    from probes import crossing_segments_probe
    scene = crossing_segments_probe(cloth, mu)
    configs = scene.configs(1)
    with cloth.Solver(scene) as solver:
        solver.step(configs, 4, 0)
        state = solver.read()
    twin = oracle.run(scene, configs, 4, 0, precision="fp32")
    p = state["particles"]
    dx = np.max(np.abs(p["positionAndInverseMass"][:, :3] - twin["positions"]))
    dv = np.max(np.abs(p["velocity"][:, :3] - twin["velocities"]))
    metrics = state["frictionStatus"]["metrics"][0]
    ratio = float(np.array([metrics[0]], np.uint32).view(np.float32)[0])
    print(f"mu {mu}: |dx| {dx:.2e}, |dv| {dv:.2e}, friction contacts gpu {metrics[3]} twin "
          f"{twin['selfFrictionContacts']}, cone ratio gpu {ratio:.7f} twin {twin['maximumConeRatio']:.7f}")
    assert state["failure"] == 0
    assert metrics[3] == twin["selfFrictionContacts"] == 1
    assert dx <= 1e-7
    assert dv <= 1e-6
    assert abs(ratio - twin["maximumConeRatio"]) <= 1e-6
