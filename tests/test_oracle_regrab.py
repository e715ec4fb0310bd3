"""Pin of the oracle's grip attachment with nearest-cuff-patch re-grab (CPU).

The reference's patch re-grab check (cmake/CheckGripTrajectory.cmake:26-37,
108-130; metrics tools/cloth_bag_metal.mm:6334-6393) replays
calibration/trajectories/synthetic-patch-regrab.csv (held here as
tests/golden/trajectories/) through the bag for one 48-substep frame and
requires: two re-grabs (attachment generations 1 -> 2 -> 3), each choosing
the cuff patch nearest the handle, the patch centre moving by 12 rings
(0 -> 12 -> 24, maximum_patch_ring_shift=12, selected_patch_center_ring=24),
every grip of the final patch on the same centre ring, grip k on cuff row
k / 5 at ring (centre + slot - 2) mod 48 (patch_topology_exact), capture
distances within the grip's capture radius and capture errors <= 2e-6 m.
The kernel is update_grip_attachment, src/metal/ClothBag.metal:577-673.

The distant-patch fixture moves the handle beyond the capture radius before
its re-grab; the reference rejects it (reattachment exceeds capture radius,
CMakeLists.txt:648-656), here the GRIP_CAPTURE failure bit.
"""

import numpy as np
import pytest

GRIP_CAPTURE = 16
AROUND = 48
PATCH_WIDTH = 5


def _replay(cloth, golden, name):
    c = cloth.cloth
    trajectory = cloth.GripTrajectory(golden / "trajectories" / name)
    scene = cloth.Scene.bag_replay(c.REPLAY_RECORDED, trajectory)
    configs = c.replay_frame_configs(scene.config, c.REPLAY_RECORDED, 0, trajectory)
    return scene, configs


def _check_patch(scene, records, centre, generation):
    particle, gen, capture_bits, centres = records.T
    assert np.all(gen == generation)
    assert np.all(centres == centre)
    for k in range(len(records)):
        row, slot = divmod(k, PATCH_WIDTH)
        ring = (centre + AROUND + slot - 2) % AROUND
        assert particle[k] == scene.cuffParticles[row * AROUND + ring]
    capture = capture_bits.astype(np.uint32).view(np.float32)
    assert np.all(capture <= scene.config["gripMaterial"][0, 0] + 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_patch_regrab_selects_rings_12_then_24(cloth, oracle, golden, precision):
    scene, configs = _replay(cloth, golden, "synthetic-patch-regrab.csv")
    generations = configs["gripControl"][:, 0]
    assert list(np.unique(generations)) == [1, 2, 3]
    first_regrab = int(np.argmax(generations == 2)) + 1      # substeps run when generation 2 is captured
    assert len(scene.grips) == 2 * PATCH_WIDTH
    mid = oracle.run(scene, configs[:first_regrab], 32, 3, precision=precision)
    assert mid["failure"] == 0
    _check_patch(scene, mid["gripRecords"], 12, 2)
    end = oracle.run(scene, configs, 32, 3, precision=precision)
    assert end["failure"] == 0
    _check_patch(scene, end["gripRecords"], 24, 3)
    assert np.max(end["gripCaptureError"]) <= 2e-6
    # regrab_count = generations - 1 = 2; ring shifts 12 and 12.
    print(f"{precision}: centre 12 after {first_regrab} substeps, 24 after {len(configs)}; "
          f"capture error {np.max(end['gripCaptureError']):.3e} m")


def test_distant_patch_regrab_is_refused(cloth, oracle, golden):
    scene, configs = _replay(cloth, golden, "invalid-distant-patch-regrab.csv")
    out = oracle.run(scene, configs, 32, 3)
    assert out["failure"] & GRIP_CAPTURE
