"""The headline configuration on the GPU: 256x256x32 voxels at 0.1 m,
2,000,000 particles + 200,000 newborn per frame, a 64x2048 LiDAR (40 m) and
an ego at 10 m/s, so that the map shifts by 10 voxels every frame.

- The initialisation frame and three normal frames, each after a map shift,
  are compared bit for bit with the serial restatement on every array the
  smaller parity tests compare (every GridCell field, particles, sort
  permutation and keys, newborn set, born masses, evidence, resampled
  indices).  The restatement's kernel inference runs on host threads; it
  gives the same bits for any thread count.
- The same with enclosing walls, so that every ray of the fan returns
  (131,072 points per frame): the initialisation frame and two normal frames.
- Twelve further frames are checked against the size-independent properties
  of tests/_invariants.py.

Reference path: src/k3dom/dom.cu:29-50 (updateGrid), src/common/dom.cu:111-297
(updatePose and the particle stages)."""
import os

import pytest

from tests._invariants import check_frame_invariants
from tests._parity import run_pair

pytestmark = pytest.mark.gpu


def _host_threads():
    return max(1, min(16, os.cpu_count() or 1))


@pytest.mark.timeout(900)
def test_config2_bit_exact_with_map_shifts(gpu):
    from k3dom_amd.scene import workload
    from oracle import orc
    orc.set_threads(_host_threads())
    try:
        w = workload(2)
        gd, od = run_pair(w, 4)
    finally:
        orc.set_threads(1)
    # ego at 10 m/s, 10 Hz, shift threshold 1 m: the map moved every frame
    assert gd.info().center_pos[0] == od.info().center_pos[0] == pytest.approx(3.0)
    assert gd.stats()["points"] > 60000


@pytest.mark.timeout(900)
def test_config2_every_ray_returns_bit_exact(gpu):
    """The headline configuration with enclosing walls, so that all 131,072
    rays of the fan return: the initialisation frame and two normal frames
    (each after a map shift) equal the serial restatement bit for bit."""
    from k3dom_amd.scene import with_walls, workload
    from oracle import orc
    orc.set_threads(_host_threads())
    try:
        gd, od = run_pair(with_walls(workload(2)), 3)
    finally:
        orc.set_threads(1)
    assert gd.stats()["points"] == 131072


@pytest.mark.timeout(600)
def test_config2_invariants_over_frames(gpu):
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, workload
    w = workload(2)
    sc = Scene(w.scene)
    gd = DOM(w.params)
    for f in range(14):
        t, pose, xs, ys, zs = sc.frame(f)
        normal = gd.info().updated_time > 0
        gd.updatePose(*pose)
        gd.updateGrid(t, xs, ys, zs)
        if f == 0 or f >= 2:
            print(f"  frame {f}: {len(xs)} points", flush=True)
            check_frame_invariants(gd, normal, what=f"cfg2 frame {f}:")


@pytest.mark.timeout(1200)
def test_config5_four_lidars_bit_exact(gpu):
    """The stress configuration: 512x512x64 voxels at 0.1 m, 8,000,000
    particles + 800,000 newborn, four 128x2048 LiDARs at the vehicle's
    corners 90 degrees apart (about 415k points per frame), ego at 15 m/s.
    The initialisation frame and one normal frame (after a map shift) equal
    the serial restatement's multi-LiDAR frames bit for bit."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, workload
    from oracle import orc
    from oracle.orc import OracleDOM
    from tests._parity import compare_state
    orc.set_threads(_host_threads())
    try:
        w = workload(5)
        sc = Scene(w.scene)
        gd, od = DOM(w.params), OracleDOM(w.params)
        for f in range(2):
            t, pose, lidars = sc.frame_lidars(f, w.extra["rig"])
            print(f"  frame {f}: {sum(len(l[2]) for l in lidars)} points", flush=True)
            normal = gd.info().updated_time > 0
            for m in (gd, od):
                m.updatePose(*pose)
                m.updateGridLidars(t, lidars)
            compare_state(gd, od, normal_frame=normal, what=f"cfg5 frame {f}:")
        assert gd.info().center_pos[0] == od.info().center_pos[0] > 0
    finally:
        orc.set_threads(1)
