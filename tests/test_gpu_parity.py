"""Bit-exact parity of the HIP update against the serial restatement.

Every GridCell field, every particle array, the sort permutation, the sorted
keys, the newborn set, the born masses, the kernel-inference evidence and the
resampled indices must be identical bit for bit after each frame."""
import numpy as np
import pytest

from tests._parity import run_pair, compare_state

pytestmark = pytest.mark.gpu


def test_small_scene_bit_exact(gpu):
    from k3dom_amd.scene import small_workload
    w = small_workload()
    run_pair(w, w.frames)


def test_small_scene_moving_ego_map_shift(gpu):
    from k3dom_amd.scene import small_workload
    # ego at 4 m/s, 0.2 m voxels, map shift threshold 5 voxels -> shifts every few frames
    w = small_workload(ego_vx=4.0, frames=8, seed=3)
    gd, od = run_pair(w, w.frames)
    assert gd.info().center_pos[0] == od.info().center_pos[0]
    assert gd.info().center_pos[0] > 0.0


def test_cells_read_between_pose_and_grid_update(gpu):
    """A map shift moves the carried fields at once and the rest only when
    needed: cells read right after a shifting updatePose, and a second shift
    before the next updateGrid, must still equal the serial restatement."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    from oracle.orc import OracleDOM
    from tests._parity import bits_equal, compare_cells, compare_state, first_diff
    w = small_workload(ego_vx=4.0, frames=10, seed=3)
    sc = Scene(w.scene)
    gd, od = DOM(w.params), OracleDOM(w.params)
    shifts = 0
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        c0 = gd.info().center_pos[0]
        for m in (gd, od):
            m.updatePose(*pose)
        if gd.info().center_pos[0] != c0:
            shifts += 1
            compare_cells(gd.copyCellsToHost(), od.copyCellsToHost(), f"after the shift of frame {f}:")
            if shifts % 2:
                a, b = gd.read_array("PARTICLE_POS"), od.read_array("PARTICLE_POS")
                assert bits_equal(a, b), f"particles after the shift of frame {f}: {first_diff(a, b)}"
            if shifts == 2:
                # a second shift with the first still pending
                far = (pose[0] + 2.0, pose[1], pose[2])
                for m in (gd, od):
                    m.updatePose(*far)
                compare_cells(gd.copyCellsToHost(), od.copyCellsToHost(), f"after a double shift in frame {f}:")
                for m in (gd, od):
                    m.updatePose(*pose)
        normal = gd.info().updated_time > 0
        for m in (gd, od):
            m.updateGrid(t, xs, ys, zs)
        compare_state(gd, od, normal_frame=normal, what=f"frame {f}:")
    assert shifts >= 2


def test_config1_bit_exact(gpu):
    from k3dom_amd.scene import workload
    w = workload(1)
    run_pair(w, 5)


def test_capacity_regrowth_is_bit_exact(gpu, monkeypatch):
    """Every candidate-list capacity starts tiny, so frames overflow, grow the
    buffers and re-run the kernel-inference stage and the frame tail; the
    results must not change."""
    from k3dom_amd.scene import small_workload
    monkeypatch.setenv("K3DOM_TEST_TINY_CAPACITY", "1")
    w = small_workload(frames=4, seed=5)
    gd, od = run_pair(w, w.frames)
    assert gd.stats()["capacity_regrowths"] >= 3


def test_empty_and_invalid_point_clouds(gpu):
    """Frames with no points, and points with NaN / inf / zero coordinates."""
    from k3dom_amd.scene import small_workload

    def mutate(f, t, pose, xs, ys, zs):
        if f == 2:
            return t, pose, xs[:0], ys[:0], zs[:0]
        if f == 3:
            bad = np.array([np.nan, np.inf, 0.0, -np.inf, 1e30], np.float32)
            return (t, pose, np.concatenate([xs, bad]), np.concatenate([ys, bad[::-1]]),
                    np.concatenate([zs, np.zeros(5, np.float32)]))
        return t, pose, xs, ys, zs
    w = small_workload(frames=5, seed=9)
    run_pair(w, w.frames, mutate=mutate)


def test_zero_time_first_frame_reinitialises(gpu):
    """A first frame at t = 0 leaves the stored time at 0, so the next frame
    initialises again (src/k3dom/dom.cu:31); both paths must agree."""
    from k3dom_amd.scene import small_workload

    def mutate(f, t, pose, xs, ys, zs):
        return t - 1.0, pose, xs, ys, zs
    w = small_workload(frames=4, seed=13)
    run_pair(w, w.frames, mutate=mutate)


def test_crowded_voxels_bit_exact(gpu):
    """60,000 particles on 2,048 voxels: many voxels hold more than 64
    particles (the wavefront-per-voxel paths of the segment sums)."""
    from k3dom_amd.scene import small_workload
    w = small_workload(particles=60000, newborn=6000, frames=4, seed=21)
    gd, od = run_pair(w, w.frames)
    cells = gd.copyCellsToHost()
    occ = cells["start_idx"] >= 0
    counts = cells["end_idx"][occ] - cells["start_idx"][occ] + 1
    assert counts.max() > 64


def test_partial_single_sort_tile_bit_exact(gpu):
    """1,000 particles: fewer than one 2,048-key sort tile, not a multiple of
    the 64-lane wavefront, so every sort pass runs one partly filled tile."""
    from k3dom_amd.scene import small_workload
    w = small_workload(particles=1000, newborn=100, frames=3, seed=29)
    run_pair(w, w.frames)


def test_very_crowded_voxels_bit_exact(gpu):
    """200,000 particles on an 8x8x4 grid: voxels with more than 64 blocks of
    64 particles (several rounds of the wavefront's block loop)."""
    from k3dom_amd.scene import small_workload
    w = small_workload(edge=8, edge_z=4, particles=200000, newborn=20000, frames=2, seed=23, beams=8, az=64)
    gd, od = run_pair(w, w.frames)
    cells = gd.copyCellsToHost()
    occ = cells["start_idx"] >= 0
    counts = cells["end_idx"][occ] - cells["start_idx"][occ] + 1
    assert counts.max() > 64 * 64


def test_graph_replay_is_bit_exact(gpu, monkeypatch):
    """Frames replayed as a captured graph (re-captured after map shifts and
    capacity regrowth) equal the serial restatement bit for bit."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    from oracle.orc import OracleDOM
    monkeypatch.setenv("K3DOM_TEST_TINY_CAPACITY", "1")
    w = small_workload(ego_vx=4.0, frames=8, seed=31)
    sc = Scene(w.scene)
    gd, od = DOM(w.params), OracleDOM(w.params)
    gd.set_graph(True)
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        print(f"  graph frame {f}: {len(xs)} points", flush=True)
        normal = gd.info().updated_time > 0
        for m in (gd, od):
            m.updatePose(*pose)
            m.updateGrid(t, xs, ys, zs)
        compare_state(gd, od, normal_frame=normal, what=f"graph frame {f}:")
    assert gd.stats()["capacity_regrowths"] >= 1
