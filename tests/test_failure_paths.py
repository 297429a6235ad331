"""Host-driven failure and edge paths of a frame (no fault is provoked on the
GPU): a repeated time stamp and a time going backwards on normal frames,
every particle leaving the grid (a zero joint weight, src/common/dom.cu:
265-289), a map shift larger than the grid (src/common/dom.cu:125-156), and
device allocations that fail (through the K3DOM_TEST_FAIL_ALLOC hook: the
n-th allocation after creation reports out-of-memory).  Each is compared
bit for bit with the serial restatement."""
import numpy as np
import pytest

from tests._parity import compare_state, run_pair

pytestmark = pytest.mark.gpu

K3DOM_ERR_OUT_OF_MEMORY = 3  # include/k3dom.h


def test_repeated_and_backward_time_stamps(gpu):
    """dt = 0 (gamma^0 = 1, no motion) and dt < 0 (gamma^dt > 1, motion
    backwards), as the reference computes them (src/k3dom/dom.cu:33-36)."""
    from k3dom_amd.scene import small_workload
    times = {}

    def mutate(f, t, pose, xs, ys, zs):
        times[f] = t
        if f == 2:
            t = times[1]              # the previous frame's time again
        elif f == 3:
            t = times[1] - 0.05       # before the previous frame
        return t, pose, xs, ys, zs
    w = small_workload(frames=6, seed=17, ego_vx=2.0)
    run_pair(w, w.frames, mutate=mutate)


@pytest.mark.parametrize("newborn", [0, 409])
def test_every_particle_leaves_the_grid(gpu, newborn):
    """A frame 1000 s after the previous one: every particle (speed at least
    the minimum) moves out of the grid and weighs 0; without newborns the
    joint weight is 0 and resampling keeps the set in order."""
    from k3dom_amd.scene import small_workload

    def mutate(f, t, pose, xs, ys, zs):
        return (t + 1000.0 if f >= 2 else t), pose, xs, ys, zs
    w = small_workload(frames=4, seed=19, newborn=newborn)
    gd, od = run_pair(w, w.frames, mutate=mutate)
    if newborn == 0:
        assert np.all(gd.read_array("SORTED_WEIGHT") == 0)
        assert np.array_equal(gd.read_array("RESAMPLE_IDX"), np.arange(gd.particle_count, dtype=np.int32))


def test_map_shift_larger_than_the_grid(gpu):
    """The ego jumps by more than the map's size: every voxel takes the
    priors, the particles move out of the grid."""
    from k3dom_amd.scene import small_workload
    w = small_workload(frames=6, seed=21, ego_vx=2.0)
    jump = 2.0 * w.params.size

    def mutate(f, t, pose, xs, ys, zs):
        return t, ((pose[0] + jump) if f >= 3 else pose[0], pose[1], pose[2]), xs, ys, zs
    gd, od = run_pair(w, w.frames, mutate=mutate)
    assert gd.info().center_pos[0] >= jump


@pytest.mark.parametrize("fail_at", list(range(1, 31)))
def test_failed_allocation_then_the_frame_again(gpu, monkeypatch, fail_at):
    """The fail_at-th device allocation after creation fails (tiny starting
    capacities, so frames regrow their lists and allocate mid-frame): the
    frame reports out-of-memory and leaves the state as it was; the same
    frame given again completes, and every frame equals the serial
    restatement's, which never failed."""
    from k3dom_amd._lib import K3domError
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    from oracle.orc import OracleDOM
    monkeypatch.setenv("K3DOM_TEST_TINY_CAPACITY", "1")
    monkeypatch.setenv("K3DOM_TEST_FAIL_ALLOC", str(fail_at))
    w = small_workload(frames=5, seed=23, ego_vx=4.0)
    sc = Scene(w.scene)
    gd, od = DOM(w.params), OracleDOM(w.params)
    failed = 0
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        normal = od.info().updated_time > 0
        for m in (gd, od):
            m.updatePose(*pose)
        od.updateGrid(t, xs, ys, zs)
        try:
            gd.updateGrid(t, xs, ys, zs)
        except K3domError as e:
            assert e.status == K3DOM_ERR_OUT_OF_MEMORY, str(e)
            failed += 1
            info = gd.info()
            assert info.iteration == f and (info.updated_time > 0) == normal
            gd.updateGrid(t, xs, ys, zs)
        compare_state(gd, od, normal_frame=normal, what=f"fail_at {fail_at} frame {f}:")
    assert failed <= 1


@pytest.mark.parametrize("fail_at", [1, 2, 3, 4])
def test_failed_export_allocation_then_export_again(gpu, monkeypatch, fail_at):
    """The MapPoint export's buffers are allocated at its first call; when
    one of them fails the call reports out-of-memory, and the next call
    allocates what is missing and exports the right records."""
    from k3dom_amd._lib import K3domError
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import small_workload
    from tests.test_mappoints import expected_records
    monkeypatch.setenv("K3DOM_TEST_FAIL_ALLOC", str(fail_at))
    w = small_workload(frames=1, seed=5)
    d = DOM(w.params)  # no frame yet: the export's allocations are the first
    with pytest.raises(K3domError) as ei:
        d.export_mappoints(visible=True, min_occ=-1.0, min_mass=-1.0)
    assert ei.value.status == K3DOM_ERR_OUT_OF_MEMORY
    got = d.export_mappoints(visible=True, min_occ=-1.0, min_mass=-1.0)
    rec, visible = expected_records(d, "k3dom")
    want = rec[visible(-1.0, -1.0)]
    assert len(got) == len(want) == d.grid_cell_count
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
