"""Point ingest on the GPU from several LiDARs (k3dom_update_grid_lidars): the
node's PCL transform (src/k3dom/k3dom_node.cpp:73-83) generalised to LiDARs
with their own extrinsics.  Each LiDAR's evidence is computed as the
reference computes its one sensor's, with the sensor at that LiDAR's origin,
and the LiDARs' evidence is added in LiDAR order.

CPU tests pin the oracle's multi-LiDAR path: one LiDAR at the pose with the
identity is exactly the single-sensor update; an exactly representable
rotation gives the same frame as pre-rotated points; the culled kernel
inference equals brute force per LiDAR; splitting one LiDAR's points into two
LiDARs at the same origin changes the evidence only by rounding.  GPU tests
hold the HIP path to the oracle bit for bit."""
import numpy as np
import pytest

EYE = np.eye(3, dtype=np.float32)
ROT90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32)  # exact in float


def _frames(w, n, rig=None):
    from k3dom_amd.scene import Scene
    sc = Scene(w.scene)
    if rig is None:
        return [sc.frame(f) for f in range(n)]
    return [sc.frame_lidars(f, rig) for f in range(n)]


def _cells_equal(a, b):
    from tests._parity import compare_cells
    compare_cells(a.copyCellsToHost(), b.copyCellsToHost(), "")


def test_oracle_identity_lidar_is_the_plain_update(oracle_lib):
    from oracle.orc import OracleDOM
    from k3dom_amd.scene import small_workload
    from tests._parity import bits_equal
    w = small_workload(ego_vx=2.0, frames=4, seed=3)
    a, b = OracleDOM(w.params), OracleDOM(w.params)
    for t, pose, xs, ys, zs in _frames(w, w.frames):
        for m in (a, b):
            m.updatePose(*pose)
        a.updateGrid(t, xs, ys, zs)
        b.updateGridLidars(t, [(EYE, (0, 0, 0), xs, ys, zs)])
        _cells_equal(a, b)
        for name in ("DEL_OCC", "DEL_FREE", "PARTICLE_POS", "PARTICLE_WEIGHT"):
            assert bits_equal(a.read_array(name), b.read_array(name)), name


def test_oracle_exact_rotation(oracle_lib):
    """Points given as (y, -x, z) with a 90-degree rotation about z are the
    points (x, y, z) exactly, so the frame equals the plain update."""
    from oracle.orc import OracleDOM
    from k3dom_amd.scene import small_workload
    from tests._parity import bits_equal
    w = small_workload(frames=3, seed=5)
    a, b = OracleDOM(w.params), OracleDOM(w.params)
    for t, pose, xs, ys, zs in _frames(w, w.frames):
        for m in (a, b):
            m.updatePose(*pose)
        a.updateGrid(t, xs, ys, zs)
        b.updateGridLidars(t, [(ROT90, (0, 0, 0), ys, -xs, zs)])
        _cells_equal(a, b)
        assert bits_equal(a.read_array("DEL_FREE"), b.read_array("DEL_FREE"))


def test_oracle_culled_equals_brute_force_per_lidar(oracle_lib):
    from oracle.orc import OracleDOM
    from k3dom_amd.scene import small_workload, small_rig
    from tests._parity import bits_equal
    w = small_workload(edge=12, edge_z=6, frames=2, seed=7, beams=6, az=64)
    a, b = OracleDOM(w.params), OracleDOM(w.params, brute=True)
    for t, pose, lidars in _frames(w, w.frames, small_rig(0.4, 3)):
        for m in (a, b):
            m.updatePose(*pose)
            m.updateGridLidars(t, lidars)
        for name in ("DEL_OCC", "DEL_FREE"):
            assert bits_equal(a.read_array(name), b.read_array(name)), name


def test_oracle_split_lidar_changes_only_rounding(oracle_lib):
    """One LiDAR's points as two LiDARs at the same origin: the same terms,
    summed in two groups."""
    from oracle.orc import OracleDOM
    from k3dom_amd.scene import small_workload
    w = small_workload(frames=1, seed=9)
    (t, pose, xs, ys, zs), = _frames(w, 1)
    h = len(xs) // 2
    o = (0.3, -0.2, 0.1)
    a, b = OracleDOM(w.params), OracleDOM(w.params)
    for m in (a, b):
        m.updatePose(*pose)
    a.updateGridLidars(t, [(EYE, o, xs, ys, zs)])
    b.updateGridLidars(t, [(EYE, o, xs[:h], ys[:h], zs[:h]), (EYE, o, xs[h:], ys[h:], zs[h:])])
    for name in ("DEL_OCC", "DEL_FREE"):
        x, y = a.read_array(name), b.read_array(name)
        assert np.allclose(x, y, rtol=1e-5, atol=1e-6), name
        assert x.max() > 0


def test_oracle_lidar_origin_moves_the_evidence(oracle_lib):
    """A LiDAR at origin o sees the scene from o: its evidence equals that of a
    single sensor at o (the same voxel-centre expression with the sensor at
    pose + o) when the map centre is the same."""
    from oracle.orc import OracleDOM
    from k3dom_amd.scene import small_workload
    from tests._parity import bits_equal
    w = small_workload(frames=1, seed=9)
    (t, pose, xs, ys, zs), = _frames(w, 1)
    o = np.array([0.4, -0.2, 0.0], np.float32)
    a, b = OracleDOM(w.params), OracleDOM(w.params)
    a.updatePose(*pose)
    a.updateGridLidars(t, [(EYE, o, xs, ys, zs)])
    # b: first pose at `pose` (same map centre), then the sensor moved by o
    # within the shift threshold before its first frame
    b.updatePose(*pose)
    b.updatePose(*(np.float32(pose) + o))
    assert list(b.info().center_pos) == list(a.info().center_pos)
    b.updateGrid(t, xs, ys, zs)
    for name in ("DEL_OCC", "DEL_FREE"):
        assert bits_equal(a.read_array(name), b.read_array(name)), name


def test_dsphdmib_takes_one_lidar(oracle_lib):
    from oracle.orc import OracleDOM
    from k3dom_amd.scene import small_workload
    w = small_workload(frames=1, method="dsphdmib")
    (t, pose, xs, ys, zs), = _frames(w, 1)
    o = OracleDOM(w.params, method="dsphdmib")
    o.updatePose(*pose)
    with pytest.raises(ValueError):
        o.updateGridLidars(t, [(EYE, (0, 0, 0), xs, ys, zs), (EYE, (0, 0, 0), xs, ys, zs)])


# ------------------------------------------------------------------ GPU parity

def _run_rig(w, rig, frames, method="k3dom", on_device=False, mutate=None):
    from k3dom_amd.dom import DOM
    from oracle.orc import OracleDOM
    from tests._parity import compare_state
    gd, od = DOM(w.params, method=method), OracleDOM(w.params, method=method)
    for f, (t, pose, lidars) in enumerate(_frames(w, frames, rig)):
        if mutate is not None:
            lidars = mutate(f, lidars)
        print(f"  rig frame {f}: {[len(l[2]) for l in lidars]} points", flush=True)
        normal = gd.info().updated_time > 0
        for m in (gd, od):
            m.updatePose(*pose)
        od.updateGridLidars(t, lidars)
        if on_device:
            import torch
            dev = [tuple(torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda() for a in l[2:5])
                   for l in lidars]
            gd.updateGridLidars(t, [(l[0], l[1], d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), len(l[2]))
                                    for l, d in zip(lidars, dev)], on_device=True)
        else:
            gd.updateGridLidars(t, lidars)
        compare_state(gd, od, normal_frame=normal, what=f"rig frame {f}:")
    return gd, od


@pytest.mark.gpu
def test_lidar_rig_bit_exact(gpu):
    from k3dom_amd.scene import small_workload, small_rig
    w = small_workload(ego_vx=4.0, frames=6, seed=3)
    _run_rig(w, small_rig(0.5, 3), w.frames)


@pytest.mark.gpu
def test_lidar_rig_device_points_and_regrowth(gpu, monkeypatch):
    """Points already in device memory; tiny list capacities (each LiDAR's
    pass overflows, the buffers grow and every pass re-runs); an empty
    LiDAR and non-finite points."""
    from k3dom_amd.scene import small_workload, small_rig
    monkeypatch.setenv("K3DOM_TEST_TINY_CAPACITY", "1")
    bad = np.array([np.nan, np.inf, 0.0], np.float32)

    def mutate(f, lidars):
        if f == 2:
            r, o = lidars[1][:2]
            lidars = [lidars[0], (r, o, bad[:0], bad[:0], bad[:0])] + list(lidars[2:])
        if f == 3:
            r, o, xs, ys, zs = lidars[0]
            lidars = [(r, o, np.concatenate([xs, bad]), np.concatenate([ys, bad]), np.concatenate([zs, bad]))] + \
                list(lidars[1:])
        return lidars
    w = small_workload(ego_vx=4.0, frames=5, seed=31)
    gd, _ = _run_rig(w, small_rig(0.4, 4), w.frames, on_device=True, mutate=mutate)
    assert gd.stats()["capacity_regrowths"] >= 1


@pytest.mark.gpu
def test_single_lidar_dsphdmib_bit_exact(gpu):
    """DS-PHD/MIB with one LiDAR off the pose and turned: the measurement grid
    with the sensor at its origin."""
    from k3dom_amd.scene import small_workload
    w = small_workload(ego_vx=4.0, frames=5, seed=3, method="dsphdmib")
    _run_rig(w, [((0.3, -0.2, 0.1), 0.7)], w.frames, method="dsphdmib")


@pytest.mark.gpu
def test_identity_lidar_equals_update_grid(gpu):
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import small_workload
    from tests._parity import compare_state
    w = small_workload(ego_vx=2.0, frames=4, seed=3)
    a, b = DOM(w.params), DOM(w.params)
    for f, (t, pose, xs, ys, zs) in enumerate(_frames(w, w.frames)):
        normal = a.info().updated_time > 0
        for m in (a, b):
            m.updatePose(*pose)
        a.updateGrid(t, xs, ys, zs)
        b.updateGridLidars(t, [(EYE, (0, 0, 0), xs, ys, zs)])
        compare_state(a, b, normal_frame=normal, what=f"frame {f}:")


@pytest.mark.gpu
def test_lidar_rig_graph_replay_equals_restatement(gpu):
    """A three-LiDAR rig replayed from captured graphs (every LiDAR's lists,
    one joint gather, the per-LiDAR combines and the evidence sums in the
    graph; the LiDARs' point slots fixed across frames) equals the serial
    restatement bit for bit over frames with map shifts, also after a frame
    whose LiDARs return more points than any before (new slots, new graph)."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import small_workload, small_rig
    from oracle.orc import OracleDOM
    from tests._parity import compare_state
    w = small_workload(ego_vx=4.0, frames=7, seed=5)
    gd, od = DOM(w.params), OracleDOM(w.params)
    gd.set_graph(True)
    for f, (t, pose, lidars) in enumerate(_frames(w, w.frames, small_rig(0.5, 3))):
        if f == 4:  # a LiDAR with more points than the slot holds
            r, o, xs, ys, zs = lidars[2]
            lidars = list(lidars[:2]) + [(r, o, np.tile(xs, 3), np.tile(ys, 3), np.tile(zs, 3))]
        print(f"  rig frame {f}: {[len(l[2]) for l in lidars]} points", flush=True)
        normal = gd.info().updated_time > 0
        for m in (gd, od):
            m.updatePose(*pose)
            m.updateGridLidars(t, lidars)
        compare_state(gd, od, normal_frame=normal, what=f"graph rig frame {f}:")
    assert gd.info().center_pos[0] == od.info().center_pos[0] > 0
