"""Helpers shared by the GPU parity tests: run the HIP path and the serial
restatement on the same frames and compare every output bit for bit."""
import numpy as np

PARTICLE_ARRAYS = ["PARTICLE_POS", "PARTICLE_VEL", "PARTICLE_WEIGHT", "PARTICLE_KEY"]
FRAME_ARRAYS = ["SORT_PERM", "SORTED_KEY", "RESAMPLE_IDX", "NEWBORN_POS", "NEWBORN_VEL", "NEWBORN_WEIGHT",
                "NEWBORN_KEY", "BORN_MASS", "DEL_OCC", "DEL_FREE", "SORTED_WEIGHT"]


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return False
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def first_diff(a, b):
    a = np.ascontiguousarray(a).ravel()
    b = np.ascontiguousarray(b).ravel()
    bad = np.nonzero(a.view(np.uint8 if a.dtype.itemsize == 1 else f"u{a.dtype.itemsize}") !=
                     b.view(np.uint8 if b.dtype.itemsize == 1 else f"u{b.dtype.itemsize}"))[0]
    if len(bad) == 0:
        return "equal"
    i = bad[0]
    return f"{len(bad)} differ, first at {i}: {a[i]!r} vs {b[i]!r}"


def compare_cells(g, o, what=""):
    for name in g.dtype.names:
        assert bits_equal(g[name], o[name]), f"{what} cell field {name}: {first_diff(g[name], o[name])}"


def compare_state(gd, od, normal_frame, what=""):
    compare_cells(gd.copyCellsToHost(), od.copyCellsToHost(), what)
    for name in PARTICLE_ARRAYS:
        a, b = gd.read_array(name), od.read_array(name)
        assert bits_equal(a, b), f"{what} {name}: {first_diff(a, b)}"
    names = FRAME_ARRAYS if normal_frame else ["BORN_MASS", "DEL_OCC", "DEL_FREE"]
    for name in names:
        a, b = gd.read_array(name), od.read_array(name)
        assert bits_equal(a, b), f"{what} {name}: {first_diff(a, b)}"


def run_pair(workload, frames, check_every=1, mutate=None, method="k3dom"):
    """Drives both implementations (mass-update method `method`) through
    `frames` frames of the workload's scene and compares after each checked
    frame."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene
    from oracle.orc import OracleDOM
    sc = Scene(workload.scene)
    gd = DOM(workload.params, method=method)
    od = OracleDOM(workload.params, method=method)
    for f in range(frames):
        t, pose, xs, ys, zs = sc.frame(f)
        if mutate is not None:
            t, pose, xs, ys, zs = mutate(f, t, pose, xs, ys, zs)
        print(f"  frame {f}: {len(xs)} points", flush=True)
        normal = gd.info().updated_time > 0
        gd.updatePose(*pose)
        od.updatePose(*pose)
        gd.updateGrid(t, xs, ys, zs)
        od.updateGrid(t, xs, ys, zs)
        if f % check_every == 0 or f == frames - 1:
            compare_state(gd, od, normal_frame=normal, what=f"{workload.name} frame {f}:")
    return gd, od
