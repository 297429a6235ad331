"""The /map_data export: MapPoint records (include/common/mappoint.h:5-28)
built on the GPU as the node's publish_mapdata builds them on the host
(src/k3dom/k3dom_node.cpp:161-184, src/dsphdmib/dsphdmib_node.cpp:153-177),
and the compaction of the voxels passing publish_map's visualisation filter
(src/k3dom/k3dom_node.cpp:137-138, src/dsphdmib/dsphdmib_node.cpp:134-135).

The expected records are computed on the host from the cells read back with
k3dom_read_cells, with the GridCell accessors' float expressions, so the GPU
kernel is checked against an independent path through the same data."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32
K3DOM_ERR_INVALID_ARG = 1  # include/k3dom.h


def test_mappoint_layout_matches_header(tmp_path):
    from k3dom_amd.dom import MAPPOINT_DTYPE
    src = tmp_path / "mp.cpp"
    fields = ["x", "y", "z", "w", "v_x", "v_y", "v_z", "occ_val", "dyn_val", "eval_aug", "dyn_aug", "pad"]
    cxx = ["#include <cstdio>", "#include <cstddef>", '#include "common/mappoint.h"', "int main(){",
           'printf("size %zu %zu %zu\\n", sizeof(k3dom_mappoint), sizeof(MapPoint), alignof(MapPoint));']
    cxx += [f'printf("{f} %zu\\n", offsetof(k3dom_mappoint, {f}));' for f in fields]
    cxx += ['printf("occ_val_cpp %zu\\n", offsetof(MapPoint, occ_val));',
            'printf("dyn_aug_cpp %zu\\n", offsetof(MapPoint, dyn_aug));', "return 0;}"]
    src.write_text("\n".join(cxx))
    exe = tmp_path / "mp"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        k, *v = line.split()
        got[k] = [int(x) for x in v]
    assert got["size"] == [48, 48, 16]
    for f in fields:
        assert got[f][0] == MAPPOINT_DTYPE.fields[f][1], f
    # the reference's field order after PCL_ADD_POINT4D (16 bytes)
    assert got["occ_val_cpp"][0] == 28 and got["dyn_aug_cpp"][0] == 40


def expected_records(d, method):
    """publish_mapdata's loop on the host, from the read-back cells."""
    cells = d.copyCellsToHost()
    info = d.info()
    gs, gsz, V = info.grid_size, info.grid_size_z, info.grid_cell_count
    res, size, size_z = f32(info.resolution), f32(d.params.size), f32(d.params.size_z)
    c = [f32(info.center_pos[k]) for k in range(3)]
    i = np.arange(V)
    two = f32(2.0)
    x = ((i % gs).astype(f32) + f32(0.5)) * res - size / two + c[0]
    y = (((i % (gs * gs)) // gs).astype(f32) + f32(0.5)) * res - size / two + c[1]
    z = ((i // (gs * gs)).astype(f32) + f32(0.5)) * res - size_z / two + c[2]
    fm, sm, dm, om = (cells[k] for k in ("free_mass", "static_mass", "dynamic_mass", "occ_mass"))
    vx, vy, vz = (cells[k] for k in ("mean_x_vel", "mean_y_vel", "mean_z_vel"))
    from k3dom_amd.dom import MAPPOINT_DTYPE
    out = np.zeros(V, MAPPOINT_DTYPE)
    out["x"], out["y"], out["z"], out["w"] = x, y, z, f32(1.0)
    out["v_x"], out["v_y"], out["v_z"] = vx * res, vy * res, vz * res
    one = f32(1.0)
    if method == "dsphdmib":
        occ_prob = om + (one - om - fm) / two
        out["occ_val"] = occ_prob
        out["dyn_val"] = (vx * vx + vy * vy + vz * vz) * res * res
        out["eval_aug"] = -(one - om - fm)
        out["dyn_aug"] = om
        visible = lambda mo, mm: (occ_prob > f32(mo)) & (om + fm > f32(mm))
    else:
        total = fm + sm + dm
        out["occ_val"] = (sm + dm) / total
        out["dyn_val"] = dm / total
        out["eval_aug"] = total
        out["dyn_aug"] = f32(1.1)
        visible = lambda mo, mm: ((sm + dm) / total > f32(mo)) & (total > f32(mm))
    return out, visible


def _same(a, b):
    a = np.ascontiguousarray(a).view(np.uint32).reshape(len(a), -1)
    b = np.ascontiguousarray(b).view(np.uint32).reshape(len(b), -1)
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["k3dom", "dsphdmib"])
def test_export_matches_publish_mapdata(gpu, method):
    """All voxels and the visible subset after frames with a moving ego (map
    centre away from the origin), bit for bit against the host loop."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    w = small_workload(ego_vx=4.0, frames=6, seed=3, method=method)
    sc = Scene(w.scene)
    d = DOM(w.params, method=method)
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        d.updatePose(*pose)
        d.updateGrid(t, xs, ys, zs)
    print(f"  {method}: exported after {w.frames} frames", flush=True)
    assert d.info().center_pos[0] != 0.0
    exp, visible = expected_records(d, method)
    got = d.export_mappoints()
    assert _same(got, exp), "all-voxel export differs from publish_mapdata"
    for mo, mm in ((0.3, 0.01), (0.6, 0.5), (0.0, 0.0)):
        sel = visible(mo, mm)
        got = d.export_mappoints(visible=True, min_occ=mo, min_mass=mm)
        assert 0 < len(got) < len(exp) or (mo == 0.0 and len(got) > 0)
        assert _same(got, exp[sel]), f"visible export ({mo}, {mm})"


@pytest.mark.gpu
def test_export_count_query_and_capacity(gpu):
    from k3dom_amd.dom import DOM, MAPPOINT_DTYPE
    from k3dom_amd.scene import Scene, small_workload
    w = small_workload(frames=3)
    sc = Scene(w.scene)
    d = DOM(w.params)
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        d.updatePose(*pose)
        d.updateGrid(t, xs, ys, zs)
    L = d._lib
    n = C.c_int64(-1)
    assert L.k3dom_export_mappoints(d._h, 1, 0.3, 0.01, None, 0, C.byref(n)) == 0
    full = d.export_mappoints(visible=True, min_occ=0.3, min_mass=0.01)
    assert n.value == len(full) > 1
    small = np.empty(n.value - 1, MAPPOINT_DTYPE)
    m = C.c_int64(-1)
    rc = L.k3dom_export_mappoints(d._h, 1, 0.3, 0.01, small.ctypes.data_as(C.c_void_p), len(small), C.byref(m))
    assert rc == K3DOM_ERR_INVALID_ARG and m.value == n.value
    assert L.k3dom_export_mappoints(d._h, 7, 0.0, 0.0, None, 0, C.byref(m)) == K3DOM_ERR_INVALID_ARG
