"""Allocation-free MapPoint export (k3dom_export_mappoints_noalloc, C ABI)
against the allocating k3dom_export_mappoints: the same records bit for bit,
for every voxel and for the visible voxels at several thresholds, on two
maps in a row (the second smaller, so the shared scratch holds the first
call's tile statuses), including a count-only call and a capacity below the
count; and the device's free memory does not drop across the calls, while
the allocating export's first call on the same map does allocate."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K3DOM_OK, K3DOM_ERR_INVALID_ARG = 0, 1
ALL, VISIBLE = 0, 1
THRESHOLDS = [(-1.0, -1.0), (0.5, 0.0), (0.3, 0.0025), (0.9, 0.01)]


def _frames(w, frames):
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene
    sc = Scene(w.scene)
    d = DOM(w.params)
    for f in range(frames):
        t, pose, xs, ys, zs = sc.frame(f)
        d.updatePose(*pose)
        d.updateGrid(t, xs, ys, zs)
    return d


def _noalloc(d, mode, min_occ, min_mass, buf, cap):
    n = C.c_int64(-1)
    ptr = C.c_void_p(buf.data_ptr()) if buf is not None else None
    st = d._lib.k3dom_export_mappoints_noalloc(d._h, mode, min_occ, min_mass, ptr, cap, C.byref(n))
    return st, n.value


def _records(buf, n):
    from k3dom_amd.dom import MAPPOINT_DTYPE
    return buf[:n * 48].cpu().numpy().view(MAPPOINT_DTYPE)


def _check_map(d, what):
    import torch
    V = d.grid_cell_count
    buf = torch.empty(V * 48, dtype=torch.uint8, device="cuda")
    ref = d.export_mappoints(visible=False)
    st, n = _noalloc(d, ALL, 0.0, 0.0, buf, V)
    assert st == K3DOM_OK and n == V, f"{what}: every voxel: status {st}, count {n}"
    assert np.array_equal(_records(buf, n).view(np.uint32), ref.view(np.uint32)), f"{what}: every voxel"
    counts = {}
    for mo, mm in THRESHOLDS:
        ref = d.export_mappoints(visible=True, min_occ=mo, min_mass=mm)
        counts[(mo, mm)] = len(ref)
        buf.fill_(0xAB)
        st, n = _noalloc(d, VISIBLE, mo, mm, buf, V)
        assert st == K3DOM_OK and n == len(ref), f"{what} {mo},{mm}: status {st}, count {n} vs {len(ref)}"
        got = _records(buf, n)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"{what} {mo},{mm}: records differ"
        st, n2 = _noalloc(d, VISIBLE, mo, mm, None, 0)
        assert st == K3DOM_OK and n2 == len(ref), f"{what} {mo},{mm}: count-only call"
        if len(ref) >= 2:
            cap = len(ref) // 2
            buf.fill_(0xAB)
            st, n3 = _noalloc(d, VISIBLE, mo, mm, buf, cap)
            assert st == K3DOM_ERR_INVALID_ARG and n3 == len(ref), f"{what} {mo},{mm}: capacity below the count"
            assert np.array_equal(_records(buf, cap).view(np.uint32), ref[:cap].view(np.uint32)), \
                f"{what} {mo},{mm}: first records under a short capacity"
            tail = buf[cap * 48:].cpu().numpy()
            assert np.all(tail == 0xAB), f"{what} {mo},{mm}: wrote past the capacity"
    return counts


def test_noalloc_export_matches_export_on_two_maps(gpu):
    import torch
    from k3dom_amd.scene import small_workload
    # 128x128x32 voxels = 128 scan tiles (several look-back windows), then a
    # 16x16x8 map of one tile
    big = _frames(small_workload(edge=128, edge_z=32, res=0.1, particles=200000, newborn=20000, beams=32, az=512,
                                 frames=3, seed=41, ego_vx=3.0, static=40, moving=6), 3)
    small = _frames(small_workload(frames=4, seed=43, ego_vx=2.0), 4)
    V = big.grid_cell_count
    buf = torch.empty(V * 48, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for mo, mm in THRESHOLDS:
        st, _ = _noalloc(big, VISIBLE, mo, mm, buf, V)
        assert st == K3DOM_OK
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    big.export_mappoints(visible=True, min_occ=0.5)  # first allocating export on this map
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 < 2 << 20, f"free device memory dropped by {free0 - free1} bytes across noalloc calls"
    assert free1 - free2 >= 20 << 20, f"the allocating export took only {free1 - free2} bytes"
    del buf
    n_big = _check_map(big, "128x128x32")
    n_small = _check_map(small, "16x16x8")
    assert n_big[(0.5, 0.0)] > 4096 and n_small[(0.5, 0.0)] > 0
