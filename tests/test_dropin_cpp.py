"""The C++ drop-in header include/k3dom/dom.h (namespace dom, DOM / DOM_c /
Params / GridCell as in the reference's include/common/dom.h:16-148 and
include/common/dom_types.h:12-58), driven by a frame-replay program that
follows the reference node's callback (src/k3dom/k3dom_node.cpp:64-97)."""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPLAY = os.path.join(ROOT, "k3dom_amd", "k3dom_replay")
REPLAY_DS = os.path.join(ROOT, "k3dom_amd", "k3dom_replay_dsphdmib")


def test_header_compiles_with_reference_names(tmp_path):
    """A caller written against the reference's class compiles unchanged
    (aggregate Params, non-const vector arguments, GridCell accessors)."""
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include "k3dom/dom.h"
int main() {
    dom::DOM::Params p{25.6f, 3.2f, 2.5f, 0.1f, 1000, 100, 0.99f, 0.1f, 1.0f, 0.02f, 1.0f, 50.0f, 1.0f,
                       0.f, 0.f, 0.f, 10, 0.001f, 0.001f, 0.001f, 0.5f, 0.1f, 0.5f, 0.99f, 40.f, 0.1f};
    static_assert(sizeof(dom::GridCell) == 68, "GridCell size");
    dom::GridCell c{};
    c.free_mass = 1.f; c.static_mass = 2.f; c.dynamic_mass = 1.f;
    if (c.occ_mean() != 0.75f || c.dyn_mean() != 0.25f) return 1;
    if (c.calcVar() != 0.5f * 0.5f / 5.0f) return 2;
    dom::GridCell z{};
    if (z.calcVar() != 1e6f) return 3;
    void (dom::DOM::*upd)(float, std::vector<float>&, std::vector<float>&, std::vector<float>&) =
        &dom::DOM::updateGrid;
    void (dom::DOM_c::*pose)(float, float, float) = &dom::DOM_c::updatePose;
    bool (dom::DOM_c::*save)(std::vector<unsigned char>&) = &dom::DOM_c::saveCheckpoint;
    bool (dom::DOM_c::*load)(const void*, size_t) = &dom::DOM_c::loadCheckpoint;
    void (dom::DOM_c::*mp)(std::vector<MapPoint>&, bool, float, float) = &dom::DOM_c::copyMapPointsToHost;
    void (dom::DOM_c::*lid)(double, const std::vector<k3dom_lidar>&, bool) = &dom::DOM_c::updateGridLidars;
    (void)upd; (void)pose; (void)p; (void)save; (void)load; (void)mp; (void)lid;
    return 0;
}
''')
    exe = tmp_path / "caller"
    lib = os.path.join(ROOT, "k3dom_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", lib, "-lk3dom",
                    "-Wl,-rpath," + lib, "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def _write_inputs(tmp_path, workload, frames):
    from k3dom_amd.scene import Scene
    ppath = tmp_path / "params.bin"
    ppath.write_bytes(bytes(workload.params))
    sc = Scene(workload.scene)
    recs = [sc.frame(f) for f in range(frames)]
    with open(tmp_path / "frames.bin", "wb") as fh:
        fh.write(b"K3FR" + struct.pack("<i", frames))
        for t, pose, xs, ys, zs in recs:
            fh.write(struct.pack("<d3fi", t, *pose, len(xs)))
            for a in (xs, ys, zs):
                fh.write(np.ascontiguousarray(a, np.float32).tobytes())
    return recs


@pytest.mark.gpu
def test_replay_through_cpp_header_matches_oracle(gpu, tmp_path):
    """dom::DOM over the C ABI, in its own process, gives the oracle's cells
    bit for bit after a map-shifting sequence of frames."""
    from k3dom_amd.dom import CELL_DTYPE
    from k3dom_amd.scene import small_workload
    from oracle.orc import OracleDOM
    from tests._parity import compare_cells
    assert os.path.exists(REPLAY), "k3dom_replay is not built (run __graft_entry__.build())"
    w = small_workload(ego_vx=4.0, frames=6, seed=4)
    recs = _write_inputs(tmp_path, w, w.frames)
    out = tmp_path / "cells.bin"
    print("replay: small workload, 6 frames, moving ego", flush=True)
    r = subprocess.run([REPLAY, str(tmp_path / "params.bin"), str(tmp_path / "frames.bin"), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    V = struct.unpack("<i", raw[:4])[0]
    cells = np.frombuffer(raw[4:], CELL_DTYPE)
    assert len(cells) == V
    od = OracleDOM(w.params)
    for t, pose, xs, ys, zs in recs:
        od.updatePose(*pose)
        od.updateGrid(t, xs, ys, zs)
    compare_cells(cells, od.copyCellsToHost(), "replay")


@pytest.mark.gpu
def test_replay_through_cpp_stage_methods_matches_oracle(gpu, tmp_path):
    """The same replay driving every frame through the reference's stage
    methods (initializeParticles, then particlePrediction ... resampling in
    the order of src/k3dom/dom.cu:29-50) gives the oracle's cells bit for bit."""
    from k3dom_amd.dom import CELL_DTYPE
    from k3dom_amd.scene import small_workload
    from oracle.orc import OracleDOM
    from tests._parity import compare_cells
    w = small_workload(ego_vx=4.0, frames=6, seed=4)
    recs = _write_inputs(tmp_path, w, w.frames)
    out = tmp_path / "cells.bin"
    r = subprocess.run([REPLAY, str(tmp_path / "params.bin"), str(tmp_path / "frames.bin"), str(out), "stages"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    cells = np.frombuffer(out.read_bytes()[4:], CELL_DTYPE)
    od = OracleDOM(w.params)
    for t, pose, xs, ys, zs in recs:
        od.updatePose(*pose)
        od.updateGrid(t, xs, ys, zs)
    compare_cells(cells, od.copyCellsToHost(), "stage replay")


def test_k3dom_header_declares_the_reference_stage_methods(tmp_path):
    """A caller of the reference's stage methods (include/common/dom.h:89-94,
    include/k3dom/dom.h:24-28) compiles unchanged against the drop-in header."""
    src = tmp_path / "caller_stages.cpp"
    src.write_text(r'''
#include "k3dom/dom.h"
int main() {
    void (dom::DOM_c::*pred)(float) = &dom::DOM_c::particlePrediction;
    void (dom::DOM_c::*stages[])() = {&dom::DOM_c::particleAssignment, &dom::DOM_c::updatePersistentParticles,
                                       &dom::DOM_c::initializeNewParticles, &dom::DOM_c::statisticalMoments,
                                       &dom::DOM_c::resampling};
    void (dom::DOM::*init)(std::vector<float>&, std::vector<float>&, std::vector<float>&) =
        &dom::DOM::initializeParticles;
    void (dom::DOM::*occ)(float, std::vector<float>&, std::vector<float>&, std::vector<float>&) =
        &dom::DOM::gridCellOccupancyUpdate;
    (void)pred; (void)stages; (void)init; (void)occ;
    return 0;
}
''')
    exe = tmp_path / "caller_stages"
    lib = os.path.join(ROOT, "k3dom_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", lib, "-lk3dom",
                    "-Wl,-rpath," + lib, "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_dsphdmib_header_compiles_with_reference_names(tmp_path):
    """A caller of the reference's DS-PHD/MIB class (include/dsphdmib/dom.h:17-41)
    compiles unchanged: same class and member names, MeasurementCell records,
    window_radius = 0.866 resolution."""
    src = tmp_path / "caller_ds.cpp"
    src.write_text(r'''
#include "dsphdmib/dom.h"
int main() {
    static_assert(sizeof(dom::MeasurementCell) == 8, "MeasurementCell size");
    dom::MeasurementCell m{0.25f, 0.5f};
    if (m.free_mass != 0.25f || m.occ_mass != 0.5f) return 1;
    dom::GridCell c{};
    c.occ_mass = 0.5f; c.free_mass = 0.25f;
    if (c.occ_prob() != 0.625f) return 2;
    void (dom::DOM::*upd)(float, std::vector<float>&, std::vector<float>&, std::vector<float>&) =
        &dom::DOM::updateGrid;
    void (dom::DOM::*meas)(std::vector<dom::MeasurementCell>&) = &dom::DOM::copyMeasGridToHost;
    float dom::DOM::*wr = &dom::DOM::window_radius;
    (void)upd; (void)meas; (void)wr;
    return 0;
}
''')
    exe = tmp_path / "caller_ds"
    lib = os.path.join(ROOT, "k3dom_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", lib, "-lk3dom",
                    "-Wl,-rpath," + lib, "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


@pytest.mark.gpu
def test_replay_through_dsphdmib_header_matches_oracle(gpu, tmp_path):
    """The DS-PHD/MIB dom::DOM, in its own process, gives the oracle's cells
    and measurement grid bit for bit after a map-shifting sequence of frames."""
    from k3dom_amd.dom import CELL_DTYPE
    from k3dom_amd.scene import small_workload
    from oracle.orc import OracleDOM
    from tests._parity import bits_equal, compare_cells, first_diff
    assert os.path.exists(REPLAY_DS), "k3dom_replay_dsphdmib is not built (run __graft_entry__.build())"
    w = small_workload(ego_vx=4.0, frames=6, seed=4, method="dsphdmib")
    recs = _write_inputs(tmp_path, w, w.frames)
    out = tmp_path / "cells.bin"
    print("DS replay: small workload, 6 frames, moving ego", flush=True)
    r = subprocess.run([REPLAY_DS, str(tmp_path / "params.bin"), str(tmp_path / "frames.bin"), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    V = struct.unpack("<i", raw[:4])[0]
    cells = np.frombuffer(raw[4:4 + 68 * V], CELL_DTYPE)
    meas = np.frombuffer(raw[4 + 68 * V:], np.float32).reshape(V, 2)
    od = OracleDOM(w.params, method="dsphdmib")
    for t, pose, xs, ys, zs in recs:
        od.updatePose(*pose)
        od.updateGrid(t, xs, ys, zs)
    compare_cells(cells, od.copyCellsToHost(), "DS replay")
    for k, name in ((0, "DEL_FREE"), (1, "DEL_OCC")):
        a, b = np.ascontiguousarray(meas[:, k]), od.read_array(name)
        assert bits_equal(a, b), f"DS replay measurement grid {name}: {first_diff(a, b)}"
