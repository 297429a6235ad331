"""The reference's public stage methods (include/common/dom.h:89-94,
include/k3dom/dom.h:24-28) driven one by one in the order DOM::updateGrid
calls them (src/k3dom/dom.cu:29-50) give the same map as updateGrid, bit
for bit, through map shifts and list regrowth; calls out of that order, a
whole frame in the middle of a staged one, and DS-PHD/MIB maps are refused
without disturbing the map."""
import numpy as np
import pytest

from tests._parity import FRAME_ARRAYS, PARTICLE_ARRAYS, bits_equal, compare_cells, first_diff

pytestmark = pytest.mark.gpu

K3DOM_ERR_STATE = 4  # include/k3dom.h


def _staged_frame(m, prev_t, t, xs, ys, zs):
    if prev_t > 0:
        dt = np.float32(t - prev_t)
        m.particlePrediction(dt)
        m.particleAssignment()
        m.gridCellOccupancyUpdate(dt, xs, ys, zs)
        m.updatePersistentParticles()
        m.initializeNewParticles()
        m.statisticalMoments()
        m.resampling()
    else:
        m.initializeParticles(xs, ys, zs)


def _compare(a, b, normal, what):
    compare_cells(a.copyCellsToHost(), b.copyCellsToHost(), what)
    for name in PARTICLE_ARRAYS + (FRAME_ARRAYS if normal else ["BORN_MASS", "DEL_OCC", "DEL_FREE"]):
        x, y = a.read_array(name), b.read_array(name)
        assert bits_equal(x, y), f"{what} {name}: {first_diff(x, y)}"


@pytest.mark.parametrize("tiny", [False, True])
def test_stage_methods_equal_update_grid(gpu, monkeypatch, tiny):
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    if tiny:
        monkeypatch.setenv("K3DOM_TEST_TINY_CAPACITY", "1")
    w = small_workload(ego_vx=4.0, frames=7, seed=29)
    sc = Scene(w.scene)
    a, b = DOM(w.params), DOM(w.params)
    prev_t = -1.0
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        normal = prev_t > 0
        for m in (a, b):
            m.updatePose(*pose)
        a.updateGrid(t, xs, ys, zs)
        _staged_frame(b, prev_t, t, xs, ys, zs)
        prev_t = t
        _compare(a, b, normal, f"frame {f}:")
        assert a.info().iteration == b.info().iteration == f + 1
    assert b.info().updated_time == -1.0  # the stages leave the stored time to updateGrid
    if tiny:
        assert b.stats()["capacity_regrowths"] >= 1


def test_stage_order_is_enforced(gpu):
    from k3dom_amd._lib import K3domError
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    w = small_workload(frames=4, seed=31)
    sc = Scene(w.scene)
    a, b = DOM(w.params), DOM(w.params)
    recs = [sc.frame(f) for f in range(w.frames)]
    for m in (a, b):
        m.updatePose(*recs[0][1])
        m.updateGrid(*([recs[0][0]] + list(recs[0][2:])))
    m = b
    dt = np.float32(recs[1][0] - recs[0][0])

    def refused(call, *args):
        with pytest.raises(K3domError) as e:
            call(*args)
        assert e.value.status == K3DOM_ERR_STATE
    refused(m.particleAssignment)
    refused(m.resampling)
    refused(m.statisticalMoments)
    for mm in (a, b):
        mm.updatePose(*recs[1][1])
    m.particlePrediction(dt)
    refused(m.particlePrediction, dt)
    refused(m.updateGrid, recs[1][0], *recs[1][2:])   # a whole frame inside a staged one
    m.particleAssignment()
    m.gridCellOccupancyUpdate(dt, *recs[1][2:])
    refused(m.initializeNewParticles)
    m.updatePersistentParticles()
    m.statisticalMoments()   # before the newborns: the moments are already computed
    m.initializeNewParticles()
    m.statisticalMoments()
    m.resampling()
    a.updateGrid(recs[1][0], *recs[1][2:])
    _compare(a, b, True, "after refused calls:")


def test_stage_methods_refuse_dsphdmib_maps(gpu):
    from k3dom_amd._lib import K3domError
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import small_workload
    w = small_workload(method="dsphdmib")
    d = DOM(w.params, method="dsphdmib")
    with pytest.raises(K3domError) as e:
        d.particlePrediction(0.1)
    assert e.value.status == K3DOM_ERR_STATE
