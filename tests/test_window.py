"""Differentiated window: the GPU adjoint against the double-precision chain
of the serial restatement (forward-mode derivatives and central differences).

The chain is pinned on the CPU first: its loss equals the loss of the float
filter's own masses (so the continuous model with frozen counts is the
filter), and its forward-mode gradient equals central differences."""
import numpy as np
import pytest

from tests._window import bce_loss, oracle_window


def _small(**kw):
    from k3dom_amd.scene import small_workload
    return small_workload(**kw)


@pytest.mark.parametrize("ego_vx", [0.0, 4.0])
def test_chain_loss_is_the_float_filter_loss(oracle_lib, ego_vx):
    from k3dom_amd.adjoint import theta_of
    w = _small(frames=6, seed=7, ego_vx=ego_vx)
    chain, recs = oracle_window(w, 6)
    if ego_vx:
        assert any(r["move"].any() for r in recs), "the moving case must shift the map"
    L = chain.loss(theta_of(w.params).astype(np.float64))
    ref = bce_loss(recs)
    assert abs(L - ref) <= 1e-5 * abs(ref), (L, ref)


@pytest.mark.parametrize("ego_vx", [0.0, 4.0])
def test_chain_forward_mode_equals_central_differences(oracle_lib, ego_vx):
    from k3dom_amd.adjoint import theta_of
    w = _small(frames=5, seed=8, ego_vx=ego_vx)
    chain, _ = oracle_window(w, 5)
    th = theta_of(w.params).astype(np.float64)
    th[1], th[2] = 0.12, 0.08  # the chain takes the two evidence weights apart
    _, g = chain.grad(th)
    assert np.all(np.isfinite(g)) and np.any(g != 0)
    # step 1e-4 of each parameter: within 1e-3 (a step that wide straddles a
    # few voxels' min/max branch switches of the mass update)
    fd = chain.central_differences(th)
    assert np.all(np.abs(g - fd) <= 1e-3 * np.abs(fd)), (g, fd)
    # narrower steps converge to the forward-mode derivative
    fd6 = chain.central_differences(th, rel_step=1e-6)
    assert np.all(np.abs(g - fd6) <= 2e-5 * np.abs(fd6)), (g, fd6)


def test_chain_with_mask(oracle_lib):
    """A voxel subset in the loss."""
    from k3dom_amd.adjoint import theta_of
    w = _small(frames=4, seed=9)
    rng = np.random.default_rng(0)
    V = int(np.prod([16, 16, 8]))
    mask = np.zeros(V, np.uint8)
    mask[rng.choice(V, 200, replace=False)] = 1
    chain, recs = oracle_window(w, 4, mask=mask)
    th = theta_of(w.params).astype(np.float64)
    L = chain.loss(th)
    assert abs(L - bce_loss(recs, mask)) <= 1e-5 * L
    _, g = chain.grad(th)
    fd = chain.central_differences(th, rel_step=1e-6)
    assert np.all(np.abs(g - fd) <= 2e-5 * np.abs(fd))


# ---------------------------------------------------------------------------- GPU
def _gpu_window(w, frames, first_frame=0, mask=None):
    from k3dom_amd.adjoint import Window
    from k3dom_amd.scene import Scene
    win = Window(w.params, frames)
    win.set_scene(Scene(w.scene), first_frame)
    win.set_mask(mask)
    return win


def _chain_from_gpu(win, w, frames, first_frame=0, mask=None):
    """The double chain fed with the GPU forward's own record (moves, counts)."""
    from k3dom_amd.adjoint import map_centers
    from k3dom_amd.scene import Scene
    from oracle.orc import WindowChain64
    p = w.params
    sc = Scene(w.scene)
    recs = [sc.frame(first_frame + f) for f in range(frames)]
    centers = map_centers(p, [r[1] for r in recs])
    srel = [np.float32(r[1]) - np.float32(c) for r, c in zip(recs, centers)]
    decays, prev = [], None
    for r in recs:
        decays.append(1.0 if prev is None else float(np.float32(float(np.float32(p.gamma)) **
                                                               float(np.float32(r[0] - prev)))))
        prev = r[0]
    gs = win.grid[0]
    gts = [sc.occupancy(first_frame + f, gs, win.grid[2], p.resolution, centers[f], p.size, p.size_z)
           for f in range(frames)]
    moves = [win.record(f, "move") for f in range(frames)]
    counts = [win.record(f, "counts") for f in range(frames)]
    return WindowChain64(p, [(r[2], r[3], r[4]) for r in recs], srel, moves, decays, counts, gts, mask)


@pytest.mark.gpu
def test_window_forward_record_is_bit_exact(gpu):
    """The window's forward frames are the filter's frames: masses and frozen
    counts equal the serial restatement's bit for bit."""
    from k3dom_amd.adjoint import theta_of
    w = _small(frames=6, seed=7, ego_vx=4.0)
    print("window: small scene, 6 frames, moving ego", flush=True)
    win = _gpu_window(w, 6)
    L = win.forward(theta_of(w.params))
    _, recs = oracle_window(w, 6)
    for f, r in enumerate(recs):
        m = win.record(f, "masses")
        for k, name in enumerate("fsd"):
            assert np.array_equal(m[k].view(np.uint32), r[name].view(np.uint32)), (f, name)
        assert np.array_equal(win.record(f, "counts"), r["counts"]), f
        assert np.array_equal(win.record(f, "move"), r["move"]), f
    assert abs(L - bce_loss(recs)) <= 1e-6 * L


def _rel_err(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("ego_vx,sig", [(0.0, (0.1, 0.1)), (4.0, (0.12, 0.08))])
def test_window_gradient_matches_fp64(gpu, ego_vx, sig):
    from k3dom_amd.adjoint import theta_of
    w = _small(frames=5, seed=11, ego_vx=ego_vx)
    th = theta_of(w.params)
    th[1], th[2] = sig
    print(f"window gradient: small scene, ego {ego_vx} m/s, sigmas {sig}", flush=True)
    win = _gpu_window(w, 5)
    L = win.forward(th)
    g = win.backward()
    chain = _chain_from_gpu(win, w, 5)
    L64, g64 = chain.grad(th.astype(np.float64))
    # a narrow step: the small scene's loss has branch switches of the mass
    # update within 1e-4 of ls (central differences converge to g64 below it)
    fd = chain.central_differences(th.astype(np.float64), rel_step=1e-6)
    print("grad", g, "fp64", g64, "fd", fd, flush=True)
    assert abs(L - L64) <= 1e-5 * abs(L64)
    assert np.all(_rel_err(g, g64) <= 1e-3), (g, g64)
    assert np.all(_rel_err(g, fd) <= 1e-3), (g, fd)


@pytest.mark.gpu
def test_window_gradient_config3(gpu):
    """The differentiated configuration: config-1 scene, ls 0.3, 8 frames,
    loss over 1,000 random voxels; GPU FP32 gradient within 1e-3 relative of
    FP64 central differences (step 1e-4 of each parameter)."""
    from k3dom_amd.adjoint import theta_of
    from k3dom_amd.scene import workload
    w = workload(3)
    V = 64 * 64 * 16
    mask = np.zeros(V, np.uint8)
    mask[np.random.default_rng(3).choice(V, 1000, replace=False)] = 1
    print("window gradient: config 3, 8 frames, 1000 voxels", flush=True)
    win = _gpu_window(w, 8, mask=mask)
    th = theta_of(w.params)
    L = win.forward(th)
    g = win.backward()
    chain = _chain_from_gpu(win, w, 8, mask=mask)
    fd = chain.central_differences(th.astype(np.float64))
    L64 = chain.loss(th.astype(np.float64))
    print("loss", L, L64, "grad", g, "fd", fd, flush=True)
    assert abs(L - L64) <= 1e-5 * abs(L64)
    assert np.all(_rel_err(g, fd) <= 1e-3), (g, fd)


@pytest.mark.gpu
def test_window_torch_autograd_and_per_voxel_priors(gpu):
    """torch.autograd over the libtorch function gives the window's gradient;
    per-voxel initial-mass gradients sum to the prior gradients (no shift)."""
    import torch
    from k3dom_amd.adjoint import theta_of
    w = _small(frames=4, seed=12)
    win = _gpu_window(w, 4)
    th = torch.tensor(theta_of(w.params), dtype=torch.float32, requires_grad=True)
    loss = win.loss(th)
    loss.backward()
    L = win.forward(theta_of(w.params))
    g, pv = win.backward(per_voxel=True)
    assert abs(loss.item() - L) <= 1e-12 * abs(L)
    assert np.allclose(th.grad.numpy().astype(np.float64), g.astype(np.float32), rtol=1e-6, atol=0)
    for k in range(3):
        assert abs(pv[k].astype(np.float64).sum() - g[3 + k]) <= 1e-4 * abs(g[3 + k]) + 1e-6


@pytest.mark.gpu
def test_window_two_losses_before_backward(gpu):
    """Two forward passes on one window before the first loss's backward
    (a second loss, or a validation loss): each loss's gradient is the one
    at its own theta, as computed alone."""
    import torch
    from k3dom_amd.adjoint import theta_of
    w = _small(frames=3, seed=14)
    win = _gpu_window(w, 3)
    th0 = np.array(theta_of(w.params), np.float32)
    th1 = th0 * np.array([1.3, 0.7, 1.2, 2.0, 0.5, 1.5], np.float32)
    a = torch.tensor(th0, requires_grad=True)
    b = torch.tensor(th1, requires_grad=True)
    la = win.loss(a)
    lb = win.loss(b)
    la.backward()
    lb.backward()
    for th, t in ((th0, a), (th1, b)):
        win.forward(th)
        g = win.backward()
        assert np.array_equal(t.grad.numpy(), g.astype(np.float32)), (t.grad.numpy(), g)
    assert not np.allclose(a.grad.numpy(), b.grad.numpy())


# ------------------------------------------------------------ several LiDARs

def _rotate32(rot, xs, ys, zs):
    """k3_lidar_rotate in float32: each row as (r0 x + r1 y) + r2 z."""
    r = np.asarray(rot, np.float32)
    xs, ys, zs = (np.asarray(a, np.float32) for a in (xs, ys, zs))
    return tuple((r[i, 0] * xs + r[i, 1] * ys) + r[i, 2] * zs for i in range(3))


def _chain_from_gpu_rig(win, w, rig, frames, mask=None):
    """The double chain of a rig window: points rotated as the GPU does,
    each LiDAR's sensor position (pose + origin) - centre as the frame uses."""
    from k3dom_amd.adjoint import map_centers
    from k3dom_amd.scene import Scene
    from oracle.orc import WindowChain64
    p = w.params
    sc = Scene(w.scene)
    recs = [sc.frame_lidars(f, rig) for f in range(frames)]
    centers = map_centers(p, [r[1] for r in recs])
    K = len(rig)
    pts, srel, decays, prev = [], [], [], None
    ng = np.full(frames, K, np.int32)
    gc = np.zeros((frames, 16), np.int32)
    gr = np.zeros((frames, 16, 3), np.float32)
    for f, (t, pose, lidars) in enumerate(recs):
        rot = [_rotate32(l[0], *l[2:5]) for l in lidars]
        pts.append(tuple(np.concatenate([r[k] for r in rot]) for k in range(3)))
        s, c = np.float32(pose), np.float32(centers[f])
        srel.append(s - c)
        for k, l in enumerate(lidars):
            gc[f, k] = len(l[2])
            gr[f, k] = (s + np.asarray(l[1], np.float32)) - c
        decays.append(1.0 if prev is None else float(np.float32(float(np.float32(p.gamma)) ** float(np.float32(t - prev)))))
        prev = t
    gs = win.grid[0]
    gts = [sc.occupancy(f, gs, win.grid[2], p.resolution, centers[f], p.size, p.size_z) for f in range(frames)]
    moves = [win.record(f, "move") for f in range(frames)]
    counts = [win.record(f, "counts") for f in range(frames)]
    return WindowChain64(p, pts, srel, moves, decays, counts, gts, mask, groups=(ng, gc, gr))


def test_chain_single_identity_group_is_the_plain_chain(oracle_lib):
    """One LiDAR at the pose with the identity: the grouped chain is the plain one."""
    from k3dom_amd.adjoint import theta_of
    from oracle.orc import WindowChain64
    w = _small(frames=4, seed=7, ego_vx=4.0)
    chain, _ = oracle_window(w, 4)
    T = chain.T
    ng = np.ones(T, np.int32)
    gc = np.zeros((T, 16), np.int32)
    gc[:, 0] = chain.npts
    gr = np.zeros((T, 16, 3), np.float32)
    gr[:, 0] = chain.sensor_rel.reshape(T, 3)
    g = WindowChain64(chain.params, [], [], [], [], [], [])
    g.__dict__.update({k: v for k, v in chain.__dict__.items()})
    g.groups = (ng, gc.reshape(-1), gr.reshape(-1))
    th = theta_of(w.params).astype(np.float64)
    assert g.loss(th) == chain.loss(th)


@pytest.mark.gpu
def test_window_rig_record_and_gradient(gpu):
    """A window over three LiDARs: its forward frames equal the serial
    restatement's multi-LiDAR frames bit for bit, and its gradient matches the
    FP64 chain with the LiDAR groups within 1e-3."""
    from k3dom_amd.adjoint import Window, theta_of
    from k3dom_amd.scene import Scene, small_rig
    from oracle.orc import OracleDOM
    from tests._parity import bits_equal
    w = _small(frames=4, seed=11, ego_vx=4.0)
    rig = small_rig(0.4, 3)
    print("window gradient: small scene, 3 LiDARs, 4 frames", flush=True)
    win = Window(w.params, 4)
    recs = win.set_scene_rig(Scene(w.scene), rig)
    th = theta_of(w.params)
    th[1], th[2] = 0.12, 0.08
    L = win.forward(th)
    # the forward frames at the map's own parameters against the restatement
    od = OracleDOM(w.params)
    win0 = Window(w.params, 4)
    win0.set_scene_rig(Scene(w.scene), rig)
    win0.forward(theta_of(w.params))
    for f, (t, pose, lidars) in enumerate(recs):
        od.updatePose(*pose)
        od.updateGridLidars(t, lidars)
        cells = od.copyCellsToHost()
        m = win0.record(f, "masses")
        for k, name in enumerate(("free_mass", "static_mass", "dynamic_mass")):
            assert bits_equal(m[k], cells[name]), (f, name)
    win0.close()
    g = win.backward()
    chain = _chain_from_gpu_rig(win, w, rig, 4)
    L64, g64 = chain.grad(th.astype(np.float64))
    fd = chain.central_differences(th.astype(np.float64), rel_step=1e-6)
    print("grad", g, "fp64", g64, "fd", fd, flush=True)
    assert abs(L - L64) <= 1e-5 * abs(L64)
    assert np.all(_rel_err(g, g64) <= 1e-3), (g, g64)
    assert np.all(_rel_err(g, fd) <= 1e-3), (g, fd)


def test_dual_credit_derivative_is_finite_at_large_masses(oracle_lib):
    """credit(m) = 1 - 2 / (1 + e^(2m/s)) = tanh(m/s); its forward-mode
    derivative is (1 - tanh^2) / s, finite and 0 where e^(2m/s) overflows
    (voxels near a sensor reach masses of 1e4)."""
    import ctypes as C
    import math
    v, d = C.c_double(), C.c_double()
    for m, s in ((0.0, 2.5), (1.0, 2.5), (3.0, 0.7), (-2.0, 2.5)):
        oracle_lib.orc_dual_credit(m, s, C.byref(v), C.byref(d))
        assert abs(v.value - math.tanh(m / s)) < 1e-14
        assert abs(d.value - (1 - math.tanh(m / s) ** 2) / s) < 1e-12
    for m in (400.0, 1e4, 1e30):
        oracle_lib.orc_dual_credit(m, 2.5, C.byref(v), C.byref(d))
        assert v.value == 1.0 and d.value == 0.0
