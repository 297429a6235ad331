"""Differentiated window over the K3DOM update (no counterpart in the
reference, whose mass update is forward-only: src/k3dom/kernel/mass_update.cu:152-183).

A window is a fixed sequence of frames replayed on a fresh map.  Its loss is
the mean over frames of the mean over the selected voxels of the binary
cross-entropy of the occupancy mean (s + d) / (f + s + d) against a 0/1 ground
truth.  The gradient is taken with respect to

    theta = (ls, sigma_occ, sigma_free, prior_free, prior_static, prior_dynamic)

(the reference's single `sigma` split into the weights of the occupied and
the free kernel terms, src/k3dom/kernel/mass_update.cu:51,70).  The forward
pass is the ordinary GPU frame pipeline; the backward pass is the window's
reverse sweep (k3_window.cpp / k3_adjoint.hip), which recomputes each frame's
kernel-inference evidence.  `Window.loss(theta)` is a torch autograd function
(libtorch C++, module `_k3window`).
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import Params, check, load

THETA = ("ls", "sigma_occ", "sigma_free", "prior_free", "prior_static", "prior_dynamic")
_HERE = os.path.dirname(os.path.abspath(__file__))


def theta_of(params):
    """theta of a parameter set (both sigmas = the reference's sigma)."""
    return np.array([params.ls, params.sigma, params.sigma, params.prior_free, params.prior_static,
                     params.prior_dynamic], np.float32)


def map_centers(params, poses):
    """Map centre at each frame for a pose sequence, with the float32 arithmetic
    of DOM_c::updatePose (src/common/dom.cu:111-158)."""
    f32 = np.float32
    res = f32(params.resolution)
    off = [f32(params.sensor_off_x), f32(params.sensor_off_y), f32(params.sensor_off_z)]
    centers = []
    c = None
    for pose in poses:
        s = [f32(v) for v in pose]
        if c is None:
            c = [s[k] - off[k] for k in range(3)]
        else:
            mv = [int(np.trunc((s[k] - c[k] - off[k]) / res)) for k in range(3)]
            if any(abs(m) >= params.map_shift_thresh for m in mv):
                c = [f32(c[k] + f32(f32(mv[k]) * res)) for k in range(3)]
        centers.append(tuple(float(v) for v in c))
    return centers


class Window:
    """A differentiated window of `frames` frames over a map of `params`."""

    def __init__(self, params, frames):
        self._lib = load()
        self.params = params
        self.frames = frames
        h = C.c_void_p()
        p = Params.from_buffer_copy(bytes(params)) if not isinstance(params, Params) else params
        status = self._lib.k3dom_window_create(C.byref(p), int(frames), C.byref(h))
        if status != 0:
            raise RuntimeError(f"k3dom_window_create failed with status {status}: "
                               f"{self._lib.k3dom_last_error(None).decode()}")
        self._h = h
        gs = int(np.float32(params.size) / np.float32(params.resolution))
        gsz = int(np.float32(params.size_z) / np.float32(params.resolution))
        self.grid = (gs, gs, gsz)
        self.V = gs * gs * gsz

    def _check(self, status, what):
        if status != 0:
            msg = self._lib.k3dom_window_last_error(self._h)
            raise RuntimeError(f"{what} failed with status {status}: {msg.decode() if msg else ''}")

    @property
    def handle(self):
        return int(self._h.value)

    def set_frame(self, f, t, pose, xs, ys, zs, gt):
        xs, ys, zs = (np.ascontiguousarray(a, np.float32) for a in (xs, ys, zs))
        gt = np.ascontiguousarray(gt, np.uint8)
        assert gt.size == self.V and len(xs) == len(ys) == len(zs)
        pz = (C.c_float * 3)(*pose)
        fp = C.POINTER(C.c_float)
        self._check(self._lib.k3dom_window_set_frame(self._h, int(f), float(t), pz, xs.ctypes.data_as(fp),
                                                     ys.ctypes.data_as(fp), zs.ctypes.data_as(fp), len(xs),
                                                     gt.ctypes.data_as(C.POINTER(C.c_uint8))), "set_frame")

    def set_scene(self, scene, first_frame=0):
        """Frames first_frame .. first_frame + frames - 1 of a synthetic Scene, with
        its analytic box occupancy as the ground truth."""
        p = self.params
        recs = [scene.frame(first_frame + f) for f in range(self.frames)]
        centers = map_centers(p, [r[1] for r in recs])
        gs, _, gsz = self.grid
        for f, (t, pose, xs, ys, zs) in enumerate(recs):
            gt = scene.occupancy(first_frame + f, gs, gsz, p.resolution, centers[f], p.size, p.size_z)
            self.set_frame(f, t, pose, xs, ys, zs, gt)
        return recs

    def set_frame_lidars(self, f, t, pose, lidars, gt):
        """Frame f from several LiDARs: lidars = [(rot 3x3, origin 3, xs, ys, zs), ...]
        with points in each LiDAR's frame (k3dom_window_set_frame_lidars)."""
        K = len(lidars)
        arr = (_lib.Lidar * K)()
        keep = []
        for k, l in enumerate(lidars):
            arr[k].rot[:] = [float(v) for v in np.asarray(l[0], np.float32).reshape(9)]
            arr[k].origin[:] = [float(v) for v in np.asarray(l[1], np.float32).reshape(3)]
            pts = [np.ascontiguousarray(a, np.float32) for a in l[2:5]]
            keep.append(pts)
            arr[k].xs, arr[k].ys, arr[k].zs = (a.ctypes.data for a in pts)
            arr[k].n = len(pts[0])
        pz = np.ascontiguousarray(pose, np.float32)
        g = np.ascontiguousarray(gt, np.uint8)
        assert g.size == self.V
        self._check(self._lib.k3dom_window_set_frame_lidars(
            self._h, int(f), float(t), pz.ctypes.data_as(C.POINTER(C.c_float)), arr, K,
            g.ctypes.data_as(C.POINTER(C.c_uint8))), "set_frame_lidars")

    def set_scene_rig(self, scene, rig, first_frame=0):
        """set_scene for a rig of LiDARs (Scene.frame_lidars)."""
        p = self.params
        recs = [scene.frame_lidars(first_frame + f, rig) for f in range(self.frames)]
        centers = map_centers(p, [r[1] for r in recs])
        gs, _, gsz = self.grid
        for f, (t, pose, lidars) in enumerate(recs):
            gt = scene.occupancy(first_frame + f, gs, gsz, p.resolution, centers[f], p.size, p.size_z)
            self.set_frame_lidars(f, t, pose, lidars, gt)
        return recs

    def set_mask(self, mask=None):
        if mask is None:
            self._check(self._lib.k3dom_window_set_mask(self._h, None), "set_mask")
        else:
            m = np.ascontiguousarray(mask, np.uint8)
            assert m.size == self.V
            self._check(self._lib.k3dom_window_set_mask(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8))), "set_mask")

    def forward(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        loss = C.c_double()
        self._check(self._lib.k3dom_window_forward(self._h, th.ctypes.data_as(C.POINTER(C.c_float)),
                                                   C.byref(loss)), "forward")
        return loss.value

    def backward(self, dloss=1.0, per_voxel=False):
        g = (C.c_double * 6)()
        pv = np.empty(3 * self.V, np.float32) if per_voxel else None
        self._check(self._lib.k3dom_window_backward(self._h, float(dloss), g,
                                                    pv.ctypes.data_as(C.POINTER(C.c_float)) if per_voxel else None),
                    "backward")
        grad = np.array(list(g))
        return (grad, pv.reshape(3, self.V)) if per_voxel else grad

    def record(self, f, which):
        """which: "masses" (3, V), "dyn_pred" (V,), "counts" (4, V) int32, "move" (3,) int32."""
        V = self.V
        spec = {"masses": (0, np.float32, (3, V)), "dyn_pred": (1, np.float32, (V,)),
                "counts": (2, np.int32, (4, V)), "move": (3, np.int32, (3,))}[which]
        out = np.empty(spec[2], spec[1])
        self._check(self._lib.k3dom_window_read_record(self._h, int(f), spec[0], out.ctypes.data, out.nbytes),
                    "read_record")
        return out

    def times(self):
        ms = (C.c_float * 2)()
        self._check(self._lib.k3dom_window_times(self._h, ms), "times")
        return ms[0], ms[1]

    def loss(self, theta):
        """torch autograd: loss tensor (float64, 0-d) differentiable in theta."""
        from . import _k3window
        return _k3window.window_loss(theta, self.handle)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.k3dom_window_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
