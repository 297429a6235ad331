"""Python mirror of the reference's filter class dom::DOM
(include/k3dom/dom.h:17-30, include/common/dom.h:16-148) over the C ABI.

Method names follow the reference (updatePose, updateGrid, getGridSize, ...);
readback replaces the caller-side cudaMemcpy of grid_cell_array
(src/k3dom/k3dom_node.cpp:131-132) with copyCellsToHost().
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import Params, Cell, Info, ARR, check

CELL_DTYPE = np.dtype([
    ("start_idx", "<i4"), ("end_idx", "<i4"), ("mu_A", "<f4"), ("mu_UA", "<f4"), ("w_A", "<f4"), ("w_UA", "<f4"),
    ("mean_x_vel", "<f4"), ("mean_y_vel", "<f4"), ("mean_z_vel", "<f4"), ("likelihood", "<f4"), ("p_A", "<f4"),
    ("free_mass", "<f4"), ("static_mass", "<f4"), ("dynamic_mass", "<f4"), ("occ_mass", "<f4"),
    ("pers_mass", "<f4"), ("pred_mass", "<f4"),
])
assert CELL_DTYPE.itemsize == C.sizeof(Cell) == 68

# MapPoint (include/common/mappoint.h:5-28): x, y, z + padding, seven floats, 16-byte aligned
MAPPOINT_DTYPE = np.dtype([(n, np.float32) for n in (
    "x", "y", "z", "w", "v_x", "v_y", "v_z", "occ_val", "dyn_val", "eval_aug", "dyn_aug", "pad")])
assert MAPPOINT_DTYPE.itemsize == 48


def params_default():
    p = Params()
    check(_lib.load().k3dom_params_default(C.byref(p)), None, "k3dom_params_default")
    return p


def params_from_yaml(method_yaml=None, dataset_yaml=None):
    """YAML -> Params with the ROS node's defaults and unit conversion
    (src/k3dom/k3dom_node.cpp:215-368)."""
    p = Params()
    err = C.create_string_buffer(512)
    st = _lib.load().k3dom_params_from_yaml(
        method_yaml.encode() if method_yaml else None, dataset_yaml.encode() if dataset_yaml else None,
        C.byref(p), err, 512)
    if st != 0:
        raise _lib.K3domError(f"k3dom_params_from_yaml: {err.value.decode()}")
    return p


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# Mass-update methods (include/k3dom.h: K3DOM_METHOD_*).
METHODS = {"k3dom": 0, "dsphdmib": 1}


class DOM:
    """One map on the current HIP device."""

    def __init__(self, params, _handle=None, _owner=None, method="k3dom", _restored=None):
        """method: "k3dom" (kernel inference) or "dsphdmib" (the reference's
        second filter, a measurement grid of occupied / free masses)."""
        self._lib = _lib.load()
        self.params = params
        self.method = method
        if _restored is not None:  # a handle created from a checkpoint
            h = _restored
            self._owned = True
        elif _handle is None:
            h = C.c_void_p()
            check(self._lib.k3dom_create_method(C.byref(params), METHODS[method], C.byref(h)), None,
                  "k3dom_create_method")
            self._owned = True
        else:  # a map owned by a batch (kept alive through _owner)
            h = _handle
            self._owned = False
            self._owner = _owner
        self._h = h
        info = self.info()
        self.grid_size = info.grid_size
        self.grid_size_z = info.grid_size_z
        self.grid_cell_count = info.grid_cell_count
        self.particle_count = info.particle_count
        self.new_born_particle_count = info.new_born_particle_count

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                self._lib.k3dom_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- reference API
    def updatePose(self, x, y, z):
        check(self._lib.k3dom_update_pose(self._h, float(x), float(y), float(z)), self._h, "updatePose")

    def updateGrid(self, t, xs, ys, zs):
        xs = np.ascontiguousarray(xs, dtype=np.float32)
        ys = np.ascontiguousarray(ys, dtype=np.float32)
        zs = np.ascontiguousarray(zs, dtype=np.float32)
        if not (len(xs) == len(ys) == len(zs)):
            raise ValueError("point arrays differ in length")
        check(self._lib.k3dom_update_grid(self._h, float(t), _fptr(xs), _fptr(ys), _fptr(zs), len(xs)),
              self._h, "updateGrid")

    def updateGridLidars(self, t, lidars, on_device=False):
        """One frame from several LiDARs: lidars = [(rot 3x3, origin 3, xs,
        ys, zs), ...] with points in each LiDAR's frame (numpy arrays, or
        device pointers with their count when on_device: (rot, origin, dx,
        dy, dz, n)) (k3dom_update_grid_lidars)."""
        K = len(lidars)
        arr = (_lib.Lidar * K)()
        keep = []
        for k, l in enumerate(lidars):
            rot = np.asarray(l[0], np.float32).reshape(9)
            org = np.asarray(l[1], np.float32).reshape(3)
            arr[k].rot[:] = [float(v) for v in rot]
            arr[k].origin[:] = [float(v) for v in org]
            if on_device:
                arr[k].xs, arr[k].ys, arr[k].zs, arr[k].n = l[2], l[3], l[4], int(l[5])
            else:
                pts = [np.ascontiguousarray(a, np.float32) for a in l[2:5]]
                keep.append(pts)
                arr[k].xs, arr[k].ys, arr[k].zs = (a.ctypes.data for a in pts)
                arr[k].n = len(pts[0])
        check(self._lib.k3dom_update_grid_lidars(self._h, float(t), arr, K, 1 if on_device else 0), self._h,
              "update_grid_lidars")

    def updateGridDevice(self, t, d_xs, d_ys, d_zs, n):
        """Points already in device memory (integer device addresses)."""
        check(self._lib.k3dom_update_grid_device(self._h, float(t), C.c_void_p(d_xs), C.c_void_p(d_ys),
                                                 C.c_void_p(d_zs), int(n)), self._h, "updateGridDevice")

    # ---------------------------------------------------------------- stage methods
    # include/common/dom.h:89-94, include/k3dom/dom.h:24-28, in the order of
    # DOM::updateGrid (src/k3dom/dom.cu:33-44); see k3dom.h.
    def particlePrediction(self, dt):
        check(self._lib.k3dom_particle_prediction(self._h, float(dt)), self._h, "particlePrediction")

    def particleAssignment(self):
        check(self._lib.k3dom_particle_assignment(self._h), self._h, "particleAssignment")

    def gridCellOccupancyUpdate(self, dt, xs, ys, zs):
        xs, ys, zs = (np.ascontiguousarray(a, dtype=np.float32) for a in (xs, ys, zs))
        if not (len(xs) == len(ys) == len(zs)):
            raise ValueError("point arrays differ in length")
        check(self._lib.k3dom_grid_cell_occupancy_update(self._h, float(dt), _fptr(xs), _fptr(ys), _fptr(zs),
                                                          len(xs)), self._h, "gridCellOccupancyUpdate")

    def updatePersistentParticles(self):
        check(self._lib.k3dom_update_persistent_particles(self._h), self._h, "updatePersistentParticles")

    def initializeNewParticles(self):
        check(self._lib.k3dom_initialize_new_particles(self._h), self._h, "initializeNewParticles")

    def statisticalMoments(self):
        check(self._lib.k3dom_statistical_moments(self._h), self._h, "statisticalMoments")

    def resampling(self):
        check(self._lib.k3dom_resampling(self._h), self._h, "resampling")

    def initializeParticles(self, xs, ys, zs):
        xs, ys, zs = (np.ascontiguousarray(a, dtype=np.float32) for a in (xs, ys, zs))
        if not (len(xs) == len(ys) == len(zs)):
            raise ValueError("point arrays differ in length")
        check(self._lib.k3dom_initialize_particles(self._h, _fptr(xs), _fptr(ys), _fptr(zs), len(xs)), self._h,
              "initializeParticles")

    def info(self):
        out = Info()
        check(self._lib.k3dom_get_info(self._h, C.byref(out)), self._h, "get_info")
        return out

    def getGridSize(self):
        return self.grid_size

    def getResolution(self):
        return self.params.resolution

    def getCenterPosition(self):
        return tuple(self.info().center_pos)

    def getSensorPosition(self):
        return tuple(self.info().sensor_pos)

    def getIteration(self):
        return self.info().iteration

    def indexToCoord(self, i):
        """indexToCoordX/Y/Z (include/common/dom.h:70-81), FP32 as in the reference."""
        gs = np.int64(self.grid_size)
        i = np.asarray(i, dtype=np.int64)
        res = np.float32(self.params.resolution)
        c = self.info().center_pos
        f32 = np.float32
        x = (f32(i % gs) + f32(0.5)) * res - f32(self.params.size) / f32(2.0) + f32(c[0])
        y = (f32((i % (gs * gs)) // gs) + f32(0.5)) * res - f32(self.params.size) / f32(2.0) + f32(c[1])
        z = (f32(i // (gs * gs)) + f32(0.5)) * res - f32(self.params.size_z) / f32(2.0) + f32(c[2])
        return x, y, z

    def copyCellsToHost(self):
        """All voxels as a GridCell-compatible structured array (index order of the reference)."""
        out = np.empty(self.grid_cell_count, dtype=CELL_DTYPE)
        check(self._lib.k3dom_read_cells(self._h, out.ctypes.data_as(C.POINTER(Cell))), self._h, "read_cells")
        return out

    # ---------------------------------------------------------------- checkpoints
    def checkpoint(self):
        """The filter state as bytes (k3dom_checkpoint_save)."""
        n = C.c_int64(0)
        check(self._lib.k3dom_checkpoint_size(self._h, C.byref(n)), self._h, "checkpoint_size")
        buf = np.empty(n.value, np.uint8)
        w = C.c_int64(0)
        check(self._lib.k3dom_checkpoint_save(self._h, buf.ctypes.data_as(C.c_void_p), n.value, C.byref(w)),
              self._h, "checkpoint_save")
        return buf[:w.value].tobytes()

    @classmethod
    def from_checkpoint(cls, data):
        """A map restored from checkpoint bytes (k3dom_checkpoint_load): same
        parameters and method, continues bit for bit."""
        lib = _lib.load()
        buf = np.frombuffer(data, np.uint8)
        h = C.c_void_p()
        check(lib.k3dom_checkpoint_load(buf.ctypes.data_as(C.c_void_p), len(buf), C.byref(h)), None,
              "checkpoint_load")
        params = _lib.Params.from_buffer_copy(bytes(buf[128:128 + C.sizeof(_lib.Params)]))
        method = {v: k for k, v in METHODS.items()}[int(np.frombuffer(bytes(buf[20:24]), np.int32)[0])]
        return cls(params, method=method, _restored=h)

    def save(self, path):
        """Writes a checkpoint file."""
        data = self.checkpoint()
        tmp = path + ".tmp"
        with open(tmp, "wb") as fh:
            fh.write(data)
        os.replace(tmp, path)

    @classmethod
    def load(cls, path):
        """A map restored from a checkpoint file."""
        with open(path, "rb") as fh:
            return cls.from_checkpoint(fh.read())

    def export_mappoints(self, visible=False, min_occ=0.0, min_mass=0.0):
        """The node's evaluation records (MapPoint) built on the GPU: every
        voxel in index order, or (visible=True) only the voxels passing the
        visualisation filter with thresholds min_occ / min_mass, compacted in
        index order (k3dom_export_mappoints)."""
        mode = 1 if visible else 0
        n = C.c_int64(0)
        out = np.empty(self.grid_cell_count, MAPPOINT_DTYPE)  # at most one record per voxel
        check(self._lib.k3dom_export_mappoints(self._h, mode, min_occ, min_mass, out.ctypes.data_as(C.c_void_p),
                                               len(out), C.byref(n)), self._h, "export_mappoints")
        return out[:n.value]

    # ---------------------------------------------------------------- extras
    def read_masses(self):
        V = self.grid_cell_count
        a = [np.empty(V, np.float32) for _ in range(6)]
        check(self._lib.k3dom_read_masses(self._h, *[_fptr(x) for x in a]), self._h, "read_masses")
        return dict(zip(["free", "static", "dynamic", "mean_vx", "mean_vy", "mean_vz"], a))

    def read_array(self, name):
        which = ARR[name]
        N, B, V = self.particle_count, self.new_born_particle_count, self.grid_cell_count
        shape, dt = {
            "PARTICLE_POS": ((3, N), np.float32), "PARTICLE_VEL": ((3, N), np.float32),
            "PARTICLE_WEIGHT": ((N,), np.float32), "PARTICLE_KEY": ((N,), np.int32),
            "SORT_PERM": ((N,), np.int32), "SORTED_KEY": ((N,), np.int32), "RESAMPLE_IDX": ((N,), np.int32),
            "NEWBORN_POS": ((3, B), np.float32), "NEWBORN_VEL": ((3, B), np.float32),
            "NEWBORN_WEIGHT": ((B,), np.float32), "NEWBORN_KEY": ((B,), np.int32),
            "BORN_MASS": ((V,), np.float32), "DEL_OCC": ((V,), np.float32), "DEL_FREE": ((V,), np.float32),
            "SORTED_WEIGHT": ((N,), np.float32),
        }[name]
        out = np.empty(shape, dt)
        check(self._lib.k3dom_read_array(self._h, which, out.ctypes.data_as(C.c_void_p), out.nbytes), self._h,
              f"read_array({name})")
        return out

    def set_timing(self, on=True):
        """True (or 1): time every stage; 2: time only the gather kernel; False: off."""
        level = 0 if not on else (2 if on == 2 else 1)
        check(self._lib.k3dom_set_timing(self._h, level), self._h, "set_timing")

    def autotune(self):
        """Times the candidate-list build for several list segment lengths and
        the gather over the last frame's items for several block and grid
        sizes, keeps the fastest; returns the chosen blocks per compute unit."""
        c = C.c_int32()
        check(self._lib.k3dom_autotune(self._h, C.byref(c)), self._h, "autotune")
        return c.value

    def launch_shape(self):
        """The launch shapes in use: gather blocks per compute unit, wavefronts
        per gather block, and the list segment lengths (top level, finer levels)."""
        s = (C.c_int32 * 4)()
        check(self._lib.k3dom_launch_shape(self._h, s), self._h, "launch_shape")
        return {"blocks_per_cu": s[0], "gather_waves": s[1], "list_seg_top": s[2], "list_seg_fine": s[3]}

    def set_launch_shape(self, blocks_per_cu, gather_waves, list_seg_top, list_seg_fine):
        """Sets the launch shapes (between frames; results do not depend on them)."""
        s = (C.c_int32 * 4)(blocks_per_cu, gather_waves, list_seg_top, list_seg_fine)
        check(self._lib.k3dom_set_launch_shape(self._h, s), self._h, "set_launch_shape")

    def set_graph(self, on=True):
        """Replay each normal frame as one captured HIP graph (same results)."""
        check(self._lib.k3dom_set_graph(self._h, 1 if on else 0), self._h, "set_graph")

    STAGES = ["predict", "sort", "lists", "bgki", "mass", "persist", "newborn", "moments", "resample", "frame",
              "bgki_kernel"]

    def stage_times(self):
        out = np.zeros(len(self.STAGES), np.float32)
        check(self._lib.k3dom_stage_times(self._h, _fptr(out), len(out)), self._h, "stage_times")
        return out

    def stats(self):
        out = _lib.Stats()
        check(self._lib.k3dom_get_stats(self._h, C.byref(out)), self._h, "get_stats")
        return {"points": out.points, "level_entries": list(out.level_entries), "gather_items": out.gather_items,
                "gather_split_items": out.gather_split_items, "capacity_regrowths": out.capacity_regrowths}

    def count_bgki_pairs(self):
        out = (C.c_uint64 * 2)()
        check(self._lib.k3dom_count_bgki_pairs(self._h, out), self._h, "count_bgki_pairs")
        return int(out[0]), int(out[1])
