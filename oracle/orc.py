"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the serial CPU restatement
(oracle/oracle.cpp -> oracle/liborc.so).  Imported by tests/, by the smoke
check and by bench.py's CPU-baseline leg, never by the product."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")

_orc = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(LIB):
            build()
        from k3dom_amd._lib import Params, Cell, Info  # struct layouts shared with the public header
        L = C.CDLL(LIB)
        fp, u32p = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
        sig = {
            "orc_create": (C.c_void_p, [C.POINTER(Params)]),
            "orc_create_method": (C.c_void_p, [C.POINTER(Params), C.c_int]),
            "orc_ds_pocc": (C.c_float, [C.c_float] * 7),
            "orc_ds_pfree": (C.c_float, [C.c_float] * 4),
            "orc_ds_update_masses": (None, [C.c_float] * 4 + [C.POINTER(C.c_float)]),
            "orc_ds_newborn_part": (C.c_float, [C.c_float] * 3),
            "orc_destroy": (None, [C.c_void_p]),
            "orc_set_brute": (None, [C.c_void_p, C.c_int]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_update_pose": (None, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
            "orc_update_grid": (C.c_int, [C.c_void_p, C.c_double, fp, fp, fp, C.c_int]),
            "orc_update_grid_lidars": (C.c_int, [C.c_void_p, C.c_double, C.c_int, fp, fp, C.POINTER(C.c_int),
                                                 fp, fp, fp]),
            "orc_get_info": (None, [C.c_void_p, C.POINTER(Info)]),
            "orc_read_cells": (None, [C.c_void_p, C.POINTER(Cell)]),
            "orc_read_array": (C.c_int64, [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]),
            "orc_stage_times": (None, [C.c_void_p, C.POINTER(C.c_double)]),
            "orc_bgki_pairs": (None, [C.c_void_p, C.POINTER(C.c_uint64)]),
            "orc_philox": (None, [u32p, u32p, u32p]),
            "orc_ln": (C.c_float, [C.c_float]),
            "orc_exp": (C.c_float, [C.c_float]),
            "orc_sincospi": (None, [C.c_float, fp, fp]),
            "orc_kernel": (C.c_float, [C.c_float, C.c_float, C.c_float]),
            "orc_seg_sum": (C.c_float, [C.POINTER(C.c_float), C.c_int32, C.c_int32]),
            "orc_read_counts": (C.c_int64, [C.c_void_p, C.POINTER(C.c_int32), C.c_int64]),
            "orc_window_chain64": (C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_int32), fp, fp,
                                             C.POINTER(C.c_int32), fp, C.POINTER(C.c_int32),
                                             C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_window_chain64_groups": (C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_int32), fp, fp,
                                                    C.POINTER(C.c_int32), fp, C.POINTER(C.c_int32),
                                                    C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                                                    C.POINTER(C.c_int32), C.POINTER(C.c_int32), fp,
                                                    C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                    C.POINTER(C.c_double)]),
            "orc_credit": (C.c_float, [C.c_float, C.c_float]),
            "orc_kernel_dls": (C.c_float, [C.c_float, C.c_float, C.c_float]),
            "orc_kernel_dls_sq": (C.c_float, [C.c_float, C.c_float, C.c_float]),
            "orc_dual_credit": (None, [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_box_muller": (None, [C.c_uint32, C.c_uint32, fp, fp]),
            "orc_update_mass": (None, [fp, C.c_float, C.c_float, C.c_float]),
            "orc_predict_mass": (None, [fp, C.c_float, C.c_float, C.c_float]),
            "orc_newborn_part": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_float]),
            "orc_key": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_int, C.c_int]),
            "orc_predict": (None, [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                   C.c_int, C.c_uint32, C.c_uint64] + [fp] * 7 + [C.POINTER(C.c_int32)]),
            "orc_sample_velocity": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float, fp]),
            "orc_moments": (None, [C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)] + [fp] * 8),
            "orc_set_cells": (None, [C.c_void_p, C.c_void_p]),
            "orc_set_particle_pos": (None, [C.c_void_p, fp]),
            "orc_set_keep64": (None, [C.c_void_p, C.c_int]),
            "orc_read64": (C.c_int64, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int64]),
            "orc_bgki": (None, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                fp, fp, fp, C.c_int, C.c_int, fp, fp, C.POINTER(C.c_uint64)]),
            "orc_bgki64": (None, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                  C.c_float, fp, fp, fp, C.c_int, C.POINTER(C.c_int32), C.c_int,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_mass_update64": (None, [fp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.POINTER(C.c_double)]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _orc = L
    return _orc


def set_threads(n):
    """Threads of the culled kernel inference in every oracle map (1 = serial;
    the results are the same bits for any count)."""
    lib().orc_set_threads(int(n))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class OracleDOM:
    """Serial restatement with the same surface as k3dom_amd.dom.DOM."""

    def __init__(self, params, brute=False, method="k3dom"):
        from k3dom_amd._lib import Info
        from k3dom_amd.dom import METHODS
        self._L = lib()
        self.params = params
        self._h = self._L.orc_create_method(C.byref(params), METHODS[method])
        if not self._h:
            raise ValueError("invalid parameters")
        if brute:
            self._L.orc_set_brute(self._h, 1)
        info = Info()
        self._L.orc_get_info(self._h, C.byref(info))
        self.grid_size, self.grid_size_z = info.grid_size, info.grid_size_z
        self.grid_cell_count = info.grid_cell_count
        self.particle_count = info.particle_count
        self.new_born_particle_count = info.new_born_particle_count

    def __del__(self):
        try:
            if self._h:
                self._L.orc_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def updatePose(self, x, y, z):
        self._L.orc_update_pose(self._h, float(x), float(y), float(z))

    def updateGrid(self, t, xs, ys, zs):
        xs = np.ascontiguousarray(xs, np.float32)
        ys = np.ascontiguousarray(ys, np.float32)
        zs = np.ascontiguousarray(zs, np.float32)
        st = self._L.orc_update_grid(self._h, float(t), _fp(xs), _fp(ys), _fp(zs), len(xs))
        if st != 0:
            raise ValueError(f"orc_update_grid status {st}")

    def updateGridLidars(self, t, lidars):
        """lidars: sequence of (rot 3x3, origin 3, xs, ys, zs), points in each LiDAR's frame."""
        K = len(lidars)
        rot = np.ascontiguousarray(np.stack([np.asarray(l[0], np.float32).reshape(9) for l in lidars]), np.float32)
        org = np.ascontiguousarray(np.stack([np.asarray(l[1], np.float32).reshape(3) for l in lidars]), np.float32)
        counts = np.array([len(l[2]) for l in lidars], np.int32)
        xs, ys, zs = (np.ascontiguousarray(np.concatenate([np.asarray(l[k], np.float32) for l in lidars]), np.float32)
                      for k in (2, 3, 4))
        st = self._L.orc_update_grid_lidars(self._h, float(t), K, _fp(rot), _fp(org),
                                            counts.ctypes.data_as(C.POINTER(C.c_int)), _fp(xs), _fp(ys), _fp(zs))
        if st != 0:
            raise ValueError(f"orc_update_grid_lidars status {st}")

    def info(self):
        from k3dom_amd._lib import Info
        info = Info()
        self._L.orc_get_info(self._h, C.byref(info))
        return info

    def copyCellsToHost(self):
        from k3dom_amd.dom import CELL_DTYPE
        from k3dom_amd._lib import Cell
        out = np.empty(self.grid_cell_count, dtype=CELL_DTYPE)
        self._L.orc_read_cells(self._h, out.ctypes.data_as(C.POINTER(Cell)))
        return out

    def read_array(self, name):
        from k3dom_amd._lib import ARR
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
        n = self._L.orc_read_array(self._h, ARR[name], out.ctypes.data_as(C.c_void_p), out.nbytes)
        if n < 0:
            raise ValueError(f"read_array({name}) failed")
        return out

    def stage_times(self):
        out = (C.c_double * 10)()
        self._L.orc_stage_times(self._h, out)
        return np.array(out[:])

    def counts(self):
        """Frozen counts of the last frame: (4, V) int32 (inside the grid, of them
        not penalised; newborns, of them not penalised)."""
        out = np.empty((4, self.grid_cell_count), np.int32)
        n = self._L.orc_read_counts(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), out.nbytes)
        if n < 0:
            raise ValueError("counts not available")
        return out

    def bgki_pairs(self):
        out = (C.c_uint64 * 2)()
        self._L.orc_bgki_pairs(self._h, out)
        return int(out[0]), int(out[1])

    def set_cells(self, cells):
        """Overwrites every cell with a GridCell-compatible structured array
        (hand-built states)."""
        from k3dom_amd.dom import CELL_DTYPE
        a = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        assert a.shape == (self.grid_cell_count,)
        self._L.orc_set_cells(self._h, a.ctypes.data_as(C.c_void_p))

    SHADOW64 = ["pre_f", "pre_s", "pre_d", "dyn", "f", "s", "d", "born", "mvx", "mvy", "mvz"]

    def keep64(self, on=True):
        """Keep an FP64 shadow of the floating-point stages of later frames."""
        self._L.orc_set_keep64(self._h, 1 if on else 0)

    def read64(self, name):
        """FP64 shadow array of the last normal frame (SHADOW64 names)."""
        out = np.empty(self.grid_cell_count, np.float64)
        n = self._L.orc_read64(self._h, self.SHADOW64.index(name), out.ctypes.data_as(C.POINTER(C.c_double)),
                               out.nbytes)
        if n < 0:
            raise ValueError(f"read64({name}): not kept")
        return out

    def set_particle_pos(self, pos):
        """Overwrites the particle positions with a (3, N) float32 array."""
        a = np.ascontiguousarray(pos, dtype=np.float32)
        assert a.shape == (3, self.particle_count)
        self._L.orc_set_particle_pos(self._h, _fp(a))


def bgki(gs, gsz, res, sensor_rel, ls, sigma, xs, ys, zs, brute=False):
    """(del_occ, del_free, (occ_pairs, free_pairs)) of one point set on a grid."""
    V = gs * gs * gsz
    o = np.empty(V, np.float32)
    f = np.empty(V, np.float32)
    pairs = (C.c_uint64 * 2)()
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    zs = np.ascontiguousarray(zs, np.float32)
    lib().orc_bgki(gs, gsz, res, sensor_rel[0], sensor_rel[1], sensor_rel[2], ls, sigma, _fp(xs), _fp(ys), _fp(zs),
                   len(xs), 1 if brute else 0, _fp(o), _fp(f), pairs)
    return o, f, (int(pairs[0]), int(pairs[1]))


def bgki64(gs, gsz, res, sensor_rel, ls, sigma, xs, ys, zs, cells):
    cells = np.ascontiguousarray(cells, np.int32)
    o = np.empty(len(cells), np.float64)
    f = np.empty(len(cells), np.float64)
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    zs = np.ascontiguousarray(zs, np.float32)
    lib().orc_bgki64(gs, gsz, res, sensor_rel[0], sensor_rel[1], sensor_rel[2], ls, sigma, _fp(xs), _fp(ys),
                     _fp(zs), len(xs), cells.ctypes.data_as(C.POINTER(C.c_int32)), len(cells),
                     o.ctypes.data_as(C.POINTER(C.c_double)), f.ctypes.data_as(C.POINTER(C.c_double)))
    return o, f


class WindowChain64:
    """The differentiated window's continuous chain in double precision, with
    the frozen decisions of a forward run: loss(theta) and, by forward-mode
    differentiation, grad(theta)."""

    def __init__(self, params, frames, sensor_rel, moves, decays, counts, gts, mask=None, groups=None):
        """groups (several LiDARs): (ngroups[T], gcount[T, 16], gsrel[T, 16, 3])
        with each group's points consecutive within its frame."""
        from k3dom_amd._lib import Params
        self.params = params if isinstance(params, Params) else Params.from_buffer_copy(bytes(params))
        self.T = len(frames)
        self.npts = np.array([len(f[0]) for f in frames], np.int32)
        self.pts = np.concatenate([np.concatenate([np.asarray(a, np.float32) for a in f]) for f in frames]).astype(
            np.float32) if self.T else np.zeros(1, np.float32)
        self.sensor_rel = np.ascontiguousarray(np.asarray(sensor_rel, np.float32).reshape(-1))
        self.moves = np.ascontiguousarray(np.asarray(moves, np.int32).reshape(-1))
        self.decays = np.ascontiguousarray(np.asarray(decays, np.float32))
        self.counts = np.ascontiguousarray(np.asarray(counts, np.int32).reshape(-1))
        self.gts = np.ascontiguousarray(np.asarray(gts, np.uint8).reshape(-1))
        self.mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        self.groups = None
        if groups is not None:
            ng, gc, gr = groups
            self.groups = (np.ascontiguousarray(ng, np.int32), np.ascontiguousarray(gc, np.int32).reshape(-1),
                           np.ascontiguousarray(gr, np.float32).reshape(-1))

    def _call(self, theta, grad):
        th = np.ascontiguousarray(theta, np.float64)
        loss = C.c_double()
        g = (C.c_double * 6)() if grad else None
        u8 = C.POINTER(C.c_uint8)
        i32 = C.POINTER(C.c_int32)
        gr = self.groups
        st = lib().orc_window_chain64_groups(
            C.byref(self.params), self.T, self.npts.ctypes.data_as(i32), _fp(self.pts),
            _fp(self.sensor_rel), self.moves.ctypes.data_as(i32), _fp(self.decays),
            self.counts.ctypes.data_as(i32), self.gts.ctypes.data_as(u8),
            None if self.mask is None else self.mask.ctypes.data_as(u8),
            None if gr is None else gr[0].ctypes.data_as(i32), None if gr is None else gr[1].ctypes.data_as(i32),
            None if gr is None else _fp(gr[2]),
            th.ctypes.data_as(C.POINTER(C.c_double)), C.byref(loss), g)
        if st != 0:
            raise ValueError(f"orc_window_chain64 status {st}")
        return (loss.value, np.array(list(g))) if grad else loss.value

    def loss(self, theta):
        return self._call(theta, False)

    def grad(self, theta):
        return self._call(theta, True)

    def central_differences(self, theta, rel_step=1e-4):
        theta = np.asarray(theta, np.float64)
        g = np.empty(6)
        for k in range(6):
            h = rel_step * abs(theta[k])
            tp, tm = theta.copy(), theta.copy()
            tp[k] += h
            tm[k] -= h
            g[k] = (self.loss(tp) - self.loss(tm)) / (tp[k] - tm[k])
        return g


def predict(px, py, pz, vx, vy, vz, w, dt, p_S, sigma_pos, sigma_vel, vclamp, gs, gsz, frame=0, seed=123456):
    """The frame's motion step on copies of the given particle arrays; returns
    (px, py, pz, vx, vy, vz, w, key)."""
    arrs = [np.array(a, dtype=np.float32, copy=True) for a in (px, py, pz, vx, vy, vz, w)]
    n = len(arrs[0])
    key = np.empty(n, np.int32)
    lib().orc_predict(n, dt, p_S, sigma_pos, sigma_vel, vclamp, gs, gsz, frame, seed, *[_fp(a) for a in arrs],
                      key.ctypes.data_as(C.POINTER(C.c_int32)))
    return (*arrs, key)


def sample_velocity(r0, r1, r2, vmax, min_vel):
    """The new-particle velocity of three Philox words."""
    out = np.empty(3, np.float32)
    lib().orc_sample_velocity(int(r0) & 0xFFFFFFFF, int(r1) & 0xFFFFFFFF, int(r2) & 0xFFFFFFFF, vmax, min_vel,
                              _fp(out))
    return out


def moments(start, end, w, vx, vy, vz, pers):
    """The frame's per-voxel mean velocities; returns (mvx, mvy, mvz)."""
    start = np.ascontiguousarray(start, np.int32)
    end = np.ascontiguousarray(end, np.int32)
    w, vx, vy, vz, pers = (np.ascontiguousarray(a, np.float32) for a in (w, vx, vy, vz, pers))
    V = len(start)
    out = [np.empty(V, np.float32) for _ in range(3)]
    i32 = C.POINTER(C.c_int32)
    lib().orc_moments(V, start.ctypes.data_as(i32), end.ctypes.data_as(i32), _fp(w), _fp(vx), _fp(vy), _fp(vz),
                      _fp(pers), *[_fp(o) for o in out])
    return tuple(out)


def mass_update64(fsd, dyn_pred, decay, del_occ, del_free, mass_scale, p_B):
    """One voxel's mass rules in double from FP32 masses (f, s, d) before
    prediction; returns (f, s, d, born) after the update."""
    fsd = np.ascontiguousarray(fsd, np.float32)
    out = (C.c_double * 4)()
    lib().orc_mass_update64(_fp(fsd), float(dyn_pred), float(decay), float(del_occ), float(del_free),
                            float(mass_scale), float(p_B), out)
    return tuple(out)
