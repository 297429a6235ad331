"""ctypes binding of the C ABI declared in include/k3dom.h.

The shared library is built in-tree (k3dom_amd/libk3dom.so).  Loading fails
loudly when it is missing: there is no CPU fallback for the update.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libk3dom.so")


class Params(C.Structure):
    """dom::DOM_c::Params (include/common/dom.h:19-52) plus the generator seed."""
    _fields_ = [
        ("size", C.c_float), ("size_z", C.c_float), ("mass_scale", C.c_float), ("resolution", C.c_float),
        ("particle_count", C.c_int32), ("new_born_particle_count", C.c_int32),
        ("persistence_prob", C.c_float), ("stddev_process_noise_position", C.c_float),
        ("stddev_process_noise_velocity", C.c_float), ("birth_prob", C.c_float),
        ("stddev_velocity", C.c_float), ("init_max_velocity", C.c_float), ("particle_min_vel", C.c_float),
        ("sensor_off_x", C.c_float), ("sensor_off_y", C.c_float), ("sensor_off_z", C.c_float),
        ("map_shift_thresh", C.c_int32),
        ("prior_free", C.c_float), ("prior_static", C.c_float), ("prior_dynamic", C.c_float),
        ("prior_occ", C.c_float), ("sigma", C.c_float), ("ls", C.c_float), ("gamma", C.c_float),
        ("max_range", C.c_float), ("meas_occ_stddev", C.c_float), ("seed", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Cell(C.Structure):
    """dom::GridCell (include/common/dom_types.h:12-45), 68 bytes."""
    _fields_ = [
        ("start_idx", C.c_int32), ("end_idx", C.c_int32), ("mu_A", C.c_float), ("mu_UA", C.c_float),
        ("w_A", C.c_float), ("w_UA", C.c_float), ("mean_x_vel", C.c_float), ("mean_y_vel", C.c_float),
        ("mean_z_vel", C.c_float), ("likelihood", C.c_float), ("p_A", C.c_float), ("free_mass", C.c_float),
        ("static_mass", C.c_float), ("dynamic_mass", C.c_float), ("occ_mass", C.c_float),
        ("pers_mass", C.c_float), ("pred_mass", C.c_float),
    ]


class Lidar(C.Structure):
    """k3dom_lidar (include/k3dom.h)."""
    _fields_ = [("rot", C.c_float * 9), ("origin", C.c_float * 3), ("xs", C.c_void_p), ("ys", C.c_void_p),
                ("zs", C.c_void_p), ("n", C.c_int32), ("pad", C.c_int32)]


class Info(C.Structure):
    _fields_ = [
        ("grid_size", C.c_int32), ("grid_size_z", C.c_int32), ("grid_cell_count", C.c_int32),
        ("particle_count", C.c_int32), ("new_born_particle_count", C.c_int32), ("iteration", C.c_int32),
        ("resolution", C.c_float), ("center_pos", C.c_float * 3), ("sensor_pos", C.c_float * 3),
        ("updated_time", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [("points", C.c_uint64), ("level_entries", C.c_uint64 * 3), ("gather_items", C.c_uint64),
                ("gather_split_items", C.c_uint64), ("capacity_regrowths", C.c_uint64)]


# numbering of k3dom_array
ARR = dict(PARTICLE_POS=0, PARTICLE_VEL=1, PARTICLE_WEIGHT=2, PARTICLE_KEY=3, SORT_PERM=4, SORTED_KEY=5,
           RESAMPLE_IDX=6, NEWBORN_POS=7, NEWBORN_VEL=8, NEWBORN_WEIGHT=9, NEWBORN_KEY=10, BORN_MASS=11,
           DEL_OCC=12, DEL_FREE=13, SORTED_WEIGHT=14)

# Every symbol the header declares, with (restype, argtypes).
_fp = C.POINTER(C.c_float)
SIGNATURES = {
    "k3dom_params_default": (C.c_int, [C.POINTER(Params)]),
    "k3dom_params_from_yaml": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(Params), C.c_char_p, C.c_int32]),
    "k3dom_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "k3dom_create_method": (C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_void_p)]),
    "k3dom_destroy": (C.c_int, [C.c_void_p]),
    "k3dom_update_pose": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    "k3dom_update_grid": (C.c_int, [C.c_void_p, C.c_double, _fp, _fp, _fp, C.c_int32]),
    "k3dom_update_grid_device": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_int32]),
    "k3dom_get_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "k3dom_read_cells": (C.c_int, [C.c_void_p, C.POINTER(Cell)]),
    "k3dom_read_masses": (C.c_int, [C.c_void_p, _fp, _fp, _fp, _fp, _fp, _fp]),
    "k3dom_read_array": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    "k3dom_update_grid_lidars": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(Lidar), C.c_int32, C.c_int32]),
    "k3dom_checkpoint_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "k3dom_checkpoint_save": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "k3dom_checkpoint_load": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]),
    "k3dom_export_mappoints": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_void_p, C.c_int64,
                                         C.POINTER(C.c_int64)]),
    "k3dom_export_mappoints_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_void_p,
                                                 C.c_int64, C.POINTER(C.c_int64)]),
    "k3dom_particle_prediction": (C.c_int, [C.c_void_p, C.c_float]),
    "k3dom_particle_assignment": (C.c_int, [C.c_void_p]),
    "k3dom_grid_cell_occupancy_update": (C.c_int, [C.c_void_p, C.c_float, _fp, _fp, _fp, C.c_int32]),
    "k3dom_update_persistent_particles": (C.c_int, [C.c_void_p]),
    "k3dom_initialize_new_particles": (C.c_int, [C.c_void_p]),
    "k3dom_statistical_moments": (C.c_int, [C.c_void_p]),
    "k3dom_resampling": (C.c_int, [C.c_void_p]),
    "k3dom_initialize_particles": (C.c_int, [C.c_void_p, _fp, _fp, _fp, C.c_int32]),
    "k3dom_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "k3dom_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "k3dom_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "k3dom_launch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "k3dom_set_launch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "k3dom_batch_create": (C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_void_p)]),
    "k3dom_batch_destroy": (C.c_int, [C.c_void_p]),
    "k3dom_batch_update": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_float), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_int32)]),
    "k3dom_batch_map": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "k3dom_batch_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "k3dom_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "k3dom_batch_last_ms": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "k3dom_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "k3dom_stage_times": (C.c_int, [C.c_void_p, _fp, C.c_int32]),
    "k3dom_count_bgki_pairs": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "k3dom_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "k3dom_last_error": (C.c_char_p, [C.c_void_p]),
    "k3dom_window_create": (C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_void_p)]),
    "k3dom_window_destroy": (C.c_int, [C.c_void_p]),
    "k3dom_window_set_frame": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, _fp, _fp, _fp, _fp, C.c_int32,
                                         C.POINTER(C.c_uint8)]),
    "k3dom_window_set_mask": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8)]),
    "k3dom_window_set_frame_lidars": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, _fp, C.POINTER(Lidar), C.c_int32,
                                                C.POINTER(C.c_uint8)]),
    "k3dom_window_forward": (C.c_int, [C.c_void_p, _fp, C.POINTER(C.c_double)]),
    "k3dom_window_forward_id": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "k3dom_window_backward": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_double), _fp]),
    "k3dom_window_read_record": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int64]),
    "k3dom_window_times": (C.c_int, [C.c_void_p, _fp]),
    "k3dom_window_last_error": (C.c_char_p, [C.c_void_p]),
}

_lib = None


def load(path=None):
    """Loads libk3dom.so (raises if it is missing or lacks a declared symbol)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m k3dom_amd.build lib` "
                           "(the update has no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class K3domError(RuntimeError):
    pass


def check(status, handle=None, what=""):
    if status != 0:
        lib = load()
        msg = lib.k3dom_last_error(handle)
        err = K3domError(f"{what} failed with status {status}: {msg.decode() if msg else ''}")
        err.status = status
        raise err
