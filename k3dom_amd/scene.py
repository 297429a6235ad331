"""Synthetic workloads (seeded box scenes + ray-cast LiDAR) and the filter
parameters of each benchmark configuration.

Filter parameters follow the K3DOM method configuration of the reference
(config/k3dom.yaml:22-44: sigma 0.1, ls 0.5 m, mass_scale 2.5, priors 0.001,
gamma 0.99, p_S 0.99, p_B 0.02, min velocity 0.1 m/s, initial max velocity
5 m/s, process noise 0.01 m / 0.1 m/s, map shift threshold 1 m) with the unit
conversion of the node (src/k3dom/k3dom_node.cpp:344-368): velocities and
stddevs in voxels, map_shift_thresh in voxels.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import Params

HERE = os.path.dirname(os.path.abspath(__file__))
SCENE_LIB = os.path.join(HERE, "libk3scene.so")


class SceneDesc(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("n_static", C.c_int32), ("n_moving", C.c_int32),
        ("side_min", C.c_float), ("side_max", C.c_float), ("speed_min", C.c_float), ("speed_max", C.c_float),
        ("xmin", C.c_float), ("xmax", C.c_float), ("ymin", C.c_float), ("ymax", C.c_float),
        ("ground_z", C.c_float), ("clearance", C.c_float),
        ("ego_x0", C.c_float), ("ego_y0", C.c_float), ("ego_z0", C.c_float), ("ego_vx", C.c_float),
        ("ego_vy", C.c_float), ("beams", C.c_int32), ("az_bins", C.c_int32),
        ("elev_min_deg", C.c_float), ("elev_max_deg", C.c_float), ("max_range", C.c_float),
        ("range_noise", C.c_float), ("frame_dt", C.c_float), ("t0", C.c_double),
        ("ground", C.c_int32), ("walls", C.c_float),
    ]


_slib = None


def _scene_lib():
    global _slib
    if _slib is None:
        if not os.path.exists(SCENE_LIB):
            raise RuntimeError(f"{SCENE_LIB} is missing: build it with `python -m k3dom_amd.build scene`")
        lib = C.CDLL(SCENE_LIB)
        lib.k3scene_create.restype = C.c_void_p
        lib.k3scene_create.argtypes = [C.POINTER(SceneDesc)]
        lib.k3scene_destroy.argtypes = [C.c_void_p]
        lib.k3scene_frame.restype = C.c_int
        lib.k3scene_frame.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.POINTER(C.c_double)]
        lib.k3scene_frame_lidar.restype = C.c_int
        lib.k3scene_frame_lidar.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_double, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        lib.k3scene_occupancy.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                          C.c_float, C.c_float, C.c_float, C.c_void_p]
        _slib = lib
    return _slib


class Scene:
    def __init__(self, desc):
        self.desc = desc
        self._lib = _scene_lib()
        self._h = self._lib.k3scene_create(C.byref(desc))
        self.cap = desc.beams * desc.az_bins

    def frame(self, f):
        """Returns (t, sensor_xyz, xs, ys, zs) of frame f."""
        xs = np.empty(self.cap, np.float32)
        ys = np.empty(self.cap, np.float32)
        zs = np.empty(self.cap, np.float32)
        pose = (C.c_double * 4)()
        n = self._lib.k3scene_frame(self._h, f, xs.ctypes.data, ys.ctypes.data, zs.ctypes.data, self.cap, pose)
        return pose[3], (pose[0], pose[1], pose[2]), xs[:n].copy(), ys[:n].copy(), zs[:n].copy()

    def frame_lidars(self, f, rig):
        """Frame f seen by a rig of LiDARs, rig = [(mount xyz from the ego in
        metres, yaw in radians), ...]: returns (t, ego position, lidars) with
        lidars = [(rot 3x3 float32 LiDAR-to-map axes, origin float32 = mount,
        xs, ys, zs in the LiDAR's frame), ...]."""
        out = []
        pose = (C.c_double * 4)()
        for mount, yaw in rig:
            xs = np.empty(self.cap, np.float32)
            ys = np.empty(self.cap, np.float32)
            zs = np.empty(self.cap, np.float32)
            m = (C.c_double * 3)(*mount)
            n = self._lib.k3scene_frame_lidar(self._h, f, m, float(yaw), xs.ctypes.data, ys.ctypes.data,
                                              zs.ctypes.data, self.cap, pose)
            c, sn = np.cos(yaw), np.sin(yaw)
            rot = np.array([[c, -sn, 0.0], [sn, c, 0.0], [0.0, 0.0, 1.0]], np.float32)
            out.append((rot, np.asarray(mount, np.float32), xs[:n].copy(), ys[:n].copy(), zs[:n].copy()))
        return pose[3], (pose[0], pose[1], pose[2]), out

    def occupancy(self, f, gs, gsz, res, center, size, size_z):
        out = np.empty(gs * gs * gsz, np.uint8)
        self._lib.k3scene_occupancy(self._h, f, gs, gsz, res, center[0], center[1], center[2], size, size_z,
                                    out.ctypes.data)
        return out

    def __del__(self):
        try:
            if self._h:
                self._lib.k3scene_destroy(self._h)
                self._h = None
        except Exception:
            pass


def _f32(x):
    return float(np.float32(x))


def grid_size_param(edge, res):
    """A `size` whose FP32 truncation size/res gives exactly `edge` voxels."""
    r = np.float32(res)
    for s in (edge * res, (edge + 0.5) * res):
        if int(np.float32(s) / r) == edge:
            return float(np.float32(s))
    raise ValueError(f"no size for edge {edge} at {res}")


def k3dom_params(res, edge_xy, edge_z, particles, newborn, seed=123456, **over):
    """Filter parameters of the K3DOM method configuration at resolution res."""
    r = np.float32(res)
    p = Params()
    p.size = grid_size_param(edge_xy, res)
    p.size_z = grid_size_param(edge_z, res)
    p.mass_scale = 2.5
    p.resolution = res
    p.particle_count = particles
    p.new_born_particle_count = newborn
    p.persistence_prob = 0.99
    p.stddev_process_noise_position = _f32(np.float32(0.01) / r)
    p.stddev_process_noise_velocity = _f32(np.float32(0.1) / r)
    p.birth_prob = 0.02
    p.stddev_velocity = _f32(np.float32(4.0) / r)
    p.init_max_velocity = _f32(np.float32(5.0) / r)
    p.particle_min_vel = _f32(np.float32(0.1) / r)
    p.sensor_off_x = p.sensor_off_y = p.sensor_off_z = 0.0
    p.map_shift_thresh = int(np.float32(1.0) / r)
    p.prior_free = p.prior_static = p.prior_dynamic = 0.001
    p.prior_occ = 0.0
    p.sigma = 0.1
    p.ls = 0.5
    p.gamma = 0.99
    p.max_range = 0.0
    p.meas_occ_stddev = 0.0
    p.seed = seed
    for k, v in over.items():
        setattr(p, k, v)
    return p


@dataclass
class Workload:
    name: str
    params: Params
    scene: SceneDesc
    frames: int
    extra: dict = field(default_factory=dict)


def _desc(**kw):
    d = SceneDesc()
    defaults = dict(side_min=0.5, side_max=3.0, speed_min=0.0, clearance=2.0, ego_x0=0.0, ego_y0=0.0,
                    ego_z0=0.0, ego_vx=0.0, ego_vy=0.0, elev_min_deg=-15.0, elev_max_deg=15.0, frame_dt=0.1,
                    t0=1.0, ground=1)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(d, k, v)
    return d


def workload(cfg):
    """The five benchmark configurations (1-5)."""
    if cfg == 1:
        # 64x64x16 @0.2 m, 65,536 particles (+10% newborn), 16x1024 LiDAR, static ego
        p = k3dom_params(0.2, 64, 16, 65536, 6553)
        s = _desc(seed=1, n_static=20, n_moving=4, speed_max=2.0, xmin=-12, xmax=12, ymin=-12, ymax=12,
                  ground_z=-1.6, beams=16, az_bins=1024, max_range=20.0, range_noise=0.02)
        return Workload("plain-small 64x64x16@0.2m N=65536", p, s, 20)
    if cfg == 2:
        # 256x256x32 @0.1 m, 2,000,000 particles + 200,000 newborn, 64x2048 LiDAR, ego 10 m/s
        p = k3dom_params(0.1, 256, 32, 2_000_000, 200_000)
        s = _desc(seed=2, n_static=300, n_moving=50, speed_max=5.0, xmin=-40, xmax=140, ymin=-40, ymax=40,
                  ground_z=-1.6, ego_vx=10.0, beams=64, az_bins=2048, max_range=40.0, range_noise=0.03)
        return Workload("plain-full 256x256x32@0.1m N=2000000", p, s, 100)
    if cfg == 3:
        p = k3dom_params(0.2, 64, 16, 65536, 6553, ls=0.3)
        s = workload(1).scene
        return Workload("differentiated config-1 scene, ls0=0.3", p, s, 8)
    if cfg == 5:
        # 512x512x64 @0.1 m, 8,000,000 particles (+10% newborn), four 128x2048 LiDARs
        # at the vehicle's corners at 90-degree yaw spacing, ego 15 m/s
        p = k3dom_params(0.1, 512, 64, 8_000_000, 800_000)
        s = _desc(seed=5, n_static=1000, n_moving=200, speed_max=15.0, xmin=-60, xmax=300, ymin=-60, ymax=60,
                  ground_z=-3.2, ego_vx=15.0, beams=128, az_bins=2048, max_range=40.0, range_noise=0.03)
        return Workload("stress 512x512x64@0.1m N=8000000, 4 LiDARs", p, s, 50, extra={"rig": CFG5_RIG})
    raise ValueError(f"unknown configuration {cfg}")


def with_walls(w, half=27.0):
    """Workload w with enclosing walls: a square of half-size `half` metres
    around the sensor, unbounded in z, that every ray hits if nothing nearer
    does (at cfg2, 27 m keeps the farthest wall return of the +-15 degree fan
    within the 40 m range, so all 64x2048 rays return)."""
    d = SceneDesc.from_buffer_copy(bytes(w.scene))
    d.walls = half
    return Workload(f"{w.name}, enclosing walls at {half:g} m", w.params, d, w.frames, dict(w.extra))


# four LiDARs at the vehicle's corners, 90 degrees apart in yaw (mount from the ego, yaw)
CFG5_RIG = [((1.0, 0.0, 0.2), 0.0), ((0.0, 0.8, 0.2), np.pi / 2), ((-1.0, 0.0, 0.2), np.pi),
            ((0.0, -0.8, 0.2), 1.5 * np.pi)]


def small_rig(spread=0.5, n=3):
    """n LiDARs on a circle of radius `spread` around the ego, yaws 360/n apart."""
    return [((spread * np.cos(2 * np.pi * k / n), spread * np.sin(2 * np.pi * k / n), 0.1 * k), 2 * np.pi * k / n)
            for k in range(n)]


def batch_workloads(n_tiles=32, seed=4, edge_min=32, edge_max=128, particles_min=16_384, particles_max=1_048_576,
                    res=0.1, beams=32, az=1024, frames=50):
    """The batched-tiles configuration: n_tiles independent cubic maps with
    edges log-uniform in [edge_min, edge_max] voxels at res, particle counts
    log-uniform in [particles_min, particles_max] (+10% newborn), each with its
    own box scene around a static ego and a beams x az LiDAR."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_tiles):
        edge = int(np.exp(rng.uniform(np.log(edge_min), np.log(edge_max + 1))))
        edge = max(edge_min, min(edge_max, edge))
        n = int(np.exp(rng.uniform(np.log(particles_min), np.log(particles_max))))
        p = k3dom_params(res, edge, edge, n, max(1, n // 10), seed=123456 + i)
        half = edge * res / 2
        s = _desc(seed=1000 * seed + i, n_static=max(4, int(edge * edge * res * res / 4)), n_moving=max(1, edge // 32),
                  speed_max=2.0, xmin=-half, xmax=half, ymin=-half, ymax=half, ground_z=-half * 0.3,
                  clearance=0.5, beams=beams, az_bins=az, max_range=1.8 * half, range_noise=0.03)
        out.append(Workload(f"tile {i} {edge}^3@{res}m N={n}", p, s, frames))
    return out


def dsphdmib_params(res, edge_xy, edge_z, particles, newborn, seed=123456, **over):
    """Filter parameters of the DS-PHD/MIB method configuration at resolution
    res (config/dsphdmib.yaml: meas_occ_stddev 0.1 m, particle_min_vel 0, zero
    priors; max_range 50 m as src/dsphdmib/dsphdmib_node.cpp:346)."""
    over = dict(dict(particle_min_vel=0.0, prior_free=0.0, prior_static=0.0, prior_dynamic=0.0, prior_occ=0.0,
                     meas_occ_stddev=0.1, max_range=50.0), **over)
    return k3dom_params(res, edge_xy, edge_z, particles, newborn, seed=seed, **over)


def small_workload(edge=16, edge_z=8, res=0.2, particles=4096, newborn=409, beams=8, az=128, frames=6, seed=11,
                   moving=2, static=6, ego_vx=0.0, method="k3dom", **over):
    """A tiny scene for fast parity tests (method "dsphdmib": that method's parameters)."""
    make = dsphdmib_params if method == "dsphdmib" else k3dom_params
    p = make(res, edge, edge_z, particles, newborn, **over)
    half = edge * res / 2
    s = _desc(seed=seed, n_static=static, n_moving=moving, speed_max=2.0, xmin=-2 * half, xmax=2 * half,
              ymin=-2 * half, ymax=2 * half, ground_z=-edge_z * res / 2, clearance=0.6, ego_vx=ego_vx,
              beams=beams, az_bins=az, max_range=3 * half, range_noise=0.02)
    return Workload(f"small {edge}x{edge}x{edge_z}@{res} N={particles}", p, s, frames)
