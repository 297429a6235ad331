"""Bit-exact parity of the HIP update against the serial restatement over
parameter sets that steer the update through different code paths:

- a grid whose edges are not multiples of the 4-voxel tile (partial tiles on
  every axis, lanes outside the grid);
- a length scale below the voxel size (candidate lists of a few points);
- a length scale of half the map (every tile sees nearly every ray: long lists
  split into 1024-point chunks and recombined in chunk order);
- a sensor offset from the pose, another kernel amplitude and mass scale;
- particle counts that fill no block evenly;
- the DS-PHD/MIB method on the odd grid with a sensor offset;
- an initialisation frame with no points, for both methods.

Every GridCell field, particle array and per-frame array is compared after
each frame (tests/_parity.py)."""
import pytest

from tests._parity import run_pair

pytestmark = pytest.mark.gpu

VARIANTS = {
    "odd_grid": dict(edge=13, edge_z=7, frames=5, seed=21, ego_vx=3.0),
    "ls_below_voxel": dict(frames=5, seed=22, ls=0.15),
    "ls_half_map": dict(frames=4, seed=23, ls=1.6, beams=16, az=256),
    "sensor_offset_sigma_scale": dict(frames=5, seed=24, ego_vx=2.0, sensor_off_x=0.3, sensor_off_y=-0.2,
                                      sensor_off_z=0.45, sigma=0.35, mass_scale=1.7),
    "uneven_particle_counts": dict(frames=5, seed=25, particles=3001, newborn=17),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_parameter_variant_bit_exact(gpu, name):
    from k3dom_amd.scene import small_workload
    w = small_workload(**VARIANTS[name])
    print(f"  variant {name}: {w.name}", flush=True)
    gd, _ = run_pair(w, w.frames)
    if name == "ls_half_map":
        assert gd.stats()["gather_split_items"] > 0, "expected tile lists longer than one chunk"


@pytest.mark.parametrize("method", ["k3dom", "dsphdmib"])
def test_first_frame_without_points_bit_exact(gpu, method):
    """The initialisation frame sees no points (particles placed from the
    priors alone); the frames after it are normal."""
    from k3dom_amd.scene import small_workload

    def mutate(f, t, pose, xs, ys, zs):
        return (t, pose, xs[:0], ys[:0], zs[:0]) if f == 0 else (t, pose, xs, ys, zs)
    w = small_workload(frames=4, seed=27, ego_vx=2.0, method=method)
    run_pair(w, w.frames, mutate=mutate, method=method)


def test_dsphdmib_odd_grid_sensor_offset_bit_exact(gpu):
    from k3dom_amd.scene import small_workload
    w = small_workload(edge=13, edge_z=7, frames=5, seed=26, ego_vx=3.0, method="dsphdmib", sensor_off_x=0.25,
                       sensor_off_z=0.4)
    print(f"  variant dsphdmib odd grid: {w.name}", flush=True)
    run_pair(w, w.frames, method="dsphdmib")
