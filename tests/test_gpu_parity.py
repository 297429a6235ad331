"""Bit-exact parity of the HIP update against the serial restatement.

Every GridCell field, every particle array, the sort permutation, the sorted
keys, the newborn set, the born masses, the kernel-inference evidence and the
resampled indices must be identical bit for bit after each frame."""
import numpy as np
import pytest

from tests._parity import run_pair, compare_state

pytestmark = pytest.mark.gpu


def test_small_scene_bit_exact(gpu):
    from k3dom_amd.scene import small_workload
    w = small_workload()
    run_pair(w, w.frames)


def test_small_scene_moving_ego_map_shift(gpu):
    from k3dom_amd.scene import small_workload
    # ego at 4 m/s, 0.2 m voxels, map shift threshold 5 voxels -> shifts every few frames
    w = small_workload(ego_vx=4.0, frames=8, seed=3)
    gd, od = run_pair(w, w.frames)
    assert gd.info().center_pos[0] == od.info().center_pos[0]
    assert gd.info().center_pos[0] > 0.0


def test_config1_bit_exact(gpu):
    from k3dom_amd.scene import workload
    w = workload(1)
    run_pair(w, 5)
