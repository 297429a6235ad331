"""The size-independent frame properties (tests/_invariants.py) hold for the
serial restatement, and each one fails when a plausible mistake is planted in
the restatement's output (a swapped sort entry, an unstable tie, a shifted
range, a misplaced newborn, a skipped or repeated draw, a non-uniform weight,
a slow or stray particle, a negative mass)."""
import numpy as np
import pytest

from tests._invariants import check_frame_invariants


class Planted:
    """The restatement's map with one array or cell field altered on read."""

    def __init__(self, od, array=None, cells=None):
        self._od, self._array, self._cells = od, array or {}, cells
        for k in ("params", "particle_count", "new_born_particle_count", "grid_cell_count", "grid_size",
                  "grid_size_z"):
            setattr(self, k, getattr(od, k))

    def copyCellsToHost(self):
        c = self._od.copyCellsToHost()
        if self._cells:
            self._cells(c)
        return c

    def read_array(self, name):
        a = self._od.read_array(name).copy()
        if name in self._array:
            self._array[name](a)
        return a


@pytest.fixture(scope="module")
def oracle_map(oracle_lib):
    from k3dom_amd.scene import Scene, small_workload
    from oracle.orc import OracleDOM
    w = small_workload(ego_vx=2.0, frames=4, seed=7, particles=8192, newborn=819)
    sc = Scene(w.scene)
    od = OracleDOM(w.params)
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        normal = od.info().updated_time > 0
        od.updatePose(*pose)
        od.updateGrid(t, xs, ys, zs)
        check_frame_invariants(od, normal, what=f"frame {f}:")
    return od


def _swap_adjacent_distinct(a):
    i = int(np.nonzero(np.diff(a) != 0)[0][0])
    a[i], a[i + 1] = a[i + 1], a[i]


def _unstable_tie(a, keys):
    i = int(np.nonzero(np.diff(keys) == 0)[0][0])
    a[i], a[i + 1] = a[i + 1], a[i]


def _plants(od):
    skey = od.read_array("SORTED_KEY")
    ridx = od.read_array("RESAMPLE_IDX")
    cells = od.copyCellsToHost()
    occ = np.nonzero(cells["start_idx"] >= 0)[0]
    v_first = occ[np.argmin(cells["start_idx"][occ])]
    born = od.read_array("BORN_MASS")
    empty_born = int(np.nonzero(born == 0)[0][0])
    heavy = int(np.argmax(np.bincount(ridx)))

    def set_cell(field, v, val):
        def f(c):
            c[field][v] = val
        return f

    def neg_mass(c):
        c["static_mass"][0] = -1e-3

    return {
        "sort order": dict(array={"SORTED_KEY": _swap_adjacent_distinct}),
        "stability": dict(array={"SORT_PERM": lambda a: _unstable_tie(a, skey)}),
        "range end": dict(cells=set_cell("end_idx", v_first, cells["end_idx"][v_first] + 1)),
        "newborn voxel": dict(array={"NEWBORN_KEY": lambda a: a.__setitem__(slice(0, 1), empty_born)}),
        "newborn count": dict(array={"NEWBORN_KEY": lambda a: a.__setitem__(slice(0, len(a) // 2), a[0])}),
        "skipped draw": dict(array={"RESAMPLE_IDX": lambda a: a.__setitem__(a == heavy, heavy + 1)}),
        "weights": dict(array={"PARTICLE_WEIGHT": lambda a: a.__setitem__(0, a[0] * 1.5)}),
        "stray particle": dict(array={"PARTICLE_POS": lambda a: a.__setitem__((0, 0), -0.5)}),
        "slow particle": dict(array={"PARTICLE_VEL": lambda a: a.__setitem__((slice(None), 0), 0.0)}),
        "key of position": dict(array={"PARTICLE_KEY": lambda a: a.__setitem__(0, (a[0] + 1) % 64)}),
        "negative mass": dict(cells=neg_mass),
    }


@pytest.mark.parametrize("name", ["sort order", "stability", "range end", "newborn voxel", "newborn count",
                                  "skipped draw", "weights", "stray particle", "slow particle", "key of position",
                                  "negative mass"])
def test_planted_mistake_is_caught(oracle_map, name):
    plant = _plants(oracle_map)[name]
    with pytest.raises(AssertionError):
        check_frame_invariants(Planted(oracle_map, **plant), True, what=name)
