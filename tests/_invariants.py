"""Size-independent properties of a map after a frame, for configurations too
large to replay on the serial restatement every frame.

Each property follows from the algorithm, not from the implementation:
- the stable sort: keys ascending, the permutation a permutation, equal keys
  in ascending predicted index (src/common/dom.cu:169-183);
- per-voxel ranges tile [0, N) and hold exactly their voxel's keys
  (src/common/kernel/particle_to_grid.cu:13-39; every key is clamped into the
  grid by src/common/kernel/predict.cu:52-57);
- newborn slots: ascending voxel keys, every one a voxel with born mass, and
  each voxel's count within one of B * born / sum(born)
  (src/common/kernel/init_new_particles.cu:14-52);
- systematic resampling: indices ascending inside [0, N + B), each entry
  drawn floor or ceil of N * w / W times, the new set of uniform weight W / N
  (src/common/kernel/resampling.cu:38-89);
- resampled particles come from entries of positive weight, so they lie in
  the grid and move at least the minimum speed (src/common/kernel/predict.cu:44-50,
  resampling.cu:13-36);
- masses non-negative and finite, pers + born = d up to one rounding
  (src/k3dom/kernel/mass_update.cu:141-150,179-181).
"""
import numpy as np


def _segments_ok(starts, ends, n):
    order = np.argsort(starts, kind="stable")
    s, e = starts[order], ends[order]
    if len(s) == 0:
        return n == 0
    return s[0] == 0 and e[-1] == n - 1 and np.all(s[1:] == e[:-1] + 1) and np.all(e >= s)


def check_frame_invariants(gd, normal, what=""):
    p = gd.params
    N, B, V = gd.particle_count, gd.new_born_particle_count, gd.grid_cell_count
    gs, gsz = gd.grid_size, gd.grid_size_z
    cells = gd.copyCellsToHost()
    f, s, d = cells["free_mass"], cells["static_mass"], cells["dynamic_mass"]
    for name, a in (("free", f), ("static", s), ("dynamic", d)):
        assert np.all(np.isfinite(a)) and np.all(a >= 0.0), f"{what} {name} mass negative or not finite"
    for name in ("mean_x_vel", "mean_y_vel", "mean_z_vel"):
        assert np.all(np.isfinite(cells[name])), f"{what} {name} not finite"
    born = gd.read_array("BORN_MASS")
    assert np.all(np.isfinite(born)) and np.all(born >= 0.0), f"{what} born mass negative or not finite"

    if normal:
        pers = cells["pers_mass"]
        assert np.all(pers >= -1e-6 * np.maximum(d, 1e-30)), f"{what} persistent mass below zero"
        np.testing.assert_allclose(pers.astype(np.float64) + born, d, rtol=2e-7, atol=1e-30,
                                   err_msg=f"{what} pers + born != dynamic mass")
        # stable sort of the predicted keys
        skey = gd.read_array("SORTED_KEY")
        perm = gd.read_array("SORT_PERM")
        assert np.all(np.diff(skey) >= 0), f"{what} sorted keys not ascending"
        assert np.all((skey >= 0) & (skey < V)), f"{what} key outside the grid"
        assert np.array_equal(np.sort(perm), np.arange(N, dtype=np.int32)), f"{what} sort order not a permutation"
        same = np.diff(skey) == 0
        assert np.all(np.diff(perm)[same] > 0), f"{what} sort not stable"
        # ranges tile [0, N) and hold their voxel's keys
        occ = cells["start_idx"] >= 0
        st, en = cells["start_idx"][occ].astype(np.int64), cells["end_idx"][occ].astype(np.int64)
        assert np.all(cells["end_idx"][~occ] == -1), f"{what} empty voxel with an end index"
        assert _segments_ok(st, en, N), f"{what} voxel ranges do not tile [0, N)"
        vox = np.nonzero(occ)[0]
        counts = en - st + 1
        by_start = np.argsort(st)
        assert np.array_equal(skey, np.repeat(vox[by_start], counts[by_start]).astype(np.int32)), \
            f"{what} a range holds another voxel's keys"
        # newborn slots in proportion to the born masses
        nkey = gd.read_array("NEWBORN_KEY")
        assert np.all(np.diff(nkey) >= 0), f"{what} newborn keys not ascending"
        tot = float(np.sum(born, dtype=np.float64))
        if tot > 0 and B > 0:
            assert np.all(born[nkey] > 0), f"{what} newborn in a voxel without born mass"
            nb = np.bincount(nkey, minlength=V).astype(np.float64)
            expect = B * born.astype(np.float64) / tot
            err = np.abs(nb - expect).max()
            assert err <= 1.0 + 1e-6 * B, f"{what} newborn counts off by {err}"
        # systematic resampling over [persistent ++ newborn]
        ridx = gd.read_array("RESAMPLE_IDX").astype(np.int64)
        assert np.all(np.diff(ridx) >= 0) and ridx[0] >= 0 and ridx[-1] < N + B, f"{what} resampled indices"
        wj = np.concatenate([gd.read_array("SORTED_WEIGHT"), gd.read_array("NEWBORN_WEIGHT")]).astype(np.float64)
        assert np.all(wj >= 0.0), f"{what} negative joint weight"
        W = wj.sum()
        if W > 0:
            taken = np.bincount(ridx, minlength=N + B).astype(np.float64)
            exp_t = N * wj / W
            assert np.all(taken >= np.floor(exp_t - 1e-6 * N - 1e-9) - 0) and \
                np.all(taken <= np.ceil(exp_t + 1e-6 * N + 1e-9)), f"{what} systematic draw counts"
            assert np.all(wj[ridx] > 0), f"{what} resampled an entry of weight zero"
            w = gd.read_array("PARTICLE_WEIGHT")
            assert np.all(w == w[0]), f"{what} resampled weights not uniform"
            np.testing.assert_allclose(float(w[0]), W / N, rtol=1e-5, err_msg=f"{what} weight != W / N")
        # resampled particles: inside the grid, at least the minimum speed, keys of their positions
        pos = gd.read_array("PARTICLE_POS")
        vel = gd.read_array("PARTICLE_VEL")
        key = gd.read_array("PARTICLE_KEY")
        assert np.all((pos[0] >= 0) & (pos[0] < gs) & (pos[1] >= 0) & (pos[1] < gs) & (pos[2] >= 0) &
                      (pos[2] < gsz)), f"{what} resampled particle outside the grid"
        v2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]
        mv = np.float32(p.particle_min_vel)
        assert np.all(v2 >= mv * mv), f"{what} resampled particle below the minimum speed"
        ix = np.clip(pos[0].astype(np.int64), 0, gs - 1)
        iy = np.clip(pos[1].astype(np.int64), 0, gs - 1)
        iz = np.clip(pos[2].astype(np.int64), 0, gsz - 1)
        assert np.array_equal(key, (ix + gs * iy + gs * gs * iz).astype(np.int32)), f"{what} keys of positions"
    else:
        w = gd.read_array("PARTICLE_WEIGHT")
        assert np.all(w == np.float32(1.0) / np.float32(N)), f"{what} initial weights != 1/N"
