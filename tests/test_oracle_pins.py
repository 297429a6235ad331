"""Pins of four functions of the serial restatement that bit-exact GPU parity
cannot pin (the GPU and the restatement share k3_math.h): the motion step,
the velocity draw of new particles, the per-voxel velocity moments and the map
shift.  Each is checked against a closed form, a statistical law or a
hand-built case, chosen so that a plausible mistake (a dropped term, a swapped
operand, a clamp in the wrong place, an off-by-one bound, a wrong index)
fails at least one of them."""
import math

import numpy as np
import pytest

f32 = np.float32


# ---------------------------------------------------------------- motion step
# src/common/kernel/predict.cu:33-55: pos += vel * dt + n_pos; vel += n_vel;
# w *= p_S; each velocity component clamped to +-5 * init_max_velocity after the
# move; weight 0 outside [0, gs)^2 x [0, gsz); key from the truncated, clamped
# position.

def _key(x, y, z, gs, gsz):
    ix = min(max(int(math.trunc(float(x))), 0), gs - 1)
    iy = min(max(int(math.trunc(float(y))), 0), gs - 1)
    iz = min(max(int(math.trunc(float(z))), 0), gsz - 1)
    return ix + gs * iy + gs * gs * iz


def test_predict_without_noise_is_constant_velocity(oracle_lib):
    from oracle import orc
    gs, gsz = 16, 8
    dt, p_S, vclamp = f32(0.1), f32(0.99), f32(25.0)
    cases = np.array([
        (1.25, 2.5, 3.75, 10.0, -5.0, 2.5),    # inside
        (15.5, 4.0, 4.0, 4.99, 0.0, 0.0),      # lands just below gs: inside
        (15.5, 4.0, 4.0, 5.0, 0.0, 0.0),       # lands on gs exactly: outside
        (0.5, 4.0, 4.0, -5.0, 0.0, 0.0),       # lands on 0 exactly: inside
        (0.5, 4.0, 4.0, -5.01, 0.0, 0.0),      # just below 0: outside, key clamped to 0
        (4.0, 4.0, 7.5, 0.0, 0.0, 6.0),        # z past gsz (below gs): outside
        (4.0, 4.0, 4.0, -40.0, 0.0, 0.0),      # moves by the unclamped velocity, then clamps
        (4.0, 15.9, 4.0, 0.0, 30.0, 0.0),      # y past gs with a clamped velocity
        (4.0, 4.0, 4.0, 0.0, 0.0, -24.0),      # inside the clamp: unchanged velocity
    ], dtype=np.float32)
    px, py, pz, vx, vy, vz = (cases[:, k].copy() for k in range(6))
    w = np.full(len(cases), 0.5, np.float32)
    npx, npy, npz, nvx, nvy, nvz, nw, key = orc.predict(px, py, pz, vx, vy, vz, w, dt, p_S, 0.0, 0.0, vclamp, gs,
                                                        gsz, frame=3)
    for i in range(len(cases)):
        ex = f32(px[i] + f32(vx[i] * dt))
        ey = f32(py[i] + f32(vy[i] * dt))
        ez = f32(pz[i] + f32(vz[i] * dt))
        assert (npx[i], npy[i], npz[i]) == (ex, ey, ez), f"case {i}: position"
        ev = [f32(min(max(v, -vclamp), vclamp)) for v in (vx[i], vy[i], vz[i])]
        assert [nvx[i], nvy[i], nvz[i]] == ev, f"case {i}: velocity"
        inside = 0 <= ex < gs and 0 <= ey < gs and 0 <= ez < gsz
        assert nw[i] == (f32(w[i] * p_S) if inside else f32(0.0)), f"case {i}: weight"
        assert key[i] == _key(ex, ey, ez, gs, gsz), f"case {i}: key"
    inside = [bool(x) for x in nw > 0]
    assert inside == [True, True, False, True, False, False, True, False, True]
    assert npx[6] == f32(0.0) and nvx[6] == -vclamp  # moved by -40 * 0.1, not by the clamped -25 * 0.1


def test_predict_noise_moments(oracle_lib):
    """With zero velocity and dt = 0 the move is the noise itself: six
    independent N(0, sigma) draws per particle, position noise with the
    position sigma and velocity noise with the velocity sigma, different for
    every particle and every frame."""
    from oracle import orc
    n = 100_000
    zero = np.zeros(n, np.float32)
    sp, sv = 0.5, 2.0
    out = orc.predict(zero, zero, zero, zero, zero, zero, np.ones(n, np.float32), 0.0, 1.0, sp, sv, 1e6, 16, 8,
                      frame=5)
    noise = np.stack(out[:6]).astype(np.float64)
    sig = np.array([sp, sp, sp, sv, sv, sv])
    z = noise / sig[:, None]
    assert np.all(np.abs(z.mean(axis=1)) < 5 / math.sqrt(n)), z.mean(axis=1)
    assert np.all(np.abs(z.std(axis=1) - 1) < 0.015), z.std(axis=1)
    corr = np.corrcoef(z)
    off = corr[~np.eye(6, dtype=bool)]
    assert np.all(np.abs(off) < 5 / math.sqrt(n)), corr
    tail = np.mean(np.abs(z) > 2, axis=1)
    assert np.all(np.abs(tail - 0.0455) < 0.004), tail
    other = orc.predict(zero, zero, zero, zero, zero, zero, np.ones(n, np.float32), 0.0, 1.0, sp, sv, 1e6, 16, 8,
                        frame=6)
    assert np.mean(other[0] == out[0]) < 1e-3  # another frame, other draws
    assert len(np.unique(out[0])) > 0.99 * n    # other particles, other draws


# ---------------------------------------------------------------- velocity draw
# src/common/kernel/init_new_particles.cu:160-176 with curand_uniform(min, max)
# of include/common/cuda_utils.h:27-31: each component min + (max - min) * u,
# u in [0, 1); a non-zero velocity slower than min_vel is rescaled to min_vel.

def test_velocity_draw_bounds_and_rescale(oracle_lib):
    from oracle import orc
    vmax, vmin = f32(50.0), f32(1.0)
    assert np.array_equal(orc.sample_velocity(0, 0, 0, vmax, vmin), [-vmax] * 3)
    top = orc.sample_velocity(0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, vmax, vmin)
    assert np.all(top < vmax) and np.all(top > vmax - 1e-4)
    # u = 1/2 exactly: the zero velocity stays zero (no division by zero)
    half = 0x80000000
    assert np.array_equal(orc.sample_velocity(half, half, half, vmax, vmin), [0.0, 0.0, 0.0])
    # a tiny non-zero velocity: rescaled to the minimum speed, same direction
    for ks in ((1, 2, -3), (5, 0, 0), (-7, 11, 13)):
        r = [half + (k << 8) for k in ks]
        raw = np.array([f32(-vmax + f32(f32(2 * vmax) * f32((rr >> 8) * 2.0 ** -24))) for rr in r], np.float64)
        got = orc.sample_velocity(*r, vmax, vmin).astype(np.float64)
        assert abs(np.linalg.norm(got) - vmin) < 2e-6 * vmin, (ks, got)
        np.testing.assert_allclose(got / np.linalg.norm(got), raw / np.linalg.norm(raw), atol=2e-6)
    # a velocity above the minimum is not rescaled
    r = [half + (400000 << 8), half, half]
    got = orc.sample_velocity(*r, vmax, vmin)
    assert got[1] == 0 and got[2] == 0 and 1.0 < got[0] and abs(got[0] - 100 * 400000 * 2.0 ** -24) < 1e-5


def test_velocity_draw_is_uniform_over_the_box(oracle_lib):
    from oracle import orc
    rng = np.random.default_rng(7)
    n = 60_000
    r = rng.integers(0, 2 ** 32, size=(n, 3), dtype=np.uint64)
    vmax = f32(50.0)
    v = np.array([orc.sample_velocity(a, b, c, vmax, 0.0) for a, b, c in r], np.float64)
    assert v.min() >= -vmax and v.max() < vmax
    sd = vmax / math.sqrt(3)
    assert np.all(np.abs(v.mean(axis=0)) < 5 * sd / math.sqrt(n))
    assert np.all(np.abs(v.var(axis=0) / sd ** 2 - 1) < 0.03)
    hist, _ = np.histogram(v[:, 0], bins=20, range=(-vmax, vmax))
    assert np.all(np.abs(hist - n / 20) < 5 * math.sqrt(n / 20))
    # with a minimum speed of 10% of vmax no non-zero draw is slower than it
    vm = np.array([orc.sample_velocity(a, b, c, vmax, 5.0) for a, b, c in r[:5000]], np.float64)
    sp = np.linalg.norm(vm, axis=1)
    assert np.all((sp >= 5.0 * (1 - 1e-6)) | (sp == 0))


# ---------------------------------------------------------------- moments
# src/common/kernel/statistical_moments.cu:12-63: per voxel with particles,
# sum over its range [start, end] of w * v, divided by the persistent mass when
# that is > 0, else 0; 0 for a voxel without particles.

def test_moments_hand_placed_particles(oracle_lib):
    from oracle import orc
    # voxel: 0 one particle; 1 two particles; 2 no particles (pers > 0);
    # 3 particles but pers = 0; 4 pers < 0; 5 200 particles (several 64-particle
    # blocks, exact sums); 6 pers != sum of weights; 7 the particles after 5
    w = [0.5, 0.25, 0.75, 0.3, 0.4]
    vx = [2.0, 4.0, 8.0, 9.0, 9.0]
    vy = [-4.0, -8.0, 0.0, 1.0, 1.0]
    vz = [6.0, 1.0, 1.0, 1.0, 1.0]
    k = np.arange(200)
    w += [1 / 256] * 200
    vx += list((k % 7 - 3).astype(float))
    vy += list((k % 5).astype(float))
    vz += list((-(k % 3)).astype(float))
    w += [0.25, 0.25]
    vx += [4.0, 12.0]
    vy += [0.0, 0.0]
    vz += [-2.0, 2.0]
    w += [0.5]
    vx += [3.0]
    vy += [3.0]
    vz += [3.0]
    start = [0, 1, -1, 3, 4, 5, 205, 207]
    end = [0, 2, -1, 3, 4, 204, 206, 207]
    pers = [0.5, 1.0, 3.0, 0.0, -1.0, 200 / 256, 2.0, 0.25]
    mx, my, mz = orc.moments(start, end, w, vx, vy, vz, pers)
    assert (mx[0], my[0], mz[0]) == (2.0, -4.0, 6.0)
    assert (mx[1], my[1], mz[1]) == (7.0, -2.0, 1.0)
    assert (mx[2], my[2], mz[2]) == (0.0, 0.0, 0.0)
    assert (mx[3], my[3], mz[3]) == (0.0, 0.0, 0.0)
    assert (mx[4], my[4], mz[4]) == (0.0, 0.0, 0.0)
    sx, sy, sz = (float(np.sum(a)) / 256 for a in ((k % 7 - 3), (k % 5), -(k % 3)))
    p5 = f32(200 / 256)
    assert (mx[5], my[5], mz[5]) == (f32(sx) / p5, f32(sy) / p5, f32(sz) / p5)
    assert (mx[6], my[6], mz[6]) == (2.0, 0.0, 0.0)   # (1 + 3) / 2, not / (0.25 + 0.25)
    assert (mx[7], my[7], mz[7]) == (6.0, 6.0, 6.0)   # 0.5 * 3 / 0.25


# ---------------------------------------------------------------- map shift
# DOM_c::updatePose (src/common/dom.cu:111-158): the move is the truncation
# toward zero of (sensor - centre - offset) / resolution per axis; when any
# |move| >= map_shift_thresh, initGridCellsKernel (src/common/kernel/init.cu:20-35)
# resets f, s, d, occ to the priors, start/end to -1, likelihood to 1 and p_A
# to 0 in every cell, moveMapKernel (ego_motion_compensation.cu:22-41) copies
# whole cells new[c] = old[c + move] where c + move is inside the grid, so the
# other fields of a cell with no source keep that cell's own old values;
# particles move by -move (:12-20) and the centre by move * resolution.

FIELDS = ["start_idx", "end_idx", "mu_A", "mu_UA", "w_A", "w_UA", "mean_x_vel", "mean_y_vel", "mean_z_vel",
          "likelihood", "p_A", "free_mass", "static_mass", "dynamic_mass", "occ_mass", "pers_mass", "pred_mass"]
KEPT = ["mu_A", "mu_UA", "w_A", "w_UA", "mean_x_vel", "mean_y_vel", "mean_z_vel", "pers_mass", "pred_mass"]


def _hand_cells(V):
    from k3dom_amd.dom import CELL_DTYPE
    c = np.zeros(V, CELL_DTYPE)
    for k, name in enumerate(FIELDS):
        c[name] = 100 * (k + 1) + np.arange(V)
    return c


def _shifted(old, move, gs, gsz, prior):
    new = old.copy()
    for name in ("free_mass", "static_mass", "dynamic_mass", "occ_mass", "likelihood", "p_A", "start_idx",
                 "end_idx"):
        new[name] = prior[name]
    for c in range(len(old)):
        x, y, z = c % gs, (c // gs) % gs, c // (gs * gs)
        sx, sy, sz = x + move[0], y + move[1], z + move[2]
        if 0 <= sx < gs and 0 <= sy < gs and 0 <= sz < gsz:
            new[c] = old[sx + gs * sy + gs * gs * sz]
    return new


def test_map_shift_hand_built_4x4x4(oracle_lib):
    from k3dom_amd.scene import k3dom_params
    from oracle.orc import OracleDOM
    p = k3dom_params(0.5, 4, 4, 8, 2, map_shift_thresh=1, prior_free=0.25, prior_static=0.5, prior_dynamic=0.75,
                     prior_occ=0.125)
    od = OracleDOM(p)
    gs = gsz = 4
    od.updatePose(0.0, 0.0, 0.0)
    assert tuple(od.info().center_pos) == (0.0, 0.0, 0.0)
    cells = _hand_cells(64)
    od.set_cells(cells)
    pos = np.stack([np.arange(8, dtype=np.float32) * 0.5, np.arange(8, dtype=np.float32) * 0.25 + 1,
                    np.full(8, 2.0, np.float32)])
    od.set_particle_pos(pos)
    prior = {"free_mass": 0.25, "static_mass": 0.5, "dynamic_mass": 0.75, "occ_mass": 0.125, "likelihood": 1.0,
             "p_A": 0.0, "start_idx": -1, "end_idx": -1}

    # move (+1, -2, 0): x 0.6 / 0.5 -> 1, y -1.1 / 0.5 -> -2, z 0.2 / 0.5 -> 0
    od.updatePose(0.6, -1.1, 0.2)
    got = od.copyCellsToHost()
    want = _shifted(cells, (1, -2, 0), gs, gsz, prior)
    for name in FIELDS:
        assert np.array_equal(got[name], want[name]), name
    # hand-read entries: new (0, 2, 0) = index 8 takes old (1, 0, 0) = index 1
    assert got["free_mass"][8] == 1200 + 1 and got["start_idx"][8] == 100 + 1 and got["pred_mass"][8] == 1700 + 1
    # new (3, 2, 0) = index 11 has no source (x + 1 = 4): priors, own kept fields
    assert got["free_mass"][11] == 0.25 and got["start_idx"][11] == -1 and got["likelihood"][11] == 1.0
    assert got["mu_A"][11] == 300 + 11 and got["pers_mass"][11] == 1600 + 11
    # new (0, 0, 0): y - 2 < 0, no source
    assert got["dynamic_mass"][0] == 0.75 and got["mean_x_vel"][0] == 700
    ppos = od.read_array("PARTICLE_POS")
    assert np.array_equal(ppos, pos - np.array([[1], [-2], [0]], np.float32))
    assert tuple(od.info().center_pos) == (f32(0.5), f32(-1.0), f32(0.0))

    # below the threshold (0.9, -0.9, 0.6 voxels): no shift
    before = od.copyCellsToHost()
    od.updatePose(0.5 + 0.45, -1.0 - 0.45, 0.3)
    after = od.copyCellsToHost()
    for name in FIELDS:
        assert np.array_equal(after[name], before[name]), name
    assert tuple(od.info().center_pos) == (f32(0.5), f32(-1.0), f32(0.0))

    # truncation toward zero: x -0.6 voxel -> 0 (not -1), y 1.2 -> 1
    od.updatePose(0.5 - 0.3, -1.0 + 0.6, 0.0)
    got = od.copyCellsToHost()
    want = _shifted(after, (0, 1, 0), gs, gsz, prior)
    for name in FIELDS:
        assert np.array_equal(got[name], want[name]), name
    assert tuple(od.info().center_pos) == (f32(0.5), f32(-0.5), f32(0.0))

    # a shift of more than the whole grid: every cell reset, kept fields stay
    prev = od.copyCellsToHost()
    prev_pos = od.read_array("PARTICLE_POS")
    od.updatePose(0.5 + 5 * 0.5 + 0.1, -0.5, 0.0)
    got = od.copyCellsToHost()
    for name, v in prior.items():
        assert np.all(got[name] == v), name
    for name in KEPT:
        assert np.array_equal(got[name], prev[name]), name
    assert np.array_equal(od.read_array("PARTICLE_POS"), prev_pos - np.array([[5], [0], [0]], np.float32))
