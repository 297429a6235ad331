"""The DS-PHD/MIB mass update (src/dsphdmib/): measurement grid, Dempster-
Shafer cell update, newborn split and weight renormalisation.

CPU tests pin the serial restatement against closed forms and hand-computed
geometry; GPU tests hold the HIP path to it bit for bit."""
import ctypes as C
import math

import numpy as np
import pytest

EPS = 1e-5          # mass clipping, src/dsphdmib/kernel/measurement_grid.cu:146
P_MIN, OCC_MAX = 0.15, 0.95


def _ds(L, name, *args):
    return getattr(L, f"orc_ds_{name}")(*args)


def _update(L, op, fp, om, fm):
    out = (C.c_float * 2)()
    L.orc_ds_update_masses(op, fp, om, fm, out)
    return out[0], out[1]


# ---------------------------------------------------------------- closed forms

def test_pocc_closed_form(oracle_lib):
    L = oracle_lib
    assert _ds(L, "pocc", 1.0, 2.0, 3.0, 1.0, 2.0, 3.0, 0.1) == np.float32(OCC_MAX)
    # r = delta: 0.95 exp(-1/2); r = 2 delta: 0.95 exp(-2)
    assert abs(_ds(L, "pocc", 0.1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.1) - OCC_MAX * math.exp(-0.5)) < 1e-6
    assert abs(_ds(L, "pocc", 0.0, 0.0, 0.5, 0.0, 0.0, 0.3, 0.1) - OCC_MAX * math.exp(-2.0)) < 1e-6
    # symmetric in voxel and point
    assert _ds(L, "pocc", 0.3, -0.2, 0.1, 0.25, -0.1, 0.0, 0.1) == _ds(L, "pocc", 0.25, -0.1, 0.0, 0.3, -0.2, 0.1, 0.1)


def test_pfree_is_linear_in_range(oracle_lib):
    L = oracle_lib
    assert _ds(L, "pfree", 0.0, 0.0, 0.0, 50.0) == np.float32(P_MIN)
    assert _ds(L, "pfree", 30.0, 40.0, 0.0, 50.0) == np.float32(1.0)   # |c| = max_range
    assert abs(_ds(L, "pfree", 0.0, 12.0, 5.0, 26.0) - (P_MIN + 0.5 * (1 - P_MIN))) < 1e-6  # |c| = 13


def test_dempster_vacuous_operands_are_identities(oracle_lib):
    L = oracle_lib
    for op, fp in [(0.3, 0.5), (0.0, 0.9), (0.999, 0.0), (0.25, 0.25)]:
        # a vacuous measurement leaves the prediction, a vacuous prediction takes the measurement
        assert _update(L, op, fp, 0.0, 0.0) == (np.float32(op), np.float32(fp))
        assert _update(L, 0.0, 0.0, op, fp) == (np.float32(op), np.float32(fp))


def test_dempster_known_combination(oracle_lib):
    # prediction (occ 0.6, free 0.3, unknown 0.1), measurement (0.5, 0.2, 0.3):
    # conflict K = 0.3*0.5 + 0.6*0.2 = 0.27, occ = 0.53/0.73, free = 0.17/0.73,
    # unknown = 0.1*0.3/0.73
    o, f = _update(oracle_lib, 0.6, 0.3, 0.5, 0.2)
    assert abs(o - 0.53 / 0.73) < 1e-6 and abs(f - 0.17 / 0.73) < 1e-6
    assert abs((1 - o - f) - 0.03 / 0.73) < 1e-6


def test_dempster_commutes_and_conserves_mass(oracle_lib):
    L = oracle_lib
    rng = np.random.default_rng(7)
    for _ in range(2000):
        a = rng.dirichlet([1, 1, 1])
        b = rng.dirichlet([1, 1, 1])
        op, fp, om, fm = (float(np.float32(v)) for v in (a[0], a[1], b[0], b[1]))
        x = _update(L, op, fp, om, fm)
        assert x == _update(L, om, fm, op, fp)            # exact: the same float terms in the same order
        K = fp * om + op * fm
        if K > 0.999:
            continue
        unknown = (1 - op - fp) * (1 - om - fm) / (1 - K)
        assert abs(x[0] + x[1] + unknown - 1.0) < 1e-5
        assert x[0] >= 0 and x[1] >= 0


def test_newborn_split_limits(oracle_lib):
    L = oracle_lib
    # nothing predicted: all updated mass is newborn; fully predicted: none is
    assert abs(_ds(L, "newborn_part", 0.0, 0.7, 0.02) - 0.7) < 1e-7
    assert _ds(L, "newborn_part", 1.0, 0.7, 0.02) == 0.0
    # pred 0.5, p_B 0.02: 0.7 * 0.02*0.5 / (0.5 + 0.01)
    assert abs(_ds(L, "newborn_part", 0.5, 0.7, 0.02) - 0.7 * 0.01 / 0.51) < 1e-7


# ------------------------------------------------------- measurement geometry

def _tiny_map(particles=64):
    from k3dom_amd.scene import dsphdmib_params
    return dsphdmib_params(0.2, 8, 4, particles, 8)


def _meas(dom_cls, pts, **kw):
    p = _tiny_map()
    d = dom_cls(p, method="dsphdmib", **kw)
    d.updatePose(0.0, 0.0, 0.0)
    xs, ys, zs = (np.ascontiguousarray(np.asarray(pts, np.float32)[:, k]) for k in range(3))
    d.updateGrid(1.0, xs, ys, zs)
    return d.read_array("DEL_OCC"), d.read_array("DEL_FREE"), d.read_array("BORN_MASS")


def _cell(i, j, k):
    # 8x8x4 grid, x fastest; voxel centre ((i + 0.5 - 4) 0.2, (j + 0.5 - 4) 0.2, (k + 0.5 - 2) 0.2)
    return i + 8 * j + 64 * k


def _free(r):
    return 1.0 - (P_MIN + r / 50.0 * (1.0 - P_MIN))


def test_meas_grid_single_ray_by_hand(oracle_lib):
    """One point at the centre (0.5, 0.1, 0.1) of voxel (6, 4, 2), sensor at
    the origin: the voxel itself is occupied, the voxels the ray crosses before
    it are free with 1 - pFree(|c|), the voxel behind it gets the clipped
    minimum, voxels off the ray or behind the sensor get nothing."""
    from oracle.orc import OracleDOM
    occ, free, born = _meas(OracleDOM, [(0.5, 0.1, 0.1)])
    assert abs(occ[_cell(6, 4, 2)] - OCC_MAX) < 1e-7 and free[_cell(6, 4, 2)] == np.float32(EPS)
    assert occ[_cell(5, 4, 2)] == np.float32(EPS) and abs(free[_cell(5, 4, 2)] - _free(math.sqrt(0.11))) < 1e-6
    assert occ[_cell(4, 4, 2)] == np.float32(EPS) and abs(free[_cell(4, 4, 2)] - _free(math.sqrt(0.03))) < 1e-6
    # behind the point: not all farther, pOcc(r = 0.2) = 0.95 e^-2 < 0.5
    assert occ[_cell(7, 4, 2)] == np.float32(EPS) and free[_cell(7, 4, 2)] == np.float32(EPS)
    # behind the sensor (negative projection) and far from the ray
    for c in (_cell(3, 4, 2), _cell(4, 7, 2), _cell(0, 0, 0)):
        assert occ[c] == 0.0 and free[c] == 0.0
    # the first frame's born masses are the occupied masses (src/dsphdmib/dom.cu:55-82)
    assert np.array_equal(born, occ)


def _ray_points(first_centre, n_far):
    """n_far points on the half-line through voxel (6, 4, 2) well beyond it
    (all farther, negligible pOcc) and one point at the voxel centre, first or
    last."""
    d = np.array([0.5, 0.1, 0.1])
    far = [tuple(d * (2.0 + 0.01 * q)) for q in range(n_far)]
    return [tuple(d)] + far if first_centre else far + [tuple(d)]


def test_meas_grid_keeps_the_first_100_rays(oracle_lib):
    """Each voxel considers only the first 100 rays passing near it in point
    order (measurement_grid.cu:82-96): a hit after 100 far rays is not seen, a
    hit as the 100th ray is."""
    from oracle.orc import OracleDOM
    c = _cell(6, 4, 2)
    free_c = _free(math.sqrt(0.27))
    occ, free, _ = _meas(OracleDOM, _ray_points(False, 100))
    assert occ[c] == np.float32(EPS) and abs(free[c] - free_c) < 1e-6
    occ, free, _ = _meas(OracleDOM, _ray_points(False, 99))
    assert abs(occ[c] - OCC_MAX) < 1e-7 and free[c] == np.float32(EPS)
    occ, free, _ = _meas(OracleDOM, _ray_points(True, 150))
    assert abs(occ[c] - OCC_MAX) < 1e-7 and free[c] == np.float32(EPS)


def test_renormalisation_occurs_in_small_scene(oracle_lib):
    """From the fourth frame of the small scene some voxels' predicted occupied
    mass exceeds 1, so their weights are divided by it (mass_update.cu:61-67):
    afterwards such a voxel's weights sum to 1 within rounding."""
    from oracle.orc import OracleDOM
    from k3dom_amd.scene import Scene, small_workload
    w = small_workload(method="dsphdmib")
    sc = Scene(w.scene)
    od = OracleDOM(w.params, method="dsphdmib")
    seen = 0
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        od.updatePose(*pose)
        od.updateGrid(t, xs, ys, zs)
        cells = od.copyCellsToHost()
        hit = np.nonzero((cells["pred_mass"] == 1.0) & (cells["start_idx"] >= 0))[0]
        if f == 0 or len(hit) == 0:
            continue
        ws = od.read_array("SORTED_WEIGHT")
        for c in hit[:4]:
            s = ws[cells["start_idx"][c]:cells["end_idx"][c] + 1].astype(np.float64).sum()
            assert abs(s - 1.0) < 1e-4
        seen += len(hit)
    assert seen > 0


# ------------------------------------------------- culling bound of the GPU path

_BOUND_CHECK = r"""
#include <cmath>
#include <cstdio>
#include <random>
#include "k3_math.h"
// For random grids, boxes of voxels and points whose line passes near a
// voxel of the box (perpendicular distance 0.9..1.1 window or within 1e-3 of
// it, at ranges from 1e-9 to 1e17 times the voxel's distance): whenever the reference's float
// test accepts the point for that voxel, the box bound must keep it.
int main() {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long long accepted = 0, kept_by_bound_only = 0;
    for (int it = 0; it < 400000; ++it) {
        const int gs = 8 + (int)(U(rng) * 600), gsz = 4 + (int)(U(rng) * 120);
        const float res = (float)(0.02 + U(rng) * 0.5);
        const float sx = (float)((U(rng) - 0.5) * gs * res), sy = (float)((U(rng) - 0.5) * gs * res),
                    sz = (float)((U(rng) - 0.5) * gsz * res);
        // supertiles, tiles, and single voxels (R = 0: no geometric slack)
        const double kind = U(rng);
        const int bx = kind < 0.3 ? 16 : (kind < 0.6 ? 4 : 1), bz = kind < 0.3 ? 8 : (kind < 0.6 ? 4 : 1);
        const int i0 = (int)(U(rng) * gs), j0 = (int)(U(rng) * gs), k0 = (int)(U(rng) * gsz);
        const int i1 = std::min(i0 + bx, gs) - 1, j1 = std::min(j0 + bx, gs) - 1, k1 = std::min(k0 + bz, gsz) - 1;
        const float wr = res * 0.866f, w2 = wr * wr;
        const K3DsBox b = k3ds_box(gs, gsz, res, sx, sy, sz, i0, i1, j0, j1, k0, k1, w2);
        const int i = i0 + (int)(U(rng) * (i1 - i0 + 1)), j = j0 + (int)(U(rng) * (j1 - j0 + 1)),
                  k = k0 + (int)(U(rng) * (k1 - k0 + 1));
        float x, y, z;
        k3_voxel_center((k * gs + j) * gs + i, gs, gsz, res, sx, sy, sz, &x, &y, &z);
        // a direction at perpendicular distance ~window from c, scaled
        double ux = U(rng) - 0.5, uy = U(rng) - 0.5, uz = U(rng) - 0.5;
        const double cn = std::sqrt((double)x * x + (double)y * y + (double)z * z);
        if (cn < 1e-6) continue;
        const double dot = (ux * x + uy * y + uz * z) / (cn * cn);
        ux -= dot * x; uy -= dot * y; uz -= dot * z;
        const double un = std::sqrt(ux * ux + uy * uy + uz * uz);
        if (un < 1e-12) continue;
        const double off = wr * (U(rng) < 0.5 ? 0.9 + 0.2 * U(rng) : 1.0 + 2e-3 * (U(rng) - 0.5)) / un;
        const double scale = std::pow(10.0, -9.0 + 26.0 * U(rng));
        const float mx = (float)((x + ux * off) * scale), my = (float)((y + uy * off) * scale),
                    mz = (float)((z + uz * off) * scale);
        const float c2 = x * x + y * y + z * z, m2 = k3ds_m2(mx, my, mz);
        // the reciprocal form agrees with the division for 1/m2 off by up to 1 ulp
        const bool exact = k3ds_ray_near(x, y, z, c2, mx, my, mz, m2, w2);
        const float r = 1.0f / m2;
        for (float rr : {r, std::nextafter(r, 0.0f), std::nextafter(r, INFINITY)})
            if (k3ds_ray_near_rcp(x, y, z, c2, mx, my, mz, m2, rr, w2) != exact) {
                std::printf("RCP gs=%d res=%g c=(%g,%g,%g) m=(%g,%g,%g) r=%a\n", gs, res, x, y, z, mx, my, mz, rr);
                return 1;
            }
        if (!exact) continue;
        ++accepted;
        if (!k3ds_may_pass(b, mx, my, mz, m2)) {
            std::printf("MISS gs=%d gsz=%d res=%g s=(%g,%g,%g) box=[%d,%d]x[%d,%d]x[%d,%d] c=(%g,%g,%g) m=(%g,%g,%g)\n",
                        gs, gsz, res, sx, sy, sz, i0, i1, j0, j1, k0, k1, x, y, z, mx, my, mz);
            return 1;
        }
        if (!(m2 >= 1e-20f && m2 <= 1e30f)) ++kept_by_bound_only;
    }
    std::printf("%lld %lld\n", accepted, kept_by_bound_only);
    return 0;
}
"""


def test_culling_bound_keeps_every_accepted_point(tmp_path):
    """The GPU measurement grid skips points by a per-box bound
    (k3ds_box / k3ds_may_pass); a point that the exact float test accepts for
    a voxel of the box must never be skipped, also at the test's rounding
    edge (points within 1e-3 of the window, single-voxel boxes) and at extreme
    ranges.  Its exact test takes 1/|m|^2 from the hardware reciprocal
    (k3ds_ray_near_rcp); with the reciprocal off by up to 1 ulp it must give
    the division's answer."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "bound.cpp"
    src.write_text(_BOUND_CHECK)
    exe = tmp_path / "bound"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v3",
                    "-I", os.path.join(root, "k3dom_amd", "csrc"), str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    accepted, _ = (int(v) for v in r.stdout.split())
    assert accepted > 50000


# ------------------------------------------------------------------ GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("short_segments", [False, True])
def test_ds_meas_grid_single_ray_gpu(gpu, monkeypatch, short_segments):
    """The hand-built single-ray cases on the GPU; with short_segments the
    candidate lists are split into segments of 64 points, so the first-100
    rule is carried across segments."""
    from k3dom_amd.dom import DOM
    from oracle.orc import OracleDOM
    from tests._parity import bits_equal, first_diff
    if short_segments:
        monkeypatch.setenv("K3DOM_TEST_TINY_CAPACITY", "1")
    for pts in ([(0.5, 0.1, 0.1)], _ray_points(False, 100), _ray_points(False, 99), _ray_points(True, 150)):
        print(f"  single-ray case with {len(pts)} points", flush=True)
        g = _meas(DOM, pts)
        o = _meas(OracleDOM, pts)
        for a, b, name in zip(g, o, ("occ", "free", "born")):
            assert bits_equal(a, b), f"{len(pts)} points, {name}: {first_diff(a, b)}"


@pytest.mark.gpu
def test_ds_small_scene_bit_exact(gpu):
    from k3dom_amd.scene import small_workload
    from tests._parity import run_pair
    w = small_workload(method="dsphdmib")
    run_pair(w, w.frames, method="dsphdmib")


@pytest.mark.gpu
def test_ds_moving_ego_map_shift(gpu):
    from k3dom_amd.scene import small_workload
    from tests._parity import run_pair
    w = small_workload(ego_vx=4.0, frames=8, seed=3, method="dsphdmib")
    gd, od = run_pair(w, w.frames, method="dsphdmib")
    assert gd.info().center_pos[0] > 0.0


@pytest.mark.gpu
def test_ds_crowded_renormalisation_bit_exact(gpu):
    from k3dom_amd.scene import small_workload
    from tests._parity import run_pair
    # renormalisation from frame 7 on; more than 64 particles per voxel
    w = small_workload(particles=60000, newborn=6000, frames=8, seed=21, method="dsphdmib")
    gd, od = run_pair(w, w.frames, method="dsphdmib")
    cells = gd.copyCellsToHost()
    occ = cells["start_idx"] >= 0
    assert (cells["end_idx"][occ] - cells["start_idx"][occ] + 1).max() > 64


@pytest.mark.gpu
def test_ds_empty_frames_and_graph_replay(gpu, monkeypatch):
    """Empty and invalid point clouds, tiny capacities and graph replay with
    map shifts, in DS-PHD/MIB mode."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    from oracle.orc import OracleDOM
    from tests._parity import compare_state
    monkeypatch.setenv("K3DOM_TEST_TINY_CAPACITY", "1")
    w = small_workload(ego_vx=4.0, frames=8, seed=31, method="dsphdmib")
    sc = Scene(w.scene)
    gd, od = DOM(w.params, method="dsphdmib"), OracleDOM(w.params, method="dsphdmib")
    gd.set_graph(True)
    bad = np.array([np.nan, np.inf, 0.0, -np.inf, 1e30], np.float32)
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        if f == 2:
            xs, ys, zs = xs[:0], ys[:0], zs[:0]
        elif f == 3:
            xs, ys, zs = (np.concatenate([xs, bad]), np.concatenate([ys, bad[::-1]]),
                          np.concatenate([zs, np.zeros(5, np.float32)]))
        print(f"  DS graph frame {f}: {len(xs)} points", flush=True)
        normal = gd.info().updated_time > 0
        for m in (gd, od):
            m.updatePose(*pose)
            m.updateGrid(t, xs, ys, zs)
        compare_state(gd, od, normal_frame=normal, what=f"DS graph frame {f}:")


def _window_edge_points(info, gs, gsz, res, rng, n):
    """Points whose line passes at 0.99..1.01 window from random voxel
    centres (sensor-relative, float32 like the voxel-centre expression), at
    several ranges."""
    sx, sy, sz = (np.float32(info.sensor_pos[k] - info.center_pos[k]) for k in range(3))
    i, j, k = rng.integers(0, gs, n), rng.integers(0, gs, n), rng.integers(0, gsz, n)
    r = np.float32(res)
    c = np.stack([(((i + np.float32(0.5)).astype(np.float32) - np.float32(gs / 2)) * r - sx),
                  (((j + np.float32(0.5)).astype(np.float32) - np.float32(gs / 2)) * r - sy),
                  (((k + np.float32(0.5)).astype(np.float32) - np.float32(gsz / 2)) * r - sz)], 1).astype(np.float64)
    u = rng.normal(size=(n, 3))
    u -= (np.sum(u * c, 1) / np.maximum(np.sum(c * c, 1), 1e-12))[:, None] * c
    u /= np.maximum(np.linalg.norm(u, axis=1), 1e-12)[:, None]
    wr = res * 0.866 * rng.uniform(0.99, 1.01, n)
    scale = rng.choice([0.5, 1.0, 1.5, 3.0], n)
    return ((c + u * wr[:, None]) * scale[:, None]).astype(np.float32)


@pytest.mark.gpu
def test_ds_culled_grid_edge_cases(gpu, monkeypatch):
    """The culled GPU measurement grid against the oracle's test of every
    point for every voxel: a 20x20x6 grid (partial supertiles and tiles), a
    sensor offset, points at the window's edge, points at extreme ranges,
    NaN / inf / zero points, and candidate-list regrowth from a tiny capacity."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    from oracle.orc import OracleDOM
    from tests._parity import compare_state
    monkeypatch.setenv("K3DOM_TEST_TINY_CAPACITY", "1")
    w = small_workload(edge=20, edge_z=6, particles=2048, newborn=200, frames=4, seed=17, method="dsphdmib",
                       sensor_off_x=0.13, sensor_off_z=-0.07)
    sc = Scene(w.scene)
    gd, od = DOM(w.params, method="dsphdmib"), OracleDOM(w.params, method="dsphdmib")
    rng = np.random.default_rng(5)
    odd = np.array([[1e-12, 2e-12, 0.0], [3e16, -1e16, 2e15], [1e20, 1e20, 1e20], [np.nan, 1.0, 1.0],
                    [np.inf, 0.0, 0.0], [0.0, 0.0, 0.0], [-1e-25, 1e-25, 0.0], [1e30, 0.0, -1e30]], np.float32)
    for f in range(w.frames):
        t, pose, xs, ys, zs = sc.frame(f)
        for m in (gd, od):
            m.updatePose(*pose)
        edge = _window_edge_points(gd.info(), 20, 6, 0.2, rng, 3000)
        pts = np.concatenate([np.stack([xs, ys, zs], 1), edge, odd])[rng.permutation(len(xs) + 3000 + len(odd))]
        xs, ys, zs = (np.ascontiguousarray(pts[:, k]) for k in range(3))
        print(f"  DS edge-case frame {f}: {len(xs)} points", flush=True)
        normal = gd.info().updated_time > 0
        for m in (gd, od):
            m.updateGrid(t, xs, ys, zs)
        compare_state(gd, od, normal_frame=normal, what=f"DS edge-case frame {f}:")
    assert gd.stats()["capacity_regrowths"] >= 1
