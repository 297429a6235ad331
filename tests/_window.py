"""Inputs of the double-precision window chain from a run of the serial
restatement: per frame the map shift, the sensor position relative to the map
centre, the mass decay, the frozen counts and the ground-truth occupancy."""
import numpy as np


def oracle_window(workload, frames, first_frame=0, theta=None, mask=None):
    """Runs the serial restatement over the frames (with theta's ls, sigma and
    priors when given; theta's two sigmas must then be equal) and returns
    (chain, records): chain a WindowChain64, records a list of per-frame dicts
    with the float masses, counts and move."""
    from k3dom_amd._lib import Params
    from k3dom_amd.adjoint import map_centers
    from k3dom_amd.scene import Scene
    from oracle.orc import OracleDOM, WindowChain64
    p = Params.from_buffer_copy(bytes(workload.params))
    if theta is not None:
        assert theta[1] == theta[2]
        p.ls, p.sigma = float(theta[0]), float(theta[1])
        p.prior_free, p.prior_static, p.prior_dynamic = (float(v) for v in theta[3:6])
    sc = Scene(workload.scene)
    recs = [sc.frame(first_frame + f) for f in range(frames)]
    centers = map_centers(p, [r[1] for r in recs])
    gs = int(np.float32(p.size) / np.float32(p.resolution))
    gsz = int(np.float32(p.size_z) / np.float32(p.resolution))
    od = OracleDOM(p)
    prev_t, prev_c = None, None
    moves, srel, decays, counts, gts, out = [], [], [], [], [], []
    for f, (t, pose, xs, ys, zs) in enumerate(recs):
        od.updatePose(*pose)
        info = od.info()
        c = np.array(info.center_pos[:], np.float32)
        s = np.array(info.sensor_pos[:], np.float32)
        mv = np.zeros(3, np.int32) if prev_c is None else np.rint((c - prev_c) / np.float32(p.resolution)).astype(
            np.int32)
        decay = 1.0 if prev_t is None else float(np.float32(float(np.float32(p.gamma)) **
                                                            float(np.float32(t - prev_t))))
        od.updateGrid(t, xs, ys, zs)
        cells = od.copyCellsToHost()
        gt = sc.occupancy(first_frame + f, gs, gsz, p.resolution, centers[f], p.size, p.size_z)
        moves.append(mv)
        srel.append(s - c)
        decays.append(decay)
        counts.append(od.counts())
        gts.append(gt)
        out.append(dict(f=cells["free_mass"].copy(), s=cells["static_mass"].copy(), d=cells["dynamic_mass"].copy(),
                        counts=counts[-1], move=mv, gt=gt))
        prev_t, prev_c = t, c
    chain = WindowChain64(p, [(r[2], r[3], r[4]) for r in recs], srel, moves, decays, counts, gts, mask)
    return chain, out


def bce_loss(records, mask=None):
    """Loss of float masses: mean over frames of the mean masked BCE."""
    tot = 0.0
    for r in records:
        f, s, d = (r[k].astype(np.float64) for k in "fsd")
        p = np.clip((s + d) / (f + s + d), 1e-6, 1 - 1e-6)
        y = r["gt"].astype(np.float64)
        b = -(y * np.log(p) + (1 - y) * np.log(1 - p))
        tot += b[mask.astype(bool)].mean() if mask is not None else b.mean()
    return tot / len(records)
