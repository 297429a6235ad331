// Differentiated window: forward frames with a per-frame record, and the
// reverse pass over the frames (declarations in include/k3dom.h).  Runs the
// ordinary frame pipeline of k3_api.cpp on its own map; the backward pass
// rebuilds each frame's candidate lists, recomputes the evidence with its
// length-scale derivative, and applies the per-voxel adjoint (k3_adjoint.hip).
#include "k3_internal.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

using namespace k3api;

namespace {

struct FrameIn {
    double t = 0.0;
    float pose[3] = {0, 0, 0};
    int n = 0;
    size_t off = 0;  // offset of the frame's x array in the point store (floats)
    bool set = false;
    // several LiDARs (0: one sensor): points per LiDAR, consecutive in the
    // frame's arrays (already rotated into map axes), and their origins
    int ngroups = 0;
    int gcount[K3DOM_MAX_LIDARS] = {};
    float gorig[K3DOM_MAX_LIDARS][3] = {};
};

struct FrameOut {
    int move[3] = {0, 0, 0};
    float sensor[3] = {0, 0, 0}, center[3] = {0, 0, 0};
    float decay = 1.0f;
    bool normal = false;
};

}  // namespace

struct k3dom_window {
    k3dom_handle* h = nullptr;
    int T = 0, V = 0;
    std::vector<FrameIn> in;
    std::vector<FrameOut> out;
    std::vector<float> host_pts;  // per frame x[n], y[n], z[n]
    float* d_pts = nullptr;
    size_t d_pts_cap = 0;
    bool pts_dirty = true;
    uint8_t* d_gt = nullptr;      // T * V
    uint8_t* d_mask = nullptr;    // V (or unused)
    bool use_mask = false;
    long long mask_count = 0;
    std::vector<WindowFrame> rec;
    double* loss_partial = nullptr;  // T * blocks
    float *del_o = nullptr, *del_f = nullptr, *dls_o = nullptr, *dls_f = nullptr;
    double* lam[2][3] = {};
    double* acc = nullptr;
    double* adj_partial = nullptr;
    FrameConsts* d_fc = nullptr;   // T: the forward frames' constants (device)
    FrameConsts* h_fc = nullptr;   // T: pinned host copies
    FrameConsts* d_gfc = nullptr;  // T * K3DOM_MAX_LIDARS: per-LiDAR constants of grouped frames
    FrameConsts* h_gfc = nullptr;
    // evidence of LiDARs after the first (and its length-scale derivative) in the backward pass
    float *g_o = nullptr, *g_f = nullptr, *g_lo = nullptr, *g_lf = nullptr;
    float theta[6] = {0, 0, 0, 0, 0, 0};
    bool forward_done = false;
    uint64_t forward_id = 0;       // forward passes run so far (k3dom_window_forward_id)
    uint32_t* overflow_any = nullptr;  // device: a frame of the backward pass overflowed
    float ms[2] = {0, 0};
    std::string err;
};

namespace {

k3dom_status wfail(k3dom_window* w, k3dom_status s, const std::string& msg) {
    if (w) w->err = msg;
    return s;
}

// propagate a handle error into the window's error text
k3dom_status from_handle(k3dom_window* w, k3dom_status s) {
    if (s != K3DOM_OK) w->err = w->h->err;
    return s;
}

#define W_HIP(w, expr)                                                                             \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) return wfail((w), K3DOM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <class T>
hipError_t walloc(k3dom_window* w, T** p, size_t count) {
    return dalloc(w->h, p, count);  // owned (and freed) by the window's map
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

k3dom_status upload_points(k3dom_window* w) {
    if (!w->pts_dirty) return K3DOM_OK;
    const size_t need = std::max<size_t>(w->host_pts.size(), 1);
    if (need > w->d_pts_cap) {
        dfree(w->h, w->d_pts);
        w->d_pts = nullptr;
        W_HIP(w, walloc(w, &w->d_pts, need));
        w->d_pts_cap = need;
    }
    if (!w->host_pts.empty())
        W_HIP(w, hipMemcpy(w->d_pts, w->host_pts.data(), sizeof(float) * w->host_pts.size(), hipMemcpyHostToDevice));
    w->pts_dirty = false;
    return K3DOM_OK;
}

}  // namespace

extern "C" {

k3dom_status k3dom_window_create(const k3dom_params* params, int32_t frames, k3dom_window** out) {
    if (!out || frames <= 0) return K3DOM_ERR_INVALID_ARG;
    *out = nullptr;
    k3dom_window* w = new (std::nothrow) k3dom_window();
    if (!w) return K3DOM_ERR_OUT_OF_MEMORY;
    k3dom_status s = k3dom_create(params, &w->h);
    if (s != K3DOM_OK) {
        delete w;
        return s;
    }
    w->h->ibuf.partial_stride = 256;  // split items carry the length-scale derivatives too
    w->T = frames;
    w->V = w->h->V;
    w->in.resize(frames);
    w->out.resize(frames);
    w->rec.resize(frames);
    const size_t V = (size_t)w->V;
    auto bail = [&](hipError_t e) {
        w->err = hipGetErrorString(e);
        k3dom_window_destroy(w);
        return e == hipErrorOutOfMemory ? K3DOM_ERR_OUT_OF_MEMORY : K3DOM_ERR_HIP;
    };
    hipError_t e;
#define WC(expr) if ((e = (expr)) != hipSuccess) return bail(e)
    for (int f = 0; f < frames; ++f) {
        WindowFrame& r = w->rec[f];
        WC(walloc(w, &r.f, V)); WC(walloc(w, &r.s, V)); WC(walloc(w, &r.d, V)); WC(walloc(w, &r.dyn, V));
        WC(walloc(w, &r.n, V)); WC(walloc(w, &r.m, V)); WC(walloc(w, &r.nb, V)); WC(walloc(w, &r.mb, V));
    }
    WC(walloc(w, &w->d_gt, (size_t)frames * V));
    WC(hipMemset(w->d_gt, 0, (size_t)frames * V));
    WC(walloc(w, &w->d_mask, V));
    WC(walloc(w, &w->loss_partial, (size_t)frames * window_blocks(w->V)));
    WC(walloc(w, &w->del_o, V)); WC(walloc(w, &w->del_f, V)); WC(walloc(w, &w->dls_o, V)); WC(walloc(w, &w->dls_f, V));
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 3; ++k) WC(walloc(w, &w->lam[b][k], V));
    WC(walloc(w, &w->acc, 8));
    WC(walloc(w, &w->overflow_any, 1));
    WC(walloc(w, &w->adj_partial, (size_t)window_blocks(w->V) * 8));
    WC(walloc(w, &w->d_fc, (size_t)frames));
    WC(hipHostMalloc((void**)&w->h_fc, sizeof(FrameConsts) * frames, hipHostMallocDefault));
    WC(walloc(w, &w->d_gfc, (size_t)frames * K3DOM_MAX_LIDARS));
    WC(hipHostMalloc((void**)&w->h_gfc, sizeof(FrameConsts) * frames * K3DOM_MAX_LIDARS, hipHostMallocDefault));
#undef WC
    *out = w;
    return K3DOM_OK;
}

k3dom_status k3dom_window_destroy(k3dom_window* w) {
    if (!w) return K3DOM_OK;
    if (w->h) (void)hipStreamSynchronize(w->h->stream);
    if (w->h_fc) (void)hipHostFree(w->h_fc);
    if (w->h_gfc) (void)hipHostFree(w->h_gfc);
    k3dom_destroy(w->h);  // frees every device buffer of the window too
    delete w;
    return K3DOM_OK;
}

k3dom_status k3dom_window_set_frame(k3dom_window* w, int32_t f, double t, const float pose[3], const float* xs,
                                    const float* ys, const float* zs, int32_t n, const uint8_t* gt_occ) {
    if (!w) return K3DOM_ERR_INVALID_ARG;
    if (f < 0 || f >= w->T) return wfail(w, K3DOM_ERR_INVALID_ARG, "frame index out of range");
    if (!pose || n < 0 || (n > 0 && (!xs || !ys || !zs)) || !gt_occ)
        return wfail(w, K3DOM_ERR_INVALID_ARG, "bad frame arrays");
    // rebuild the host point store with this frame replaced
    std::vector<float> pts;
    pts.reserve(w->host_pts.size() + 3 * (size_t)n);
    for (int g = 0; g < w->T; ++g) {
        FrameIn& fi = w->in[g];
        const size_t off = pts.size();
        if (g == f) {
            pts.insert(pts.end(), xs, xs + n);
            pts.insert(pts.end(), ys, ys + n);
            pts.insert(pts.end(), zs, zs + n);
            fi.n = n;
            fi.t = t;
            for (int k = 0; k < 3; ++k) fi.pose[k] = pose[k];
            fi.set = true;
            fi.ngroups = 0;
        } else if (fi.set) {
            pts.insert(pts.end(), w->host_pts.begin() + fi.off, w->host_pts.begin() + fi.off + 3 * (size_t)fi.n);
        }
        fi.off = off;
    }
    w->host_pts.swap(pts);
    w->pts_dirty = true;
    W_HIP(w, hipMemcpy(w->d_gt + (size_t)f * w->V, gt_occ, (size_t)w->V, hipMemcpyHostToDevice));
    w->forward_done = false;
    return K3DOM_OK;
}

k3dom_status k3dom_window_set_frame_lidars(k3dom_window* w, int32_t f, double t, const float pose[3],
                                          const k3dom_lidar* lidars, int32_t n_lidars, const uint8_t* gt_occ) {
    if (!w) return K3DOM_ERR_INVALID_ARG;
    if (!lidars || n_lidars < 1 || n_lidars > K3DOM_MAX_LIDARS)
        return wfail(w, K3DOM_ERR_INVALID_ARG, "1 to K3DOM_MAX_LIDARS LiDARs");
    long long total = 0;
    for (int k = 0; k < n_lidars; ++k) {
        if (lidars[k].n < 0 || (lidars[k].n > 0 && (!lidars[k].xs || !lidars[k].ys || !lidars[k].zs)))
            return wfail(w, K3DOM_ERR_INVALID_ARG, "bad point arrays");
        total += lidars[k].n;
    }
    if (total > 0x7fffffff) return wfail(w, K3DOM_ERR_INVALID_ARG, "too many points");
    // the frame's points rotated into map axes, LiDAR after LiDAR (the same
    // float expression as the ingest kernel)
    const int n = (int)total;
    std::vector<float> x(n), y(n), z(n);
    int at = 0;
    for (int k = 0; k < n_lidars; ++k) {
        const k3dom_lidar& L = lidars[k];
        for (int q = 0; q < L.n; ++q) k3_lidar_rotate(L.rot, L.xs[q], L.ys[q], L.zs[q], &x[at + q], &y[at + q], &z[at + q]);
        at += L.n;
    }
    k3dom_status s = k3dom_window_set_frame(w, f, t, pose, x.data(), y.data(), z.data(), n, gt_occ);
    if (s != K3DOM_OK) return s;
    FrameIn& fi = w->in[f];
    fi.ngroups = n_lidars;
    for (int k = 0; k < n_lidars; ++k) {
        fi.gcount[k] = lidars[k].n;
        for (int i = 0; i < 3; ++i) fi.gorig[k][i] = lidars[k].origin[i];
    }
    if (n_lidars > 1) {
        const size_t V = (size_t)w->V;
        // each allocated while missing: a failed allocation leaves the others in place
        for (float** p : {&w->g_o, &w->g_f, &w->g_lo, &w->g_lf, &w->h->gdel_occ, &w->h->gdel_free})
            if (!*p) W_HIP(w, walloc(w, p, V));
    }
    return K3DOM_OK;
}

k3dom_status k3dom_window_set_mask(k3dom_window* w, const uint8_t* mask) {
    if (!w) return K3DOM_ERR_INVALID_ARG;
    if (!mask) {
        w->use_mask = false;
        w->mask_count = w->V;
    } else {
        long long c = 0;
        for (int v = 0; v < w->V; ++v) c += mask[v] != 0;
        if (c == 0) return wfail(w, K3DOM_ERR_INVALID_ARG, "mask selects no voxel");
        W_HIP(w, hipMemcpy(w->d_mask, mask, (size_t)w->V, hipMemcpyHostToDevice));
        w->use_mask = true;
        w->mask_count = c;
    }
    w->forward_done = false;
    return K3DOM_OK;
}

k3dom_status k3dom_window_forward(k3dom_window* w, const float theta[6], double* loss) {
    if (!w || !theta || !loss) return K3DOM_ERR_INVALID_ARG;
    for (int f = 0; f < w->T; ++f)
        if (!w->in[f].set) return wfail(w, K3DOM_ERR_STATE, "every frame must be set before a forward pass");
    if (!(theta[0] >= 1e-6f) || !(theta[1] > 0.0f) || !(theta[2] > 0.0f))
        return wfail(w, K3DOM_ERR_INVALID_ARG, "theta: ls >= 1e-6 and sigmas > 0 required");
    if (!w->use_mask) w->mask_count = w->V;
    k3dom_handle* h = w->h;
    const double t0 = now_ms();
    k3dom_status s = upload_points(w);
    if (s != K3DOM_OK) return s;
    for (int k = 0; k < 6; ++k) w->theta[k] = theta[k];
    h->p.ls = theta[0];
    h->sigma_occ = theta[1];
    h->sigma_free = theta[2];
    h->p.prior_free = theta[3];
    h->p.prior_static = theta[4];
    h->p.prior_dynamic = theta[5];
    if ((s = from_handle(w, reset_map(h))) != K3DOM_OK) return s;
    const int nb = window_blocks(w->V);
    for (int f = 0; f < w->T; ++f) {
        const FrameIn& fi = w->in[f];
        FrameOut& fo = w->out[f];
        if ((s = from_handle(w, update_pose(h, fi.pose[0], fi.pose[1], fi.pose[2], fo.move))) != K3DOM_OK) return s;
        fo.normal = h->updated_time > 0;
        fo.decay = fo.normal ? (float)std::pow((double)h->p.gamma, (double)(float)(fi.t - h->updated_time)) : 1.0f;
        for (int k = 0; k < 3; ++k) {
            fo.sensor[k] = h->sensor[k];
            fo.center[k] = h->center[k];
        }
        if ((s = from_handle(w, ensure_points(h, std::max(fi.n, 1)))) != K3DOM_OK) return s;
        const float* px = w->d_pts + fi.off;
        h->ngroups = fi.ngroups;
        for (int g = 0, at = 0; g < fi.ngroups; at += fi.gcount[g], ++g) {
            h->gstart[g] = at;
            h->gcount[g] = fi.gcount[g];
            for (int k = 0; k < 3; ++k) h->gorig[g][k] = fi.gorig[g][k];
        }
        s = from_handle(w, frame(h, fi.t, px, px + fi.n, px + 2 * fi.n, fi.n));
        h->ngroups = 0;
        if (s != K3DOM_OK) return s;
        w->h_fc[f] = *h->h_fc;
        for (int g = 0; g < fi.ngroups; ++g) w->h_gfc[(size_t)f * K3DOM_MAX_LIDARS + g] = h->h_gfc[g];
        W_HIP(w, launch_window_record(h->stream, w->V, h->gs, h->gsz, fo.normal, h->p.particle_min_vel,
                                      h->vox[h->vcur], h->P[1], h->N, h->NB, h->B, w->rec[f],
                                      w->d_gt + (size_t)f * w->V, w->use_mask ? w->d_mask : nullptr,
                                      w->loss_partial + (size_t)f * nb));
    }
    std::vector<double> lp((size_t)w->T * nb);
    W_HIP(w, hipMemcpyAsync(lp.data(), w->loss_partial, sizeof(double) * lp.size(), hipMemcpyDeviceToHost, h->stream));
    W_HIP(w, hipStreamSynchronize(h->stream));
    double total = 0.0;
    for (double x : lp) total += x;
    *loss = total / ((double)w->T * (double)w->mask_count);
    w->forward_done = true;
    w->forward_id++;
    w->ms[0] = (float)(now_ms() - t0);
    return K3DOM_OK;
}

k3dom_status k3dom_window_backward(k3dom_window* w, double dloss, double grad[6], float* grad_prior_voxels) {
    if (!w || !grad) return K3DOM_ERR_INVALID_ARG;
    if (!w->forward_done) return wfail(w, K3DOM_ERR_STATE, "backward needs a forward pass first");
    k3dom_handle* h = w->h;
    const double t0 = now_ms();
    k3dom_status s;
    W_HIP(w, launch_window_clear(h->stream, w->acc, 8));
    W_HIP(w, launch_flag_clear(h->stream, w->overflow_any));
    const double loss_scale = dloss / ((double)w->T * (double)w->mask_count);
    int cur = 0;  // lam[cur] receives this frame's pre-state adjoint; lam[1 - cur] holds the next frame's
    for (int f = w->T - 1; f >= 0; --f) {
        const FrameIn& fi = w->in[f];
        const FrameOut& fo = w->out[f];
        // evidence of frame f and its length-scale derivative, from the frame's constants
        const FrameConsts* fc = w->d_fc + f;
        W_HIP(w, hipMemcpyAsync(w->d_fc + f, w->h_fc + f, sizeof(FrameConsts), hipMemcpyHostToDevice, h->stream));
        W_HIP(w, launch_reset_scalars(h->stream, h->sc));
        const float* px = w->d_pts + fi.off;
        if (fi.ngroups == 0) {
            if ((s = from_handle(w, build_lists(h, fc, px, px + fi.n, px + 2 * fi.n))) != K3DOM_OK) return s;
            W_HIP(w, launch_bgki_grad(h->stream, h->gather_blocks, fc, h->tg, h->p.ls, h->sigma_occ, h->sigma_free,
                                      h->ibuf, h->lbuf[2], h->rec, w->del_o, w->del_f, w->dls_o, w->dls_f, h->sc,
                                      h->gather_waves));
            W_HIP(w, launch_bgki_combine_grad(h->stream, h->ntiles, w->h_fc[f].g, h->tg, h->ibuf, h->p.ls,
                                              h->sigma_occ, h->sigma_free, w->del_o, w->del_f, w->dls_o, w->dls_f,
                                              h->sc));
        } else {
            // one pass per LiDAR with the sensor at its origin, added in LiDAR order
            const size_t g0 = (size_t)f * K3DOM_MAX_LIDARS;
            W_HIP(w, hipMemcpyAsync(w->d_gfc + g0, w->h_gfc + g0, sizeof(FrameConsts) * fi.ngroups,
                                    hipMemcpyHostToDevice, h->stream));
            for (int g = 0, at = 0; g < fi.ngroups; at += fi.gcount[g], ++g) {
                const FrameConsts* gfc = w->d_gfc + g0 + g;
                const float* gx = px + at;
                if ((s = from_handle(w, build_lists(h, gfc, gx, gx + fi.n, gx + 2 * fi.n))) != K3DOM_OK) return s;
                float* to = g == 0 ? w->del_o : w->g_o;
                float* tf = g == 0 ? w->del_f : w->g_f;
                float* tlo = g == 0 ? w->dls_o : w->g_lo;
                float* tlf = g == 0 ? w->dls_f : w->g_lf;
                W_HIP(w, launch_bgki_grad(h->stream, h->gather_blocks, gfc, h->tg, h->p.ls, h->sigma_occ,
                                          h->sigma_free, h->ibuf, h->lbuf[2], h->rec, to, tf, tlo, tlf, h->sc,
                                          h->gather_waves));
                W_HIP(w, launch_bgki_combine_grad(h->stream, h->ntiles, w->h_gfc[g0 + g].g, h->tg, h->ibuf, h->p.ls,
                                                  h->sigma_occ, h->sigma_free, to, tf, tlo, tlf, h->sc));
                if (g > 0) {
                    W_HIP(w, launch_add_evidence(h->stream, &h->sc->overflow, w->del_o, w->del_f, w->g_o, w->g_f,
                                                 w->V));
                    W_HIP(w, launch_add_evidence(h->stream, &h->sc->overflow, w->dls_o, w->dls_f, w->g_lo, w->g_lf,
                                                 w->V));
                }
            }
        }
        // the next frame's scalar reset clears the flag: keep it
        W_HIP(w, launch_flag_or(h->stream, w->overflow_any, &h->sc->overflow));
        AdjointFrameArgs a{};
        a.V = w->V;
        a.gs = h->gs;
        a.gsz = h->gsz;
        a.post = w->rec[f];
        if (f > 0) a.prev = w->rec[f - 1];
        for (int k = 0; k < 3; ++k) {
            a.move[k] = fo.move[k];
            a.move_next[k] = f + 1 < w->T ? w->out[f + 1].move[k] : 0;
        }
        if (f + 1 < w->T) {
            a.lam_next_f = w->lam[1 - cur][0];
            a.lam_next_s = w->lam[1 - cur][1];
            a.lam_next_d = w->lam[1 - cur][2];
        }
        a.lam_f = w->lam[cur][0];
        a.lam_s = w->lam[cur][1];
        a.lam_d = w->lam[cur][2];
        a.del_occ = w->del_o;
        a.del_free = w->del_f;
        a.dls_occ = w->dls_o;
        a.dls_free = w->dls_f;
        a.gt = w->d_gt + (size_t)f * w->V;
        a.mask = w->use_mask ? w->d_mask : nullptr;
        a.loss_scale = loss_scale;
        a.prior_f = h->p.prior_free;
        a.prior_s = h->p.prior_static;
        a.prior_d = h->p.prior_dynamic;
        a.decay = fo.decay;
        a.mass_scale = h->p.mass_scale;
        a.p_B = h->p.birth_prob;
        a.sigma_occ = h->sigma_occ;
        a.sigma_free = h->sigma_free;
        a.w_coef = h->p.persistence_prob / (float)h->N;
        a.has_w = f >= 1;
        a.has_w_prev = f >= 2;
        W_HIP(w, launch_window_adjoint(h->stream, a, w->acc, w->adj_partial));
        cur = 1 - cur;
    }
    double acc[8];
    uint32_t overflow = 0;
    W_HIP(w, hipMemcpyAsync(acc, w->acc, sizeof(acc), hipMemcpyDeviceToHost, h->stream));
    W_HIP(w, hipMemcpyAsync(&overflow, w->overflow_any, sizeof(overflow), hipMemcpyDeviceToHost, h->stream));
    std::vector<double> pv;
    if (grad_prior_voxels) {
        // the adjoint of frame 0's pre-state sits in lam[1 - cur] after the last swap
        pv.resize(3 * (size_t)w->V);
        for (int k = 0; k < 3; ++k)
            W_HIP(w, hipMemcpyAsync(pv.data() + (size_t)k * w->V, w->lam[1 - cur][k], sizeof(double) * w->V,
                                    hipMemcpyDeviceToHost, h->stream));
    }
    W_HIP(w, hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < pv.size(); ++i) grad_prior_voxels[i] = (float)pv[i];
    if (overflow) return wfail(w, K3DOM_ERR_STATE, "candidate lists overflowed in the backward pass");
    for (int k = 0; k < 6; ++k) grad[k] = acc[1 + k];
    w->ms[1] = (float)(now_ms() - t0);
    return K3DOM_OK;
}

k3dom_status k3dom_window_forward_id(const k3dom_window* w, uint64_t* id) {
    if (!w || !id) return K3DOM_ERR_INVALID_ARG;
    *id = w->forward_id;
    return K3DOM_OK;
}

k3dom_status k3dom_window_read_record(k3dom_window* w, int32_t f, int32_t which, void* dst, int64_t capacity) {
    if (!w || !dst) return K3DOM_ERR_INVALID_ARG;
    if (f < 0 || f >= w->T) return wfail(w, K3DOM_ERR_INVALID_ARG, "frame index out of range");
    if (!w->forward_done) return wfail(w, K3DOM_ERR_STATE, "no forward pass recorded");
    const size_t V = (size_t)w->V;
    const WindowFrame& r = w->rec[f];
    auto cp = [&](const void* src, size_t bytes, size_t at) -> k3dom_status {
        if ((int64_t)(at + bytes) > capacity) return wfail(w, K3DOM_ERR_INVALID_ARG, "capacity too small");
        W_HIP(w, hipMemcpy((char*)dst + at, src, bytes, hipMemcpyDeviceToHost));
        return K3DOM_OK;
    };
    k3dom_status s = K3DOM_OK;
    switch (which) {
    case 0:
        if ((s = cp(r.f, V * 4, 0)) || (s = cp(r.s, V * 4, V * 4)) || (s = cp(r.d, V * 4, 2 * V * 4))) return s;
        break;
    case 1: s = cp(r.dyn, V * 4, 0); break;
    case 2:
        if ((s = cp(r.n, V * 4, 0)) || (s = cp(r.m, V * 4, V * 4)) || (s = cp(r.nb, V * 4, 2 * V * 4)) ||
            (s = cp(r.mb, V * 4, 3 * V * 4)))
            return s;
        break;
    case 3:
        if (capacity < 12) return wfail(w, K3DOM_ERR_INVALID_ARG, "capacity too small");
        std::memcpy(dst, w->out[f].move, 12);
        break;
    default: return wfail(w, K3DOM_ERR_INVALID_ARG, "unknown record id");
    }
    return s;
}

k3dom_status k3dom_window_times(const k3dom_window* w, float ms[2]) {
    if (!w || !ms) return K3DOM_ERR_INVALID_ARG;
    ms[0] = w->ms[0];
    ms[1] = w->ms[1];
    return K3DOM_OK;
}

const char* k3dom_window_last_error(const k3dom_window* w) { return w ? w->err.c_str() : ""; }

}  // extern "C"
