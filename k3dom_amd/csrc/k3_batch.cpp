// Many independent maps ("tiles") updated together (declarations in
// include/k3dom.h; no counterpart in the reference, which runs one map per
// process).  Each map keeps its own buffers and stream; a batch frame runs
// every map's prediction, sort and candidate lists concurrently on the map
// streams, then the kernel-inference gather of all maps in ONE persistent
// launch over a joint work queue, then every map's tail concurrently.  The
// normal batch frame can be captured once as a HIP graph (fork/join over the
// map streams) and replayed.  Every map's results are those of the same map
// updated alone.
#include "k3_internal.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

using namespace k3api;

struct k3dom_batch {
    std::vector<k3dom_handle*> maps;
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr, gathered = nullptr;
    std::vector<hipEvent_t> head_done, tail_done;
    BgArgs* d_args = nullptr;   // device argument sets of the joint gather
    BgArgs* h_args = nullptr;   // pinned staging
    uint32_t* d_prefix = nullptr;
    uint32_t* d_qheads = nullptr;  // work-queue heads of the joint gather (BG_QWORDS)
    void* d_mem = nullptr;
    int gather_blocks = 0;
    int gather_waves = 4;
    int num_cus = 256;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<int> graph_caps, graph_vcur;
    std::vector<int> m_now;     // points of each map in the current frame
    float ms_last = 0.0f;
    std::string err;
};

namespace {

k3dom_status bfail(k3dom_batch* b, k3dom_status s, const std::string& msg) {
    if (b) b->err = msg;
    return s;
}

#define B_HIP(b, expr)                                                                              \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) return bfail((b), K3DOM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

k3dom_status from_map(k3dom_batch* b, int i, k3dom_status s) {
    if (s != K3DOM_OK) b->err = "map " + std::to_string(i) + ": " + b->maps[i]->err;
    return s;
}

void drop_batch_graph(k3dom_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

const float* pts(k3dom_handle* h, int k) { return h->d_pts + (size_t)k * h->m_cap; }

// Enqueues one normal batch frame over every map (points already staged in
// each map's pinned buffer, constants in each map's pinned h_fc).
k3dom_status enqueue_batch(k3dom_batch* b) {
    const int n = (int)b->maps.size();
    B_HIP(b, hipEventRecord(b->fork, b->main));
    for (int i = 0; i < n; ++i) {
        k3dom_handle* h = b->maps[i];
        B_HIP(b, hipStreamWaitEvent(h->stream, b->fork, 0));
        const size_t cap = (size_t)h->m_cap;
        B_HIP(b, hipMemcpyAsync(h->d_pts, h->h_pts, sizeof(float) * 3 * cap, hipMemcpyHostToDevice, h->stream));
        k3dom_status s = enqueue_head(h, true);
        if (s == K3DOM_OK) s = build_lists(h, h->fc, pts(h, 0), pts(h, 1), pts(h, 2));
        if (s != K3DOM_OK) return from_map(b, i, s);
        B_HIP(b, hipEventRecord(b->head_done[i], h->stream));
        B_HIP(b, hipStreamWaitEvent(b->main, b->head_done[i], 0));
    }
    B_HIP(b, hipMemcpyAsync(b->d_args, b->h_args, sizeof(BgArgs) * n, hipMemcpyHostToDevice, b->main));
    B_HIP(b, launch_bgki_batch(b->main, b->gather_blocks, b->d_args, n, b->d_prefix, b->d_qheads, b->gather_waves));
    B_HIP(b, hipEventRecord(b->gathered, b->main));
    for (int i = 0; i < n; ++i) {
        k3dom_handle* h = b->maps[i];
        B_HIP(b, hipStreamWaitEvent(h->stream, b->gathered, 0));
        k3dom_status s = enqueue_after_gather(h, true);
        if (s != K3DOM_OK) return from_map(b, i, s);
        B_HIP(b, hipEventRecord(b->tail_done[i], h->stream));
        B_HIP(b, hipStreamWaitEvent(b->main, b->tail_done[i], 0));
    }
    return K3DOM_OK;
}

k3dom_status capture_batch(k3dom_batch* b) {
    drop_batch_graph(b);
    B_HIP(b, hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
    k3dom_status s = enqueue_batch(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->main, &g);
    b->graph = g;
    if (s != K3DOM_OK) {
        drop_batch_graph(b);
        return s;
    }
    B_HIP(b, e);
    B_HIP(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    const int n = (int)b->maps.size();
    b->graph_caps.assign(n, 0);
    b->graph_vcur.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        b->graph_caps[i] = b->maps[i]->caps_version;
        b->graph_vcur[i] = b->maps[i]->vcur;
    }
    return K3DOM_OK;
}

bool graph_current(const k3dom_batch* b) {
    if (!b->gexec || b->graph_caps.size() != b->maps.size()) return false;
    for (size_t i = 0; i < b->maps.size(); ++i)
        if (b->graph_caps[i] != b->maps[i]->caps_version || b->graph_vcur[i] != b->maps[i]->vcur) return false;
    return true;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

k3dom_status k3dom_batch_create(const k3dom_params* params, int32_t n_maps, k3dom_batch** out) {
    if (!out || !params || n_maps <= 0) return K3DOM_ERR_INVALID_ARG;
    *out = nullptr;
    k3dom_batch* b = new (std::nothrow) k3dom_batch();
    if (!b) return K3DOM_ERR_OUT_OF_MEMORY;
    for (int i = 0; i < n_maps; ++i) {
        k3dom_handle* h = nullptr;
        k3dom_status s = k3dom_create(params + i, &h);
        if (s != K3DOM_OK) {
            k3dom_batch_destroy(b);
            return s;
        }
        b->maps.push_back(h);
    }
    b->num_cus = b->maps[0]->num_cus;
    b->gather_blocks = b->num_cus * 8;
    b->head_done.assign(n_maps, nullptr);
    b->tail_done.assign(n_maps, nullptr);
    b->m_now.assign(n_maps, 0);
    auto bail = [&](hipError_t e) {
        b->err = hipGetErrorString(e);
        k3dom_batch_destroy(b);
        return K3DOM_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking)) != hipSuccess) return bail(e);
    if ((e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess) return bail(e);
    if ((e = hipEventCreateWithFlags(&b->gathered, hipEventDisableTiming)) != hipSuccess) return bail(e);
    for (int i = 0; i < n_maps; ++i) {
        if ((e = hipEventCreateWithFlags(&b->head_done[i], hipEventDisableTiming)) != hipSuccess) return bail(e);
        if ((e = hipEventCreateWithFlags(&b->tail_done[i], hipEventDisableTiming)) != hipSuccess) return bail(e);
    }
    const size_t bytes = sizeof(BgArgs) * n_maps + sizeof(uint32_t) * (n_maps + 1 + BG_QWORDS);
    if ((e = hipMalloc(&b->d_mem, bytes)) != hipSuccess) return bail(e);
    b->d_args = (BgArgs*)b->d_mem;
    b->d_prefix = (uint32_t*)((char*)b->d_mem + sizeof(BgArgs) * n_maps);
    b->d_qheads = b->d_prefix + n_maps + 1;
    if ((e = hipHostMalloc((void**)&b->h_args, sizeof(BgArgs) * n_maps, hipHostMallocDefault)) != hipSuccess)
        return bail(e);
    *out = b;
    return K3DOM_OK;
}

k3dom_status k3dom_batch_destroy(k3dom_batch* b) {
    if (!b) return K3DOM_OK;
    if (b->main) (void)hipStreamSynchronize(b->main);
    drop_batch_graph(b);
    for (k3dom_handle* h : b->maps) k3dom_destroy(h);
    for (hipEvent_t ev : b->head_done)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : b->tail_done)
        if (ev) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->gathered) (void)hipEventDestroy(b->gathered);
    if (b->d_mem) (void)hipFree(b->d_mem);
    if (b->h_args) (void)hipHostFree(b->h_args);
    if (b->main) (void)hipStreamDestroy(b->main);
    delete b;
    return K3DOM_OK;
}

k3dom_status k3dom_batch_map(k3dom_batch* b, int32_t i, k3dom_handle** out) {
    if (!b || !out || i < 0 || i >= (int)b->maps.size()) return K3DOM_ERR_INVALID_ARG;
    *out = b->maps[i];
    return K3DOM_OK;
}

k3dom_status k3dom_batch_set_graph(k3dom_batch* b, int32_t enable) {
    if (!b) return K3DOM_ERR_INVALID_ARG;
    b->use_graph = enable != 0;
    if (!b->use_graph) drop_batch_graph(b);
    return K3DOM_OK;
}

k3dom_status k3dom_batch_update(k3dom_batch* b, double t, const float* poses, const float* const* xs,
                                const float* const* ys, const float* const* zs, const int32_t* n) {
    if (!b || !poses || !xs || !ys || !zs || !n) return K3DOM_ERR_INVALID_ARG;
    const int nm = (int)b->maps.size();
    const double t0 = now_ms();
    bool all_normal = true;
    for (int i = 0; i < nm; ++i) {
        if (n[i] < 0 || (n[i] > 0 && (!xs[i] || !ys[i] || !zs[i])))
            return bfail(b, K3DOM_ERR_INVALID_ARG, "bad point arrays of map " + std::to_string(i));
        all_normal = all_normal && b->maps[i]->updated_time > 0;
    }
    // poses (map shifts run at once on each map's stream), point staging, frame constants
    for (int i = 0; i < nm; ++i) {
        k3dom_handle* h = b->maps[i];
        int moved[3];
        k3dom_status s = update_pose(h, poses[3 * i], poses[3 * i + 1], poses[3 * i + 2], moved);
        if (s == K3DOM_OK) s = ensure_points(h, std::max(n[i], 1));
        if (s != K3DOM_OK) return from_map(b, i, s);
        // the batch frame runs on the batch's streams: let a map's shift finish first
        if (moved[0] || moved[1] || moved[2]) B_HIP(b, hipStreamSynchronize(h->stream));
        const size_t cap = (size_t)h->m_cap;
        if (n[i] > 0) {
            std::memcpy(h->h_pts, xs[i], sizeof(float) * n[i]);
            std::memcpy(h->h_pts + cap, ys[i], sizeof(float) * n[i]);
            std::memcpy(h->h_pts + 2 * cap, zs[i], sizeof(float) * n[i]);
        }
        b->m_now[i] = n[i];
    }
    if (!all_normal) {
        // a map in its first frame: every map runs its own frame
        for (int i = 0; i < nm; ++i) {
            k3dom_handle* h = b->maps[i];
            const size_t cap = (size_t)h->m_cap;
            B_HIP(b, hipMemcpyAsync(h->d_pts, h->h_pts, sizeof(float) * 3 * cap, hipMemcpyHostToDevice, h->stream));
            k3dom_status s = frame(h, t, pts(h, 0), pts(h, 1), pts(h, 2), n[i]);
            if (s != K3DOM_OK) return from_map(b, i, s);
        }
        b->ms_last = (float)(now_ms() - t0);
        return K3DOM_OK;
    }
    for (int i = 0; i < nm; ++i) {
        k3dom_handle* h = b->maps[i];
        k3dom_status s = begin_frame_voxels(h, true);
        if (s != K3DOM_OK) return from_map(b, i, s);
        fill_frame_consts(h, h->h_fc, t, n[i]);
        b->h_args[i] = make_bg_args(h->fc, h->tg, h->p.ls, h->sigma_occ, h->sigma_free, h->ibuf, h->lbuf[2], h->rec,
                                    h->del_occ, h->del_free, nullptr, nullptr, h->sc);
    }
    if (b->use_graph) {
        if (!graph_current(b)) {
            k3dom_status s = capture_batch(b);
            if (s != K3DOM_OK) return s;
        }
        B_HIP(b, hipGraphLaunch(b->gexec, b->main));
    } else {
        k3dom_status s = enqueue_batch(b);
        if (s != K3DOM_OK) return s;
    }
    B_HIP(b, hipStreamSynchronize(b->main));
    for (int i = 0; i < nm; ++i) {
        k3dom_handle* h = b->maps[i];
        k3dom_status s = resolve_overflow(h, true, pts(h, 0), pts(h, 1), pts(h, 2), !h->h_sc->overflow, true);
        if (s != K3DOM_OK) return from_map(b, i, s);
        finish_frame(h, t, n[i]);
    }
    b->ms_last = (float)(now_ms() - t0);
    return K3DOM_OK;
}

k3dom_status k3dom_batch_autotune(k3dom_batch* b, int32_t* blocks_per_cu) {
    if (!b) return K3DOM_ERR_INVALID_ARG;
    for (k3dom_handle* h : b->maps)
        if (h->normal_here < 1 || !h->last_frame_ok || point_buffers_missing(h))
            return bfail(b, K3DOM_ERR_STATE, "autotune needs a completed normal frame on every map first");
    const int n = (int)b->maps.size();
    // the last frame's lists are still in place: re-run the joint gather over
    // them for each launch shape (the gather only rewrites the same evidence)
    B_HIP(b, hipMemcpyAsync(b->d_args, b->h_args, sizeof(BgArgs) * n, hipMemcpyHostToDevice, b->main));
    EventPair ev;  // destroyed on every return below
    B_HIP(b, ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    // block size (wavefronts) and grid size, at most 32 wavefronts per compute unit
    float best = 1e30f;
    int best_c = b->gather_blocks / b->num_cus, best_w = b->gather_waves;
    for (int wv : {4, 8, 16})
        for (int c : {1, 2, 3, 4, 5, 6, 8}) {
            if (wv * c > 32) continue;
            float tot = 0.0f;
            for (int rep = 0; rep < 4; ++rep) {
                B_HIP(b, hipEventRecord(e0, b->main));
                B_HIP(b, launch_bgki_batch(b->main, c * b->num_cus, b->d_args, n, b->d_prefix, b->d_qheads, wv));
                B_HIP(b, hipEventRecord(e1, b->main));
                B_HIP(b, hipEventSynchronize(e1));
                float ms = 0.0f;
                B_HIP(b, hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) tot += ms;
            }
            if (tot < best) {
                best = tot;
                best_c = c;
                best_w = wv;
            }
        }
    b->gather_blocks = best_c * b->num_cus;
    b->gather_waves = best_w;
    drop_batch_graph(b);
    if (blocks_per_cu) *blocks_per_cu = best_c;
    return K3DOM_OK;
}

k3dom_status k3dom_batch_last_ms(const k3dom_batch* b, float* ms) {
    if (!b || !ms) return K3DOM_ERR_INVALID_ARG;
    *ms = b->ms_last;
    return K3DOM_OK;
}

const char* k3dom_batch_last_error(const k3dom_batch* b) { return b ? b->err.c_str() : ""; }

}  // extern "C"
