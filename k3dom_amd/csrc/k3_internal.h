// Internal state of a k3dom_handle and the frame-level helpers shared by the
// C ABI (k3_api.cpp) and the differentiated-window driver (k3_window.cpp).
#pragma once

#include "../../include/k3dom.h"
#include "k3_kernels.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

using namespace k3;

enum { ST_PREDICT = 0, ST_SORT, ST_LISTS, ST_BGKI, ST_MASS, ST_PERSIST, ST_NEWBORN, ST_MOMENTS, ST_RESAMPLE,
       ST_FRAME, ST_BGKI_KERNEL, ST_COUNT };
// extra events: frame start, end of the gather kernel
enum { EV_FRAME_START = ST_COUNT, EV_BGKI_END, EV_COUNT };

struct k3dom_handle {
    k3dom_params p{};
    // evidence weights of the occupied and free terms (both p.sigma unless a
    // differentiated window sets them apart)
    float sigma_occ = 0.0f, sigma_free = 0.0f;
    int gs = 0, gsz = 0, V = 0, N = 0, B = 0, key_bits = 1;
    int ntx = 0, nty = 0, ntz = 0, ntiles = 0;
    int device = 0, num_cus = 256;
    hipStream_t stream = nullptr;
    // point uploads of eager frames run on cstream and are waited for (event
    // pts_ready) just before the points are first read, after predict and sort
    hipStream_t cstream = nullptr;
    hipEvent_t pts_ready = nullptr;
    bool pts_pending = false;

    std::vector<void*> allocs;
    VoxelArrays vox[2]{};
    int vcur = 0;
    int method = K3DOM_METHOD_K3DOM;  // DS-PHD/MIB maps keep the occupied mass in the s slot
    // A map shift moves the SHIFT_KEPT fields at once; the SHIFT_REWRITTEN ones
    // of vox[vcur] stay due from vox[1 - vcur] moved by pend_move until a full
    // frame rewrites them or something reads them (settle_shift).
    bool shift_pending = false;
    int pend_move[3] = {0, 0, 0};
    // The particle part of a map shift waits for the next frame's predict
    // (FrameConsts::pshift), or for a read of the particles (settle_particles).
    bool pshift_pending = false;
    int pshift[3] = {0, 0, 0};
    float *born = nullptr, *del_occ = nullptr, *del_free = nullptr;
    ParticleArrays P[2]{};  // 0: current set, 1: sorted set
    ParticleArrays NB{};
    SortScratch sort{};
    float* wsort = nullptr;  // N: sorted particle weights before the persistent update
    float4* prec = nullptr;  // 2N: predicted particle state, one 32-byte record per particle
    float* usum = nullptr;   // 3N: per-unit sums of the particle-parallel voxel sums
    uint32_t* maxparts = nullptr;  // per-block maxima feeding the fixed-point CDF exponents
    float* ds_norm = nullptr;      // V (DS-PHD/MIB): sums above 1 of the voxels' weights, else 0
    DsLists ds{};                  // DS-PHD/MIB: measurement-grid candidate lists
    // MapPoint export (allocated at the first export): 3 float4 per voxel,
    // visibility flags, their scan and its scratch
    float4* mp_out = nullptr;
    // LiDAR groups of the current frame (k3dom_update_grid_lidars; 0: one
    // sensor at the pose): point ranges, origins, per-group frame constants
    // (device array and pinned staging), and the evidence of groups after the first
    int ngroups = 0;
    int gstart[K3DOM_MAX_LIDARS] = {}, gcount[K3DOM_MAX_LIDARS] = {};
    float gorig[K3DOM_MAX_LIDARS][3] = {};
    FrameConsts *gfc = nullptr, *h_gfc = nullptr;
    float *gdel_occ = nullptr, *gdel_free = nullptr;  // (the differentiated window's per-LiDAR passes)
    // Kernel-inference buffers of LiDAR groups 1.. (group 0 uses lbuf, ibuf
    // and sc): every group builds its lists and work items into its own
    // buffers, and one persistent gather launch covers all groups' items.
    struct GroupKi {
        LevelBufs lbuf[3]{};
        ItemBufs ibuf{};
        float *del_occ = nullptr, *del_free = nullptr;
    };
    GroupKi gk[K3DOM_MAX_LIDARS]{};
    int ngk = 1;                                      // groups with buffers (group 0 always)
    DevScalars* gsc = nullptr;                        // K3DOM_MAX_LIDARS scalar blocks ([0] unused)
    BgArgs *d_gargs = nullptr, *h_gargs = nullptr;    // the joint gather's argument sets
    uint32_t* d_gq = nullptr;                         // its item prefix (K3DOM_MAX_LIDARS + 1) and queue head
    int gstride = 0;                                  // point-buffer slot of each LiDAR (update_grid_lidars)
    bool last_grouped = false;  // the last frame ran per-LiDAR kernel inference
    uint32_t *mp_flags = nullptr, *mp_off = nullptr, *mp_scratch = nullptr;
    const uint32_t* last_perm = nullptr;
    const uint32_t* last_skeys = nullptr;
    uint64_t *cdf_v = nullptr, *cdf_j = nullptr, *blocksums = nullptr;  // blocksums: fixed-point scan scratch
    uint32_t* bnd = nullptr;
    int32_t* ridx = nullptr;
    DevScalars* sc = nullptr;

    int m_cap = 0;
    float* d_pts = nullptr;
    K3PointRec* rec = nullptr;
    float4* pxyz = nullptr;
    float* h_pts = nullptr;

    TileGeom tg{};
    LevelDesc ldesc[3]{};
    LevelBufs lbuf[3]{};
    ItemBufs ibuf{};
    bool tiny_caps = false;  // test hook: start every capacity small to exercise the regrowth path
    // test hook (K3DOM_TEST_FAIL_ALLOC=n at creation): the n-th device
    // allocation after creation reports out-of-memory
    int fail_alloc_at = 0, alloc_count = 0;
    int overflow_retries = 0;
    DevScalars* h_sc = nullptr;  // pinned mirror

    uint32_t frame = 0;
    // frames (and normal frames) this handle has processed since it was
    // created, reset or restored from a checkpoint: the lists and per-frame
    // arrays are its own only after one
    uint32_t frames_here = 0, normal_here = 0;
    // the last frame completed (its lists and work items are whole); false
    // after a frame that returned an error
    bool last_frame_ok = false;
    int stage = 0;  // the stage methods' position in the frame (0: between frames)
    bool first_pose = false;
    float sensor[3] = {0, 0, 0}, center[3] = {0, 0, 0};
    double updated_time = -1.0;
    std::string err;

    // frame constants (device copy read by the kernels, pinned host staging)
    FrameConsts* fc = nullptr;
    FrameConsts* h_fc = nullptr;
    int gather_blocks = 0;     // blocks of the kernel-inference gather (autotuned)
    int gather_waves = 4;      // wavefronts per gather block (autotuned: 4, 8 or 16)
    // captured normal-frame graphs, one per voxel buffer (vcur), each
    // re-captured when buffers move
    bool use_graph = false;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int caps_version = 0;
    int graph_caps[2] = {-1, -1};
    // the LiDAR layout each graph was captured for (0 groups: one sensor)
    int graph_ngroups[2] = {0, 0};
    int graph_gstart[2][K3DOM_MAX_LIDARS] = {};

    int timing = 0;  // 0 off, 1 every stage, 2 the kernel-inference gather only
    hipEvent_t ev[EV_COUNT]{};
    k3dom_stats stats{};
    GridGeom last_geom{};
    unsigned long long* pair_counts = nullptr;
    float stage_ms[ST_COUNT]{};
};

namespace k3api {

k3dom_status fail(k3dom_handle* h, k3dom_status s, const std::string& msg);

#define HIP_TRY(h, expr)                                                                          \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return k3api::fail((h), _e == hipErrorOutOfMemory ? K3DOM_ERR_OUT_OF_MEMORY : K3DOM_ERR_HIP, \
                               std::string(#expr) + ": " + hipGetErrorString(_e));                \
    } while (0)

template <class T>
hipError_t dalloc(k3dom_handle* h, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = sizeof(T) * (count ? count : 1);
    if (h->fail_alloc_at > 0 && ++h->alloc_count == h->fail_alloc_at) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(&q, bytes);
    if (e == hipSuccess) {
        h->allocs.push_back(q);
        *p = static_cast<T*>(q);
    }
    return e;
}
// dalloc, then zero the memory on the map's stream (ahead of its later work)
template <class T>
hipError_t dalloc_zero(k3dom_handle* h, T** p, size_t count) {
    hipError_t e = dalloc(h, p, count);
    if (e == hipSuccess) e = hipMemsetAsync(*p, 0, sizeof(T) * (count ? count : 1), h->stream);
    return e;
}
void dfree(k3dom_handle* h, void* p);
// Two timing events destroyed on every exit path of their scope.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t create() {
        hipError_t e = hipEventCreate(&e0);
        return e == hipSuccess ? hipEventCreate(&e1) : e;
    }
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};
k3dom_status ensure_points(k3dom_handle* h, int m);
// True when a buffer sized from the point capacity is missing (an allocation
// of an earlier frame failed part way).
bool point_buffers_missing(const k3dom_handle* h);
GridGeom grid_geom(const k3dom_handle* h);
// One filter frame on device-resident points (DOM::updateGrid).
k3dom_status frame(k3dom_handle* h, double t, const float* dx, const float* dy, const float* dz, int m);
// Candidate lists of the points of frame constants fc (point preparation, the
// three list levels, work items); the gather itself is launched by the caller.
k3dom_status build_lists(k3dom_handle* h, const FrameConsts* fc, const float* xs, const float* ys, const float* zs);
// Pieces of a frame for drivers that interleave several maps: the head
// (constants upload, prediction, sort, ranges), the part after the gather,
// overflow handling with re-runs, and the bookkeeping at the end.
k3dom_status enqueue_head(k3dom_handle* h, bool normal);
k3dom_status enqueue_after_gather(k3dom_handle* h, bool normal);
k3dom_status resolve_overflow(k3dom_handle* h, bool normal, const float* dx, const float* dy, const float* dz,
                              bool done, bool checked);
void finish_frame(k3dom_handle* h, double t, int m);
// The counters and statistics of a completed frame (finish_frame without the time).
void count_frame(k3dom_handle* h, bool normal, int m);
// Host copy of the frame constants of a frame at time t with m points, or of
// a frame of time step dt (normal) or an initialisation frame (dt 0).
void fill_frame_consts(k3dom_handle* h, FrameConsts* c, double t, int m);
void fill_frame_consts_dt(k3dom_handle* h, FrameConsts* c, bool normal, float dt, int m);
// Stages of the frame tail (stage order of DOM::updateGrid) and the growth
// of overflowed kernel-inference capacities before a re-run.
k3dom_status tail_mass(k3dom_handle* h, bool normal);
k3dom_status tail_persistent(k3dom_handle* h);
k3dom_status tail_newborn(k3dom_handle* h, bool normal);
k3dom_status tail_resample(k3dom_handle* h);
k3dom_status grow_after_overflow(k3dom_handle* h);
void drop_graph(k3dom_handle* h);
// State after construction (priors, no pose, frame count 0).
k3dom_status reset_map(k3dom_handle* h);
// DOM_c::updatePose; *shifted receives the map shift in voxels (0 when none).
k3dom_status update_pose(k3dom_handle* h, float x, float y, float z, int shifted[3]);
// Completes a deferred map shift (before reads of the voxel state and before
// frames that do not rewrite every voxel).
k3dom_status settle_shift(k3dom_handle* h);
// Prior of the second mass slot: static mass (K3DOM) or occupied mass (DS-PHD/MIB).
inline float prior_s_slot(const k3dom_handle* h) {
    return h->method == K3DOM_METHOD_DSPHDMIB ? h->p.prior_occ : h->p.prior_static;
}
// Applies a pending particle shift (before the particle arrays are read).
k3dom_status settle_particles(k3dom_handle* h);
// Voxel state at the start of a frame: a full frame makes a deferred shift moot.
k3dom_status begin_frame_voxels(k3dom_handle* h, bool normal);
// Whether the next frame runs as a captured graph.
bool frame_uses_graph(const k3dom_handle* h);
// Looks up the allocation-free export's device scratch on a GPU (once per
// device; false when the lookup failed, which the export then reports).
bool noalloc_prepare(int device);

}  // namespace k3api
