// C ABI of the K3DOM update on MI355X: handle lifetime, device buffers, the
// per-frame launch sequence and readback.  Declarations and the reference
// interface each entry point replaces are in include/k3dom.h.
#include "k3_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace k3;

namespace {
thread_local std::string g_create_error;
}  // namespace


namespace k3api {

k3dom_status fail(k3dom_handle* h, k3dom_status s, const std::string& msg) {
    if (h) h->err = msg;
    else g_create_error = msg;
    return s;
}

void dfree(k3dom_handle* h, void* p) {
    if (!p) return;
    auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
    if (it != h->allocs.end()) h->allocs.erase(it);
    (void)hipFree(p);
}

k3dom_status validate(const k3dom_params* p, std::string* why) {
    if (!p) { *why = "params is NULL"; return K3DOM_ERR_INVALID_ARG; }
    if (!(p->resolution > 0.0f)) { *why = "resolution must be > 0"; return K3DOM_ERR_INVALID_ARG; }
    int gs = (int)(p->size / p->resolution), gsz = (int)(p->size_z / p->resolution);
    if (gs < 1 || gsz < 1) { *why = "size and size_z must be >= resolution"; return K3DOM_ERR_INVALID_ARG; }
    if ((double)gs * gs * gsz > 2147483647.0 / 2) { *why = "grid too large"; return K3DOM_ERR_INVALID_ARG; }
    if (p->particle_count <= 0) { *why = "particle_count must be > 0"; return K3DOM_ERR_INVALID_ARG; }
    if (p->new_born_particle_count < 0) { *why = "new_born_particle_count must be >= 0"; return K3DOM_ERR_INVALID_ARG; }
    if (!(p->ls >= 1e-6f) || !std::isfinite(p->ls)) { *why = "ls must be >= 1e-6 and finite"; return K3DOM_ERR_INVALID_ARG; }
    if (!std::isfinite(p->sigma)) { *why = "sigma must be finite"; return K3DOM_ERR_INVALID_ARG; }
    if (!(p->mass_scale != 0.0f)) { *why = "mass_scale must be non-zero"; return K3DOM_ERR_INVALID_ARG; }
    return K3DOM_OK;
}

// Hierarchy of candidate-list boxes over the tile grid: level 0 boxes of
// 8x8x4 tiles (32x32x16 voxels; doubled in x/y until there are at most
// MAX_CHILDREN of them), level 1 of 2x2x2 tiles, level 2 single tiles.
void setup_levels(k3dom_handle* h) {
    TileGeom& tg = h->tg;
    int sizes[3][3] = {{8, 8, 4}, {2, 2, 2}, {1, 1, 1}};
    auto nb = [&](int t, int b) { return (t + b - 1) / b; };
    while (nb(tg.ntx, sizes[0][0]) * nb(tg.nty, sizes[0][1]) * nb(tg.ntz, sizes[0][2]) > MAX_CHILDREN) {
        sizes[0][0] *= 2;
        sizes[0][1] *= 2;
    }
    const int segs[3] = {256, 512, 512};
    for (int l = 0; l < 3; ++l) {
        LevelDesc& L = h->ldesc[l];
        L.bx = sizes[l][0]; L.by = sizes[l][1]; L.bz = sizes[l][2];
        L.nbx = nb(tg.ntx, L.bx); L.nby = nb(tg.nty, L.by); L.nbz = nb(tg.ntz, L.bz);
        L.seg = segs[l];
        if (l == 0) {
            L.nparents = 1;
            L.pnbx = L.pnby = 1;
            L.ncx = L.nbx; L.ncy = L.nby; L.ncz = L.nbz;
            L.nch = L.nbx * L.nby * L.nbz;
        } else {
            const LevelDesc& P = h->ldesc[l - 1];
            L.nparents = P.nbx * P.nby * P.nbz;
            L.pnbx = P.nbx; L.pnby = P.nby;
            L.ncx = P.bx / L.bx; L.ncy = P.by / L.by; L.ncz = P.bz / L.bz;
            L.nch = L.ncx * L.ncy * L.ncz;
        }
    }
}

k3dom_status ensure_caps(k3dom_handle* h, const long long want[3], long long want_heavy);

// DS-PHD/MIB measurement-grid candidate list of at least want entries, with
// the per-segment buffers its segment count implies.
k3dom_status ensure_ds_list(k3dom_handle* h, long long want) {
    if (want <= h->ds.cap && h->ds.list) return K3DOM_OK;
    for (void* p : {(void*)h->ds.list, (void*)h->ds.segcnt, (void*)h->ds.part}) dfree(h, p);
    h->ds.list = nullptr;
    h->ds.segcnt = nullptr;
    h->ds.part = nullptr;
    h->ds.cap = 0;
    h->caps_version++;
    const size_t slots = (size_t)ds_segments_max(h->gs, h->gsz, want, h->ds.seg_len) * 32 * 64;
    HIP_TRY(h, dalloc(h, &h->ds.list, (size_t)want));
    HIP_TRY(h, dalloc(h, &h->ds.segcnt, slots));
    HIP_TRY(h, dalloc(h, &h->ds.part, slots));
    h->ds.cap = want;
    return K3DOM_OK;
}

// Kernel-inference buffers and scalars of LiDAR group g (group 0: the handle's own).
LevelBufs* group_lbuf(k3dom_handle* h, int g) { return g == 0 ? h->lbuf : h->gk[g].lbuf; }
const LevelBufs* group_lbuf(const k3dom_handle* h, int g) { return g == 0 ? h->lbuf : h->gk[g].lbuf; }
ItemBufs& group_ibuf(k3dom_handle* h, int g) { return g == 0 ? h->ibuf : h->gk[g].ibuf; }
const ItemBufs& group_ibuf(const k3dom_handle* h, int g) { return g == 0 ? h->ibuf : h->gk[g].ibuf; }
DevScalars* group_sc(k3dom_handle* h, int g) { return g == 0 ? h->sc : h->gsc + g; }

// True when a buffer sized from the point capacity is missing: a previous
// frame's allocation failed part way (that frame returned the error).
bool point_buffers_missing(const k3dom_handle* h) {
    if (!h->d_pts || !h->rec || !h->pxyz || !h->h_pts) return true;
    for (int g = 0; g < h->ngk; ++g) {
        const ItemBufs& I = group_ibuf(h, g);
        if (!I.items || !I.partial) return true;
        for (int l = 0; l < 3; ++l) {
            const LevelBufs& B = group_lbuf(h, g)[l];
            if (!B.list || !B.counts || !B.slot_off || !B.bsum || !B.mask || !B.ictx) return true;
        }
    }
    return h->method == K3DOM_METHOD_DSPHDMIB && (!h->ds.list || !h->ds.segcnt || !h->ds.part);
}

k3dom_status ensure_points(k3dom_handle* h, int m) {
    if (m <= h->m_cap && !point_buffers_missing(h)) return K3DOM_OK;
    int cap = std::max(m, h->m_cap * 3 / 2);
    dfree(h, h->d_pts); dfree(h, h->rec); dfree(h, h->pxyz);
    if (h->h_pts) { (void)hipHostFree(h->h_pts); h->h_pts = nullptr; }
    h->d_pts = nullptr; h->rec = nullptr; h->pxyz = nullptr;
    h->m_cap = 0;
    HIP_TRY(h, dalloc(h, &h->d_pts, (size_t)3 * cap));
    HIP_TRY(h, dalloc(h, &h->rec, (size_t)cap));
    HIP_TRY(h, dalloc(h, &h->pxyz, (size_t)cap));
    HIP_TRY(h, hipHostMalloc((void**)&h->h_pts, sizeof(float) * 3 * (size_t)cap, hipHostMallocDefault));
    h->m_cap = cap;
    h->caps_version++;
    // first estimate of the candidate-list sizes; frames that need more grow them
    const long long mc = cap;
    long long want[3] = {8 * mc + 4096, 48 * mc + 16384, 160 * mc + 65536};
    long long want_heavy = 4096;
    if (h->tiny_caps) {
        want[0] = want[1] = want[2] = 64;
        want_heavy = 1;
    }
    if (h->method == K3DOM_METHOD_DSPHDMIB) {
        k3dom_status s = ensure_ds_list(h, h->tiny_caps ? 64 : 32 * mc + 65536);
        if (s != K3DOM_OK) return s;
    }
    return ensure_caps(h, want, want_heavy);
}

// (Re)allocates the buffers of one candidate level whose sizes follow from
// its list capacity and its parent's.
// (every group with buffers gets the same capacities)
k3dom_status alloc_level(k3dom_handle* h, int l, long long list_cap) {
    const LevelDesc& L = h->ldesc[l];
    h->caps_version++;
    for (int g = 0; g < h->ngk; ++g) {
        LevelBufs& B = group_lbuf(h, g)[l];
        for (void* p : {(void*)B.counts, (void*)B.slot_off, (void*)B.list, (void*)B.bsum, (void*)B.mask, (void*)B.ictx})
            dfree(h, p);
        B.counts = B.slot_off = B.list = B.bsum = nullptr;
        B.mask = nullptr;
        B.ictx = nullptr;
        const long long parent_cap = l == 0 ? (long long)std::max(h->m_cap, 1) : group_lbuf(h, g)[l - 1].list_cap;
        B.items_cap = L.nparents + (parent_cap + L.seg - 1) / L.seg;
        B.slot_cap = B.items_cap * L.nch;
        B.list_cap = list_cap;
        B.mask_stride = parent_cap;
        HIP_TRY(h, dalloc(h, &B.mask, (size_t)parent_cap * (size_t)level_child_groups(L)));
        HIP_TRY(h, dalloc(h, &B.ictx, (size_t)B.items_cap));
        HIP_TRY(h, dalloc(h, &B.counts, (size_t)B.slot_cap));
        HIP_TRY(h, dalloc(h, &B.slot_off, (size_t)B.slot_cap + 1));
        HIP_TRY(h, dalloc(h, &B.list, (size_t)std::max(1ll, list_cap)));
        HIP_TRY(h, dalloc_zero(h, &B.bsum,
                               (size_t)scan_scratch_words(std::max<long long>(B.slot_cap, L.nparents + 1))));
    }
    return K3DOM_OK;
}

k3dom_status alloc_heavy(k3dom_handle* h, long long heavy_cap) {
    h->caps_version++;
    for (int g = 0; g < h->ngk; ++g) {
        ItemBufs& I = group_ibuf(h, g);
        dfree(h, I.items);
        dfree(h, I.partial);
        I.items = nullptr;
        I.partial = nullptr;
        I.heavy_cap = heavy_cap;
        HIP_TRY(h, dalloc(h, &I.items, (size_t)(h->ntiles + heavy_cap)));
        HIP_TRY(h, dalloc(h, &I.partial, (size_t)std::max(1ll, heavy_cap) * I.partial_stride));
    }
    return K3DOM_OK;
}

// The capacity-independent buffers of a group's lists and work items (as the
// handle's own, allocated at creation).
k3dom_status alloc_group_fixed(k3dom_handle* h, k3dom_handle::GroupKi& G) {
    for (int l = 0; l < 3; ++l) {
        LevelBufs& LB = G.lbuf[l];
        const LevelDesc& L = h->ldesc[l];
        const size_t nboxes = (size_t)L.nbx * L.nby * L.nbz;
        HIP_TRY(h, dalloc(h, &LB.nseg, (size_t)L.nparents));
        HIP_TRY(h, dalloc(h, &LB.item_base, (size_t)L.nparents + 1));
        HIP_TRY(h, dalloc(h, &LB.box_off, nboxes));
        HIP_TRY(h, dalloc(h, &LB.box_len, nboxes));
        if (l > 0) HIP_TRY(h, dalloc(h, &LB.ctab, 2 * (size_t)L.nparents * (size_t)L.nch));
    }
    const size_t T = (size_t)h->ntiles;
    ItemBufs& I = G.ibuf;
    HIP_TRY(h, dalloc(h, &I.heavy_cnt, 2 * T));
    I.light_cnt = I.heavy_cnt + T;
    HIP_TRY(h, dalloc(h, &I.heavy_off, 2 * T + 1));
    I.light_off = I.heavy_off + T;
    HIP_TRY(h, dalloc_zero(h, &I.bsum, (size_t)scan_scratch_words(2 * (long long)T)));
    HIP_TRY(h, dalloc(h, &I.tile_partial, 2 * T));
    HIP_TRY(h, dalloc(h, &I.qheads, BG_QWORDS));
    HIP_TRY(h, dalloc(h, &G.del_occ, (size_t)h->V));
    HIP_TRY(h, dalloc(h, &G.del_free, (size_t)h->V));
    return K3DOM_OK;
}

// Buffers for n LiDAR groups: groups 1.. get their own lists, work items and
// evidence with the handle's current capacities, and the joint gather its
// argument sets, once.
k3dom_status ensure_groups(k3dom_handle* h, int n) {
    // (each allocated while missing: a failed allocation leaves the others in place)
    if (!h->gsc) HIP_TRY(h, dalloc_zero(h, &h->gsc, (size_t)K3DOM_MAX_LIDARS));
    if (!h->d_gargs) HIP_TRY(h, dalloc(h, &h->d_gargs, (size_t)K3DOM_MAX_LIDARS));
    if (!h->d_gq) HIP_TRY(h, dalloc(h, &h->d_gq, (size_t)K3DOM_MAX_LIDARS + 1 + BG_QWORDS));
    if (!h->h_gargs)
        HIP_TRY(h, hipHostMalloc((void**)&h->h_gargs, sizeof(BgArgs) * K3DOM_MAX_LIDARS, hipHostMallocDefault));
    if (n <= h->ngk) return K3DOM_OK;
    for (int g = h->ngk; g < n; ++g) {
        k3dom_handle::GroupKi& G = h->gk[g];
        G = k3dom_handle::GroupKi{};
        G.ibuf.partial_stride = 128;
        k3dom_status s = alloc_group_fixed(h, G);
        if (s != K3DOM_OK) return s;
    }
    // capacities as group 0's: the new groups' missing buffers make
    // ensure_caps allocate every level (alloc_level / alloc_heavy cover
    // every group)
    h->ngk = n;
    h->caps_version++;
    long long want[3];
    for (int l = 0; l < 3; ++l) want[l] = h->lbuf[l].list_cap;
    return ensure_caps(h, want, h->ibuf.heavy_cap);
}


// Capacities for a frame of m points: grown for larger point sets, and by the
// overflow path when a frame needs more than estimated.
k3dom_status ensure_caps(k3dom_handle* h, const long long want[3], long long want_heavy) {
    for (int l = 0; l < 3; ++l) {
        const long long parent_cap = l == 0 ? (long long)std::max(h->m_cap, 1) : h->lbuf[l - 1].list_cap;
        const long long items_need = h->ldesc[l].nparents + (parent_cap + h->ldesc[l].seg - 1) / h->ldesc[l].seg;
        bool grow = false;
        for (int g = 0; g < h->ngk; ++g) {
            const LevelBufs& B = group_lbuf(h, g)[l];
            // (a buffer still missing after a failed allocation is allocated again)
            const bool missing = !B.list || !B.counts || !B.slot_off || !B.bsum || !B.mask || !B.ictx;
            grow = grow || want[l] > B.list_cap || items_need > B.items_cap || parent_cap > B.mask_stride || missing;
        }
        if (grow) {
            k3dom_status s = alloc_level(h, l, std::max(want[l], h->lbuf[l].list_cap));
            if (s != K3DOM_OK) return s;
        }
    }
    bool grow = want_heavy > h->ibuf.heavy_cap;
    for (int g = 0; g < h->ngk; ++g) grow = grow || !group_ibuf(h, g).items || !group_ibuf(h, g).partial;
    if (grow) return alloc_heavy(h, std::max(want_heavy, h->ibuf.heavy_cap));
    return K3DOM_OK;
}

k3dom_status sync_scalars(k3dom_handle* h) {
    HIP_TRY(h, hipMemcpyAsync(h->h_sc, h->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return K3DOM_OK;
}

// Timing level 1 records every stage event, level 2 only the two around the
// kernel-inference gather.
void mark(k3dom_handle* h, int stage) {
    if (h->timing == 1 || (h->timing == 2 && (stage == ST_BGKI_KERNEL || stage == EV_BGKI_END)))
        (void)hipEventRecord(h->ev[stage], h->stream);
}

GridGeom grid_geom(const k3dom_handle* h) {
    GridGeom g;
    g.gs = h->gs;
    g.gsz = h->gsz;
    g.res = h->p.resolution;
    g.sx = h->sensor[0] - h->center[0];
    g.sy = h->sensor[1] - h->center[1];
    g.sz = h->sensor[2] - h->center[2];
    return g;
}

// Point records and candidate lists of group g's points (fc->m of them at
// xs, ys, zs), into the group's buffers; the records go to rec + off.
k3dom_status build_lists_group(k3dom_handle* h, int g, size_t off, const FrameConsts* fc, const float* xs,
                               const float* ys, const float* zs) {
    if (h->pts_pending) {
        HIP_TRY(h, hipStreamWaitEvent(h->stream, h->pts_ready, 0));
        h->pts_pending = false;
    }
    LevelBufs* LB = group_lbuf(h, g);
    DevScalars* sc = group_sc(h, g);
    float4* pxyz = h->pxyz + off;
    HIP_TRY(h, launch_point_prep(h->stream, fc, xs, ys, zs, h->m_cap, h->p.resolution, h->p.ls, h->rec + off, pxyz));
    HIP_TRY(h, launch_build_levels(h->stream, fc, h->tg, h->ldesc, LB, group_ibuf(h, g), h->p.ls, pxyz, sc));
    HIP_TRY(h, launch_build_items(h->stream, h->ntiles, LB[2], group_ibuf(h, g), sc));
    return K3DOM_OK;
}

k3dom_status build_lists(k3dom_handle* h, const FrameConsts* fc, const float* xs, const float* ys, const float* zs) {
    return build_lists_group(h, 0, 0, fc, xs, ys, zs);
}

void fill_frame_consts(k3dom_handle* h, FrameConsts* c, double t, int m) {
    const bool normal = h->updated_time > 0;
    fill_frame_consts_dt(h, c, normal, normal ? (float)(t - h->updated_time) : 0.0f, m);
}

void fill_frame_consts_dt(k3dom_handle* h, FrameConsts* c, bool normal, float dt, int m) {
    c->g = grid_geom(h);
    c->decay = normal ? (float)std::pow((double)h->p.gamma, (double)dt) : 1.0f;
    c->frame = h->frame;
    c->m = m;
    c->pad = 0;
    // the pending particle shift goes to this frame's predict (an
    // initialisation frame draws new particles, so it is dropped there)
    for (int k = 0; k < 3; ++k) c->pshift[k] = h->pshift_pending ? (float)h->pshift[k] : 0.0f;
    h->pshift_pending = false;
    K3PredictConst& pc = c->pc;
    pc.dt = dt;
    pc.p_S = h->p.persistence_prob;
    pc.sigma_pos = h->p.stddev_process_noise_position;
    pc.sigma_vel = h->p.stddev_process_noise_velocity;
    pc.vclamp = 5 * h->p.init_max_velocity;
    pc.gs = h->gs;
    pc.gsz = h->gsz;
    pc.frame = h->frame;
    pc.key0 = (uint32_t)h->p.seed;
    pc.key1 = (uint32_t)(h->p.seed >> 32);
}

// The frame's tail in the stage order of DOM::updateGrid (src/k3dom/dom.cu:
// 38-43): tail_mass (the part of gridCellOccupancyUpdate after the
// evidence), tail_persistent (updatePersistentParticles, with
// statisticalMoments and the static penalty of resampling fused into it: the
// moments read the weights after the persistent update and before the
// penalty, and the newborn stage between them changes neither), tail_newborn
// (initializeNewParticles, or the first frame's particle placement) and
// tail_resample (resampling).  Tail kernels return at once when the overflow
// flag is set, so the frame can be re-run.
namespace {
// per-block maxima: born masses first, then the joint particle weights
uint32_t* born_parts(k3dom_handle* h) { return h->maxparts; }
uint32_t* joint_parts(k3dom_handle* h) { return h->maxparts + mass_update_parts(h->V); }
}  // namespace

k3dom_status tail_mass(k3dom_handle* h, bool normal) {
    VoxelArrays& v = h->vox[h->vcur];
    const uint32_t* ab = &h->sc->overflow;
    const bool ds = h->method == K3DOM_METHOD_DSPHDMIB;
    if (!ds && h->ngroups == 0)
        HIP_TRY(h, launch_bgki_combine(h->stream, h->ntiles, grid_geom(h), h->tg, h->ibuf, h->sigma_occ,
                                       h->sigma_free, h->del_occ, h->del_free, h->sc));
    mark(h, ST_MASS);
    if (normal) {
        if (ds)
            HIP_TRY(h, launch_ds_mass_update(h->stream, ab, v, h->V, h->wsort, h->P[1].key, h->N, h->usum,
                                             h->del_occ, h->del_free, h->p.birth_prob, h->born, h->ds_norm,
                                             born_parts(h)));
        else
            HIP_TRY(h, launch_mass_update(h->stream, ab, v, h->V, h->wsort, h->P[1].key, h->N, h->usum, h->del_occ,
                                          h->del_free, h->fc, h->p.mass_scale, h->p.birth_prob, h->born,
                                          born_parts(h)));
    } else {
        if (ds)
            HIP_TRY(h, launch_ds_born_from_meas(h->stream, h->del_occ, h->born, h->V, born_parts(h)));
        else
            HIP_TRY(h, launch_mass_update(h->stream, ab, v, h->V, nullptr, nullptr, 0, h->usum, h->del_occ,
                                          h->del_free, nullptr, h->p.mass_scale, h->p.birth_prob, h->born,
                                          born_parts(h)));
    }
    return K3DOM_OK;
}

k3dom_status tail_persistent(k3dom_handle* h) {
    mark(h, ST_PERSIST);
    HIP_TRY(h, launch_voxel_particles(h->stream, &h->sc->overflow, h->vox[h->vcur], h->V, h->P[1], h->wsort, h->N,
                                      h->p.particle_min_vel, h->usum, joint_parts(h)));
    return K3DOM_OK;
}

k3dom_status tail_newborn(k3dom_handle* h, bool normal) {
    VoxelArrays& v = h->vox[h->vcur];
    const uint32_t k0 = (uint32_t)h->p.seed, k1 = (uint32_t)(h->p.seed >> 32);
    const uint32_t* ab = &h->sc->overflow;
    if (!normal) mark(h, ST_PERSIST);
    mark(h, ST_NEWBORN);
    HIP_TRY(h, launch_fixed_scan(h->stream, h->born, h->V, nullptr, 0, born_parts(h), mass_update_parts(h->V),
                                 &h->sc->E_born, h->cdf_v, h->blocksums, &h->sc->total_born));
    if (normal) {
        HIP_TRY(h, launch_bounds(h->stream, ab, h->cdf_v, &h->sc->total_born, h->V, (uint32_t)h->B, h->born, h->bnd,
                                 v.wua));
        HIP_TRY(h, launch_newborn(h->stream, ab, h->bnd, h->V, h->B, h->gs, v.wua, h->NB, h->p.init_max_velocity,
                                  h->p.particle_min_vel, h->fc, k0, k1, joint_parts(h) + voxel_particles_parts(h->N)));
    } else {
        HIP_TRY(h, launch_bounds(h->stream, ab, h->cdf_v, &h->sc->total_born, h->V, (uint32_t)h->N, h->born, h->bnd,
                                 nullptr));
        HIP_TRY(h, launch_init_particles(h->stream, ab, h->bnd, h->V, h->N, h->gs, h->P[0], h->p.init_max_velocity,
                                         h->p.particle_min_vel, h->fc, k0, k1));
    }
    mark(h, ST_MOMENTS);
    return K3DOM_OK;
}

k3dom_status tail_resample(k3dom_handle* h) {
    const uint32_t k0 = (uint32_t)h->p.seed, k1 = (uint32_t)(h->p.seed >> 32);
    mark(h, ST_RESAMPLE);
    HIP_TRY(h, launch_fixed_scan(h->stream, h->P[1].w, h->N, h->NB.w, h->B, joint_parts(h),
                                 voxel_particles_parts(h->N) + newborn_parts(h->B), &h->sc->E_joint, h->cdf_j,
                                 h->blocksums, &h->sc->total_joint));
    HIP_TRY(h, launch_resample(h->stream, &h->sc->overflow, h->cdf_j, h->N + h->B, &h->sc->total_joint,
                               &h->sc->E_joint, h->P[1], h->NB, h->P[0], h->N, h->fc, k0, k1, h->ridx));
    return K3DOM_OK;
}

// Everything after the gather, then the frame's scalars to the host mirror.
k3dom_status enqueue_after_gather(k3dom_handle* h, bool normal) {
    k3dom_status s = tail_mass(h, normal);
    if (s == K3DOM_OK && normal) s = tail_persistent(h);
    if (s == K3DOM_OK) s = tail_newborn(h, normal);
    if (s == K3DOM_OK && normal) s = tail_resample(h);
    if (s != K3DOM_OK) return s;
    if (!normal) mark(h, ST_RESAMPLE);
    mark(h, ST_FRAME);
    HIP_TRY(h, hipMemcpyAsync(h->h_sc, h->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, h->stream));
    return K3DOM_OK;
}

// Kernel inference (lists, gather), or the DS-PHD/MIB measurement grid, and
// everything after it.
k3dom_status enqueue_rest(k3dom_handle* h, bool normal, const float* dx, const float* dy, const float* dz) {
    if (h->method == K3DOM_METHOD_DSPHDMIB) {
        mark(h, ST_LISTS);
        if (h->pts_pending) {
            HIP_TRY(h, hipStreamWaitEvent(h->stream, h->pts_ready, 0));
            h->pts_pending = false;
        }
        mark(h, ST_BGKI);
        mark(h, ST_BGKI_KERNEL);
        // one LiDAR at most: with the sensor at its origin
        const FrameConsts* dfc = h->ngroups > 0 ? h->gfc : h->fc;
        HIP_TRY(h, launch_ds_meas_grid(h->stream, &h->sc->overflow, dfc, h->gs, h->gsz, dx, dy, dz,
                                       h->p.resolution, h->p.meas_occ_stddev, h->p.max_range, h->ds, h->sc,
                                       h->del_occ, h->del_free));
        mark(h, EV_BGKI_END);
        return enqueue_after_gather(h, normal);
    }
    if (h->ngroups > 0) {
        // every LiDAR's lists and work items (sensor at its origin, its own
        // points, its own buffers), then one persistent gather over all
        // LiDARs' items, each into its own evidence; later LiDARs' evidence
        // is added to the first's in LiDAR order
        mark(h, ST_LISTS);
        for (int g = 0; g < h->ngroups; ++g) {
            if (g > 0) HIP_TRY(h, launch_reset_scalars(h->stream, group_sc(h, g)));
            const size_t o = (size_t)h->gstart[g];
            k3dom_status s = build_lists_group(h, g, o, h->gfc + g, dx + o, dy + o, dz + o);
            if (s != K3DOM_OK) return s;
        }
        mark(h, ST_BGKI);
        mark(h, ST_BGKI_KERNEL);
        for (int g = 0; g < h->ngroups; ++g)
            h->h_gargs[g] = make_bg_args(h->gfc + g, h->tg, h->p.ls, h->sigma_occ, h->sigma_free, group_ibuf(h, g),
                                         group_lbuf(h, g)[2], h->rec + h->gstart[g],
                                         g == 0 ? h->del_occ : h->gk[g].del_occ,
                                         g == 0 ? h->del_free : h->gk[g].del_free, nullptr, nullptr, group_sc(h, g));
        HIP_TRY(h, hipMemcpyAsync(h->d_gargs, h->h_gargs, sizeof(BgArgs) * h->ngroups, hipMemcpyHostToDevice,
                                  h->stream));
        HIP_TRY(h, launch_bgki_batch(h->stream, h->gather_blocks, h->d_gargs, h->ngroups, h->d_gq,
                                     h->d_gq + K3DOM_MAX_LIDARS + 1, h->gather_waves));
        for (int g = 0; g < h->ngroups; ++g) {
            GridGeom gg = grid_geom(h);
            gg.sx = h->h_gfc[g].g.sx;
            gg.sy = h->h_gfc[g].g.sy;
            gg.sz = h->h_gfc[g].g.sz;
            HIP_TRY(h, launch_bgki_combine(h->stream, h->ntiles, gg, h->tg, group_ibuf(h, g), h->sigma_occ,
                                           h->sigma_free, g == 0 ? h->del_occ : h->gk[g].del_occ,
                                           g == 0 ? h->del_free : h->gk[g].del_free, group_sc(h, g)));
        }
        for (int g = 1; g < h->ngroups; ++g)
            HIP_TRY(h, launch_add_evidence(h->stream, &h->sc->overflow, h->del_occ, h->del_free, h->gk[g].del_occ,
                                           h->gk[g].del_free, h->V));
        if (h->ngroups > 1) HIP_TRY(h, launch_merge_group_scalars(h->stream, h->sc, h->gsc, h->ngroups));
        mark(h, EV_BGKI_END);
        return enqueue_after_gather(h, normal);
    }
    mark(h, ST_LISTS);
    k3dom_status s = build_lists(h, h->fc, dx, dy, dz);
    if (s != K3DOM_OK) return s;
    mark(h, ST_BGKI);
    mark(h, ST_BGKI_KERNEL);
    // (build_lists left the work-queue heads zeroed)
    HIP_TRY(h, launch_bgki(h->stream, h->gather_blocks, h->fc, h->tg, h->p.ls, h->sigma_occ, h->sigma_free, h->ibuf,
                           h->lbuf[2], h->rec, h->del_occ, h->del_free, h->sc, h->gather_waves, true));
    mark(h, EV_BGKI_END);
    return enqueue_after_gather(h, normal);
}

// The part of a frame before kernel inference (prediction, sort, ranges),
// after the frame constants reach the device.
k3dom_status enqueue_head(k3dom_handle* h, bool normal) {
    VoxelArrays& v = h->vox[h->vcur];
    mark(h, EV_FRAME_START);
    HIP_TRY(h, hipMemcpyAsync(h->fc, h->h_fc, sizeof(FrameConsts), hipMemcpyHostToDevice, h->stream));
    if (h->ngroups > 0)
        HIP_TRY(h, hipMemcpyAsync(h->gfc, h->h_gfc, sizeof(FrameConsts) * h->ngroups, hipMemcpyHostToDevice,
                                  h->stream));
    HIP_TRY(h, launch_reset_scalars(h->stream, h->sc));
    if (normal) {
        mark(h, ST_PREDICT);
        HIP_TRY(h, launch_predict(h->stream, h->fc, h->P[0], h->prec, h->N));
        mark(h, ST_SORT);
        HIP_TRY(h, launch_radix_sort(h->stream, h->P[0].key, h->N, h->key_bits, h->sort, &h->last_skeys,
                                     &h->last_perm, v.start, v.end, h->V));
        HIP_TRY(h, launch_gather_sorted(h->stream, h->prec, h->P[1], h->wsort, h->last_perm, h->last_skeys, h->N,
                                        v.start, v.end));
    }
    return K3DOM_OK;
}

static void drop_graph_of(k3dom_handle* h, int p) {
    if (h->gexec[p]) (void)hipGraphExecDestroy(h->gexec[p]);
    if (h->graph[p]) (void)hipGraphDestroy(h->graph[p]);
    h->gexec[p] = nullptr;
    h->graph[p] = nullptr;
    h->graph_caps[p] = -1;
}

void drop_graph(k3dom_handle* h) {
    drop_graph_of(h, 0);
    drop_graph_of(h, 1);
}

// Captures the whole normal frame (constants upload to scalar readback) as a
// graph over the handle's own point buffer and the current voxel buffer.
k3dom_status capture_graph(k3dom_handle* h) {
    const int p = h->vcur;
    drop_graph_of(h, p);
    const size_t cap = (size_t)h->m_cap;
    const int timing = h->timing;
    h->timing = 0;  // no event records inside the graph
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    k3dom_status s = K3DOM_OK;
    if (e == hipSuccess) {
        s = enqueue_head(h, true);
        if (s == K3DOM_OK) s = enqueue_rest(h, true, h->d_pts, h->d_pts + cap, h->d_pts + 2 * cap);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(h->stream, &g);
        h->graph[p] = g;
        if (e2 != hipSuccess && e == hipSuccess) e = e2;
    }
    h->timing = timing;
    if (s != K3DOM_OK) {
        drop_graph_of(h, p);
        return s;
    }
    HIP_TRY(h, e);
    HIP_TRY(h, hipGraphInstantiate(&h->gexec[p], h->graph[p], nullptr, nullptr, 0));
    h->graph_caps[p] = h->caps_version;
    h->graph_ngroups[p] = h->ngroups;
    for (int g = 0; g < K3DOM_MAX_LIDARS; ++g) h->graph_gstart[p][g] = g < h->ngroups ? h->gstart[g] : 0;
    return K3DOM_OK;
}

// The graph of voxel buffer p was captured for this frame's LiDAR layout
// (group count and point offsets, which its kernels take as arguments).
bool graph_layout_matches(const k3dom_handle* h, int p) {
    if (h->graph_ngroups[p] != h->ngroups) return false;
    for (int g = 0; g < h->ngroups; ++g)
        if (h->graph_gstart[p][g] != h->gstart[g]) return false;
    return true;
}

// Grows what the last run of the kernel-inference stage overflowed (its need
// is known up to that stage, from the scalars read back) and clears the flag
// for a re-run.
k3dom_status grow_after_overflow(k3dom_handle* h) {
    long long want[3], want_heavy = h->ibuf.heavy_cap;
    for (int l = 0; l < 3; ++l) {
        want[l] = h->lbuf[l].list_cap;
        if (h->h_sc->overflow_mask & (1u << l))
            want[l] = std::max(want[l] * 2, (long long)h->h_sc->level_total[l] * 3 / 2 + 1024);
    }
    if (h->h_sc->overflow_mask & (1u << 4))
        want_heavy = std::max(want_heavy * 2, (long long)h->h_sc->n_heavy * 3 / 2 + 256);
    if (h->h_sc->overflow_mask & (1u << 5)) {
        const long long need = (long long)h->h_sc->ds_total * 5 / 4 + 1024;
        k3dom_status s = ensure_ds_list(h, std::max(h->ds.cap * 2, need));
        if (s != K3DOM_OK) return s;
    }
    k3dom_status s = ensure_caps(h, want, want_heavy);
    if (s != K3DOM_OK) return s;
    HIP_TRY(h, launch_clear_overflow(h->stream, h->sc));
    h->overflow_retries++;
    return K3DOM_OK;
}

// Runs kernel inference and the frame tail until no capacity overflows.
// checked: the frame already ran once and `done` is its outcome (graph path);
// otherwise the first iteration enqueues the kernel-inference stage itself.
k3dom_status resolve_overflow(k3dom_handle* h, bool normal, const float* dx, const float* dy, const float* dz,
                              bool done, bool checked) {
    for (int attempt = 0; !done; ++attempt) {
        if (checked || attempt > 0) {
            if (attempt >= 8) return fail(h, K3DOM_ERR_OUT_OF_MEMORY, "kernel-inference capacities keep overflowing");
            k3dom_status s = grow_after_overflow(h);
            if (s != K3DOM_OK) return s;
        }
        k3dom_status s = enqueue_rest(h, normal, dx, dy, dz);
        if (s != K3DOM_OK) return s;
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        done = !h->h_sc->overflow;
    }
    return K3DOM_OK;
}

void finish_frame(k3dom_handle* h, double t, int m) {
    count_frame(h, h->updated_time > 0, m);
    h->updated_time = t;
}

void count_frame(k3dom_handle* h, bool normal, int m) {
    h->last_geom = h->h_fc->g;
    h->stats.points = (uint64_t)m;
    for (int l = 0; l < 3; ++l) h->stats.level_entries[l] = h->h_sc->level_total[l];
    if (h->method == K3DOM_METHOD_DSPHDMIB) h->stats.level_entries[0] = h->h_sc->ds_total;
    h->stats.gather_items = (uint64_t)h->h_sc->n_light + h->h_sc->n_heavy;
    h->stats.gather_split_items = h->h_sc->n_heavy;
    h->stats.capacity_regrowths = (uint64_t)h->overflow_retries;
    if (normal) h->normal_here++;
    h->frames_here++;
    h->last_grouped = h->ngroups > 0;
    h->last_frame_ok = true;
    h->frame++;
}

bool frame_uses_graph(const k3dom_handle* h) {
    return h->updated_time > 0 && h->use_graph && !h->timing;
}

k3dom_status frame_body(k3dom_handle* h, double t, const float* dx, const float* dy, const float* dz, int m);

// A frame that fails (an allocation for a grown capacity, a HIP error) leaves
// the state a later frame reads as it was: the particle set, the masses, the
// pose, the time and the frame counter are only written by a frame that
// completes, and the map shift still due on the particles and on the deferred
// voxel fields stays due.  (The ranges, keys and evidence of the failed
// frame's first stages may be left in the per-frame arrays until the next frame.)
k3dom_status frame(k3dom_handle* h, double t, const float* dx, const float* dy, const float* dz, int m) {
    if (h->stage != 0)
        return fail(h, K3DOM_ERR_STATE, "a frame driven by the stage methods is unfinished (resampling ends it)");
    const bool pshift_pending = h->pshift_pending, shift_pending = h->shift_pending;
    const int pshift[3] = {h->pshift[0], h->pshift[1], h->pshift[2]};
    const k3dom_status s = frame_body(h, t, dx, dy, dz, m);
    if (s != K3DOM_OK) {
        (void)hipStreamSynchronize(h->stream);
        h->last_frame_ok = false;
        h->pshift_pending = pshift_pending;
        for (int k = 0; k < 3; ++k) h->pshift[k] = pshift[k];
        // (re-applying a deferred shift that the failed frame settled is idempotent)
        h->shift_pending = shift_pending;
    }
    return s;
}

k3dom_status frame_body(k3dom_handle* h, double t, const float* dx, const float* dy, const float* dz, int m) {
    const bool normal = h->updated_time > 0;
    if (h->ngroups > 0) {
        k3dom_status st = ensure_groups(h, h->ngroups);
        if (st != K3DOM_OK) return st;
    }
    {
        k3dom_status st = begin_frame_voxels(h, normal);
        if (st != K3DOM_OK) return st;
    }
    if (h->timing) for (float& x : h->stage_ms) x = 0.0f;
    fill_frame_consts(h, h->h_fc, t, m);
    for (int g = 0; g < h->ngroups; ++g) {
        // a LiDAR's pass: the sensor at its origin, its own points
        FrameConsts& c = h->h_gfc[g];
        c = *h->h_fc;
        c.g.sx = (h->sensor[0] + h->gorig[g][0]) - h->center[0];
        c.g.sy = (h->sensor[1] + h->gorig[g][1]) - h->center[1];
        c.g.sz = (h->sensor[2] + h->gorig[g][2]) - h->center[2];
        c.m = h->gcount[g];
    }
    const bool use_graph = frame_uses_graph(h);
    bool done = false;
    if (use_graph) {
        const size_t cap = (size_t)h->m_cap;
        if (dx != h->d_pts) {  // the graph reads the handle's own point buffer
            for (int k = 0; k < 3; ++k) {
                const float* src = k == 0 ? dx : (k == 1 ? dy : dz);
                if (m > 0)
                    HIP_TRY(h, hipMemcpyAsync(h->d_pts + k * cap, src, sizeof(float) * m, hipMemcpyDeviceToDevice,
                                              h->stream));
            }
        }
        const int p = h->vcur;
        if (!h->gexec[p] || h->graph_caps[p] != h->caps_version || !graph_layout_matches(h, p)) {
            k3dom_status s = capture_graph(h);
            if (s != K3DOM_OK) return s;
        }
        HIP_TRY(h, hipGraphLaunch(h->gexec[p], h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        done = !h->h_sc->overflow;
        dx = h->d_pts;
        dy = h->d_pts + cap;
        dz = h->d_pts + 2 * cap;
    } else {
        k3dom_status s = enqueue_head(h, normal);
        if (s != K3DOM_OK) return s;
    }
    k3dom_status s = resolve_overflow(h, normal, dx, dy, dz, done, use_graph);
    if (s != K3DOM_OK) return s;
    if (h->timing == 1) {
        const int order[] = {ST_PREDICT, ST_SORT, ST_LISTS, ST_BGKI, ST_MASS, ST_PERSIST, ST_NEWBORN, ST_MOMENTS,
                             ST_RESAMPLE, ST_FRAME};
        for (int k = 0; k + 1 < 10; ++k) {
            int a = order[k], b = order[k + 1];
            if (!normal && (a == ST_PREDICT || a == ST_SORT)) continue;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, h->ev[a], h->ev[b]) == hipSuccess) h->stage_ms[a] = ms;
        }
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, h->ev[EV_FRAME_START], h->ev[ST_FRAME]) == hipSuccess) h->stage_ms[ST_FRAME] = ms;
    }
    if (h->timing) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, h->ev[ST_BGKI_KERNEL], h->ev[EV_BGKI_END]) == hipSuccess)
            h->stage_ms[ST_BGKI_KERNEL] = ms;
    }
    finish_frame(h, t, m);
    return K3DOM_OK;
}

// ---------------------------------------------------------------- YAML params
// Back to the state after construction: priors in every voxel, no pose, no
// frame (the random counters restart with the frame count).
k3dom_status reset_map(k3dom_handle* h) {
    for (int b = 0; b < 2; ++b)
        HIP_TRY(h, launch_fill_voxels(h->stream, h->vox[b], h->V, h->p.prior_free, prior_s_slot(h),
                                      h->p.prior_dynamic));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    h->vcur = 0;
    h->shift_pending = false;
    h->pshift_pending = false;
    h->frame = 0;
    h->frames_here = h->normal_here = 0;
    h->first_pose = false;
    h->updated_time = -1.0;
    for (int i = 0; i < 3; ++i) h->sensor[i] = h->center[i] = 0.0f;
    return K3DOM_OK;
}

k3dom_status settle_shift(k3dom_handle* h) {
    if (!h->shift_pending) return K3DOM_OK;
    HIP_TRY(h, launch_shift_voxels(h->stream, h->vox[h->vcur], h->vox[1 - h->vcur], h->gs, h->gsz, h->pend_move[0],
                                   h->pend_move[1], h->pend_move[2], h->p.prior_free, prior_s_slot(h),
                                   h->p.prior_dynamic, SHIFT_REWRITTEN));
    h->shift_pending = false;
    return K3DOM_OK;
}

k3dom_status settle_particles(k3dom_handle* h) {
    if (!h->pshift_pending) return K3DOM_OK;
    HIP_TRY(h, launch_shift_particles(h->stream, h->P[0], h->N, h->pshift[0], h->pshift[1], h->pshift[2]));
    h->pshift_pending = false;
    return K3DOM_OK;
}

k3dom_status begin_frame_voxels(k3dom_handle* h, bool normal) {
    if (normal) {
        h->shift_pending = false;
        return K3DOM_OK;
    }
    return settle_shift(h);
}

k3dom_status update_pose(k3dom_handle* h, float x, float y, float z, int shifted[3]) {
    shifted[0] = shifted[1] = shifted[2] = 0;
    h->sensor[0] = x;
    h->sensor[1] = y;
    h->sensor[2] = z;
    if (!h->first_pose) {
        h->first_pose = true;
        h->center[0] = h->sensor[0] - h->p.sensor_off_x;
        h->center[1] = h->sensor[1] - h->p.sensor_off_y;
        h->center[2] = h->sensor[2] - h->p.sensor_off_z;
        return K3DOM_OK;
    }
    const int mx = static_cast<int>((h->sensor[0] - h->center[0] - h->p.sensor_off_x) / h->p.resolution);
    const int my = static_cast<int>((h->sensor[1] - h->center[1] - h->p.sensor_off_y) / h->p.resolution);
    const int mz = static_cast<int>((h->sensor[2] - h->center[2] - h->p.sensor_off_z) / h->p.resolution);
    if (std::abs(mx) >= h->p.map_shift_thresh || std::abs(my) >= h->p.map_shift_thresh ||
        std::abs(mz) >= h->p.map_shift_thresh) {
        k3dom_status st = settle_shift(h);
        if (st != K3DOM_OK) return st;
        const int nxt = 1 - h->vcur;
        HIP_TRY(h, launch_shift_voxels(h->stream, h->vox[nxt], h->vox[h->vcur], h->gs, h->gsz, mx, my, mz,
                                       h->p.prior_free, prior_s_slot(h), h->p.prior_dynamic, SHIFT_KEPT));
        st = settle_particles(h);
        if (st != K3DOM_OK) return st;
        h->pshift_pending = true;
        h->pshift[0] = mx;
        h->pshift[1] = my;
        h->pshift[2] = mz;
        h->shift_pending = true;
        h->pend_move[0] = mx;
        h->pend_move[1] = my;
        h->pend_move[2] = mz;
        h->vcur = nxt;
        h->center[0] += mx * h->p.resolution;
        h->center[1] += my * h->p.resolution;
        h->center[2] += mz * h->p.resolution;
        shifted[0] = mx;
        shifted[1] = my;
        shifted[2] = mz;
    }
    return K3DOM_OK;
}

}  // namespace k3api

// ============================================================== C ABI
using namespace k3api;

extern "C" {

k3dom_status k3dom_create(const k3dom_params* params, k3dom_handle** out) {
    return k3dom_create_method(params, K3DOM_METHOD_K3DOM, out);
}

k3dom_status k3dom_create_method(const k3dom_params* params, int32_t method, k3dom_handle** out) {
    if (!out) return fail(nullptr, K3DOM_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (method != K3DOM_METHOD_K3DOM && method != K3DOM_METHOD_DSPHDMIB)
        return fail(nullptr, K3DOM_ERR_INVALID_ARG, "unknown method");
    std::string why;
    k3dom_status vs = validate(params, &why);
    if (vs != K3DOM_OK) return fail(nullptr, vs, why);
    k3dom_handle* h = new (std::nothrow) k3dom_handle();
    if (!h) return fail(nullptr, K3DOM_ERR_OUT_OF_MEMORY, "host allocation failed");
    h->method = method;
    auto bail = [&](k3dom_status s) {
        g_create_error = h->err;
        k3dom_destroy(h);
        return s;
    };
#define CREATE_TRY(expr)                                                                           \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            h->err = std::string(#expr) + ": " + hipGetErrorString(_e);                            \
            return bail(_e == hipErrorOutOfMemory ? K3DOM_ERR_OUT_OF_MEMORY : K3DOM_ERR_HIP);      \
        }                                                                                          \
    } while (0)
    h->p = *params;
    h->sigma_occ = h->sigma_free = h->p.sigma;
    h->gs = (int)(h->p.size / h->p.resolution);
    h->gsz = (int)(h->p.size_z / h->p.resolution);
    h->V = h->gs * h->gs * h->gsz;
    h->N = h->p.particle_count;
    h->B = h->p.new_born_particle_count;
    h->key_bits = 1;
    while ((1ll << h->key_bits) < (long long)h->V) ++h->key_bits;
    h->ntx = (h->gs + 3) / 4;
    h->nty = h->ntx;
    h->ntz = (h->gsz + 3) / 4;
    h->ntiles = h->ntx * h->nty * h->ntz;
    h->tg.ntx = h->ntx;
    h->tg.nty = h->nty;
    h->tg.ntz = h->ntz;
    CREATE_TRY(hipGetDevice(&h->device));
    hipDeviceProp_t prop;
    CREATE_TRY(hipGetDeviceProperties(&prop, h->device));
    h->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    (void)noalloc_prepare(h->device);
    CREATE_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    CREATE_TRY(hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking));
    CREATE_TRY(hipEventCreateWithFlags(&h->pts_ready, hipEventDisableTiming));
    const size_t V = (size_t)h->V, N = (size_t)h->N, B = (size_t)std::max(h->B, 1);
    for (int b = 0; b < 2; ++b) {
        VoxelArrays& v = h->vox[b];
        for (float** f : {&v.f, &v.s, &v.d, &v.pers, &v.pred, &v.mu, &v.wua, &v.mvx, &v.mvy, &v.mvz})
            CREATE_TRY(dalloc(h, f, V));
        CREATE_TRY(dalloc(h, &v.start, V));
        CREATE_TRY(dalloc(h, &v.end, V));
    }
    CREATE_TRY(dalloc(h, &h->born, V));
    if (h->method == K3DOM_METHOD_DSPHDMIB) {
        const size_t nst = (size_t)ds_supertiles(h->gs, h->gsz);
        CREATE_TRY(dalloc(h, &h->ds_norm, V));
        CREATE_TRY(dalloc(h, &h->ds.counts, nst));
        CREATE_TRY(dalloc(h, &h->ds.off, nst + 1));
        CREATE_TRY(dalloc_zero(h, &h->ds.scratch, (size_t)scan_scratch_words((long long)nst)));
        CREATE_TRY(dalloc(h, &h->ds.nseg, nst));
        CREATE_TRY(dalloc(h, &h->ds.seg_off, nst + 1));
        CREATE_TRY(dalloc_zero(h, &h->ds.seg_scratch, (size_t)scan_scratch_words((long long)nst)));
    }
    CREATE_TRY(dalloc(h, &h->del_occ, V));
    CREATE_TRY(dalloc(h, &h->del_free, V));
    for (int s = 0; s < 2; ++s) {
        ParticleArrays& P = h->P[s];
        for (float** f : {&P.px, &P.py, &P.pz, &P.vx, &P.vy, &P.vz, &P.w}) CREATE_TRY(dalloc(h, f, N));
        CREATE_TRY(dalloc(h, &P.key, N));
    }
    for (float** f : {&h->NB.px, &h->NB.py, &h->NB.pz, &h->NB.vx, &h->NB.vy, &h->NB.vz, &h->NB.w})
        CREATE_TRY(dalloc(h, f, B));
    CREATE_TRY(dalloc(h, &h->NB.key, B));
    CREATE_TRY(dalloc(h, &h->wsort, N));
    CREATE_TRY(dalloc(h, &h->prec, 2 * N));
    CREATE_TRY(dalloc(h, &h->usum, 3 * N));
    CREATE_TRY(dalloc(h, &h->maxparts, (size_t)mass_update_parts(h->V) + voxel_particles_parts(h->N) +
                                           newborn_parts(h->B)));
    CREATE_TRY(dalloc(h, &h->sort.k0, N));
    CREATE_TRY(dalloc(h, &h->sort.k1, N));
    CREATE_TRY(dalloc(h, &h->sort.i0, N));
    CREATE_TRY(dalloc(h, &h->sort.i1, N));
    CREATE_TRY(dalloc_zero(h, &h->sort.hist, (size_t)sort_hist_words(h->N)));
    CREATE_TRY(dalloc(h, &h->cdf_v, V));
    CREATE_TRY(dalloc(h, &h->cdf_j, N + B));
    CREATE_TRY(dalloc(h, &h->blocksums, (size_t)fixed_scan_scratch_words((long long)std::max(V, N + B))));
    CREATE_TRY(dalloc(h, &h->bnd, V));
    CREATE_TRY(dalloc(h, &h->ridx, N));
    CREATE_TRY(dalloc(h, &h->sc, 1));
    setup_levels(h);
    for (int l = 0; l < 3; ++l) {
        LevelBufs& LB = h->lbuf[l];
        const LevelDesc& L = h->ldesc[l];
        const size_t nboxes = (size_t)L.nbx * L.nby * L.nbz;
        CREATE_TRY(dalloc(h, &LB.nseg, (size_t)L.nparents));
        CREATE_TRY(dalloc(h, &LB.item_base, (size_t)L.nparents + 1));
        CREATE_TRY(dalloc(h, &LB.box_off, nboxes));
        CREATE_TRY(dalloc(h, &LB.box_len, nboxes));
        if (l > 0) CREATE_TRY(dalloc(h, &LB.ctab, 2 * (size_t)L.nparents * (size_t)L.nch));
    }
    const size_t T = (size_t)h->ntiles;
    ItemBufs& I = h->ibuf;
    CREATE_TRY(dalloc(h, &I.heavy_cnt, 2 * T));
    I.light_cnt = I.heavy_cnt + T;
    CREATE_TRY(dalloc(h, &I.heavy_off, 2 * T + 1));
    I.light_off = I.heavy_off + T;
    CREATE_TRY(dalloc_zero(h, &I.bsum, (size_t)scan_scratch_words(2 * (long long)T)));
    CREATE_TRY(dalloc(h, &I.tile_partial, 2 * T));
    CREATE_TRY(dalloc(h, &I.qheads, BG_QWORDS));
    {
        const char* tiny = std::getenv("K3DOM_TEST_TINY_CAPACITY");
        h->tiny_caps = tiny && tiny[0] == '1';
        // measurement-grid list segments: the power of two at or below sqrt(V)
        // within [256, 1024] (short segments shorten the longest wavefront,
        // which bounds small grids; long ones cost less per entry, which bounds
        // large grids); 64 under the test hook (many segments per list)
        int seg = 256;
        while (seg < 1024 && (long long)(2 * seg) * (2 * seg) <= (long long)h->V) seg *= 2;
        h->ds.seg_len = h->tiny_caps ? 64 : seg;
    }
    CREATE_TRY(hipHostMalloc((void**)&h->h_sc, sizeof(DevScalars), hipHostMallocDefault));
    CREATE_TRY(hipHostMalloc((void**)&h->h_fc, sizeof(FrameConsts), hipHostMallocDefault));
    CREATE_TRY(dalloc(h, &h->fc, 1));
    CREATE_TRY(hipHostMalloc((void**)&h->h_gfc, sizeof(FrameConsts) * K3DOM_MAX_LIDARS, hipHostMallocDefault));
    CREATE_TRY(dalloc(h, &h->gfc, K3DOM_MAX_LIDARS));
    h->gather_blocks = h->num_cus * 8;
    for (int e = 0; e < EV_COUNT; ++e) CREATE_TRY(hipEventCreate(&h->ev[e]));
    CREATE_TRY(dalloc(h, &h->pair_counts, 2));
    // prior fill and zeroed particle state (src/common/kernel/init.cu:20-35)
    for (int b = 0; b < 2; ++b)
        CREATE_TRY(launch_fill_voxels(h->stream, h->vox[b], h->V, h->p.prior_free, prior_s_slot(h),
                                      h->p.prior_dynamic));
    for (int s = 0; s < 2; ++s) {
        ParticleArrays& P = h->P[s];
        for (float* f : {P.px, P.py, P.pz, P.vx, P.vy, P.vz, P.w})
            CREATE_TRY(hipMemsetAsync(f, 0, N * sizeof(float), h->stream));
        CREATE_TRY(hipMemsetAsync(P.key, 0, N * sizeof(int32_t), h->stream));
    }
    CREATE_TRY(hipMemsetAsync(h->born, 0, V * sizeof(float), h->stream));
    CREATE_TRY(hipMemsetAsync(h->del_occ, 0, V * sizeof(float), h->stream));
    CREATE_TRY(hipMemsetAsync(h->del_free, 0, V * sizeof(float), h->stream));
    CREATE_TRY(hipMemsetAsync(h->ridx, 0, N * sizeof(int32_t), h->stream));
    CREATE_TRY(hipStreamSynchronize(h->stream));
#undef CREATE_TRY
    {
        // test hook: the n-th device allocation after creation fails once
        const char* fa = std::getenv("K3DOM_TEST_FAIL_ALLOC");
        h->fail_alloc_at = fa ? std::atoi(fa) : 0;
        h->alloc_count = 0;
    }
    *out = h;
    return K3DOM_OK;
}

k3dom_status k3dom_destroy(k3dom_handle* h) {
    if (!h) return K3DOM_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    if (h->h_pts) (void)hipHostFree(h->h_pts);
    if (h->h_sc) (void)hipHostFree(h->h_sc);
    if (h->h_fc) (void)hipHostFree(h->h_fc);
    if (h->h_gfc) (void)hipHostFree(h->h_gfc);
    if (h->h_gargs) (void)hipHostFree(h->h_gargs);
    drop_graph(h);
    for (int e = 0; e < EV_COUNT; ++e)
        if (h->ev[e]) (void)hipEventDestroy(h->ev[e]);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->cstream) (void)hipStreamDestroy(h->cstream);
    if (h->pts_ready) (void)hipEventDestroy(h->pts_ready);
    delete h;
    return K3DOM_OK;
}

// src/common/dom.cu:111-158
k3dom_status k3dom_update_pose(k3dom_handle* h, float x, float y, float z) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    int shifted[3];
    return update_pose(h, x, y, z, shifted);
}

k3dom_status k3dom_update_grid(k3dom_handle* h, double t, const float* xs, const float* ys, const float* zs,
                               int32_t n) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    if (h->stage != 0)
        return fail(h, K3DOM_ERR_STATE, "a frame driven by the stage methods is unfinished (resampling ends it)");
    if (n < 0 || (n > 0 && (!xs || !ys || !zs))) return fail(h, K3DOM_ERR_INVALID_ARG, "bad point arrays");
    k3dom_status s = ensure_points(h, std::max(n, 1));
    if (s != K3DOM_OK) return s;
    if (frame_uses_graph(h)) {
        // the graph reads the points at fixed offsets of the handle's buffer
        const size_t cap = (size_t)h->m_cap;
        if (n > 0) {
            std::memcpy(h->h_pts, xs, sizeof(float) * n);
            std::memcpy(h->h_pts + cap, ys, sizeof(float) * n);
            std::memcpy(h->h_pts + 2 * cap, zs, sizeof(float) * n);
            HIP_TRY(h, hipMemcpyAsync(h->d_pts, h->h_pts, sizeof(float) * 3 * cap, hipMemcpyHostToDevice, h->stream));
        }
        return frame(h, t, h->d_pts, h->d_pts + cap, h->d_pts + 2 * cap, n);
    }
    // eager frame: the n points packed, uploaded beside predict and sort
    const size_t m = (size_t)n;
    if (n > 0) {
        std::memcpy(h->h_pts, xs, sizeof(float) * m);
        std::memcpy(h->h_pts + m, ys, sizeof(float) * m);
        std::memcpy(h->h_pts + 2 * m, zs, sizeof(float) * m);
        HIP_TRY(h, hipMemcpyAsync(h->d_pts, h->h_pts, sizeof(float) * 3 * m, hipMemcpyHostToDevice, h->cstream));
        HIP_TRY(h, hipEventRecord(h->pts_ready, h->cstream));
        h->pts_pending = true;
    }
    return frame(h, t, h->d_pts, h->d_pts + m, h->d_pts + 2 * m, n);
}

k3dom_status k3dom_update_grid_device(k3dom_handle* h, double t, const float* d_xs, const float* d_ys,
                                      const float* d_zs, int32_t n) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    if (h->stage != 0)
        return fail(h, K3DOM_ERR_STATE, "a frame driven by the stage methods is unfinished (resampling ends it)");
    if (n < 0 || (n > 0 && (!d_xs || !d_ys || !d_zs))) return fail(h, K3DOM_ERR_INVALID_ARG, "bad point arrays");
    k3dom_status s = ensure_points(h, std::max(n, 1));
    if (s != K3DOM_OK) return s;
    return frame(h, t, d_xs, d_ys, d_zs, n);
}

// The node's PCL transform (src/k3dom/k3dom_node.cpp:73-83) on the GPU, for
// several LiDARs kept apart by their origins.
k3dom_status k3dom_update_grid_lidars(k3dom_handle* h, double t, const k3dom_lidar* lidars, int32_t n_lidars,
                                      int32_t points_on_device) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    if (h->stage != 0)
        return fail(h, K3DOM_ERR_STATE, "a frame driven by the stage methods is unfinished (resampling ends it)");
    if (!lidars || n_lidars < 1 || n_lidars > K3DOM_MAX_LIDARS)
        return fail(h, K3DOM_ERR_INVALID_ARG, "update_grid_lidars: 1 to K3DOM_MAX_LIDARS LiDARs");
    if (h->method == K3DOM_METHOD_DSPHDMIB && n_lidars > 1)
        return fail(h, K3DOM_ERR_INVALID_ARG, "update_grid_lidars: DS-PHD/MIB maps take one LiDAR");
    long long total = 0;
    for (int k = 0; k < n_lidars; ++k) {
        const k3dom_lidar& L = lidars[k];
        if (L.n < 0 || (L.n > 0 && (!L.xs || !L.ys || !L.zs)))
            return fail(h, K3DOM_ERR_INVALID_ARG, "update_grid_lidars: bad point arrays");
        for (float v : L.rot)
            if (!(v - v == 0.0f)) return fail(h, K3DOM_ERR_INVALID_ARG, "update_grid_lidars: non-finite rotation");
        for (float v : L.origin)
            if (!(v - v == 0.0f)) return fail(h, K3DOM_ERR_INVALID_ARG, "update_grid_lidars: non-finite origin");
        total += L.n;
    }
    if (total > 0x7fffffff) return fail(h, K3DOM_ERR_INVALID_ARG, "update_grid_lidars: too many points");
    const int m = (int)total;
    // LiDAR k's points at k * gstride of the handle's buffer: the slot only
    // grows, so frames of one rig keep their offsets (and a captured graph)
    int most = 1;
    for (int k = 0; k < n_lidars; ++k) most = std::max(most, lidars[k].n);
    if ((long long)most * n_lidars > 0x7fffffff)
        return fail(h, K3DOM_ERR_INVALID_ARG, "update_grid_lidars: too many points");
    const int stride = std::max(h->gstride, most);
    k3dom_status s = ensure_points(h, stride * n_lidars);
    if (s != K3DOM_OK) return s;
    h->gstride = stride;
    // the frame's points, rotated into map axes in the handle's buffer (host
    // points go through the pinned staging buffer)
    const size_t cap = (size_t)h->m_cap;
    float* dx = h->d_pts;
    float* dy = h->d_pts + cap;
    float* dz = h->d_pts + 2 * cap;
    for (int k = 0; k < n_lidars; ++k) {
        const k3dom_lidar& L = lidars[k];
        const int at = k * stride;
        if (!points_on_device && L.n > 0) {
            std::memcpy(h->h_pts + at, L.xs, sizeof(float) * L.n);
            std::memcpy(h->h_pts + cap + at, L.ys, sizeof(float) * L.n);
            std::memcpy(h->h_pts + 2 * cap + at, L.zs, sizeof(float) * L.n);
            for (int c = 0; c < 3; ++c)
                HIP_TRY(h, hipMemcpyAsync(h->d_pts + c * cap + at, h->h_pts + c * cap + at, sizeof(float) * L.n,
                                          hipMemcpyHostToDevice, h->stream));
        }
        h->gstart[k] = at;
        h->gcount[k] = L.n;
        for (int i = 0; i < 3; ++i) h->gorig[k][i] = L.origin[i];
    }
    for (int k = 0; k < n_lidars; ++k) {
        const k3dom_lidar& L = lidars[k];
        const size_t o = (size_t)h->gstart[k];
        const float *sx = points_on_device ? L.xs : dx + o, *sy = points_on_device ? L.ys : dy + o,
                    *sz = points_on_device ? L.zs : dz + o;
        HIP_TRY(h, launch_lidar_rotate(h->stream, L.rot, sx, sy, sz, dx + o, dy + o, dz + o, L.n));
    }
    h->ngroups = n_lidars;
    s = frame(h, t, dx, dy, dz, m);
    h->ngroups = 0;
    return s;
}

// ---------------------------------------------------------------- stage methods
// The reference's public stage methods (include/common/dom.h:89-94,
// include/k3dom/dom.h:24-28), in the order DOM::updateGrid calls them
// (src/k3dom/dom.cu:33-44).  Each runs its part of the frame's kernels and
// returns when it is done; a normal frame driven through them equals the
// same frame of k3dom_update_grid bit for bit.  Like the reference's stages
// they leave the stored time of the last frame to updateGrid.
namespace {

enum { STG_IDLE = 0, STG_PREDICTED, STG_ASSIGNED, STG_OCCUPANCY, STG_PERSISTENT, STG_NEWBORN };

k3dom_status stage_entry(k3dom_handle* h, int need, const char* what) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    if (h->method != K3DOM_METHOD_K3DOM)
        return fail(h, K3DOM_ERR_STATE, std::string(what) + ": the stage methods drive K3DOM maps");
    if (h->stage != need)
        return fail(h, K3DOM_ERR_STATE, std::string(what) + ": called out of the order of DOM::updateGrid "
                                        "(particlePrediction, particleAssignment, gridCellOccupancyUpdate, "
                                        "updatePersistentParticles, initializeNewParticles, statisticalMoments, "
                                        "resampling)");
    return K3DOM_OK;
}

// A failed stage abandons the frame: the next call starts a new one.
k3dom_status stage_done(k3dom_handle* h, k3dom_status s, int next) {
    if (s == K3DOM_OK) s = hipStreamSynchronize(h->stream) == hipSuccess
                               ? K3DOM_OK : fail(h, K3DOM_ERR_HIP, "stage: stream synchronisation failed");
    if (s != K3DOM_OK) {
        h->stage = STG_IDLE;
        h->last_frame_ok = false;
        return s;
    }
    h->stage = next;
    return K3DOM_OK;
}

// Points of a stage (host arrays) into the handle's point buffer; their count
// into the frame constants.
k3dom_status stage_points(k3dom_handle* h, const float* xs, const float* ys, const float* zs, int n) {
    if (n < 0 || (n > 0 && (!xs || !ys || !zs))) return fail(h, K3DOM_ERR_INVALID_ARG, "bad point arrays");
    k3dom_status s = ensure_points(h, std::max(n, 1));
    if (s != K3DOM_OK) return s;
    const size_t cap = (size_t)h->m_cap;
    for (int k = 0; k < 3 && n > 0; ++k)
        HIP_TRY(h, hipMemcpyAsync(h->d_pts + k * cap, k == 0 ? xs : (k == 1 ? ys : zs), sizeof(float) * n,
                                  hipMemcpyHostToDevice, h->stream));
    h->h_fc->m = n;
    return K3DOM_OK;
}

// Kernel inference over the handle's points and the mass update (normal) or
// the first frame's born masses and particle placement, re-run after growing
// what overflowed.
k3dom_status stage_evidence(k3dom_handle* h, bool normal) {
    const size_t cap = (size_t)h->m_cap;
    HIP_TRY(h, hipMemcpyAsync(h->fc, h->h_fc, sizeof(FrameConsts), hipMemcpyHostToDevice, h->stream));
    for (int attempt = 0;; ++attempt) {
        if (attempt > 0) {
            if (attempt >= 8) return fail(h, K3DOM_ERR_OUT_OF_MEMORY, "kernel-inference capacities keep overflowing");
            k3dom_status s = grow_after_overflow(h);
            if (s != K3DOM_OK) return s;
        }
        k3dom_status s = build_lists(h, h->fc, h->d_pts, h->d_pts + cap, h->d_pts + 2 * cap);
        if (s != K3DOM_OK) return s;
        HIP_TRY(h, launch_bgki(h->stream, h->gather_blocks, h->fc, h->tg, h->p.ls, h->sigma_occ, h->sigma_free,
                               h->ibuf, h->lbuf[2], h->rec, h->del_occ, h->del_free, h->sc, h->gather_waves, true));
        s = tail_mass(h, normal);
        if (s == K3DOM_OK && !normal) s = tail_newborn(h, false);
        if (s != K3DOM_OK) return s;
        HIP_TRY(h, hipMemcpyAsync(h->h_sc, h->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        if (!h->h_sc->overflow) return K3DOM_OK;
    }
}

}  // namespace

k3dom_status k3dom_particle_prediction(k3dom_handle* h, float dt) {
    k3dom_status s = stage_entry(h, STG_IDLE, "particlePrediction");
    if (s != K3DOM_OK) return s;
    const bool pshift_pending = h->pshift_pending;
    if ((s = begin_frame_voxels(h, true)) != K3DOM_OK) return s;
    fill_frame_consts_dt(h, h->h_fc, true, dt, 0);
    if (hipMemcpyAsync(h->fc, h->h_fc, sizeof(FrameConsts), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        launch_reset_scalars(h->stream, h->sc) != hipSuccess ||
        launch_predict(h->stream, h->fc, h->P[0], h->prec, h->N) != hipSuccess) {
        h->pshift_pending = pshift_pending;
        return stage_done(h, fail(h, K3DOM_ERR_HIP, "particlePrediction: launch failed"), STG_IDLE);
    }
    return stage_done(h, K3DOM_OK, STG_PREDICTED);
}

k3dom_status k3dom_particle_assignment(k3dom_handle* h) {
    k3dom_status s = stage_entry(h, STG_PREDICTED, "particleAssignment");
    if (s != K3DOM_OK) return s;
    VoxelArrays& v = h->vox[h->vcur];
    if (launch_radix_sort(h->stream, h->P[0].key, h->N, h->key_bits, h->sort, &h->last_skeys, &h->last_perm,
                          v.start, v.end, h->V) != hipSuccess ||
        launch_gather_sorted(h->stream, h->prec, h->P[1], h->wsort, h->last_perm, h->last_skeys, h->N, v.start,
                             v.end) != hipSuccess)
        s = fail(h, K3DOM_ERR_HIP, "particleAssignment: launch failed");
    return stage_done(h, s, STG_ASSIGNED);
}

k3dom_status k3dom_grid_cell_occupancy_update(k3dom_handle* h, float dt, const float* xs, const float* ys,
                                              const float* zs, int32_t n) {
    k3dom_status s = stage_entry(h, STG_ASSIGNED, "gridCellOccupancyUpdate");
    if (s != K3DOM_OK) return s;
    s = stage_points(h, xs, ys, zs, n);
    if (s == K3DOM_OK) {
        h->h_fc->decay = (float)std::pow((double)h->p.gamma, (double)dt);
        s = stage_evidence(h, true);
    }
    return stage_done(h, s, STG_OCCUPANCY);
}

k3dom_status k3dom_update_persistent_particles(k3dom_handle* h) {
    k3dom_status s = stage_entry(h, STG_OCCUPANCY, "updatePersistentParticles");
    if (s != K3DOM_OK) return s;
    return stage_done(h, tail_persistent(h), STG_PERSISTENT);
}

k3dom_status k3dom_initialize_new_particles(k3dom_handle* h) {
    k3dom_status s = stage_entry(h, STG_PERSISTENT, "initializeNewParticles");
    if (s != K3DOM_OK) return s;
    return stage_done(h, tail_newborn(h, true), STG_NEWBORN);
}

k3dom_status k3dom_statistical_moments(k3dom_handle* h) {
    // computed with the persistent update (same weights, same sums)
    k3dom_status s = stage_entry(h, h && h->stage == STG_PERSISTENT ? STG_PERSISTENT : STG_NEWBORN,
                                 "statisticalMoments");
    return s;
}

k3dom_status k3dom_resampling(k3dom_handle* h) {
    k3dom_status s = stage_entry(h, STG_NEWBORN, "resampling");
    if (s != K3DOM_OK) return s;
    s = tail_resample(h);
    if (s == K3DOM_OK && hipMemcpyAsync(h->h_sc, h->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, h->stream) !=
                             hipSuccess)
        s = fail(h, K3DOM_ERR_HIP, "resampling: scalar readback failed");
    s = stage_done(h, s, STG_IDLE);
    if (s == K3DOM_OK) count_frame(h, true, h->h_fc->m);
    return s;
}

k3dom_status k3dom_initialize_particles(k3dom_handle* h, const float* xs, const float* ys, const float* zs,
                                        int32_t n) {
    k3dom_status s = stage_entry(h, STG_IDLE, "initializeParticles");
    if (s != K3DOM_OK) return s;
    if ((s = begin_frame_voxels(h, false)) != K3DOM_OK) return s;
    const bool pshift_pending = h->pshift_pending;
    const int pshift[3] = {h->pshift[0], h->pshift[1], h->pshift[2]};
    fill_frame_consts_dt(h, h->h_fc, false, 0.0f, 0);
    s = stage_points(h, xs, ys, zs, n);
    if (s == K3DOM_OK && launch_reset_scalars(h->stream, h->sc) != hipSuccess)
        s = fail(h, K3DOM_ERR_HIP, "initializeParticles: launch failed");
    if (s == K3DOM_OK) s = stage_evidence(h, false);
    s = stage_done(h, s, STG_IDLE);
    if (s == K3DOM_OK) {
        count_frame(h, false, n);
    } else {
        h->pshift_pending = pshift_pending;
        for (int k = 0; k < 3; ++k) h->pshift[k] = pshift[k];
    }
    return s;
}

k3dom_status k3dom_get_info(const k3dom_handle* h, k3dom_info* out) {
    if (!h || !out) return K3DOM_ERR_INVALID_ARG;
    out->grid_size = h->gs;
    out->grid_size_z = h->gsz;
    out->grid_cell_count = h->V;
    out->particle_count = h->N;
    out->new_born_particle_count = h->B;
    out->iteration = (int32_t)h->frame;
    out->resolution = h->p.resolution;
    for (int i = 0; i < 3; ++i) {
        out->center_pos[i] = h->center[i];
        out->sensor_pos[i] = h->sensor[i];
    }
    out->updated_time = h->updated_time;
    return K3DOM_OK;
}

k3dom_status k3dom_read_masses(k3dom_handle* h, float* f, float* s, float* d, float* mvx, float* mvy, float* mvz) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    {
        k3dom_status st = settle_shift(h);
        if (st != K3DOM_OK) return st;
    }
    const VoxelArrays& v = h->vox[h->vcur];
    const size_t bytes = sizeof(float) * (size_t)h->V;
    const float* src[6] = {v.f, v.s, v.d, v.mvx, v.mvy, v.mvz};
    float* dst[6] = {f, s, d, mvx, mvy, mvz};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) HIP_TRY(h, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return K3DOM_OK;
}

k3dom_status k3dom_read_cells(k3dom_handle* h, k3dom_cell* out) {
    if (!h || !out) return K3DOM_ERR_INVALID_ARG;
    {
        k3dom_status st = settle_shift(h);
        if (st != K3DOM_OK) return st;
    }
    const VoxelArrays& v = h->vox[h->vcur];
    const size_t V = (size_t)h->V;
    std::vector<float> buf(V * 10);
    std::vector<int32_t> ib(V * 2);
    const float* src[10] = {v.f, v.s, v.d, v.pers, v.pred, v.mu, v.wua, v.mvx, v.mvy, v.mvz};
    for (int k = 0; k < 10; ++k)
        HIP_TRY(h, hipMemcpyAsync(buf.data() + k * V, src[k], V * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(ib.data(), v.start, V * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(ib.data() + V, v.end, V * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (size_t c = 0; c < V; ++c) {
        k3dom_cell& o = out[c];
        o.start_idx = ib[c];
        o.end_idx = ib[V + c];
        o.mu_A = buf[5 * V + c];
        o.mu_UA = buf[5 * V + c];
        o.w_A = 0.0f;
        o.w_UA = buf[6 * V + c];
        o.mean_x_vel = buf[7 * V + c];
        o.mean_y_vel = buf[8 * V + c];
        o.mean_z_vel = buf[9 * V + c];
        o.likelihood = 1.0f;
        o.p_A = 0.0f;
        o.free_mass = buf[0 * V + c];
        if (h->method == K3DOM_METHOD_DSPHDMIB) {
            // the second mass slot holds the occupied mass; static and dynamic keep their priors
            o.static_mass = h->p.prior_static;
            o.dynamic_mass = h->p.prior_dynamic;
            o.occ_mass = buf[1 * V + c];
        } else {
            o.static_mass = buf[1 * V + c];
            o.dynamic_mass = buf[2 * V + c];
            o.occ_mass = h->p.prior_occ;
        }
        o.pers_mass = buf[3 * V + c];
        o.pred_mass = buf[4 * V + c];
    }
    return K3DOM_OK;
}

// src/k3dom/k3dom_node.cpp:161-184 / src/dsphdmib/dsphdmib_node.cpp:153-177
// (all voxels), :128-150 / :127-141 (the visualisation filter)
k3dom_status k3dom_export_mappoints(k3dom_handle* h, int32_t mode, float min_occ, float min_mass,
                                   k3dom_mappoint* out, int64_t capacity, int64_t* count) {
    if (!h || !count || (mode != K3DOM_MAPPOINTS_ALL && mode != K3DOM_MAPPOINTS_VISIBLE) || capacity < 0 ||
        (!out && capacity > 0))
        return K3DOM_ERR_INVALID_ARG;
    {
        k3dom_status st = settle_shift(h);
        if (st != K3DOM_OK) return st;
    }
    const size_t V = (size_t)h->V;
    // each buffer on its own while missing: an allocation that failed in an
    // earlier call leaves the others in place and is tried again here
    if (!h->mp_out) HIP_TRY(h, dalloc(h, &h->mp_out, 3 * V));
    if (!h->mp_flags) HIP_TRY(h, dalloc(h, &h->mp_flags, V));
    if (!h->mp_off) HIP_TRY(h, dalloc(h, &h->mp_off, V + 1));
    if (!h->mp_scratch) HIP_TRY(h, dalloc_zero(h, &h->mp_scratch, (size_t)scan_scratch_words((long long)V)));
    if (!h->mp_out || !h->mp_flags || !h->mp_off || !h->mp_scratch)
        return fail(h, K3DOM_ERR_STATE, "export_mappoints: a buffer is missing");
    MapPointArgs a;
    a.gs = h->gs;
    a.gsz = h->gsz;
    a.res = h->p.resolution;
    a.size = h->p.size;
    a.size_z = h->p.size_z;
    a.cx = h->center[0];
    a.cy = h->center[1];
    a.cz = h->center[2];
    a.ds = h->method == K3DOM_METHOD_DSPHDMIB ? 1 : 0;
    a.prior_s = h->p.prior_static;
    a.prior_d = h->p.prior_dynamic;
    a.prior_occ = h->p.prior_occ;
    a.min_occ = min_occ;
    a.min_mass = min_mass;
    HIP_TRY(h, launch_export_mappoints(h->stream, h->vox[h->vcur], h->V, a, mode, h->mp_flags, h->mp_off,
                                       h->mp_scratch, &h->sc->mp_total, h->mp_out));
    int64_t n = (int64_t)V;
    if (mode == K3DOM_MAPPOINTS_VISIBLE) {
        uint32_t total = 0;
        HIP_TRY(h, hipMemcpyAsync(&total, &h->sc->mp_total, 4, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        n = total;
    }
    *count = n;
    if (!out && capacity == 0) return K3DOM_OK;
    if (capacity < n) return fail(h, K3DOM_ERR_INVALID_ARG, "export_mappoints: capacity below the record count");
    static_assert(sizeof(k3dom_mappoint) == 3 * sizeof(float4), "MapPoint record layout");
    if (n > 0)
        HIP_TRY(h, hipMemcpyAsync(out, h->mp_out, (size_t)n * sizeof(k3dom_mappoint), hipMemcpyDeviceToHost,
                                  h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return K3DOM_OK;
}

k3dom_status k3dom_read_array(k3dom_handle* h, int32_t which, void* dst, int64_t capacity) {
    if (!h || !dst) return K3DOM_ERR_INVALID_ARG;
    {
        k3dom_status st = settle_particles(h);
        if (st != K3DOM_OK) return st;
    }
    const size_t N = (size_t)h->N, B = (size_t)h->B, V = (size_t)h->V;
    auto cp = [&](const void* src, size_t bytes, size_t at) -> k3dom_status {
        if ((int64_t)(at + bytes) > capacity) return fail(h, K3DOM_ERR_INVALID_ARG, "capacity too small");
        if (!src) return fail(h, K3DOM_ERR_STATE, "array not available before the first normal frame");
        HIP_TRY(h, hipMemcpyAsync((char*)dst + at, src, bytes, hipMemcpyDeviceToHost, h->stream));
        return K3DOM_OK;
    };
    k3dom_status s = K3DOM_OK;
    const ParticleArrays& P = h->P[0];
    switch (which) {
    case K3DOM_ARR_PARTICLE_POS:
        if ((s = cp(P.px, N * 4, 0)) || (s = cp(P.py, N * 4, N * 4)) || (s = cp(P.pz, N * 4, 2 * N * 4))) return s;
        break;
    case K3DOM_ARR_PARTICLE_VEL:
        if ((s = cp(P.vx, N * 4, 0)) || (s = cp(P.vy, N * 4, N * 4)) || (s = cp(P.vz, N * 4, 2 * N * 4))) return s;
        break;
    case K3DOM_ARR_PARTICLE_WEIGHT: s = cp(P.w, N * 4, 0); break;
    case K3DOM_ARR_PARTICLE_KEY: s = cp(P.key, N * 4, 0); break;
    case K3DOM_ARR_SORT_PERM: s = cp(h->last_perm, N * 4, 0); break;
    case K3DOM_ARR_SORTED_KEY: s = cp(h->last_skeys, N * 4, 0); break;
    case K3DOM_ARR_RESAMPLE_IDX: s = cp(h->ridx, N * 4, 0); break;
    case K3DOM_ARR_NEWBORN_POS:
        if ((s = cp(h->NB.px, B * 4, 0)) || (s = cp(h->NB.py, B * 4, B * 4)) || (s = cp(h->NB.pz, B * 4, 2 * B * 4)))
            return s;
        break;
    case K3DOM_ARR_NEWBORN_VEL:
        if ((s = cp(h->NB.vx, B * 4, 0)) || (s = cp(h->NB.vy, B * 4, B * 4)) || (s = cp(h->NB.vz, B * 4, 2 * B * 4)))
            return s;
        break;
    case K3DOM_ARR_NEWBORN_WEIGHT: s = cp(h->NB.w, B * 4, 0); break;
    case K3DOM_ARR_NEWBORN_KEY: s = cp(h->NB.key, B * 4, 0); break;
    case K3DOM_ARR_BORN_MASS: s = cp(h->born, V * 4, 0); break;
    case K3DOM_ARR_DEL_OCC: s = cp(h->del_occ, V * 4, 0); break;
    case K3DOM_ARR_DEL_FREE: s = cp(h->del_free, V * 4, 0); break;
    case K3DOM_ARR_SORTED_WEIGHT: s = cp(h->P[1].w, N * 4, 0); break;
    default: return fail(h, K3DOM_ERR_INVALID_ARG, "unknown array id");
    }
    if (s != K3DOM_OK) return s;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return K3DOM_OK;
}

k3dom_status k3dom_set_timing(k3dom_handle* h, int32_t enable) {
    if (!h || enable < 0 || enable > 2) return K3DOM_ERR_INVALID_ARG;
    h->timing = enable;
    return K3DOM_OK;
}

k3dom_status k3dom_autotune(k3dom_handle* h, int32_t* blocks_per_cu) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    if (h->method != K3DOM_METHOD_K3DOM) return fail(h, K3DOM_ERR_STATE, "autotune tunes the K3DOM gather");
    if (h->normal_here < 1) return fail(h, K3DOM_ERR_STATE, "autotune needs a normal frame first");
    if (!h->last_frame_ok || point_buffers_missing(h))
        return fail(h, K3DOM_ERR_STATE, "autotune needs the last frame to have completed");
    if (h->last_grouped) return fail(h, K3DOM_ERR_STATE, "autotune needs a single-sensor frame last");
    EventPair ev;  // destroyed on every return below
    HIP_TRY(h, ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    // Median-free timing: 3 runs after one warm-up, summed.
    auto timed = [&](auto&& enqueue, float* out) -> k3dom_status {
        float tot = 0.0f;
        for (int rep = 0; rep < 4; ++rep) {
            HIP_TRY(h, hipEventRecord(e0, h->stream));
            k3dom_status s = enqueue();
            if (s != K3DOM_OK) return s;
            HIP_TRY(h, hipEventRecord(e1, h->stream));
            HIP_TRY(h, hipEventSynchronize(e1));
            float ms = 0.0f;
            HIP_TRY(h, hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) tot += ms;
        }
        *out = tot;
        return K3DOM_OK;
    };
    // 1. list segment lengths: the last frame's candidate lists are rebuilt
    // from its point records (h->pxyz, h->fc) for each choice; the lists
    // (ascending per box) and the work items do not depend on it
    auto rebuild_lists = [&]() -> k3dom_status {
        HIP_TRY(h, launch_build_levels(h->stream, h->fc, h->tg, h->ldesc, h->lbuf, h->ibuf, h->p.ls, h->pxyz, h->sc));
        HIP_TRY(h, launch_build_items(h->stream, h->ntiles, h->lbuf[2], h->ibuf, h->sc));
        return K3DOM_OK;
    };
    auto set_segs = [&](int s0, int s12) -> k3dom_status {
        h->ldesc[0].seg = s0;
        h->ldesc[1].seg = h->ldesc[2].seg = s12;
        long long want[3];
        for (int l = 0; l < 3; ++l) want[l] = h->lbuf[l].list_cap;
        return ensure_caps(h, want, h->ibuf.heavy_cap);
    };
    const int segs0[] = {128, 256, 512};
    const int segs12[] = {256, 512, 1024, 2048};
    float best = 1e30f;
    int b0 = h->ldesc[0].seg, b12 = h->ldesc[2].seg;
    for (int s0 : segs0)
        for (int s12 : segs12) {
            k3dom_status s = set_segs(s0, s12);
            if (s != K3DOM_OK) return s;
            float ms = 0.0f;
            if ((s = timed(rebuild_lists, &ms)) != K3DOM_OK) return s;
            if (ms < best) {
                best = ms;
                b0 = s0;
                b12 = s12;
            }
        }
    {
        k3dom_status s = set_segs(b0, b12);
        if (s == K3DOM_OK) s = rebuild_lists();
        if (s != K3DOM_OK) return s;
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
    // 2. the gather's block size (wavefronts) and grid size, over the same
    // items (the gather only rewrites the same evidence); at most 32
    // wavefronts per compute unit
    best = 1e30f;
    int best_w = h->gather_waves, best_c = h->gather_blocks / h->num_cus;
    for (int wv : {4, 8, 16})
        for (int c : {1, 2, 3, 4, 5, 6, 8}) {
            if (wv * c > 32) continue;
            auto gather = [&]() -> k3dom_status {
                HIP_TRY(h, launch_bgki(h->stream, c * h->num_cus, h->fc, h->tg, h->p.ls, h->sigma_occ,
                                       h->sigma_free, h->ibuf, h->lbuf[2], h->rec, h->del_occ, h->del_free, h->sc,
                                       wv));
                return K3DOM_OK;
            };
            float ms = 0.0f;
            k3dom_status s = timed(gather, &ms);
            if (s != K3DOM_OK) return s;
            if (ms < best) {
                best = ms;
                best_w = wv;
                best_c = c;
            }
        }
    h->gather_waves = best_w;
    h->gather_blocks = best_c * h->num_cus;
    drop_graph(h);
    if (blocks_per_cu) *blocks_per_cu = best_c;
    return K3DOM_OK;
}

k3dom_status k3dom_set_launch_shape(k3dom_handle* h, const int32_t shape[4]) {
    if (!h || !shape) return K3DOM_ERR_INVALID_ARG;
    if (shape[0] < 1 || shape[0] > 16 || (shape[1] != 4 && shape[1] != 8 && shape[1] != 16) || shape[2] < 32 ||
        shape[2] > 8192 || shape[3] < 32 || shape[3] > 8192)
        return fail(h, K3DOM_ERR_INVALID_ARG,
                    "set_launch_shape: blocks per CU 1-16, 4/8/16 wavefronts, list segments 32-8192");
    if (h->stage != 0) return fail(h, K3DOM_ERR_STATE, "set_launch_shape between frames only");
    h->gather_blocks = shape[0] * h->num_cus;
    h->gather_waves = shape[1];
    h->ldesc[0].seg = shape[2];
    h->ldesc[1].seg = h->ldesc[2].seg = shape[3];
    drop_graph(h);
    if (h->m_cap == 0) return K3DOM_OK;  // sized with the first frame's points
    long long want[3];
    for (int l = 0; l < 3; ++l) want[l] = h->lbuf[l].list_cap;
    return ensure_caps(h, want, h->ibuf.heavy_cap);
}

k3dom_status k3dom_launch_shape(const k3dom_handle* h, int32_t shape[4]) {
    if (!h || !shape) return K3DOM_ERR_INVALID_ARG;
    shape[0] = h->gather_blocks / std::max(1, h->num_cus);
    shape[1] = h->gather_waves;
    shape[2] = h->ldesc[0].seg;
    shape[3] = h->ldesc[2].seg;
    return K3DOM_OK;
}

k3dom_status k3dom_set_graph(k3dom_handle* h, int32_t enable) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    h->use_graph = enable != 0;
    if (!h->use_graph) drop_graph(h);
    return K3DOM_OK;
}

k3dom_status k3dom_stage_times(k3dom_handle* h, float* ms, int32_t count) {
    if (!h || !ms || count < 0) return K3DOM_ERR_INVALID_ARG;
    for (int i = 0; i < count && i < ST_COUNT; ++i) ms[i] = h->stage_ms[i];
    return K3DOM_OK;
}

k3dom_status k3dom_get_stats(const k3dom_handle* h, k3dom_stats* out) {
    if (!h || !out) return K3DOM_ERR_INVALID_ARG;
    *out = h->stats;
    return K3DOM_OK;
}

k3dom_status k3dom_count_bgki_pairs(k3dom_handle* h, uint64_t* out2) {
    if (!h || !out2) return K3DOM_ERR_INVALID_ARG;
    if (h->method != K3DOM_METHOD_K3DOM) return fail(h, K3DOM_ERR_STATE, "pairs are counted for K3DOM maps");
    if (h->frames_here == 0) return fail(h, K3DOM_ERR_STATE, "no frame processed yet");
    if (h->last_grouped) return fail(h, K3DOM_ERR_STATE, "the pair counter covers single-sensor frames");
    if (!h->last_frame_ok || point_buffers_missing(h))
        return fail(h, K3DOM_ERR_STATE, "the pair counter needs the last frame to have completed");
    HIP_TRY(h, hipMemsetAsync(h->pair_counts, 0, 2 * sizeof(unsigned long long), h->stream));
    const uint32_t n_items = h->h_sc->n_light + h->h_sc->n_heavy;
    HIP_TRY(h, launch_bgki_count(h->stream, h->num_cus * 8, h->last_geom, h->tg, h->p.ls, h->ibuf, n_items,
                                 h->lbuf[2], h->rec, h->pair_counts));
    unsigned long long c[2];
    HIP_TRY(h, hipMemcpyAsync(c, h->pair_counts, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    out2[0] = c[0];
    out2[1] = c[1];
    return K3DOM_OK;
}

const char* k3dom_last_error(const k3dom_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

}  // extern "C"
