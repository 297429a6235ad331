// Kernel-inference (BGKI) evidence of one frame on gfx950, replacing the
// reference's all-pairs loop (src/k3dom/kernel/mass_update.cu:32-77, called
// once per voxel at :176).
//
// 1. k_point_prep: one 48-byte record per point (k3_point_prep).
// 2. Candidate lists: for a hierarchy of boxes of 4x4x4-voxel tiles, the
//    ascending list of points whose segment [sensor, point] passes within
//    ls + box radius of the box (a superset of the points with a non-zero
//    term for any voxel of the box).  A work item is a segment of a parent's
//    list; its block tests each candidate against every child box once.
//    count -> exclusive scan -> ordered fill keeps every list ascending.
// 3. k_bgki: one wavefront per work item, lane = voxel of a tile; candidates
//    are staged in LDS and visited in ascending order by all lanes.  Short
//    lists are one item; long lists are split into one item per aligned
//    2^K3_BGKI_CHUNK_LOG2-point chunk whose partial sums k_bgki_combine adds
//    in chunk order.  Both realise the two-level association of k3_math.h,
//    so the sums are bit-identical to the serial restatement.
// Capacities are fixed per frame; when a list or the item array would
// overflow, the kernels set DevScalars::overflow and the host grows the
// buffers and re-runs this stage and the frame tail.
#include "k3_kernels.h"

#include <algorithm>

namespace k3 {

namespace {
constexpr int TPB = 256;
constexpr int BG_WAVES = TPB / 64;
// the gather runs blocks of W wavefronts (4, 8 or 16; picked by the
// autotuner, 4 by default): one kernel table in LDS per block
constexpr uint32_t BG_HEAVY_LEN = 1u << K3_BGKI_CHUNK_LOG2;
constexpr uint32_t BG_DIRECT = 0xffffffffu;

inline unsigned blocks_for(long long n, int per_block = TPB) {
    long long b = (n + per_block - 1) / per_block;
    return (unsigned)(b < 1 ? 1 : b);
}
}  // namespace

__device__ __forceinline__ uint32_t bg_lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ void bg_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// A value equal in every lane, held in a scalar register.
__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
// The word of lane q (q the same in every lane) in every lane.
__device__ __forceinline__ uint32_t lane_word(uint32_t v, uint32_t q) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)q);
}
__device__ __forceinline__ float lane_float(float v, int q) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), q));
}

// ============================================================== points
__global__ __launch_bounds__(TPB) void k_point_prep(const FrameConsts* __restrict__ fc,
                                                    const float* __restrict__ xs, const float* __restrict__ ys,
                                                    const float* __restrict__ zs, float res, float sk,
                                                    K3PointRec* __restrict__ rec, float4* __restrict__ pxyz) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= fc->m) return;
    const float x = xs[i], y = ys[i], z = zs[i];
    K3PointRec r = k3_point_prep(x, y, z, res, sk);
    rec[i] = r;
    pxyz[i] = make_float4(x, y, z, r.valid != 0.0f ? sqrtf(fmaf(z, z, fmaf(y, y, x * x))) : -1.0f);
}
hipError_t launch_point_prep(hipStream_t st, const FrameConsts* fc, const float* xs, const float* ys,
                             const float* zs, int m_cap, float res, float ls, K3PointRec* rec, float4* pxyz) {
    if (m_cap <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_point_prep, dim3(blocks_for(m_cap)), dim3(TPB), 0, st, fc, xs, ys, zs, res, k3_kernel_sk(ls),
                       rec, pxyz);
    return hipGetLastError();
}

// ============================================================== geometry
// Sensor-relative coordinate of the centre of voxel index i along an axis of
// n voxels (the per-axis part of k3_voxel_center).
__device__ __forceinline__ float axis_center(int i, int n, float res, float s) {
    return (((float)i + 0.5f) - (float)n / 2.0f) * res - s;
}

// Bounding sphere of the voxel centres of tiles [t0, t1) per axis, as
// (centre, ls + radius + slack); every voxel centre of the box lies within the
// radius of the centre.
__device__ float4 box_sphere(const GridGeom& g, float ls, int tx0, int ty0, int tz0, int tx1, int ty1, int tz1) {
    int i0 = tx0 * 4, i1 = min(tx1 * 4, g.gs) - 1;
    int j0 = ty0 * 4, j1 = min(ty1 * 4, g.gs) - 1;
    int k0 = tz0 * 4, k1 = min(tz1 * 4, g.gsz) - 1;
    float x0 = axis_center(i0, g.gs, g.res, g.sx), x1 = axis_center(i1, g.gs, g.res, g.sx);
    float y0 = axis_center(j0, g.gs, g.res, g.sy), y1 = axis_center(j1, g.gs, g.res, g.sy);
    float z0 = axis_center(k0, g.gsz, g.res, g.sz), z1 = axis_center(k1, g.gsz, g.res, g.sz);
    float cx = 0.5f * (x0 + x1), cy = 0.5f * (y0 + y1), cz = 0.5f * (z0 + z1);
    float hx = 0.5f * (x1 - x0), hy = 0.5f * (y1 - y0), hz = 0.5f * (z1 - z0);
    float R = sqrtf(hx * hx + hy * hy + hz * hz);
    float cn = fabsf(cx) + fabsf(cy) + fabsf(cz);
    return make_float4(cx, cy, cz, ls + R + 2e-3f + 1e-5f * cn);
}

// A candidate point of a list level: the segment [sensor, p] and its slack.
// An invalid candidate (or a lane past the segment's end) gets NaN
// coordinates, so every distance test on it is false without a branch.
struct Cand {
    float x, y, z, mm, half_inv_mm, slack;
};
__device__ __forceinline__ Cand make_cand(float4 p, bool in_range) {
    Cand c;
    const bool valid = in_range && p.w >= 0.0f;
    const float nan = __builtin_nanf("");
    c.x = valid ? p.x : nan;
    c.y = valid ? p.y : nan;
    c.z = valid ? p.z : nan;
    c.mm = c.x * c.x + c.y * c.y + c.z * c.z;
    c.half_inv_mm = c.mm > 0.0f ? 0.5f * __builtin_amdgcn_rcpf(c.mm) : 0.0f;
    c.slack = 1e-5f * p.w;
    return c;
}
// Per-child constants of the distance test: twice the sphere centre b of
// the box, |b|^2, the radius w, and w^2 plus the rounding margin below (two
// float4: the children of a parent sit contiguously in the per-level table,
// read with scalar loads).
struct __align__(16) ChildBox {
    float bx2, by2, bz2, bb, w, wsq, pad0, pad1;
};
__device__ __forceinline__ ChildBox make_child(float4 b, bool real) {
    ChildBox k;
    k.bx2 = real ? 2.0f * b.x : 0.0f;
    k.by2 = real ? 2.0f * b.y : 0.0f;
    k.bz2 = real ? 2.0f * b.z : 0.0f;
    k.bb = real ? fmaf(b.z, b.z, fmaf(b.y, b.y, b.x * b.x)) : __builtin_inff();
    k.w = real ? b.w : 0.0f;
    k.wsq = real ? fmaf(k.bb, 0x1p-18f, b.w * b.w) : 0.0f;
    k.pad0 = k.pad1 = 0.0f;
    return k;
}
// Squared distance from a box's sphere centre b to the segment [0, p] of
// candidate c against the candidate's threshold thr2: with
// t = clamp(b . p / |p|^2, 0, 1) the nearest point is t p and
// |b - t p|^2 = |b|^2 - t (2 b . p - t |p|^2).  The expansion cancels: its
// rounding error stays below about 40 * 2^-24 |b|^2 (the dot product's, the
// reciprocal's and the final sums'), so the threshold carries 2^-18 |b|^2 on
// top of (w + slack)^2 (in wsq), and the test stays a superset of the pairs
// with a non-zero term.  A child outside the tile grid has |b|^2 = +inf and
// never passes.
__device__ __forceinline__ bool cand_hits(const ChildBox& k, const Cand& c, float thr2) {
    const float dp2 = fmaf(k.bz2, c.z, fmaf(k.by2, c.y, k.bx2 * c.x));  // 2 b . p
    const float t = __builtin_amdgcn_fmed3f(dp2 * c.half_inv_mm, 0.0f, 1.0f);
    const float d2 = fmaf(-t, fmaf(-t, c.mm, dp2), k.bb);
    return d2 < thr2;
}
// The threshold of candidate c shared by the children of a group: with the
// group's largest radius and largest w^2 + margin it is at least each
// child's own (w + slack)^2 + 2^-18 |b|^2, so the shared test only widens.
__device__ __forceinline__ float cand_thr2(const Cand& c, float wmax, float wsqmax) {
    return fmaf(c.slack, fmaf(2.0f, wmax, c.slack), wsqmax);
}

// Child c of parent P at this level (level 0: the box with id c).  Returns the
// box id, or -1 when the child lies outside the tile grid.
__device__ int child_box(int level, const LevelDesc& L, const TileGeom& tg, int P, int c, int* t0, int* t1) {
    int bxi, byi, bzi;
    if (level == 0) {
        bxi = c % L.nbx;
        byi = (c / L.nbx) % L.nby;
        bzi = c / (L.nbx * L.nby);
    } else {
        int pxi = P % L.pnbx, pyi = (P / L.pnbx) % L.pnby, pzi = P / (L.pnbx * L.pnby);
        int cx = c % L.ncx, cy = (c / L.ncx) % L.ncy, cz = c / (L.ncx * L.ncy);
        bxi = pxi * L.ncx + cx;
        byi = pyi * L.ncy + cy;
        bzi = pzi * L.ncz + cz;
        if (bxi >= L.nbx || byi >= L.nby || bzi >= L.nbz) return -1;
    }
    t0[0] = bxi * L.bx; t0[1] = byi * L.by; t0[2] = bzi * L.bz;
    t1[0] = min(t0[0] + L.bx, tg.ntx); t1[1] = min(t0[1] + L.by, tg.nty); t1[2] = min(t0[2] + L.bz, tg.ntz);
    return bxi + L.nbx * (byi + L.nby * bzi);
}

// ============================================================== candidate lists
struct ItemCtx {
    int P;
    uint32_t seg, nsegP, ib, begin, end, poff;
};
// Work items of level 0 (one parent: every point): ceil(m / seg).
__device__ __forceinline__ uint32_t level0_items(const LevelDesc& L, int m) {
    return ((uint32_t)m + (uint32_t)L.seg - 1u) / (uint32_t)L.seg;
}
// The parent of work item `item` (the same in every lane): the largest P
// with item_base[P] <= item.  Each round the wavefront samples 64 evenly
// spaced entries of the bracket [lo, lo + span) and keeps the last sample not
// above item, so nparents <= 4096 takes two dependent loads instead of the
// twelve of a binary search.  item_base[lo] <= item holds throughout (lane
// 0 samples lo), so the ballot is never empty.
__device__ __forceinline__ int find_parent(const uint32_t* __restrict__ item_base, int nparents, uint32_t item) {
    const int lane = (int)(threadIdx.x & 63);
    int lo = 0, span = nparents;
    while (span > 1) {
        const int step = (span + 63) >> 6;
        const int j = lo + lane * step;
        const bool ok = lane * step < span && item_base[j] <= item;
        const uint64_t bal = __ballot(ok);
        const int hi = 63 - __builtin_clzll(bal);
        lo += hi * step;
        span = min(step, span - hi * step);
    }
    return lo;
}

__device__ __forceinline__ ItemCtx item_ctx(int level, const LevelDesc& L, uint32_t item,
                                            const uint32_t* __restrict__ item_base,
                                            const uint32_t* __restrict__ parent_off,
                                            const uint32_t* __restrict__ parent_len, int m) {
    ItemCtx x;
    if (level == 0) {
        x.P = 0;
        x.ib = 0;
        x.nsegP = level0_items(L, m);
    } else {
        x.P = find_parent(item_base, L.nparents, item);
        x.ib = item_base[x.P];
        x.nsegP = item_base[x.P + 1] - x.ib;
    }
    x.seg = item - x.ib;
    uint32_t plen = level ? parent_len[x.P] : (uint32_t)m;
    x.poff = level ? parent_off[x.P] : 0u;
    x.begin = x.seg * (uint32_t)L.seg;
    x.end = min(plen, x.begin + (uint32_t)L.seg);
    return x;
}

// Children are handled in groups of LV_CG, one wavefront per group, so the
// children of a box with many (level 0: all boxes) are spread over several
// wavefronts: task t takes item t / ncg and children [LV_CG * (t % ncg), ...)
// of it.  Eight children per wavefront keep each wavefront's serial work
// short and give every level thousands of wavefronts to hide load latency.
constexpr int LV_CG = 8;
static_assert(LV_CG <= 8, "a candidate's hits per child group are one mask byte");
__host__ __device__ inline int child_groups(const LevelDesc& L) { return (L.nch + LV_CG - 1) / LV_CG; }
int level_child_groups(const LevelDesc& L) { return child_groups(L); }
// blocks of the count and fill launches: one wavefront per (item, child group)
unsigned level_task_blocks(long long items_cap, const LevelDesc& L) {
    return blocks_for(std::max(1ll, items_cap * child_groups(L)), BG_WAVES);
}

// Point index of candidate e of an item's segment, e clamped to the last
// one (an item holds at least one candidate).  The load is unconditional,
// so a load of the next group stays in flight until its value is used.
__device__ __forceinline__ uint32_t cand_index(int level, const uint32_t* __restrict__ parent_list, const ItemCtx& x,
                                               uint32_t e) {
    const uint32_t ec = min(e, x.end - 1u);
    return level ? parent_list[x.poff + ec] : ec;
}

// One wavefront per task: a work item and one group of at most LV_CG of its
// children (task = item * ncg + group).  Without block-wide barriers the
// four wavefronts of a block work on four tasks independently, so many items
// are in flight per CU while each waits on its chain of dependent loads
// (parent, bounds, candidate indices, positions).
__device__ __forceinline__ uint32_t lv_task() { return blockIdx.x * (uint32_t)BG_WAVES + (threadIdx.x >> 6); }

// Per-task, per-child counts of the candidates of the item's segment that
// pass the child's test, and per candidate the children it passes (for the
// fill).  NC (= LV_CG) bounds the group's children, nc of them real.  Lane
// c sets up child c's test constants, which every lane then reads from lane
// c: wavefront-uniform, they stay in scalar registers (as do the per-child
// counts), which keeps the kernel's vector registers low for occupancy.
template <int NC>
__global__ __launch_bounds__(TPB) void k_level_count(int level, const FrameConsts* __restrict__ fc, TileGeom tg,
                                                     LevelDesc L, float ls, const float4* __restrict__ pxyz,
                                                     const uint32_t* __restrict__ parent_off,
                                                     const uint32_t* __restrict__ parent_len,
                                                     const uint32_t* __restrict__ parent_list,
                                                     uint32_t* __restrict__ item_base,
                                                     uint32_t* __restrict__ counts, uint8_t* __restrict__ masks,
                                                     long long mask_stride, uint4* __restrict__ ictx,
                                                     const ChildBox* __restrict__ ctab, DevScalars* sc) {
    const int m = fc->m;
    if (level == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        // the single parent's items, for the slot scan, the fill and the boxes
        const uint32_t n0 = level0_items(L, m);
        item_base[0] = 0;
        item_base[1] = n0;
        sc->level_items[0] = n0;
    }
    if (sc->overflow) return;
    const int ncg = child_groups(L);
    const uint32_t task = uniform(lv_task()), item = task / (uint32_t)ncg, grp = task % (uint32_t)ncg;
    if (item >= (level ? item_base[L.nparents] : level0_items(L, m))) return;
    const int lane = threadIdx.x & 63;
    const int c0 = (int)grp * LV_CG, nc = min(NC, L.nch - c0);
    const ItemCtx x = item_ctx(level, L, item, item_base, parent_off, parent_len, m);
    // the first groups' point indices and positions are in flight during the
    // child set-up; afterwards each group's positions load while the one
    // before is tested
    uint32_t i_next = cand_index(level, parent_list, x, x.begin + 64u + lane);
    float4 p_next = pxyz[cand_index(level, parent_list, x, x.begin + lane)];
    ChildBox cb[NC];
    if (level == 0) {
        // one parent: lane c sets up child c, every lane reads it from lane c
        float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        bool real = false;
        if (lane < nc) {
            int t0[3], t1[3];
            const int box = child_box(level, L, tg, x.P, c0 + lane, t0, t1);
            if (box >= 0) {
                b = box_sphere(fc->g, ls, t0[0], t0[1], t0[2], t1[0], t1[1], t1[2]);
                real = true;
            }
        }
        const ChildBox k = make_child(b, real);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            cb[c].bx2 = lane_float(k.bx2, c);
            cb[c].by2 = lane_float(k.by2, c);
            cb[c].bz2 = lane_float(k.bz2, c);
            cb[c].bb = lane_float(k.bb, c);
            cb[c].w = lane_float(k.w, c);
            cb[c].wsq = lane_float(k.wsq, c);
        }
    } else {
        // the previous level's boxes kernel wrote the parent's children in order
        const ChildBox* tp = ctab + (size_t)x.P * (size_t)L.nch + (size_t)c0;
#pragma unroll
        for (int c = 0; c < NC; ++c) cb[c] = tp[c];
    }
    float wmax = cb[0].w, wsqmax = cb[0].wsq;
#pragma unroll
    for (int c = 1; c < NC; ++c) {
        wmax = fmaxf(wmax, cb[c].w);
        wsqmax = fmaxf(wsqmax, cb[c].wsq);
    }
    uint32_t cnt[NC];  // candidates passing child c (wavefront-uniform)
#pragma unroll
    for (int c = 0; c < NC; ++c) cnt[c] = 0;
    uint8_t* gmask = masks + (size_t)grp * (size_t)mask_stride + (level ? x.poff : 0u);
    for (uint32_t base = x.begin; base < x.end; base += 64u) {
        const float4 p = p_next;
        const bool in = base + lane < x.end;
        p_next = pxyz[i_next];
        i_next = cand_index(level, parent_list, x, base + 128u + lane);
        const Cand cd = make_cand(p, in);
        const float thr2 = cand_thr2(cd, wmax, wsqmax);
        uint32_t hits = 0;  // bit c: passes child c (kept for the fill)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const bool hit = cand_hits(cb[c], cd, thr2);
            hits |= hit ? (1u << c) : 0u;
            cnt[c] += (uint32_t)__popcll(__ballot(hit));
        }
        if (in) gmask[base + lane] = (uint8_t)hits;
    }
    uint32_t mine = 0;  // lane c: the count of child c
#pragma unroll
    for (int c = 0; c < NC; ++c) mine = lane == c ? cnt[c] : mine;
    if (lane < nc) counts[x.ib * L.nch + (c0 + lane) * x.nsegP + x.seg] = mine;
    // the item's context for the fill (slot of child 0, segments of the
    // parent, first parent-list position, candidates)
    if (grp == 0 && lane == 0)
        ictx[item] = make_uint4(x.ib * L.nch + x.seg, x.nsegP, (level ? x.poff : 0u) + x.begin, x.end - x.begin);
}

// Ordered fill: each candidate passing child c goes to that child's list at
// the scanned offset of (item, c) plus the number of earlier passing
// candidates of the item.  The task's wavefront walks the segment in order,
// 64 candidates at a time; a candidate's position is the running offset of
// child c (wavefront-uniform) plus its rank in the group's ballot.  The
// item's bounds come from the context the count kernel recorded, loaded
// together with the item count and the overflow flag, so the candidates'
// loads start after two round trips to memory.
// LISTED: the candidates come from a parent list (levels 1, 2), else they
// are the points themselves (level 0).
template <int NC, bool LISTED>
__global__ __launch_bounds__(TPB) void k_level_fill(int level, LevelDesc L, const uint32_t* __restrict__ parent_list,
                                                    const uint32_t* __restrict__ item_base,
                                                    const uint4* __restrict__ ictx, long long items_cap,
                                                    const uint32_t* __restrict__ slot_off,
                                                    const uint8_t* __restrict__ masks, long long mask_stride,
                                                    uint32_t* __restrict__ list, long long list_cap,
                                                    DevScalars* sc) {
    const int ncg = child_groups(L);
    const uint32_t task = lv_task(), item = task / (uint32_t)ncg, grp = task % (uint32_t)ncg;
    const uint32_t overflow = sc->overflow;
    const uint32_t n_items = item_base[L.nparents];
    const uint4 cx = (long long)item < items_cap ? ictx[item] : make_uint4(0u, 0u, 0u, 0u);
    if (overflow) return;
    const uint32_t total = slot_off[n_items * (uint32_t)L.nch];
    const bool live = item < n_items;
    const int lane = threadIdx.x & 63;
    const int c0 = (int)grp * LV_CG, nc = min(NC, L.nch - c0);
    // cx: slot of (item, child 0), segments of the parent, first parent-list
    // position, candidates
    // (a task past the last item reads position 0, in bounds, and stores nothing)
    const uint32_t s0 = cx.x, nsegP = cx.y, abs0 = live ? cx.z : 0u, len = live ? cx.w : 0u;
    // the children each candidate passes, as the count kernel found them; the
    // next group's indices and masks load while this one is placed
    const uint8_t* gmask = masks + (size_t)grp * (size_t)mask_stride + abs0;
    const uint32_t* gidx = parent_list + abs0;
    // candidate i of the item, i clamped to the last (a live item holds one
    // at least); unconditional loads stay in flight until their use
    const uint32_t last = len ? len - 1u : 0u;
    uint32_t idx_next = LISTED ? gidx[min((uint32_t)lane, last)] : abs0 + (uint32_t)lane;
    uint32_t hits_next = gmask[min((uint32_t)lane, last)];
    const uint32_t at0 = live && lane < nc ? slot_off[s0 + (uint32_t)(c0 + lane) * nsegP] : 0u;
    if ((long long)total > list_cap) {
        if (task == 0 && lane == 0) {
            sc->level_total[level] = total;
            atomicOr(&sc->overflow_mask, 1u << level);
            atomicExch(&sc->overflow, 1u);
        }
        return;
    }
    if (task == 0 && lane == 0) sc->level_total[level] = total;
    if (!live) return;
    uint32_t at[NC];  // where child c's next candidate goes (wavefront-uniform)
#pragma unroll
    for (int c = 0; c < NC; ++c) at[c] = lane_word(at0, (uint32_t)c);
    for (uint32_t i = 0; i < len; i += 64u) {
        const uint32_t idx = idx_next, hits = i + lane < len ? hits_next : 0u;
        const uint32_t en = min(i + 64u + lane, last);
        idx_next = LISTED ? gidx[en] : abs0 + en;
        hits_next = gmask[en];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const bool hit = (hits >> c) & 1u;
            const uint64_t bal = __ballot(hit);
            if (hit) list[at[c] + bg_lanes_below(bal)] = idx;
            at[c] += (uint32_t)__popcll(bal);
        }
    }
}

// Each box's list (offset and length in the level's list).  A box of level
// l is a parent of level l + 1, so the kernel also writes that level's work
// items per parent (next_nseg: ceil(len / next_seg)); at the tile level it
// writes each tile's gather items instead: one for a list of at most
// BG_HEAVY_LEN candidates, else one per point chunk its list spans.
// Below the tile level, thread (box, c) also writes the test constants of
// child c of the box at the next level (next_tab, parent-major) for that
// level's count kernel; the box's own work is done by its thread c = 0.
__global__ __launch_bounds__(TPB) void k_level_boxes(int level, LevelDesc L, const uint32_t* __restrict__ item_base,
                                                     const uint32_t* __restrict__ slot_off,
                                                     const uint32_t* __restrict__ list,
                                                     uint32_t* __restrict__ box_off, uint32_t* __restrict__ box_len,
                                                     uint32_t* __restrict__ next_nseg, int next_seg,
                                                     uint32_t* __restrict__ heavy_cnt, uint32_t* __restrict__ light_cnt,
                                                     const FrameConsts* __restrict__ fc, TileGeom tg, LevelDesc Ln,
                                                     float ls, ChildBox* __restrict__ next_tab,
                                                     const DevScalars* sc) {
    const int nchn = next_tab ? Ln.nch : 1;
    const long long gid = (long long)blockIdx.x * TPB + threadIdx.x;
    const int box = (int)(gid / nchn), cn = (int)(gid % nchn);
    if (box >= L.nbx * L.nby * L.nbz) return;
    if (next_tab) {
        int t0[3], t1[3];
        const int cb = child_box(level + 1, Ln, tg, box, cn, t0, t1);
        const float4 b = cb >= 0 ? box_sphere(fc->g, ls, t0[0], t0[1], t0[2], t1[0], t1[1], t1[2])
                                 : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        next_tab[gid] = make_child(b, cb >= 0);
        if (cn != 0) return;
    }
    uint32_t off = 0, len = 0;
    if (!sc->overflow) {
        int P, c;
        if (level == 0) {
            P = 0;
            c = box;
        } else {
            int bxi = box % L.nbx, byi = (box / L.nbx) % L.nby, bzi = box / (L.nbx * L.nby);
            P = (bxi / L.ncx) + L.pnbx * ((byi / L.ncy) + L.pnby * (bzi / L.ncz));
            c = (bxi % L.ncx) + L.ncx * ((byi % L.ncy) + L.ncy * (bzi % L.ncz));
        }
        const uint32_t ib = item_base[P], ns = item_base[P + 1] - ib;
        if (ns > 0) {
            uint32_t s0 = ib * L.nch + c * ns;
            off = slot_off[s0];
            len = slot_off[s0 + ns] - off;
        }
    }
    box_off[box] = off;
    box_len[box] = len;
    if (next_nseg) next_nseg[box] = (len + (uint32_t)next_seg - 1u) / (uint32_t)next_seg;
    if (heavy_cnt) {
        if (len > BG_HEAVY_LEN) {
            heavy_cnt[box] = (list[off + len - 1u] >> K3_BGKI_CHUNK_LOG2) - (list[off] >> K3_BGKI_CHUNK_LOG2) + 1u;
            light_cnt[box] = 0;
        } else {
            heavy_cnt[box] = 0;
            light_cnt[box] = 1;
        }
    }
}

hipError_t launch_build_levels(hipStream_t st, const FrameConsts* fc, const TileGeom& tg, const LevelDesc* L,
                               LevelBufs* B, const ItemBufs& I, float ls, const float4* pxyz, DevScalars* sc) {
    for (int level = 0; level < 3; ++level) {
        const LevelDesc& Ll = L[level];
        const LevelBufs& Bl = B[level];
        const LevelBufs* parent = level ? &B[level - 1] : nullptr;
        const uint32_t* poff = parent ? parent->box_off : nullptr;
        const uint32_t* plen = parent ? parent->box_len : nullptr;
        const uint32_t* plist = parent ? parent->list : nullptr;
        // level 0's single parent: its items come from the point count (k_level_count);
        // the other levels' per-parent items were written by the previous level's boxes
        if (level > 0)
            (void)launch_scan_u32(st, Bl.nseg, Bl.item_base, Bl.bsum, Ll.nparents, nullptr, 1, &sc->level_items[level]);
        const unsigned grid = level_task_blocks(Bl.items_cap, Ll);
        hipLaunchKernelGGL(k_level_count<LV_CG>, dim3(grid), dim3(TPB), 0, st, level, fc, tg, Ll, ls, pxyz, poff, plen,
                           plist, Bl.item_base, Bl.counts, Bl.mask, Bl.mask_stride, Bl.ictx,
                           reinterpret_cast<const ChildBox*>(Bl.ctab), sc);
        (void)launch_scan_u32(st, Bl.counts, Bl.slot_off, Bl.bsum, Bl.slot_cap, Bl.item_base + Ll.nparents,
                              (uint32_t)Ll.nch, nullptr);
        auto fill = level ? k_level_fill<LV_CG, true> : k_level_fill<LV_CG, false>;
        hipLaunchKernelGGL(fill, dim3(grid), dim3(TPB), 0, st, level, Ll, plist, Bl.item_base, Bl.ictx, Bl.items_cap,
                           Bl.slot_off, Bl.mask, Bl.mask_stride, Bl.list, Bl.list_cap, sc);
        const bool last = level == 2;
        const LevelDesc& Ln = L[last ? level : level + 1];
        const long long nthreads = (long long)Ll.nbx * Ll.nby * Ll.nbz * (last ? 1 : Ln.nch);
        hipLaunchKernelGGL(k_level_boxes, dim3(blocks_for(nthreads)), dim3(TPB), 0, st, level, Ll, Bl.item_base,
                           Bl.slot_off, Bl.list, Bl.box_off, Bl.box_len, last ? nullptr : B[level + 1].nseg,
                           last ? 1 : Ln.seg, last ? I.heavy_cnt : nullptr, last ? I.light_cnt : nullptr, fc, tg, Ln,
                           ls, last ? nullptr : reinterpret_cast<ChildBox*>(B[level + 1].ctab), sc);
    }
    return hipGetLastError();
}

// ============================================================== work items
// Number of leading elements of the sorted a[lo, hi) that satisfy below(x)
// (a prefix of the range), found with 8-way rounds: 7 independent probes per
// round, so the dependent loads are log8 of the length instead of log2.
template <class Below>
__device__ __forceinline__ uint32_t partition_point_u32(const uint32_t* a, uint32_t lo, uint32_t hi, Below below) {
    while (hi - lo > 8) {
        const uint32_t step = (hi - lo) / 8;
        uint32_t x[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) x[j] = a[lo + (j + 1) * step];
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 7; ++j) m += below(x[j]) ? 1u : 0u;
        // the answer is past probe m (if any) and at most probe m + 1 (if any)
        const uint32_t nlo = m ? lo + m * step + 1 : lo;
        const uint32_t nhi = m < 7 ? lo + (m + 1) * step : hi;
        lo = nlo;
        hi = nhi;
    }
    uint32_t x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = a[min(lo + j, hi - 1)];  // (hi > lo here unless the range is empty)
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) m += (lo + j < hi && below(x[j])) ? 1u : 0u;
    return lo + m;
}
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t v) {
    if (lo >= hi) return lo;
    return partition_point_u32(a, lo, hi, [v](uint32_t x) { return x < v; });
}

// heavy_off / light_off: one exclusive scan over the tiles' split-item
// counts followed by their whole-tile counts, so light_off[t] already counts
// the ntiles' split items before it (they come first in items[]).
// k_emit_light also publishes both totals and zeroes the gather's queue
// heads (so a gather launched right after needs no reset launch).
__global__ __launch_bounds__(TPB) void k_emit_light(int ntiles, const uint32_t* __restrict__ toff,
                                                    const uint32_t* __restrict__ tlen,
                                                    const uint32_t* __restrict__ light_cnt,
                                                    const uint32_t* __restrict__ heavy_off,
                                                    const uint32_t* __restrict__ light_off, uint4* __restrict__ items,
                                                    uint32_t* __restrict__ tile_partial, long long heavy_cap,
                                                    uint32_t* __restrict__ qheads, DevScalars* sc) {
    const int t = blockIdx.x * TPB + threadIdx.x;
    const uint32_t H = heavy_off[ntiles];
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            sc->n_heavy = H;
            sc->n_light = light_off[ntiles] - H;
        }
        if (threadIdx.x < BG_NQ) qheads[threadIdx.x * BG_QSTRIDE] = 0u;
    }
    // items[] holds ntiles + heavy_cap entries: with more split items than
    // that, k_emit_heavy flags the overflow and nothing may be written here
    if (t >= ntiles || sc->overflow || (long long)H > heavy_cap) return;
    if (light_cnt[t]) {
        uint32_t o = toff[t];
        items[light_off[t]] = make_uint4((uint32_t)t, o, o + tlen[t], BG_DIRECT);
        tile_partial[2 * t] = 0;
        tile_partial[2 * t + 1] = 0;
    }
}

// Split item h = (tile t, its k-th point chunk c): candidates [b, e) of the
// tile's list, b the first entry of chunk c or later.  Its end e is the next
// item's b (or the list's end for the tile's last chunk), so each item
// searches once and writes its predecessor's end; k = 0 starts the list.
__global__ __launch_bounds__(TPB) void k_emit_heavy(int ntiles, const uint32_t* __restrict__ toff,
                                                    const uint32_t* __restrict__ tlen,
                                                    const uint32_t* __restrict__ tlist,
                                                    const uint32_t* __restrict__ heavy_cnt,
                                                    const uint32_t* __restrict__ heavy_off, uint4* __restrict__ items,
                                                    uint32_t* __restrict__ tile_partial, long long heavy_cap,
                                                    DevScalars* sc) {
    if (sc->overflow) return;
    const uint32_t H = heavy_off[ntiles];
    if ((long long)H > heavy_cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicOr(&sc->overflow_mask, 1u << 4);
            atomicExch(&sc->overflow, 1u);
        }
        return;
    }
    const uint32_t h = blockIdx.x * TPB + threadIdx.x;
    if (h >= H) return;
    // the last tile whose offset is <= h (heavy_off[0] = 0 <= h)
    const int t = (int)partition_point_u32(heavy_off, 0u, (uint32_t)ntiles, [h](uint32_t x) { return x <= h; }) - 1;
    const uint32_t k = h - heavy_off[t], cnt = heavy_cnt[t];
    const uint32_t o = toff[t], len = tlen[t];
    const uint32_t c = (tlist[o] >> K3_BGKI_CHUNK_LOG2) + k;
    const uint32_t b = k ? lower_bound_u32(tlist, o, o + len, c << K3_BGKI_CHUNK_LOG2) : o;
    uint32_t* it = reinterpret_cast<uint32_t*>(items + h);
    *reinterpret_cast<uint2*>(it) = make_uint2((uint32_t)t, b);
    it[3] = h;
    if (k) it[-2] = b;  // the previous item's end (items[h - 1].z)
    if (k + 1 == cnt) it[2] = o + len;
    if (k == 0) {
        tile_partial[2 * t] = h;
        tile_partial[2 * t + 1] = cnt;
    }
}

hipError_t launch_build_items(hipStream_t st, int ntiles, const LevelBufs& T, const ItemBufs& I, DevScalars* sc) {
    // (the tiles' item counts come from the last level's boxes; light_cnt
    // follows heavy_cnt and light_off heavy_off, so one scan covers both)
    (void)launch_scan_u32(st, I.heavy_cnt, I.heavy_off, I.bsum, 2ll * ntiles, nullptr, 1, nullptr);
    hipLaunchKernelGGL(k_emit_light, dim3(blocks_for(ntiles)), dim3(TPB), 0, st, ntiles, T.box_off, T.box_len,
                       I.light_cnt, I.heavy_off, I.light_off, I.items, I.tile_partial, I.heavy_cap, I.qheads, sc);
    hipLaunchKernelGGL(k_emit_heavy, dim3(blocks_for(std::max(1ll, I.heavy_cap))), dim3(TPB), 0, st, ntiles,
                       T.box_off, T.box_len, T.list, I.heavy_cnt, I.heavy_off, I.items, I.tile_partial, I.heavy_cap,
                       sc);
    return hipGetLastError();
}

// ============================================================== gather
// Per-candidate classes, decided once per (tile, point) by the staging lane
// from the tile's bounding box with a safety margin, so that inside a class
// the skipped branches are provably dead for every lane of the tile:
//   C_GENERAL  the full three-case free distance,
//   C_MAIN     every voxel projects strictly inside (0, mg): perpendicular case,
//   C_BEHIND   every voxel projects to <= 0: free distance = |voxel|,
//   C_NOFREE   every voxel projects beyond the point: no free term;
// plus F_OCC when a voxel of the tile may lie within ls of the point.
constexpr uint32_t C_GENERAL = 0, C_MAIN = 1, C_BEHIND = 2, C_NOFREE = 3, F_OCC = 4;

// Kernel term of squared distance a2 (the table of k3_kernel_table.h, staged
// in LDS as float4 entries) and, when GRAD, its derivative in the length scale.
template <bool GRAD>
struct BgTerm {
    float v, g;
};
// The table in LDS as three arrays (a, b, c coefficients): three 4-byte
// reads per lookup, lanes in different banks unless their indices differ by a
// multiple of the bank count.
// With GRAD the derivative-factor table (k3_ktab_dls) rides along the same way.
struct BgTableV {
    float2 ab[K3_KTAB_N + 1];
    float c[K3_KTAB_N + 1];
};
template <bool GRAD>
struct BgTable : BgTableV {};
template <>
struct BgTable<true> : BgTableV {
    float2 gab[K3_KTAB_N + 1];
    float gc[K3_KTAB_N + 1];
};
// The kernel bracket (sigma = 1) of table argument s (in table units, see
// K3PointRec); the sums are multiplied by the sigma factor once per voxel.
// s may be slightly negative (a rounded difference of squares standing for
// 0) or beyond N: it is clamped with one med3, equal to min(max(s, 0), N) of
// k3_kernel_index_s for every non-NaN s, and the fraction is taken with
// fract, which equals c - (float)(int)c exactly for 0 <= c < 2^24, so the
// result is that of k3_kernel_bracket_s bit for bit.  The length-scale
// derivative (GRAD) is G(u) from the second table at the same index and
// fraction where the bracket is positive (k3_kernel_dls_s without its
// sigma / ls factor).
template <bool GRAD>
__device__ __forceinline__ BgTerm<GRAD> bg_term(float s_raw, const BgTable<GRAD>* stab) {
    BgTerm<GRAD> t;
    const float s = __builtin_amdgcn_fmed3f(s_raw, 0.0f, (float)K3_KTAB_N);
    const int i = (int)s;
    const float f = __builtin_amdgcn_fractf(s);
    // both entries addressed from 4 i (byte offsets into the two arrays)
    const uint32_t i4 = (uint32_t)i << 2;
    const float2 ab = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(stab->ab) + 2u * i4);
    const float cc = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(stab->c) + i4);
    // max(0, poly) as a clamp to [0, 1], which folds into the last FMA's
    // clamp bit (one instruction fewer per term): equal, since no entry's
    // quadratic exceeds 1 on [0, 1] (the largest value is the bracket's k(0) = 1,
    // entry 0 at f = 0; pinned by tests/test_math.py)
    t.v = __builtin_amdgcn_fmed3f(k3_kernel_poly_raw(ab.x, ab.y, cc, f), 0.0f, 1.0f);
    t.g = 0.0f;
    if constexpr (GRAD) {
        const float2 g = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(stab->gab) + 2u * i4);
        const float gc = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(stab->gc) + i4);
        t.g = t.v > 0.0f ? k3_kernel_poly_raw(g.x, g.y, gc, f) : 0.0f;
    }
    return t;
}

// Copies the kernel table(s) into LDS (every thread of the block takes part).
template <bool GRAD>
__device__ __forceinline__ void bg_load_table(BgTable<GRAD>* stab) {
    for (int i = threadIdx.x; i <= K3_KTAB_N; i += blockDim.x) {
        stab->ab[i] = make_float2(k3_ktab[4 * i], k3_ktab[4 * i + 1]);
        stab->c[i] = k3_ktab[4 * i + 2];
        if constexpr (GRAD) {
            stab->gab[i] = make_float2(k3_ktab_dls[4 * i], k3_ktab_dls[4 * i + 1]);
            stab->gc[i] = k3_ktab_dls[4 * i + 2];
        }
    }
    __syncthreads();
}

// Partial sums of a split item: occupied, free (and, when GRAD, their
// length-scale derivatives), 64 lanes each.
template <bool GRAD>
constexpr int bg_partial_stride() { return GRAD ? 256 : 128; }

// The first 16 bytes of a staged record (direction u and mgs) are read as
// one ds_read_b128: the empty asm statements of bg_keep_w keep the fourth
// word live, since a 12-byte read of the three words used measured slower.
// They stand after the values' first use, so that every read of a step is in
// flight before the one wait for them (an asm statement right after a read
// makes the compiler wait for that read there).
__device__ __forceinline__ void bg_keep_w(float w) { asm volatile("" ::"v"(w)); }

// Gather of one work item (a whole tile, or one 1024-point chunk of a long
// tile list) by one wavefront: lane = voxel of the 4x4x4 tile.  Distances
// are in table units (K3PointRec): per pair, the table argument is s =
// |v - m|^2 kscale (occupied) or a free distance of k3_free_s.
template <bool GRAD>
__device__ __forceinline__ void bgki_item(const BgArgs& a, uint32_t it, int lane, float4* srec,
                                          uint32_t* smeta, const BgTable<GRAD>* stab) {
    const GridGeom g = a.fc->g;
    const uint4 item = a.items[it];
    const uint32_t tile = uniform(item.x), ib = uniform(item.y), ie = uniform(item.z), slot = uniform(item.w);
    const int tx = (int)(tile % (uint32_t)a.tg.ntx), ty = (int)((tile / (uint32_t)a.tg.ntx) % (uint32_t)a.tg.nty);
    const int tz = (int)(tile / (uint32_t)(a.tg.ntx * a.tg.nty));
    const int i = tx * 4 + (lane & 3), j = ty * 4 + ((lane >> 2) & 3), k = tz * 4 + (lane >> 4);
    const float vx = axis_center(i, g.gs, g.res, g.sx);
    const float vy = axis_center(j, g.gs, g.res, g.sy);
    const float vz = axis_center(k, g.gsz, g.res, g.sz);
    const float dc2s = fmaf(vz, vz, fmaf(vy, vy, vx * vx)) * a.kscale;
    // free term of a voxel for every point in front of which it lies
    // (the kernel is an exact zero from distance ls on)
    const BgTerm<GRAD> kc = bg_term<GRAD>(dc2s, stab);
    // bounding box of the tile's 64 voxel centres (lanes outside the grid
    // included); uniform values are kept in scalar registers
    const float tcx =
        uniform_f(0.5f * (axis_center(tx * 4, g.gs, g.res, g.sx) + axis_center(tx * 4 + 3, g.gs, g.res, g.sx)));
    const float tcy =
        uniform_f(0.5f * (axis_center(ty * 4, g.gs, g.res, g.sy) + axis_center(ty * 4 + 3, g.gs, g.res, g.sy)));
    const float tcz =
        uniform_f(0.5f * (axis_center(tz * 4, g.gsz, g.res, g.sz) + axis_center(tz * 4 + 3, g.gsz, g.res, g.sz)));
    const float trad = uniform_f(1.5f * g.res * 1.7320508f);
    const float marg = 1e-3f + 1e-5f * (fabsf(tcx) + fabsf(tcy) + fabsf(tcz) + trad);
    const float rr = uniform_f(trad + marg);
    const float occ_r = uniform_f(a.ls + rr);
    // the same radius along the ray in table units: a lane's projection v . u
    // lies within rrs of the tile centre's (|u| = sk up to rounding, which
    // the margin covers)
    const float rrs = uniform_f(rr * a.sk * (1.0f + 1e-6f));
    // A candidate whose every term is exactly zero for the whole tile is
    // dropped before the loop.  Free and occupied distances of a voxel are at
    // least its distance to the ray's line, and over the tile's box of voxel
    // centres (half-width hb) that distance is at least
    //   B = |c_perp| - hb (|n_x| + |n_y| + |n_z|),  n = c_perp / |c_perp|,
    // c_perp the tile centre's offset from the line.  The lanes' rounded
    // table arguments are within 16 2^-24 |v|^2 kscale of the exact ones, so
    // B^2 >= ls^2 (1 + 1e-5) + lerr with lerr = 1e-5 (|c| + trad)^2 + 1e-6
    // leaves every lane at or beyond the kernel's support.  A behind-sensor
    // candidate without an occupied term is dropped when no lane of the tile
    // has a free term (kc is zero in every lane), and one lying wholly beyond
    // its point likewise.
    const float hb = uniform_f(1.5f * g.res * (1.0f + 1e-6f));
    // (these bounds carry relative margins of 1e-5, far above the 2^-22 of the
    // hardware square root and reciprocal used for them)
    const float tc_n = __builtin_amdgcn_sqrtf(fmaf(tcz, tcz, fmaf(tcy, tcy, tcx * tcx))) + trad;
    const float lim2 = uniform_f(fmaf(a.ls * a.ls, 1.0f + 1e-5f, fmaf(1e-5f * tc_n, tc_n, 1e-6f)));
    const bool kc_live = __any(kc.v != 0.0f);
    float tot_o = 0.0f, tot_f = 0.0f, cur_o = 0.0f, cur_f = 0.0f;
    float gsum_o = 0.0f, gsum_f = 0.0f;  // length-scale derivatives (GRAD)
    uint32_t last_chunk = 0xffffffffu;  // chunk of the last candidate kept so far
    for (uint32_t b = ib; b < ie; b += 64) {
        const uint32_t cnt = uniform(min(64u, ie - b));
        bool keep = false;
        uint32_t meta = 0;
        float4 q0, q1, q2;
        if ((uint32_t)lane < cnt) {
            const uint32_t p = a.tlist[b + lane];
            const float4* src = reinterpret_cast<const float4*>(a.rec + p);
            // record: q0 = (ux, uy, uz, mgs), q1 = (x, y, z, dms), q2 = (ex, ey, ez, valid)
            q0 = src[0];
            q1 = src[1];
            q2 = src[2];
            const float proj = fmaf(tcz, q0.z, fmaf(tcy, q0.y, tcx * q0.x));
            uint32_t cls = C_GENERAL;
            if (proj + rrs <= 0.0f) cls = C_BEHIND;
            else if (proj - rrs > 0.0f && proj + rrs < q0.w) cls = C_MAIN;
            else if (proj - rrs >= q1.w) cls = C_NOFREE;
            const float ox = tcx - q1.x, oy = tcy - q1.y, oz = tcz - q1.z;
            const bool occ = !(ox * ox + oy * oy + oz * oz >= occ_r * occ_r);
            if (occ) cls |= F_OCC;
            // the tile centre's offset from the ray's line (unit direction u / sk)
            const float pl = proj * a.inv_sk2;
            const float px = fmaf(-pl, q0.x, tcx);
            const float py = fmaf(-pl, q0.y, tcy);
            const float pz = fmaf(-pl, q0.z, tcz);
            const float dl = __builtin_amdgcn_sqrtf(fmaf(pz, pz, fmaf(py, py, px * px)));
            const float bnd = dl - hb * (fabsf(px) + fabsf(py) + fabsf(pz)) * __builtin_amdgcn_rcpf(dl);
            const bool far = dl > 0.0f && bnd > 0.0f && bnd * bnd >= lim2;
            const bool free_dead = cls == C_NOFREE || (cls == C_BEHIND && !kc_live);
            keep = !far && (occ || !free_dead);
            meta = (p << 3) | cls;
        }
        // kept candidates, in list order, to the front of the staging area
        const uint64_t km = __ballot(keep);
        const uint32_t nk = (uint32_t)__popcll(km);
        if (keep) {
            const uint32_t at = bg_lanes_below(km);
            srec[at * 3 + 0] = q0;
            srec[at * 3 + 1] = q1;
            srec[at * 3 + 2] = q2;
            smeta[at] = meta;
        }
        bg_wave_sync();
        // Per-candidate bits as wave masks (bit q = kept candidate q), read on
        // the scalar unit: a chunk starts at q, q is a pure perpendicular-free
        // candidate, q has an occupied term, and the two free-class bits.
        const bool live = (uint32_t)lane < nk;
        const uint32_t mt = live ? smeta[lane] : 0u;
        const uint32_t cls = mt & 7u;
        const uint32_t chunk = mt >> (3 + K3_BGKI_CHUNK_LOG2);
        uint32_t prev = (uint32_t)__shfl_up((int)chunk, 1);
        if (lane == 0) prev = last_chunk;
        const uint64_t m_start = __ballot(live && chunk != prev);
        const uint64_t m_main = __ballot(live && cls == C_MAIN);
        const uint64_t m_occ = __ballot(live && (cls & F_OCC) != 0u);
        const uint64_t m_b0 = __ballot(live && (cls & 1u) != 0u);
        const uint64_t m_b1 = __ballot(live && (cls & 2u) != 0u);
        if (nk) last_chunk = lane_word(chunk, nk - 1u);
        const uint32_t cnt_k = uniform(nk);
        uint32_t q = 0;
        while (q < cnt_k) {
            if ((m_start >> q) & 1ull) {
                tot_o += cur_o;
                tot_f += cur_f;
                cur_o = 0.0f;
                cur_f = 0.0f;
            }
            if ((m_main >> q) & 1ull) {
                // A run of candidates whose free term is the perpendicular
                // distance for every lane and that have no occupied term, up
                // to the next other candidate or chunk start: two terms are
                // evaluated together (independent table reads), then added
                // in order.  Only the direction u is used.
                const uint64_t stop = q < 63u ? ((~m_main | m_start) >> (q + 1u)) : 0ull;
                const uint32_t r = min(cnt_k, stop ? q + 1u + (uint32_t)__builtin_ctzll(stop) : 64u);
                for (; q + 2u <= r; q += 2u) {
                    const float4 u0 = srec[3u * q], u1 = srec[3u * q + 3u];
                    const float d0 = fmaf(vz, u0.z, fmaf(vy, u0.y, vx * u0.x));
                    const float d1 = fmaf(vz, u1.z, fmaf(vy, u1.y, vx * u1.x));
                    bg_keep_w(u0.w);
                    bg_keep_w(u1.w);
                    // squared distances clamped at 0 by bg_term
                    const BgTerm<GRAD> t0 = bg_term<GRAD>(fmaf(-d0, d0, dc2s), stab);
                    const BgTerm<GRAD> t1 = bg_term<GRAD>(fmaf(-d1, d1, dc2s), stab);
                    cur_f += t0.v;
                    cur_f += t1.v;
                    if (GRAD) {
                        gsum_f += t0.g;
                        gsum_f += t1.g;
                    }
                }
                if (q < r) {
                    const float4 u0 = srec[3u * q];
                    const float d0 = fmaf(vz, u0.z, fmaf(vy, u0.y, vx * u0.x));
                    bg_keep_w(u0.w);
                    const BgTerm<GRAD> t0 = bg_term<GRAD>(fmaf(-d0, d0, dc2s), stab);
                    cur_f += t0.v;
                    if (GRAD) gsum_f += t0.g;
                    ++q;
                }
                continue;
            }
            // record: r0 = (ux, uy, uz, mgs), r1 = (x, y, z, dms), r2 = (ex, ey, ez, valid)
            const float4 r0 = srec[q * 3];
            const float4 r1 = srec[q * 3 + 1];
            if ((m_occ >> q) & 1ull) {  // occupied term (k3_occ_s)
                const float ox = vx - r1.x, oy = vy - r1.y, oz = vz - r1.z;
                const float so = fmaf(oz, oz, fmaf(oy, oy, ox * ox)) * a.kscale;
                const bool co = so < (float)K3_KTAB_N;
                if (__any(co)) {
                    const BgTerm<GRAD> t = bg_term<GRAD>(so, stab);
                    cur_o += t.v;
                    if (GRAD) gsum_o += t.g;
                }
            }
            const uint32_t fc = (uint32_t)((m_b0 >> q) & 1ull) | ((uint32_t)((m_b1 >> q) & 1ull) << 1);
            if (fc == C_MAIN) {  // 0 < dp < mgs for every lane
                const float dp = fmaf(vz, r0.z, fmaf(vy, r0.y, vx * r0.x));
                const float sf = fmaf(-dp, dp, dc2s);  // clamped at 0 by bg_term
                const bool cf = sf < (float)K3_KTAB_N;
                if (__any(cf)) {
                    const BgTerm<GRAD> t = bg_term<GRAD>(sf, stab);
                    cur_f += t.v;
                    if (GRAD) gsum_f += t.g;
                }
            } else if (fc == C_BEHIND) {  // dp <= 0 for every lane
                cur_f += kc.v;
                if (GRAD) gsum_f += kc.g;
            } else if (fc == C_GENERAL) {  // k3_free_s, the three cases as selects
                const float4 r2 = srec[q * 3 + 2];
                const float dp = fmaf(vz, r0.z, fmaf(vy, r0.y, vx * r0.x));
                const bool c1 = dp <= 0.0f;
                const bool c2 = !c1 && dp < r0.w;
                const bool c3 = !c1 && !c2 && dp < r1.w;
                const float sp = fmaf(-dp, dp, dc2s);  // clamped at 0 by bg_term
                const float ex = vx - r2.x, ey = vy - r2.y, ez = vz - r2.z;
                const float se = fmaf(ez, ez, fmaf(ey, ey, ex * ex)) * a.kscale;
                const float sa = c2 ? sp : se;
                const bool cs = (c2 || c3) && sa < (float)K3_KTAB_N;
                if (__any(cs)) {
                    const BgTerm<GRAD> t = bg_term<GRAD>(sa, stab);
                    cur_f += cs ? t.v : (c1 ? kc.v : 0.0f);
                    if (GRAD) gsum_f += cs ? t.g : (c1 ? kc.g : 0.0f);
                } else {
                    cur_f += c1 ? kc.v : 0.0f;
                    if (GRAD) gsum_f += c1 ? kc.g : 0.0f;
                }
            }
            ++q;
        }
        bg_wave_sync();
    }
    tot_o += cur_o;
    tot_f += cur_f;
    if (slot == BG_DIRECT) {
        if (i < g.gs && j < g.gs && k < g.gsz) {
            const int cell = i + g.gs * j + g.gs * g.gs * k;
            const float fo = k3_sigma_factor(a.sigma_o), ff = k3_sigma_factor(a.sigma_f);
            a.del_occ[cell] = fo * tot_o;
            a.del_free[cell] = ff * tot_f;
            if (GRAD) {
                a.dls_occ[cell] = (fo * a.inv_ls) * gsum_o;
                a.dls_free[cell] = (ff * a.inv_ls) * gsum_f;
            }
        }
    } else {
        float* pp = a.partial + (size_t)slot * bg_partial_stride<GRAD>();
        pp[lane] = tot_o;
        pp[64 + lane] = tot_f;
        if (GRAD) {
            pp[128 + lane] = gsum_o;
            pp[192 + lane] = gsum_f;
        }
    }
}

// Work distribution of the gather.  One work counter in device memory that
// every wavefront increments is slow (device-scope atomics from all XCDs on
// one address are serialised), and a static deal leaves wavefronts idle
// (items range from a few to 1024 candidates).  So the items are spread over
// BG_NQ queues, item i being position i / BG_NQ of queue i % BG_NQ, each with
// its own head (one 128-byte line each).  A wavefront claims the items of its
// home queue one at a time with one atomic each; once the home queue is
// empty it moves on to the following queues, reading a queue's head before
// claiming from it so that an emptied queue costs no atomic.  The heads are
// zeroed before every launch: by the gather's launcher (k_reset_queues), or,
// right after the work items are emitted, by k_emit_light.
__device__ __forceinline__ uint32_t bg_queue_len(uint32_t n_items, uint32_t q) {
    return n_items > q ? (n_items - q + BG_NQ - 1u) / BG_NQ : 0u;
}
// Claims a position of queue q (home: claim without looking first); returns
// len (or more) when the queue is empty.
__device__ __forceinline__ uint32_t bg_claim(uint32_t* qheads, uint32_t q, uint32_t len, bool home, int lane) {
    uint32_t pos = len;
    if (lane == 0) {
        uint32_t* head = qheads + q * BG_QSTRIDE;
        if (home || __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < len)
            pos = atomicAdd(head, 1u);
    }
    return uniform(__shfl(pos, 0));
}

__global__ void k_reset_queues(uint32_t* qheads) {
    if (threadIdx.x < BG_NQ) qheads[threadIdx.x * BG_QSTRIDE] = 0u;
}

// The gather is held to 64 VGPRs, so that a SIMD runs 8 of its wavefronts
// (a few per-item values spill; the candidate loops do not).  With the
// length-scale derivative (GRAD) the bound holds in blocks of 8 or 16
// wavefronts, which is how launch_bgki_grad runs it.
template <bool GRAD, int W>
__global__ __launch_bounds__(W * 64, 8) void k_bgki(BgArgs a) {
    __shared__ float4 srec[W][64 * 3];
    __shared__ uint32_t smeta[W][64];
    __shared__ BgTable<GRAD> stab_s;
    BgTable<GRAD>* stab = &stab_s;
    DevScalars* sc = a.sc;
    if (sc->overflow) return;
    bg_load_table<GRAD>(stab);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t n_items = sc->n_light + sc->n_heavy;
    const uint32_t home = (blockIdx.x * W + (uint32_t)w) % BG_NQ;
    for (uint32_t k = 0; k < BG_NQ; ++k) {
        const uint32_t q = (home + k) % BG_NQ, len = bg_queue_len(n_items, q);
        while (true) {
            const uint32_t pos = bg_claim(a.qheads, q, len, k == 0, lane);
            if (pos >= len) break;
            bgki_item<GRAD>(a, q + pos * BG_NQ, lane, srec[w], smeta[w], stab);
        }
    }
}

BgArgs make_bg_args(const FrameConsts* fc, const TileGeom& tg, float ls, float sigma_o, float sigma_f,
                    const ItemBufs& I, const LevelBufs& T, const K3PointRec* rec, float* del_occ, float* del_free,
                    float* dls_occ, float* dls_free, DevScalars* sc) {
    BgArgs a;
    a.fc = fc;
    a.tg = tg;
    a.ls = ls;
    a.kscale = k3_kernel_scale(ls);
    a.sk = k3_kernel_sk(ls);
    a.inv_sk2 = (float)(1.0 / ((double)a.sk * (double)a.sk));
    a.inv_ls = 1.0f / ls;
    a.sigma_o = sigma_o;
    a.sigma_f = sigma_f;
    a.items = I.items;
    a.qheads = I.qheads;
    a.tlist = T.list;
    a.rec = rec;
    a.del_occ = del_occ;
    a.del_free = del_free;
    a.dls_occ = dls_occ;
    a.dls_free = dls_free;
    a.partial = I.partial;
    a.sc = sc;
    return a;
}

template <bool GRAD>
static hipError_t launch_bgki_w(hipStream_t st, int grid_blocks, int waves, const BgArgs& a, bool queues_zeroed) {
    if (!queues_zeroed) hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(BG_NQ), 0, st, a.qheads);
    if (waves == 4) hipLaunchKernelGGL((k_bgki<GRAD, 4>), dim3(grid_blocks), dim3(256), 0, st, a);
    else if (waves == 16) hipLaunchKernelGGL((k_bgki<GRAD, 16>), dim3(grid_blocks), dim3(1024), 0, st, a);
    else hipLaunchKernelGGL((k_bgki<GRAD, 8>), dim3(grid_blocks), dim3(512), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_bgki(hipStream_t st, int grid_blocks, const FrameConsts* fc, const TileGeom& tg, float ls,
                       float sigma_o, float sigma_f, const ItemBufs& I, const LevelBufs& T, const K3PointRec* rec,
                       float* del_occ, float* del_free, DevScalars* sc, int waves, bool queues_zeroed) {
    return launch_bgki_w<false>(st, grid_blocks, waves,
                                make_bg_args(fc, tg, ls, sigma_o, sigma_f, I, T, rec, del_occ, del_free, nullptr,
                                             nullptr, sc),
                                queues_zeroed);
}
hipError_t launch_bgki_grad(hipStream_t st, int grid_blocks, const FrameConsts* fc, const TileGeom& tg, float ls,
                            float sigma_o, float sigma_f, const ItemBufs& I, const LevelBufs& T,
                            const K3PointRec* rec, float* del_occ, float* del_free, float* dls_occ, float* dls_free,
                            DevScalars* sc, int waves) {
    // the same wavefronts per CU in blocks of at least 8: at 4 wavefronts per
    // block the derivative gather keeps 75 VGPRs (6 wavefronts per SIMD), and
    // the four-LiDAR window backward at 512x512x64 took 16.7 ms per frame
    // against 14.4 ms in blocks of 8
    if (waves < 8) {
        grid_blocks = std::max(1, grid_blocks * waves / 8);
        waves = 8;
    }
    return launch_bgki_w<true>(st, grid_blocks, waves,
                               make_bg_args(fc, tg, ls, sigma_o, sigma_f, I, T, rec, del_occ, del_free, dls_occ,
                                            dls_free, sc),
                               false);
}

// ---------------------------------------------------------------- many maps
// Work items of every map of a batch in one joint numbering: prefix[m] is the
// first item of map m (a map whose lists overflowed contributes none).  The
// joint items are spread over the work queues as in the single-map gather;
// their heads (qheads, BG_QWORDS words) are zeroed here.
__global__ void k_bgki_batch_prefix(const BgArgs* __restrict__ maps, int nmaps, uint32_t* __restrict__ prefix,
                                    uint32_t* __restrict__ qheads) {
    if (blockIdx.x != 0) return;
    if (threadIdx.x < BG_NQ) qheads[threadIdx.x * BG_QSTRIDE] = 0u;
    if (threadIdx.x != 0) return;
    uint32_t acc = 0;
    for (int m = 0; m < nmaps; ++m) {
        prefix[m] = acc;
        const DevScalars* sc = maps[m].sc;
        if (!sc->overflow) acc += sc->n_light + sc->n_heavy;
    }
    prefix[nmaps] = acc;
}

// One persistent launch over the items of all maps.
template <int W>
__global__ __launch_bounds__(W * 64) void k_bgki_batch(const BgArgs* __restrict__ maps, int nmaps,
                                                       const uint32_t* __restrict__ prefix, uint32_t* qheads) {
    __shared__ float4 srec[W][64 * 3];
    __shared__ uint32_t smeta[W][64];
    __shared__ BgTable<false> stab_s;
    BgTable<false>* stab = &stab_s;
    bg_load_table<false>(stab);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t total = prefix[nmaps];
    const uint32_t home = (blockIdx.x * W + (uint32_t)w) % BG_NQ;
    for (uint32_t k = 0; k < BG_NQ; ++k) {
        const uint32_t q = (home + k) % BG_NQ, len = bg_queue_len(total, q);
        while (true) {
            const uint32_t pos = bg_claim(qheads, q, len, k == 0, lane);
            if (pos >= len) break;
            const uint32_t it = q + pos * BG_NQ;
            int lo = 0, hi = nmaps - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (prefix[mid] <= it) lo = mid;
                else hi = mid - 1;
            }
            const BgArgs a = maps[lo];
            bgki_item<false>(a, it - prefix[lo], lane, srec[w], smeta[w], stab);
        }
    }
}

hipError_t launch_bgki_batch(hipStream_t st, int grid_blocks, const BgArgs* maps, int nmaps, uint32_t* prefix,
                             uint32_t* qheads, int waves) {
    hipLaunchKernelGGL(k_bgki_batch_prefix, dim3(1), dim3(BG_NQ), 0, st, maps, nmaps, prefix, qheads);
    if (waves == 4)
        hipLaunchKernelGGL(k_bgki_batch<4>, dim3(grid_blocks), dim3(256), 0, st, maps, nmaps, prefix, qheads);
    else if (waves == 16)
        hipLaunchKernelGGL(k_bgki_batch<16>, dim3(grid_blocks), dim3(1024), 0, st, maps, nmaps, prefix, qheads);
    else
        hipLaunchKernelGGL(k_bgki_batch<8>, dim3(grid_blocks), dim3(512), 0, st, maps, nmaps, prefix, qheads);
    return hipGetLastError();
}

// The frame's scalars after the joint gather of LiDAR groups: group 0's
// scalars (sc) take every other group's (gsc[1 .. ng - 1]) overflow flag and
// bits, so that an overflow anywhere stops the frame tail; the largest list
// totals and item counts per level, which size every group's regrowth; and
// the groups' summed items.
__global__ void k_merge_group_scalars(DevScalars* sc, const DevScalars* __restrict__ gsc, int ng) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DevScalars m = *sc;
    for (int g = 1; g < ng; ++g) {
        const DevScalars& o = gsc[g];
        m.overflow |= o.overflow;
        m.overflow_mask |= o.overflow_mask;
        m.n_light += o.n_light;
        m.n_heavy += o.n_heavy;
        for (int l = 0; l < 3; ++l) {
            m.level_total[l] = max(m.level_total[l], o.level_total[l]);
            m.level_items[l] = max(m.level_items[l], o.level_items[l]);
        }
    }
    *sc = m;
}
hipError_t launch_merge_group_scalars(hipStream_t st, DevScalars* sc, const DevScalars* gsc, int ng) {
    hipLaunchKernelGGL(k_merge_group_scalars, dim3(1), dim3(64), 0, st, sc, gsc, ng);
    return hipGetLastError();
}

template <bool GRAD>
__global__ __launch_bounds__(TPB) void k_bgki_combine(int ntiles, GridGeom g, TileGeom tg,
                                                      const uint32_t* __restrict__ tile_partial,
                                                      const float* __restrict__ partial, float* del_occ,
                                                      float* del_free, float* dls_occ, float* dls_free,
                                                      float fo, float ff, float inv_ls, const DevScalars* sc) {
    if (sc->overflow) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tile = blockIdx.x * BG_WAVES + w;
    if (tile >= ntiles) return;
    const uint32_t first = tile_partial[2 * tile], cnt = tile_partial[2 * tile + 1];
    if (cnt == 0) return;
    const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
    const int i = tx * 4 + (lane & 3), j = ty * 4 + ((lane >> 2) & 3), k = tz * 4 + (lane >> 4);
    if (!(i < g.gs && j < g.gs && k < g.gsz)) return;
    float ro = 0.0f, rf = 0.0f, go = 0.0f, gf = 0.0f;
    // chunk sums added in chunk order; loads of up to CB chunks issued together
    constexpr uint32_t CB = 8;
    for (uint32_t c0 = 0; c0 < cnt; c0 += CB) {
        const uint32_t nc = min(CB, cnt - c0);
        float vo[CB], vf[CB], wo[CB], wf[CB];
#pragma unroll
        for (uint32_t c = 0; c < CB; ++c) {
            if (c < nc) {
                const float* pp = partial + (size_t)(first + c0 + c) * bg_partial_stride<GRAD>();
                vo[c] = pp[lane];
                vf[c] = pp[64 + lane];
                if (GRAD) {
                    wo[c] = pp[128 + lane];
                    wf[c] = pp[192 + lane];
                }
            }
        }
#pragma unroll
        for (uint32_t c = 0; c < CB; ++c) {
            if (c < nc) {
                ro += vo[c];
                rf += vf[c];
                if (GRAD) {
                    go += wo[c];
                    gf += wf[c];
                }
            }
        }
    }
    const int cell = i + g.gs * j + g.gs * g.gs * k;
    del_occ[cell] = fo * ro;
    del_free[cell] = ff * rf;
    if (GRAD) {
        dls_occ[cell] = (fo * inv_ls) * go;
        dls_free[cell] = (ff * inv_ls) * gf;
    }
}
hipError_t launch_bgki_combine(hipStream_t st, int ntiles, const GridGeom& g, const TileGeom& tg, const ItemBufs& I,
                               float sigma_o, float sigma_f, float* del_occ, float* del_free, DevScalars* sc) {
    hipLaunchKernelGGL(k_bgki_combine<false>, dim3(blocks_for(ntiles, BG_WAVES)), dim3(TPB), 0, st, ntiles, g, tg,
                       I.tile_partial, I.partial, del_occ, del_free, (float*)nullptr, (float*)nullptr,
                       k3_sigma_factor(sigma_o), k3_sigma_factor(sigma_f), 0.0f, sc);
    return hipGetLastError();
}
hipError_t launch_bgki_combine_grad(hipStream_t st, int ntiles, const GridGeom& g, const TileGeom& tg,
                                    const ItemBufs& I, float ls, float sigma_o, float sigma_f, float* del_occ,
                                    float* del_free, float* dls_occ, float* dls_free, DevScalars* sc) {
    hipLaunchKernelGGL(k_bgki_combine<true>, dim3(blocks_for(ntiles, BG_WAVES)), dim3(TPB), 0, st, ntiles, g, tg,
                       I.tile_partial, I.partial, del_occ, del_free, dls_occ, dls_free, k3_sigma_factor(sigma_o),
                       k3_sigma_factor(sigma_f), 1.0f / ls, sc);
    return hipGetLastError();
}

// Pair counting over the same items (diagnostic): one wave per item.
__global__ __launch_bounds__(TPB) void k_bgki_count(GridGeom g, TileGeom tg, float kscale,
                                                    const uint4* __restrict__ items, uint32_t n_items,
                                                    const uint32_t* __restrict__ tlist,
                                                    const K3PointRec* __restrict__ rec,
                                                    unsigned long long* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    unsigned long long co = 0, cf = 0;
    for (uint32_t it = blockIdx.x * BG_WAVES + (threadIdx.x >> 6); it < n_items; it += gridDim.x * BG_WAVES) {
        const uint4 item = items[it];
        const int tile = (int)item.x;
        const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
        const int i = tx * 4 + (lane & 3), j = ty * 4 + ((lane >> 2) & 3), k = tz * 4 + (lane >> 4);
        if (!(i < g.gs && j < g.gs && k < g.gsz)) continue;
        const float vx = axis_center(i, g.gs, g.res, g.sx);
        const float vy = axis_center(j, g.gs, g.res, g.sy);
        const float vz = axis_center(k, g.gsz, g.res, g.sz);
        const float dc2s = fmaf(vz, vz, fmaf(vy, vy, vx * vx)) * kscale;
        for (uint32_t e = item.y; e < item.z; ++e) {
            const K3PointRec r = rec[tlist[e]];
            co += k3_occ_s(vx, vy, vz, r, kscale) >= 0.0f;
            cf += k3_free_s(vx, vy, vz, dc2s, r, kscale) >= 0.0f;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        co += __shfl_xor(co, o);
        cf += __shfl_xor(cf, o);
    }
    if (lane == 0) {
        atomicAdd(&counts[0], co);
        atomicAdd(&counts[1], cf);
    }
}
hipError_t launch_bgki_count(hipStream_t st, int grid_blocks, const GridGeom& g, const TileGeom& tg, float ls,
                             const ItemBufs& I, uint32_t n_items, const LevelBufs& T, const K3PointRec* rec,
                             unsigned long long* counts) {
    hipLaunchKernelGGL(k_bgki_count, dim3(grid_blocks), dim3(TPB), 0, st, g, tg, k3_kernel_scale(ls), I.items,
                       n_items, T.list, rec, counts);
    return hipGetLastError();
}

}  // namespace k3
