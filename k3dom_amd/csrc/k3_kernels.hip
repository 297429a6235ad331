// gfx950 kernels of the K3DOM per-frame dynamic occupancy update.
//
// Layout: every per-voxel and per-particle quantity is a separate array
// (structure of arrays), so each stage streams only the fields it uses with
// fully coalesced 4-byte lanes.  All reductions whose result is floating point
// are done in a fixed order (ascending particle index per voxel, two-level
// chunked point order in BGKI); all reductions that feed an integer decision
// are done in u64 fixed point; so every result is bit-identical to the serial
// restatement regardless of launch geometry.
//
// Reference kernels these replace (youngjae-min/k3dom):
//   k_predict            src/common/kernel/predict.cu:12-59
//   k_radix_* / k_gather src/common/dom.cu:169-183 (thrust::sort_by_key)
//   k_gather (ranges)    src/common/kernel/particle_to_grid.cu:13-39
//   k_point_prep, k_list_level, k_bgki, k_bgki_combine
//                        src/k3dom/kernel/mass_update.cu:32-77 (BGKI)
//   k_unit_weight_sums, k_mass_update
//                        src/k3dom/kernel/mass_update.cu:152-183
//   k_unit_persist, k_voxel_particles
//                        src/common/kernel/update_persistent_particles.cu:17-55,
//                        src/common/kernel/statistical_moments.cu:22-63,
//                        src/common/kernel/resampling.cu:13-36 (persistent part)
//   k_fixed_*, k_bounds  src/common/dom.cu:213-219 + init_new_particles.cu:44-52 (CDF)
//   k_newborn            src/common/kernel/init_new_particles.cu:106-183
//   k_init_particles     src/common/kernel/init_new_particles.cu:54-104
//   k_resample           src/common/kernel/resampling.cu:38-89
//   k_shift_*            src/common/kernel/ego_motion_compensation.cu:12-41
#include "k3_kernels.h"
#include "k3_mappoint.h"

#include <algorithm>

namespace k3 {

constexpr int TPB = 256;

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
static inline unsigned blocks_for(long long n, int per_block = TPB) {
    long long b = (n + per_block - 1) / per_block;
    return (unsigned)(b < 1 ? 1 : b);
}

// ============================================================== particles
// The predicted state goes to one 32-byte record per particle
// (px, py, pz, vx | vy, vz, w, 0), read only by the sort's gather, which then
// fetches each particle's state in two 16-byte loads; the keys go to P.key.
__global__ __launch_bounds__(TPB) void k_predict(const FrameConsts* __restrict__ fc, ParticleArrays P,
                                                 float4* __restrict__ prec, int n) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const K3PredictConst c = fc->pc;
    // a pending map shift first, as k_shift_particles would have applied it
    float px = P.px[i] - fc->pshift[0], py = P.py[i] - fc->pshift[1], pz = P.pz[i] - fc->pshift[2];
    float vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
    float w = P.w[i];
    int key;
    k3_predict_particle((uint32_t)i, c, &px, &py, &pz, &vx, &vy, &vz, &w, &key);
    prec[2 * i] = make_float4(px, py, pz, vx);
    prec[2 * i + 1] = make_float4(vy, vz, w, 0.0f);
    P.key[i] = key;
}
hipError_t launch_predict(hipStream_t st, const FrameConsts* fc, ParticleArrays P, float4* prec, int n) {
    hipLaunchKernelGGL(k_predict, dim3(blocks_for(n)), dim3(TPB), 0, st, fc, P, prec, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_shift_particles(ParticleArrays P, int n, int mx, int my, int mz) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    P.px[i] -= (float)mx;
    P.py[i] -= (float)my;
    P.pz[i] -= (float)mz;
}
hipError_t launch_shift_particles(hipStream_t st, ParticleArrays P, int n, int mx, int my, int mz) {
    hipLaunchKernelGGL(k_shift_particles, dim3(blocks_for(n)), dim3(TPB), 0, st, P, n, mx, my, mz);
    return hipGetLastError();
}

// new[c] = old[c + move] when in range; otherwise the init kernel's values for
// the fields it sets (masses = priors, ranges = -1) and the cell's own old value
// for the fields it leaves alone (src/common/dom.cu:132-151).
__global__ __launch_bounds__(TPB) void k_shift_voxels(VoxelArrays dst, VoxelArrays src, int gs, int gsz, int mx,
                                                      int my, int mz, float pf, float ps, float pd, int fields) {
    int V = gs * gs * gsz;
    int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= V) return;
    int x = c % gs, y = (c % (gs * gs)) / gs, z = c / (gs * gs);
    int nx = x + mx, ny = y + my, nz = z + mz;
    bool in = nx >= 0 && nx < gs && ny >= 0 && ny < gs && nz >= 0 && nz < gsz;
    int sc = in ? nx + gs * ny + gs * gs * nz : c;
    // every field loaded (unconditionally: sc is always a voxel) before the
    // first store, so that the loads are all in flight together
    if (fields & SHIFT_KEPT) {
        const float f = src.f[sc], s = src.s[sc], d = src.d[sc], mu = src.mu[sc];
        dst.f[c] = in ? f : pf;
        dst.s[c] = in ? s : ps;
        dst.d[c] = in ? d : pd;
        dst.mu[c] = mu;
    }
    if (fields & SHIFT_REWRITTEN) {
        const int32_t st = src.start[sc], en = src.end[sc];
        const float pers = src.pers[sc], pred = src.pred[sc], wua = src.wua[sc];
        const float mvx = src.mvx[sc], mvy = src.mvy[sc], mvz = src.mvz[sc];
        dst.start[c] = in ? st : -1;
        dst.end[c] = in ? en : -1;
        dst.pers[c] = pers;
        dst.pred[c] = pred;
        dst.wua[c] = wua;
        dst.mvx[c] = mvx;
        dst.mvy[c] = mvy;
        dst.mvz[c] = mvz;
    }
}
hipError_t launch_shift_voxels(hipStream_t st, VoxelArrays dst, VoxelArrays src, int gs, int gsz, int mx, int my,
                               int mz, float pf, float ps, float pd, int fields) {
    hipLaunchKernelGGL(k_shift_voxels, dim3(blocks_for((long long)gs * gs * gsz)), dim3(TPB), 0, st, dst, src, gs,
                       gsz, mx, my, mz, pf, ps, pd, fields);
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_fill_voxels(VoxelArrays v, int V, float pf, float ps, float pd) {
    int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= V) return;
    v.f[c] = pf; v.s[c] = ps; v.d[c] = pd;
    v.start[c] = -1; v.end[c] = -1;
    v.pers[c] = 0.0f; v.pred[c] = 0.0f; v.mu[c] = 0.0f; v.wua[c] = 0.0f;
    v.mvx[c] = 0.0f; v.mvy[c] = 0.0f; v.mvz[c] = 0.0f;
}
hipError_t launch_fill_voxels(hipStream_t st, VoxelArrays v, int V, float pf, float ps, float pd) {
    hipLaunchKernelGGL(k_fill_voxels, dim3(blocks_for(V)), dim3(TPB), 0, st, v, V, pf, ps, pd);
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_reset_ranges(int32_t* start, int32_t* end, int V) {
    int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= V) return;
    start[c] = -1;
    end[c] = -1;
}

// exclusive block scan of one value per thread (256 threads)
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* lds4, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(x, o);
        if (lane >= o) x += u;
    }
    if (lane == 63) lds4[w] = x;
    __syncthreads();
    uint32_t add = 0;
    for (int ww = 0; ww < w; ++ww) add += lds4[ww];
    *total = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return x - v + add;
}

// ============================================================== radix sort
// Stable LSD radix sort with digits of db <= 8 bits (the key bits split evenly
// over the passes).  Per pass two kernels: the digit counts of every tile of
// RS_TILE keys (and, by atomics, of every group of S consecutive tiles), then
// a scatter in which each tile adds up its own digit offsets from those counts
// (ceil(ntiles / S) + S - 1 rows of 256 words at most, read from L2) and ranks
// equal digits in input order with wavefront ballots.  Each pass has its own
// group counts, zero on entry: the histogram kernel of pass p clears those of
// pass p + 1 (of pass 0, for the last pass: read by a finished scatter), so a
// sort leaves them as it found them (at least two passes, so a pass never
// clears its own).
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = TPB * RS_ITEMS;
constexpr int RS_BATCH = 16;

int sort_blocks(int n) { return (n + RS_TILE - 1) / RS_TILE; }
// tiles per group, about sqrt(ntiles), so that both sums of the scatter's
// prologue stay short
static int sort_group(int nblk) {
    int S = 1;
    while ((long long)S * S < nblk) ++S;
    return S;
}
constexpr int RS_MAX_PASSES = 4;  // 32-bit keys
// counts per tile, and per group for each pass
long long sort_hist_words(int n) {
    const int nblk = sort_blocks(n);
    const int S = sort_group(nblk);
    return 256ll * nblk + RS_MAX_PASSES * 256ll * ((nblk + S - 1) / S);
}

// The peers of each lane: the valid lanes whose digit (db bits) equals its own.
__device__ __forceinline__ uint64_t digit_peers(uint32_t dg, int db, bool valid) {
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < db; ++b) {
        const bool bit = (dg >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

// cnt[b * 256 + d]: keys of digit d in tile b; grp[g * 256 + d] += the same
// over the tiles of group g (zero on entry); clear[0, nclear) = 0; and
// fa, fb[0, nfill) = -1.
__global__ __launch_bounds__(TPB) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift, int db,
                                                    int S, uint32_t* __restrict__ cnt, uint32_t* __restrict__ grp,
                                                    uint32_t* __restrict__ clear, int nclear, int32_t* __restrict__ fa,
                                                    int32_t* __restrict__ fb, int nfill) {
    __shared__ uint32_t h[256];
    const uint32_t dmask = (1u << db) - 1u;
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    uint32_t k[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) k[r] = keys[min(base + r * TPB + t, n - 1)];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r)
        if (base + r * TPB + t < n) atomicAdd(&h[(k[r] >> shift) & dmask], 1u);
    __syncthreads();
    if (t <= (int)dmask) {
        const uint32_t v = h[t];
        cnt[(size_t)blockIdx.x * 256 + t] = v;
        if (v) atomicAdd(&grp[(size_t)(blockIdx.x / S) * 256 + t], v);
    }
    // (the clears last: a load's wait issued after stores also waits for them)
    if (blockIdx.x * 256 + t < nclear) clear[blockIdx.x * 256 + t] = 0u;
    for (long long i = (long long)blockIdx.x * 256 + t; i < nfill; i += (long long)gridDim.x * 256) {
        fa[i] = -1;
        fb[i] = -1;
    }
}

// Each wavefront ranks a contiguous quarter of the tile (RS_ITEMS rounds of 64
// keys, digit counters private to the wavefront, so no block barrier per
// round); the tile is then staged in LDS in (digit, input order) order and
// written out in runs of equal digits, so consecutive lanes store to
// consecutive addresses.
__global__ __launch_bounds__(TPB) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ iin, int n, int shift, int db,
                                                       const uint32_t* __restrict__ cnt,
                                                       const uint32_t* __restrict__ grp, int nblk, int S, int ngrp,
                                                       uint32_t* __restrict__ kout, uint32_t* __restrict__ iout) {
    __shared__ uint32_t wc[4][256];
    __shared__ uint32_t gbase[256], lbase[256];
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t sk[RS_TILE], si[RS_TILE];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int R = 1 << db;
    const uint32_t dmask = (uint32_t)R - 1u;
    const int tile = blockIdx.x, g = tile / S;
    const int base = tile * RS_TILE;
    const int wbase = base + w * 64 * RS_ITEMS + lane;
    uint32_t k[RS_ITEMS], id[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    // (clamped, not predicated, loads: all in flight at once)
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int e = min(wbase + r * 64, n - 1);
        k[r] = kin[e];
        id[r] = iin ? iin[e] : (uint32_t)e;
    }
    // The tile's start for digit d: the keys of digits below d, plus the keys
    // of digit d in earlier groups (threads d < R) and in earlier tiles of
    // this group (threads R + d, or thread d itself when 2R > TPB).
    // (loads in batches of RS_BATCH, all in flight before the first add)
    uint32_t total = 0, before = 0, inside = 0;
    if (t < R) {
        const uint32_t* gp = grp + t;
        for (int j0 = 0; j0 < ngrp; j0 += RS_BATCH) {
            uint32_t v[RS_BATCH];
#pragma unroll
            for (int q = 0; q < RS_BATCH; ++q) v[q] = gp[(size_t)min(j0 + q, ngrp - 1) * 256];
#pragma unroll
            for (int q = 0; q < RS_BATCH; ++q) {
                total += j0 + q < ngrp ? v[q] : 0u;
                before += j0 + q < g ? v[q] : 0u;
            }
        }
    }
    const int di = 2 * R <= TPB ? t - R : t;
    if (di >= 0 && di < R) {
        const uint32_t* cp = cnt + di;
        for (int b0 = g * S; b0 < tile; b0 += RS_BATCH) {
            uint32_t v[RS_BATCH];
#pragma unroll
            for (int q = 0; q < RS_BATCH; ++q) v[q] = cp[(size_t)min(b0 + q, tile - 1) * 256];
#pragma unroll
            for (int q = 0; q < RS_BATCH; ++q) inside += b0 + q < tile ? v[q] : 0u;
        }
    }
    if (2 * R <= TPB && di >= 0 && di < R) gbase[di] = inside;
    wc[0][t] = 0; wc[1][t] = 0; wc[2][t] = 0; wc[3][t] = 0;
    uint32_t all;
    // (its barriers also order the gbase and wc stores above before the reads below)
    const uint32_t below_d = block_excl_scan_u32(total, lds4, &all);
    const uint32_t gb = t < R ? below_d + before + (2 * R <= TPB ? gbase[t] : inside) : 0u;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const bool valid = wbase + r * 64 < n;
        const uint32_t dg = (k[r] >> shift) & dmask;
        const uint64_t peers = digit_peers(dg, db, valid);
        const uint32_t below = lanes_below(peers);
        uint32_t old = 0;
        if (valid) old = wc[w][dg];
        rank[r] = old + below;
        if (valid && below == 0) wc[w][dg] = old + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // per digit: the wavefronts' offsets within the digit, and the tile's
    // exclusive digit offsets
    uint32_t c0 = wc[0][t], c1 = wc[1][t], c2 = wc[2][t], c3 = wc[3][t];
    uint32_t tot;
    const uint32_t lb = block_excl_scan_u32(c0 + c1 + c2 + c3, lds4, &tot);
    wc[0][t] = lb;
    wc[1][t] = lb + c0;
    wc[2][t] = lb + c0 + c1;
    wc[3][t] = lb + c0 + c1 + c2;
    gbase[t] = gb;
    lbase[t] = lb;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        if (wbase + r * 64 < n) {
            const uint32_t p = wc[w][(k[r] >> shift) & dmask] + rank[r];
            sk[p] = k[r];
            si[p] = id[r];
        }
    }
    __syncthreads();
    const int ntile = min(RS_TILE, n - base);
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int i = r * TPB + t;
        if (i < ntile) {
            const uint32_t kk = sk[i];
            const uint32_t d = (kk >> shift) & dmask;
            const uint32_t p = gbase[d] + (uint32_t)i - lbase[d];
            kout[p] = kk;
            iout[p] = si[i];
        }
    }
}

// Exclusive scans of u32.  The element count is n_host, or (*n_dev) * n_mul
// when n_dev is given (a count produced on the device); the grid is sized for
// n_max.  The total is also written to out[n] and, if given, to *total.  in
// may equal out.

// (a count read from device memory never exceeds the capacity the launch covers)
__device__ __forceinline__ uint32_t scan_count(const uint32_t* n_dev, uint32_t n_mul, uint32_t n_host) {
    return n_dev ? min((*n_dev) * n_mul, n_host) : n_host;
}

// One-workgroup scan for up to SC_SMALL elements (one launch instead of three).
constexpr int SC_SMALL_ITEMS = 16;
constexpr int SC_ITEMS = 16;  // elements per thread of both scan kernels
constexpr int SC_SMALL = 1024 * SC_SMALL_ITEMS;
// SC_ITEMS consecutive elements from b0, loaded unconditionally (as 16-byte
// loads when in is 16-byte aligned) with indices clamped to the capacity cap.
__device__ __forceinline__ void sc_load(const uint32_t* in, uint32_t b0, uint32_t cap, bool vec,
                                        uint32_t (&v)[SC_ITEMS]) {
    if (vec && b0 + SC_ITEMS <= cap) {
        const uint4* p = reinterpret_cast<const uint4*>(in + b0);
#pragma unroll
        for (int k = 0; k < SC_ITEMS / 4; ++k) {
            const uint4 q = p[k];
            v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SC_ITEMS; ++k) v[k] = in[min(b0 + k, cap - 1u)];
    }
}

__global__ __launch_bounds__(1024) void k_scan_small_u32(const uint32_t* in, const uint32_t* __restrict__ n_dev,
                                                         uint32_t n_mul, uint32_t n_host, uint32_t* out,
                                                         uint32_t* __restrict__ total, bool vec) {
    __shared__ uint32_t wsum[16];
    // (the count and the data in flight together)
    const uint32_t n = scan_count(n_dev, n_mul, n_host);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t b0 = (uint32_t)t * SC_SMALL_ITEMS;
    uint32_t v[SC_SMALL_ITEMS] = {};
    if (n_host) sc_load(in, b0, n_host, vec, v);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_SMALL_ITEMS; ++k) {
        v[k] = b0 + k < n ? v[k] : 0u;
        s += v[k];
    }
    uint32_t x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(x, o);
        if (lane >= o) x += u;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t add = 0, all = 0;
    for (int ww = 0; ww < 16; ++ww) {
        add += ww < w ? wsum[ww] : 0u;
        all += wsum[ww];
    }
    uint32_t run = x - s + add;
#pragma unroll
    for (int k = 0; k < SC_SMALL_ITEMS; ++k) {
        if (b0 + k < n) out[b0 + k] = run;
        run += v[k];
    }
    if (t == 0) {
        out[n] = all;
        if (total) *total = all;
    }
}

// Single-pass exclusive scan (decoupled look-back).  Tiles of LB_TILE
// elements are taken in the order blocks start (a ticket counter), so a block
// only waits for tiles whose blocks are already running.  Each block
// publishes its tile sum at once, then wavefront 0 reads the status words of
// the preceding tiles, 64 at a time, adding tile sums back to the nearest
// tile that has published its inclusive prefix, and publishes its own.  A
// status word is (epoch << 34) | (flag << 32) | value: flag 1 = tile sum,
// 2 = inclusive prefix; a word of another launch's epoch counts as not yet
// written.  The 64-bit ticket counter holds (epoch << 32) | next tile; the
// block taking a launch's last ticket moves it to the next epoch, tile 0
// (every other block of the launch has taken its ticket by then), so nothing
// needs clearing between launches (the scratch must be zero before the
// first one).
// Scratch (u32 words): [0, 1] the ticket counter, [2..] u64 statuses.
constexpr int LB_ITEMS = 16;
constexpr int LB_TILE = TPB * LB_ITEMS;
static_assert(LB_ITEMS == SC_ITEMS && SC_SMALL_ITEMS == SC_ITEMS, "both scans load with sc_load");
constexpr uint64_t LB_SUM = 1ull << 32, LB_PREFIX = 2ull << 32;
constexpr int LB_EPOCH_SHIFT = 34;
constexpr uint64_t LB_EPOCH_MASK = (1ull << 30) - 1ull;

static inline long long lb_tiles(long long n_max) { return std::max(1ll, (n_max + LB_TILE - 1) / LB_TILE); }

__global__ __launch_bounds__(TPB) void k_scan_lookback_u32(const uint32_t* in, const uint32_t* __restrict__ n_dev,
                                                           uint32_t n_mul, uint32_t n_host, uint32_t* out,
                                                           uint32_t* __restrict__ total, uint32_t* __restrict__ scr,
                                                           uint32_t nblocks, bool vec) {
    __shared__ uint32_t s_tile, s_excl;
    __shared__ uint64_t s_tag;
    __shared__ uint32_t lds4[4];
    unsigned long long* ticket = reinterpret_cast<unsigned long long*>(scr);
    uint64_t* status = reinterpret_cast<uint64_t*>(scr + 2);
    const int t = threadIdx.x, lane = t & 63;
    // The device count and the data of tile blockIdx.x (what the ticket
    // almost always hands out) are loaded while the ticket is taken.
    const uint32_t n = scan_count(n_dev, n_mul, n_host);
    uint32_t v[LB_ITEMS];
    sc_load(in, blockIdx.x * LB_TILE + t * LB_ITEMS, n_host, vec, v);
    if (t == 0) {
        const unsigned long long tk = atomicAdd(ticket, 1ull);
        const uint64_t epoch = tk >> 32;
        const uint32_t tile = (uint32_t)tk;
        if (tile == nblocks - 1u) atomicExch(ticket, (unsigned long long)((epoch + 1ull) << 32));
        s_tile = tile;
        s_tag = (epoch & LB_EPOCH_MASK) << LB_EPOCH_SHIFT;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t tag = s_tag;
    const uint32_t ntiles = (n + LB_TILE - 1) / LB_TILE;
    if (tile < ntiles) {
        const uint32_t b0 = tile * LB_TILE + t * LB_ITEMS;
        if (tile != blockIdx.x) sc_load(in, b0, n_host, vec, v);  // (block-uniform)
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < LB_ITEMS; ++k) {
            v[k] = b0 + k < n ? v[k] : 0u;
            sum += v[k];
        }
        uint32_t tot;
        const uint32_t x = block_excl_scan_u32(sum, lds4, &tot);
        if (t == 0)
            __hip_atomic_store(&status[tile], tag | (tile == 0 ? LB_PREFIX : LB_SUM) | tot, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (t < 64) {
            uint32_t excl = 0;
            for (int j = (int)tile - 1; j >= 0; j -= 64) {
                const int jj = j - lane;
                uint64_t st = tag | LB_PREFIX;  // before tile 0: an empty prefix
                if (jj >= 0) {
                    st = __hip_atomic_load(&status[jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    // the tile of jj was taken before this one, so its sum is on its way
                    while ((st & ~((1ull << 32) - 1ull)) != (tag | LB_SUM) &&
                           (st & ~((1ull << 32) - 1ull)) != (tag | LB_PREFIX)) {
                        __builtin_amdgcn_s_sleep(1);
                        st = __hip_atomic_load(&status[jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                const uint64_t pm = __ballot((st & (3ull << 32)) == LB_PREFIX);
                const int stop = pm ? __builtin_ctzll(pm) : 64;
                uint32_t add = lane <= stop ? (uint32_t)st : 0u;
                for (int o = 32; o > 0; o >>= 1) add += __shfl_xor(add, o);
                excl += add;
                if (pm) break;
            }
            if (t == 0) {
                if (tile > 0)
                    __hip_atomic_store(&status[tile], tag | LB_PREFIX | (excl + tot), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                s_excl = excl;
            }
        }
        __syncthreads();
        uint32_t run = s_excl + x;
#pragma unroll
        for (int k = 0; k < LB_ITEMS; ++k) {
            if (b0 + k < n) out[b0 + k] = run;
            run += v[k];
        }
        if (tile == ntiles - 1 && t == 0) {
            out[n] = s_excl + tot;
            if (total) *total = s_excl + tot;
        }
    } else if (ntiles == 0 && tile == 0 && t == 0) {
        out[0] = 0;
        if (total) *total = 0;
    }
}

long long scan_scratch_words(long long n_max) { return 2 + 2 * lb_tiles(n_max); }

hipError_t launch_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t* scratch, long long n_max,
                           const uint32_t* n_dev, uint32_t n_mul, uint32_t* total) {
    const bool vec = (reinterpret_cast<uintptr_t>(in) & 15u) == 0;
    if (n_max <= SC_SMALL) {
        hipLaunchKernelGGL(k_scan_small_u32, dim3(1), dim3(1024), 0, st, in, n_dev, n_mul, (uint32_t)n_max, out, total,
                           vec);
        return hipGetLastError();
    }
    const uint32_t nb = (uint32_t)lb_tiles(n_max);
    hipLaunchKernelGGL(k_scan_lookback_u32, dim3(nb), dim3(TPB), 0, st, in, n_dev, n_mul, (uint32_t)n_max, out,
                       total, scratch, nb, vec);
    return hipGetLastError();
}

hipError_t launch_radix_sort(hipStream_t st, const int32_t* in_keys, int n, int key_bits, SortScratch s,
                             const uint32_t** out_keys, const uint32_t** out_idx, int32_t* fill_a, int32_t* fill_b,
                             int nfill) {
    if (n <= 0) {
        *out_keys = (const uint32_t*)in_keys;
        *out_idx = s.i0;
        if (nfill > 0) {
            hipLaunchKernelGGL(k_reset_ranges, dim3(blocks_for(nfill)), dim3(TPB), 0, st, fill_a, fill_b, nfill);
            return hipGetLastError();
        }
        return hipSuccess;
    }
    const int nblk = sort_blocks(n);
    const int passes = std::max(2, (key_bits + 7) / 8);
    if (passes > RS_MAX_PASSES) return hipErrorInvalidValue;
    const int db = std::max(1, (key_bits + passes - 1) / passes);
    const int S = sort_group(nblk);
    const int ngrp = (nblk + S - 1) / S;
    uint32_t* cnt = s.hist;
    uint32_t* grp0 = cnt + 256ll * nblk;
    const uint32_t* kin = (const uint32_t*)in_keys;
    const uint32_t* iin = nullptr;
    uint32_t* kb[2] = {s.k0, s.k1};
    uint32_t* ib[2] = {s.i0, s.i1};
    for (int p = 0; p < passes; ++p) {
        int shift = db * p;
        uint32_t* grp = grp0 + 256ll * ngrp * p;
        uint32_t* clear = grp0 + 256ll * ngrp * ((p + 1) % passes);
        hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(TPB), 0, st, kin, n, shift, db, S, cnt, grp, clear,
                           256 * ngrp, fill_a, fill_b, p == 0 ? nfill : 0);
        uint32_t* ko = kb[p & 1];
        uint32_t* io = ib[p & 1];
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(TPB), 0, st, kin, iin, n, shift, db, cnt, grp, nblk,
                           S, ngrp, ko, io);
        kin = ko;
        iin = io;
    }
    *out_keys = kin;
    *out_idx = iin;
    return hipGetLastError();
}

// The sorted particles from their records, and each voxel's range of sorted
// positions (start / end of its run of equal keys; -1 on entry).
__global__ __launch_bounds__(TPB) void k_gather(const float4* __restrict__ prec, ParticleArrays dst,
                                                float* __restrict__ wsorted,
                                                const uint32_t* __restrict__ perm,
                                                const uint32_t* __restrict__ skeys, int n, int32_t* __restrict__ start,
                                                int32_t* __restrict__ end) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    // (the neighbours' keys loaded with the others, before any store: a wait
    // for a load issued after the stores would also wait for the stores)
    const uint32_t j = perm[i];
    const uint32_t k = skeys[i], kprev = skeys[max(i - 1, 0)], knext = skeys[min(i + 1, n - 1)];
    const float4 a = prec[2 * j], b = prec[2 * j + 1];
    dst.px[i] = a.x; dst.py[i] = a.y; dst.pz[i] = a.z;
    dst.vx[i] = a.w; dst.vy[i] = b.x; dst.vz[i] = b.y;
    wsorted[i] = b.z;
    dst.key[i] = (int32_t)k;
    if (i == 0 || kprev != k) start[k] = i;
    if (i == n - 1 || knext != k) end[k] = i;
}
hipError_t launch_gather_sorted(hipStream_t st, const float4* prec, ParticleArrays dst, float* wsorted,
                                const uint32_t* perm, const uint32_t* sorted_keys, int n, int32_t* start,
                                int32_t* end) {
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(n)), dim3(TPB), 0, st, prec, dst, wsorted, perm, sorted_keys, n,
                       start, end);
    return hipGetLastError();
}

// ============================================================== mass update
// Largest positive finite v over the block (TPB threads, all of which make
// this call) to *slot as float bits, 0 when there is none.  The per-block
// slots are reduced by the consumer (k_fixed_exponent), so no two blocks
// write one address: atomics on a single word from every wave serialise.
__device__ void block_max_store(float v, uint32_t* slot) {
    __shared__ float wm[TPB / 64];
    v = (v > 0.0f && v - v == 0.0f) ? v : 0.0f;
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = wm[0];
        for (int i = 1; i < TPB / 64; ++i) m = fmaxf(m, wm[i]);
        *slot = k3_f2u(m);
    }
}

// Per-voxel particle sums.  A voxel's particles are one contiguous segment of
// the sorted arrays, summed with the two-level association of K3_SEG_BLOCK:
// the units (blocks of K3_SEG_BLOCK particles from the segment start) are each
// summed left to right from 0; a voxel of one unit takes that sum, and a voxel
// of several adds its unit sums, in order, to 0.
//
// Unit sums are computed particle-parallel, so the work per wavefront does not
// depend on how the particles crowd into voxels: a wavefront owns UNIT_WAVE
// consecutive sorted positions and lane l the UNIT_LANE of them from
// UNIT_LANE * l.  The wave copies its positions, and the K3_SEG_BLOCK after
// them (the farthest a unit runs past its owner), into LDS with coalesced
// loads; each lane then sums, from LDS, every unit that starts in its
// positions and stores the sum at the unit's start position.  A voxel pass
// adds each voxel's unit sums with coalesced access to the voxel arrays.
constexpr int UNIT_LANE = 4;
constexpr int UNIT_WAVE = 64 * UNIT_LANE;
constexpr int UNIT_SPAN = UNIT_WAVE + K3_SEG_BLOCK;  // staged positions per wave, a multiple of 64
constexpr int UNIT_TPB = 256;
constexpr int UNIT_WAVES = UNIT_TPB / 64;

static inline unsigned unit_blocks(int n) { return (unsigned)((n + UNIT_TPB / 64 * UNIT_WAVE - 1) / (UNIT_TPB / 64 * UNIT_WAVE)); }

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float lane_value(float v, int q) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), q));
}

// First unit start at or after position q of voxel k (k holds q); one past
// the voxel's segment when no unit of it starts there.
__device__ __forceinline__ int first_unit_start_in(int q, int s0, int e0) {
    const int u = q + (K3_SEG_BLOCK - (q - s0) % K3_SEG_BLOCK) % K3_SEG_BLOCK;
    return u > e0 ? e0 + 1 : u;
}
__device__ __forceinline__ int first_unit_start(int q, int k, const int32_t* __restrict__ start,
                                                const int32_t* __restrict__ end) {
    return first_unit_start_in(q, start[k], end[k]);
}

// End (exclusive, local) of the unit starting at local position u: the first
// position after u with another key, at most K3_SEG_BLOCK after u, or m.
// Keys are sorted, so one key compare covers four positions.
__device__ __forceinline__ int unit_end(const int32_t* lk, int u, int m) {
    const int kk = lk[u];
    const int lim = min(u + K3_SEG_BLOCK, m);
    int i = u + 1;
    while (i + 4 <= lim && lk[i + 3] == kk) i += 4;
    while (i < lim && lk[i] == kk) ++i;
    return i;
}

// Two-level sum of a voxel's unit sums u[st + K3_SEG_BLOCK * j] (one lane).
__device__ __forceinline__ float chain_sum(const float* __restrict__ u, int st, int nb) {
    if (nb == 1) return u[st];
    float s = 0.0f;
    int j = 0;
    for (; j + 4 <= nb; j += 4) {
        const float a0 = u[st + j * K3_SEG_BLOCK], a1 = u[st + (j + 1) * K3_SEG_BLOCK];
        const float a2 = u[st + (j + 2) * K3_SEG_BLOCK], a3 = u[st + (j + 3) * K3_SEG_BLOCK];
        s += a0; s += a1; s += a2; s += a3;
    }
    for (; j < nb; ++j) s += u[st + j * K3_SEG_BLOCK];
    return s;
}

// The same sum by the whole wave, for a voxel of many units (st, nb uniform):
// 64 unit sums are loaded at once and added in order via readlane.
__device__ __forceinline__ float wave_chain_sum(const float* __restrict__ u, int st, int nb) {
    const int lane = threadIdx.x & 63;
    float s = 0.0f;
    for (int b0 = 0; b0 < nb; b0 += 64) {
        const float a = b0 + lane < nb ? u[st + (b0 + lane) * K3_SEG_BLOCK] : 0.0f;
        const int cnt = min(64, nb - b0);
        for (int q = 0; q < cnt; ++q) s += lane_value(a, q);
    }
    return s;
}

// Units whose voxel has more than this many are summed by the whole wave.
constexpr int WAVE_CHAIN_UNITS = 64;

__device__ __forceinline__ int unit_count(int st, int en) { return (en - st + K3_SEG_BLOCK) / K3_SEG_BLOCK; }

// Unit sums of the sorted weights.
__global__ __launch_bounds__(UNIT_TPB) void k_unit_weight_sums(const uint32_t* __restrict__ abort_flag,
                                                               const float* __restrict__ w,
                                                               const int32_t* __restrict__ key,
                                                               const int32_t* __restrict__ start,
                                                               const int32_t* __restrict__ end, int n,
                                                               float* __restrict__ usum) {
    __shared__ float lw_all[UNIT_WAVES][UNIT_SPAN];
    __shared__ int32_t lk_all[UNIT_WAVES][UNIT_SPAN];
    if (*abort_flag) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * UNIT_WAVES + wv) * UNIT_WAVE;
    if (r0 >= n) return;
    float* lw = lw_all[wv];
    int32_t* lk = lk_all[wv];
    const int m = min(n - r0, UNIT_SPAN);
    // (all loads in flight together: unconditional, clamped)
    constexpr int NK = UNIT_SPAN / 64;
    float xw[NK];
    int32_t xk[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int i = min(r0 + k * 64 + lane, n - 1);
        xw[k] = w[i];
        xk[k] = key[i];
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int j = k * 64 + lane;
        if (j < m) {
            lw[j] = xw[k];
            lk[j] = xk[k];
        }
    }
    wave_lds_sync();
    const int q0 = UNIT_LANE * lane, q1 = min(q0 + UNIT_LANE, m);
    if (q0 >= q1) return;
    for (int u = first_unit_start(r0 + q0, lk[q0], start, end) - r0; u < q1;) {
        const int e = unit_end(lk, u, m);
        float s = 0.0f;
        int i = u;
        for (; i + 4 <= e; i += 4) {
            const float a0 = lw[i], a1 = lw[i + 1], a2 = lw[i + 2], a3 = lw[i + 3];
            s += a0; s += a1; s += a2; s += a3;
        }
        for (; i < e; ++i) s += lw[i];
        usum[r0 + u] = s;
        u = e;
    }
}

// prediction, evidence update and newborn split of one voxel
// (src/k3dom/kernel/mass_update.cu:163-181); returns the born mass
// (f, s, d and the two evidence values: the voxel's words, loaded by the caller)
__device__ __forceinline__ float mass_voxel(const VoxelArrays& v, int c, float dyn_pred, float f, float s, float d,
                                            float docc, float dfree, float decay, float mass_scale, float p_B,
                                            float* __restrict__ born) {
    const float pred_total = (f + d) * decay;
    k3_predict_mass(&f, &s, &d, dyn_pred, pred_total, decay);
    k3_update_mass(&f, &s, &d, docc, dfree, mass_scale);
    const float b = k3_newborn_part(dyn_pred, pred_total, d, p_B);
    v.f[c] = f;
    v.s[c] = s;
    v.d[c] = d;
    born[c] = b;
    v.pers[c] = d - b;
    v.pred[c] = dyn_pred;
    return b;
}

// Mass update of every voxel (src/k3dom/kernel/mass_update.cu:152-183), the
// predicted dynamic mass being the sum of the voxel's particle weights.
__global__ __launch_bounds__(TPB) void k_mass_update(const uint32_t* __restrict__ abort_flag, VoxelArrays v, int V,
                                                     const float* __restrict__ usum,
                                                     const float* __restrict__ del_occ,
                                                     const float* __restrict__ del_free,
                                                     const FrameConsts* __restrict__ fc, float mass_scale, float p_B,
                                                     float* __restrict__ born, uint32_t* __restrict__ born_parts) {
    if (*abort_flag) return;
    const float decay = fc ? fc->decay : 1.0f;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * TPB + threadIdx.x;
    // (the voxel's seven words in one round trip: unconditional, clamped)
    const int cc = min(c, V - 1);
    const int st0 = v.start[cc], en0 = v.end[cc];
    const float f0 = v.f[cc], s0 = v.s[cc], d0 = v.d[cc], docc = del_occ[cc], dfree = del_free[cc];
    int st = -1, nb = 0;
    if (c < V) {
        st = st0;
        if (st != -1) nb = unit_count(st, en0);
    }
    const bool wide = nb > WAVE_CHAIN_UNITS;
    float dyn = (st != -1 && !wide) ? chain_sum(usum, st, nb) : 0.0f;
    for (uint64_t hm = __ballot(wide); hm; hm &= hm - 1) {
        const int h = __builtin_ctzll(hm);
        const float s = wave_chain_sum(usum, __builtin_amdgcn_readlane(st, h), __builtin_amdgcn_readlane(nb, h));
        if (lane == h) dyn = s;
    }
    float b = 0.0f;
    if (c < V) b = mass_voxel(v, c, dyn, f0, s0, d0, docc, dfree, decay, mass_scale, p_B, born);
    block_max_store(b, born_parts + blockIdx.x);
}

int mass_update_parts(int V) { return (int)blocks_for(V); }

hipError_t launch_mass_update(hipStream_t st, const uint32_t* abort_flag, VoxelArrays v, int V,
                              const float* weights, const int32_t* keys, int n, float* usum, const float* del_occ,
                              const float* del_free, const FrameConsts* fc, float mass_scale, float p_B, float* born,
                              uint32_t* born_parts) {
    if (n > 0)
        hipLaunchKernelGGL(k_unit_weight_sums, dim3(unit_blocks(n)), dim3(UNIT_TPB), 0, st, abort_flag, weights,
                           keys, v.start, v.end, n, usum);
    hipLaunchKernelGGL(k_mass_update, dim3(blocks_for(V)), dim3(TPB), 0, st, abort_flag, v, V, usum, del_occ,
                       del_free, fc, mass_scale, p_B, born, born_parts);
    return hipGetLastError();
}

// ============================================================== LiDAR ingest
struct Rot9 {
    float r[9];
};
__global__ __launch_bounds__(TPB) void k_lidar_rotate(Rot9 R, const float* sx, const float* sy, const float* sz,
                                                      float* dx, float* dy, float* dz, int n) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    float x, y, z;
    k3_lidar_rotate(R.r, sx[i], sy[i], sz[i], &x, &y, &z);
    dx[i] = x;
    dy[i] = y;
    dz[i] = z;
}
hipError_t launch_lidar_rotate(hipStream_t st, const float* rot9, const float* sx, const float* sy, const float* sz,
                               float* dx, float* dy, float* dz, int n) {
    if (n <= 0) return hipSuccess;
    Rot9 R;
    for (int k = 0; k < 9; ++k) R.r[k] = rot9[k];
    hipLaunchKernelGGL(k_lidar_rotate, dim3(blocks_for(n)), dim3(TPB), 0, st, R, sx, sy, sz, dx, dy, dz, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_add_evidence(const uint32_t* __restrict__ abort_flag,
                                                      float* __restrict__ acc_occ, float* __restrict__ acc_free,
                                                      const float* __restrict__ add_occ,
                                                      const float* __restrict__ add_free, int V) {
    if (*abort_flag) return;
    const int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= V) return;
    acc_occ[c] = acc_occ[c] + add_occ[c];
    acc_free[c] = acc_free[c] + add_free[c];
}
hipError_t launch_add_evidence(hipStream_t st, const uint32_t* abort_flag, float* acc_occ, float* acc_free,
                               const float* add_occ, const float* add_free, int V) {
    hipLaunchKernelGGL(k_add_evidence, dim3(blocks_for(V)), dim3(TPB), 0, st, abort_flag, acc_occ, acc_free,
                       add_occ, add_free, V);
    return hipGetLastError();
}

// ============================================================== MapPoint export
// Records and filter: k3_mappoint.h.
__global__ __launch_bounds__(TPB) void k_mappoint_flags(VoxelArrays v, int V, MapPointArgs a,
                                                        uint32_t* __restrict__ flags) {
    const int c = blockIdx.x * TPB + threadIdx.x;
    if (c < V) flags[c] = mp_visible(mp_cell(v, c, a), a) ? 1u : 0u;
}

__global__ __launch_bounds__(TPB) void k_mappoint_fill(VoxelArrays v, int V, MapPointArgs a,
                                                       const uint32_t* __restrict__ flags,
                                                       const uint32_t* __restrict__ off, float4* __restrict__ out) {
    const int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= V || (flags && !flags[c])) return;
    mp_store(out, off ? off[c] : (uint32_t)c, c, mp_cell(v, c, a), a);
}

hipError_t launch_export_mappoints(hipStream_t st, const VoxelArrays& v, int V, const MapPointArgs& a, int mode,
                                   uint32_t* flags, uint32_t* off, uint32_t* scratch, uint32_t* total,
                                   float4* out) {
    if (mode == 0) {
        hipLaunchKernelGGL(k_mappoint_fill, dim3(blocks_for(V)), dim3(TPB), 0, st, v, V, a, nullptr, nullptr, out);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_mappoint_flags, dim3(blocks_for(V)), dim3(TPB), 0, st, v, V, a, flags);
    hipError_t e = launch_scan_u32(st, flags, off, scratch, V, nullptr, 0, total);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mappoint_fill, dim3(blocks_for(V)), dim3(TPB), 0, st, v, V, a, flags, off, out);
    return hipGetLastError();
}

// ============================================================== DS-PHD/MIB method
// Measurement grid (src/dsphdmib/kernel/measurement_grid.cu:64-160): every
// voxel scans the points in index order and keeps the first
// K3DS_MAX_NEIGHBORS whose half-line from the sensor passes within the window
// of its centre (k3ds_ray_near); the masses go to meas_occ / meas_free (the
// evidence arrays of the K3DOM path).
//
// Culled in two levels without changing any result.  A supertile of
// DS_ST_XY x DS_ST_XY x DS_ST_Z voxels keeps the ordered list of the points
// whose line may pass near any of its voxel centres; a tile of 4x4x4 voxels
// (one wavefront, a voxel per lane) filters its supertile's list with the
// same bound for its own box and applies the exact test to what remains, in
// list order.  The bound accepts every point the exact float
// test can accept, so each voxel sees the same accepted points in the same
// order, hence the same first K3DS_MAX_NEIGHBORS (bound: k3ds_box,
// k3ds_may_pass).
constexpr int DS_TPB = 256;
constexpr int DS_ST_XY = 16, DS_ST_Z = 8;
constexpr int DS_TILE = 4;
constexpr int DS_TILES_XY = DS_ST_XY / DS_TILE, DS_TILES_Z = DS_ST_Z / DS_TILE;
constexpr int DS_WAVES_PER_ST = DS_TILES_XY * DS_TILES_XY * DS_TILES_Z;
constexpr int DS_BLOCKS_PER_ST = DS_WAVES_PER_ST / (DS_TPB / 64);
static_assert(DS_WAVES_PER_ST % (DS_TPB / 64) == 0, "whole blocks per supertile");
static_assert(DS_TILE * DS_TILE * DS_TILE == 64, "one voxel per lane");

int ds_supertiles(int gs, int gsz) {
    const int nxy = (gs + DS_ST_XY - 1) / DS_ST_XY;
    return nxy * nxy * ((gsz + DS_ST_Z - 1) / DS_ST_Z);
}

struct DsSuper {
    int i0, j0, k0;  // first voxel
};
__device__ __forceinline__ DsSuper ds_super(int st, int gs) {
    const int nxy = (gs + DS_ST_XY - 1) / DS_ST_XY;
    DsSuper s;
    s.i0 = (st % nxy) * DS_ST_XY;
    s.j0 = ((st / nxy) % nxy) * DS_ST_XY;
    s.k0 = (st / (nxy * nxy)) * DS_ST_Z;
    return s;
}

long long ds_segments_max(int gs, int gsz, long long cap, int seg_len) {
    return ds_supertiles(gs, gsz) + cap / seg_len + 1;
}

// Candidate points of each supertile, one block per supertile, in point
// order: with off == nullptr the count to counts[st], else the indices to
// list[off[st]...] (off: exclusive scan of the counts, off[nst] the total;
// above cap the overflow flag is raised with the total instead) and the
// number of seg_len-entry segments of the list (at least 1) to nseg[st].
__global__ __launch_bounds__(DS_TPB) void k_ds_super_lists(const FrameConsts* __restrict__ fc,
                                                           const float* __restrict__ xs,
                                                           const float* __restrict__ ys,
                                                           const float* __restrict__ zs, float w2, int nst,
                                                           const uint32_t* __restrict__ off,
                                                           uint32_t* __restrict__ counts, int32_t* __restrict__ list,
                                                           uint32_t* __restrict__ nseg, int seg_len,
                                                           long long cap, DevScalars* __restrict__ sc) {
    __shared__ uint32_t wsum[DS_TPB / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const GridGeom g = fc->g;
    const int n = fc->m;
    const int st = blockIdx.x;
    uint32_t pos = 0;
    if (off) {
        const uint32_t total = off[nst];
        if ((long long)total > cap) {
            if (st == 0 && t == 0) {
                sc->ds_total = total;
                atomicOr(&sc->overflow_mask, 1u << 5);
                atomicExch(&sc->overflow, 1u);
            }
            return;
        }
        pos = off[st];
    }
    const DsSuper s = ds_super(st, g.gs);
    const K3DsBox b = k3ds_box(g.gs, g.gsz, g.res, g.sx, g.sy, g.sz, s.i0, min(s.i0 + DS_ST_XY, g.gs) - 1, s.j0,
                               min(s.j0 + DS_ST_XY, g.gs) - 1, s.k0, min(s.k0 + DS_ST_Z, g.gsz) - 1, w2);
    uint32_t running = 0;
    for (int base = 0; base < n; base += DS_TPB) {
        const int q = base + t;
        bool pass = false;
        if (q < n) {
            const float mx = xs[q], my = ys[q], mz = zs[q];
            pass = k3ds_may_pass(b, mx, my, mz, k3ds_m2(mx, my, mz));
        }
        const uint64_t bits = __ballot(pass);
        if (lane == 0) wsum[wave] = (uint32_t)__popcll(bits);
        __syncthreads();
        uint32_t before = 0, chunk = 0;
        for (int k = 0; k < DS_TPB / 64; ++k) {
            before += k < wave ? wsum[k] : 0u;
            chunk += wsum[k];
        }
        if (off && pass) {
            const uint32_t rank = (uint32_t)__popcll(bits & ((1ull << lane) - 1ull));
            list[pos + running + before + rank] = q;
        }
        running += chunk;
        __syncthreads();
    }
    if (t == 0) {
        if (off)
            nseg[st] = running > 0 ? (running + (uint32_t)seg_len - 1) / (uint32_t)seg_len : 1u;
        else
            counts[st] = running;
    }
}

// One voxel's state while it collects its first K3DS_MAX_NEIGHBORS rays.
struct DsVoxel {
    float x, y, z, c2, max_occ;
    int nb;
    bool all_farther, done;
};

// Applies point p (x, y, z, |m|^2) to the voxel when near is its exact test
// (measurement_grid.cu:86-120, in point order); COUNT: only counts.
template <bool COUNT>
__device__ __forceinline__ void ds_take(DsVoxel& v, bool near, const float4& p, float delta) {
    if (v.done || !near) return;
    if (!COUNT) {
        if (v.all_farther && v.c2 > p.w) v.all_farther = false;
        const float o = k3ds_pocc(v.x, v.y, v.z, p.x, p.y, p.z, delta);
        if (o > v.max_occ) v.max_occ = o;
    }
    if (++v.nb == K3DS_MAX_NEIGHBORS) v.done = true;
}

// A wavefront's 4x4x4 tile (lane = voxel): its bound, the bounds of its eight
// 2x2x2 sub-boxes (q = x half | y half << 1 | z half << 2, lane bits 1, 3, 5)
// and the lane's voxel.  False when the tile lies outside the grid.
struct DsTile {
    K3DsBox tb, sb[8];
    uint64_t sub_lanes[8];
    int c;
    bool valid;
};
__device__ __forceinline__ bool ds_tile(const GridGeom& g, int st, int wl, int lane, float w2, DsTile& T,
                                        DsVoxel& v) {
    const DsSuper s = ds_super(st, g.gs);
    const int i0 = s.i0 + (wl % DS_TILES_XY) * DS_TILE;
    const int j0 = s.j0 + ((wl / DS_TILES_XY) % DS_TILES_XY) * DS_TILE;
    const int k0 = s.k0 + (wl / (DS_TILES_XY * DS_TILES_XY)) * DS_TILE;
    if (i0 >= g.gs || j0 >= g.gs || k0 >= g.gsz) return false;
    T.tb = k3ds_box(g.gs, g.gsz, g.res, g.sx, g.sy, g.sz, i0, min(i0 + DS_TILE, g.gs) - 1, j0,
                    min(j0 + DS_TILE, g.gs) - 1, k0, min(k0 + DS_TILE, g.gsz) - 1, w2);
    const int my_sub = ((lane >> 1) & 1) | ((lane >> 2) & 2) | ((lane >> 3) & 4);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = i0 + 2 * (q & 1), b = j0 + ((q >> 1) & 1) * 2, c = k0 + ((q >> 2) & 1) * 2;
        T.sb[q] = k3ds_box(g.gs, g.gsz, g.res, g.sx, g.sy, g.sz, a, min(a + 2, g.gs) - 1, b, min(b + 2, g.gs) - 1,
                           c, min(c + 2, g.gsz) - 1, w2);
        T.sub_lanes[q] = __ballot(my_sub == q);
    }
    const int i = i0 + (lane & 3), j = j0 + ((lane >> 2) & 3), k = k0 + (lane >> 4);
    T.valid = i < g.gs && j < g.gs && k < g.gsz;
    T.c = (k * g.gs + j) * g.gs + i;
    v.x = v.y = v.z = 0.0f;
    if (T.valid) k3_voxel_center(T.c, g.gs, g.gsz, g.res, g.sx, g.sy, g.sz, &v.x, &v.y, &v.z);
    v.c2 = v.x * v.x + v.y * v.y + v.z * v.z;
    v.max_occ = 0.0f;
    v.nb = 0;
    v.all_farther = true;
    v.done = !T.valid;
    return true;
}

// The tile's voxels take the listed points list[b..e) in order.  Chunks of 64
// are loaded one per lane and filtered by the tile's bound, then by the bounds
// of the sub-boxes that still hold a voxel short of its rays; the survivors
// are staged in LDS (sw: 64 float4, sr: 64 reciprocals of |m|^2, of this
// wavefront) in list order and every lane runs the exact test on them, four
// at a time, applying hits in order.
template <bool COUNT>
__device__ __forceinline__ void ds_run(DsVoxel& v, const DsTile& T, float4* sw, float* sr,
                                       const int32_t* __restrict__ list,
                                       uint32_t b, uint32_t e, const float* __restrict__ xs,
                                       const float* __restrict__ ys, const float* __restrict__ zs, float w2,
                                       float delta, int lane) {
    for (uint32_t base = b; base < e; base += 64) {
        const uint64_t open = __ballot(!v.done);
        if (open == 0) break;
        uint32_t active = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) active |= (open & T.sub_lanes[q]) ? (1u << q) : 0u;
        const uint32_t at = base + (uint32_t)lane;
        float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        bool pass = false;
        if (at < e) {
            const int qi = list[at];
            p.x = xs[qi];
            p.y = ys[qi];
            p.z = zs[qi];
            p.w = k3ds_m2(p.x, p.y, p.z);
            if (k3ds_may_pass(T.tb, p.x, p.y, p.z, p.w)) {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if ((active >> q) & 1u) pass = pass || k3ds_may_pass(T.sb[q], p.x, p.y, p.z, p.w);
            }
        }
        const uint64_t bits = __ballot(pass);
        if (bits == 0) continue;
        const int nc = __popcll(bits);
        if (pass) {
            const int at_s = __popcll(bits & ((1ull << lane) - 1ull));
            sw[at_s] = p;
            sr[at_s] = __builtin_amdgcn_rcpf(p.w);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int t = 0; t < nc; t += 4) {
            const int t1 = min(t + 1, nc - 1), t2 = min(t + 2, nc - 1), t3 = min(t + 3, nc - 1);
            const float4 p0 = sw[t], p1 = sw[t1], p2 = sw[t2], p3 = sw[t3];
            const float r0 = sr[t], r1 = sr[t1], r2 = sr[t2], r3 = sr[t3];
            const bool n0 = k3ds_ray_near_rcp(v.x, v.y, v.z, v.c2, p0.x, p0.y, p0.z, p0.w, r0, w2);
            const bool n1 = t + 1 < nc && k3ds_ray_near_rcp(v.x, v.y, v.z, v.c2, p1.x, p1.y, p1.z, p1.w, r1, w2);
            const bool n2 = t + 2 < nc && k3ds_ray_near_rcp(v.x, v.y, v.z, v.c2, p2.x, p2.y, p2.z, p2.w, r2, w2);
            const bool n3 = t + 3 < nc && k3ds_ray_near_rcp(v.x, v.y, v.z, v.c2, p3.x, p3.y, p3.z, p3.w, r3, w2);
            ds_take<COUNT>(v, n0, p0, delta);
            ds_take<COUNT>(v, n1, p1, delta);
            ds_take<COUNT>(v, n2, p2, delta);
            ds_take<COUNT>(v, n3, p3, delta);
        }
        // every lane has read the staged points before the next chunk overwrites them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Work item of a segment pass: block -> (supertile, segment), wavefront ->
// tile.  Segments are numbered supertile by supertile (seg_off: exclusive
// scan of nseg, seg_off[nst] the total); false past the last segment.
struct DsItem {
    int st, seg, nseg, wl;
    uint32_t gseg, seg0;
};
__device__ __forceinline__ bool ds_item(const uint32_t* __restrict__ seg_off, int nst, DsItem& it) {
    it.gseg = blockIdx.x / DS_BLOCKS_PER_ST;
    if (it.gseg >= seg_off[nst]) return false;
    int lo = 0, hi = nst - 1;  // last st with seg_off[st] <= gseg (every supertile has a segment)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg_off[mid] <= it.gseg)
            lo = mid;
        else
            hi = mid - 1;
    }
    it.st = lo;
    it.seg0 = seg_off[lo];
    it.seg = (int)(it.gseg - it.seg0);
    it.nseg = (int)(seg_off[lo + 1] - it.seg0);
    it.wl = (blockIdx.x % DS_BLOCKS_PER_ST) * (DS_TPB / 64) + (threadIdx.x >> 6);
    return true;
}

// Pass A (lists of more than one segment): hits of each voxel in every
// segment but the last, capped at K3DS_MAX_NEIGHBORS, to segcnt.
__global__ __launch_bounds__(DS_TPB) void k_ds_seg_count(const uint32_t* __restrict__ abort_flag,
                                                         const FrameConsts* __restrict__ fc,
                                                         const float* __restrict__ xs, const float* __restrict__ ys,
                                                         const float* __restrict__ zs,
                                                         const uint32_t* __restrict__ off,
                                                         const int32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ seg_off, int nst, int seg_len,
                                                         float w2, uint8_t* __restrict__ segcnt) {
    __shared__ float4 stage[DS_TPB / 64][64];
    __shared__ float stage_r[DS_TPB / 64][64];
    if (*abort_flag) return;
    DsItem it;
    if (!ds_item(seg_off, nst, it) || it.seg == it.nseg - 1) return;
    const int lane = threadIdx.x & 63;
    const GridGeom g = fc->g;
    DsTile T;
    DsVoxel v;
    if (!ds_tile(g, it.st, it.wl, lane, w2, T, v)) return;
    const uint32_t b = off[it.st] + (uint32_t)it.seg * (uint32_t)seg_len;
    const uint32_t e = min(b + (uint32_t)seg_len, off[it.st + 1]);
    ds_run<true>(v, T, stage[threadIdx.x >> 6], stage_r[threadIdx.x >> 6], list, b, e, xs, ys, zs, w2, 0.0f, lane);
    segcnt[((size_t)it.gseg * DS_WAVES_PER_ST + it.wl) * 64 + lane] = (uint8_t)v.nb;
}

// Pass B: every segment, each voxel starting from its hits in the earlier
// segments (so it takes exactly its first K3DS_MAX_NEIGHBORS in point order).
// A one-segment list writes the masses; otherwise (nb | all_farther << 8,
// max pOcc) go to part for pass C.
__global__ __launch_bounds__(DS_TPB) void k_ds_meas_tiles(const uint32_t* __restrict__ abort_flag,
                                                          const FrameConsts* __restrict__ fc,
                                                          const float* __restrict__ xs,
                                                          const float* __restrict__ ys,
                                                          const float* __restrict__ zs,
                                                          const uint32_t* __restrict__ off,
                                                          const int32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ seg_off, int nst, int seg_len,
                                                          const uint8_t* __restrict__ segcnt, float w2, float delta,
                                                          float max_range, uint2* __restrict__ part,
                                                          float* __restrict__ meas_occ,
                                                          float* __restrict__ meas_free) {
    __shared__ float4 stage[DS_TPB / 64][64];
    __shared__ float stage_r[DS_TPB / 64][64];
    if (*abort_flag) return;
    DsItem it;
    if (!ds_item(seg_off, nst, it)) return;
    const int lane = threadIdx.x & 63;
    const GridGeom g = fc->g;
    DsTile T;
    DsVoxel v;
    if (!ds_tile(g, it.st, it.wl, lane, w2, T, v)) return;
    int prior = 0;
    for (int s = 0; s < it.seg; ++s) prior += segcnt[((size_t)(it.seg0 + s) * DS_WAVES_PER_ST + it.wl) * 64 + lane];
    v.nb = min(prior, K3DS_MAX_NEIGHBORS);
    v.done = v.done || v.nb == K3DS_MAX_NEIGHBORS;
    const uint32_t b = off[it.st] + (uint32_t)it.seg * (uint32_t)seg_len;
    const uint32_t e = min(b + (uint32_t)seg_len, off[it.st + 1]);
    ds_run<false>(v, T, stage[threadIdx.x >> 6], stage_r[threadIdx.x >> 6], list, b, e, xs, ys, zs, w2, delta, lane);
    if (it.nseg == 1) {
        if (T.valid)
            k3ds_meas_masses(v.nb, v.all_farther, v.max_occ, v.x, v.y, v.z, max_range, &meas_occ[T.c],
                             &meas_free[T.c]);
    } else {
        part[((size_t)it.gseg * DS_WAVES_PER_ST + it.wl) * 64 + lane] =
            make_uint2((uint32_t)v.nb | (v.all_farther ? 256u : 0u), __float_as_uint(v.max_occ));
    }
}

// Pass C: voxels of lists of more than one segment combine their segments:
// the count after the last, all_farther over all, the largest pOcc.
__global__ __launch_bounds__(DS_TPB) void k_ds_seg_combine(const uint32_t* __restrict__ abort_flag,
                                                           const FrameConsts* __restrict__ fc,
                                                           const uint32_t* __restrict__ seg_off, float w2,
                                                           float max_range, const uint2* __restrict__ part,
                                                           float* __restrict__ meas_occ,
                                                           float* __restrict__ meas_free) {
    if (*abort_flag) return;
    const int st = blockIdx.x / DS_BLOCKS_PER_ST;
    const uint32_t seg0 = seg_off[st];
    const int nseg = (int)(seg_off[st + 1] - seg0);
    if (nseg == 1) return;
    const int lane = threadIdx.x & 63;
    const int wl = (blockIdx.x % DS_BLOCKS_PER_ST) * (DS_TPB / 64) + (threadIdx.x >> 6);
    const GridGeom g = fc->g;
    DsTile T;
    DsVoxel v;
    if (!ds_tile(g, st, wl, lane, w2, T, v) || !T.valid) return;
    int nb = 0;
    bool all_farther = true;
    float max_occ = 0.0f;
    for (int s = 0; s < nseg; ++s) {
        const uint2 r = part[((size_t)(seg0 + s) * DS_WAVES_PER_ST + wl) * 64 + lane];
        nb = (int)(r.x & 255u);
        all_farther = all_farther && (r.x & 256u);
        max_occ = fmaxf(max_occ, __uint_as_float(r.y));
    }
    k3ds_meas_masses(nb, all_farther, max_occ, v.x, v.y, v.z, max_range, &meas_occ[T.c], &meas_free[T.c]);
}

hipError_t launch_ds_meas_grid(hipStream_t st, const uint32_t* abort_flag, const FrameConsts* fc, int gs, int gsz,
                               const float* xs, const float* ys, const float* zs, float resolution,
                               float meas_occ_stddev, float max_range, const DsLists& L, DevScalars* sc,
                               float* meas_occ, float* meas_free) {
    const float wr = resolution * 0.866f;  // window_radius, src/dsphdmib/dom.cu:21-22
    const float w2 = wr * wr;
    const int nst = ds_supertiles(gs, gsz);
    const unsigned seg_blocks = (unsigned)(ds_segments_max(gs, gsz, L.cap, L.seg_len) * DS_BLOCKS_PER_ST);
    hipLaunchKernelGGL(k_ds_super_lists, dim3(nst), dim3(DS_TPB), 0, st, fc, xs, ys, zs, w2, nst, nullptr, L.counts,
                       nullptr, nullptr, L.seg_len, 0ll, sc);
    hipError_t e = launch_scan_u32(st, L.counts, L.off, L.scratch, nst, nullptr, 0, &sc->ds_total);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ds_super_lists, dim3(nst), dim3(DS_TPB), 0, st, fc, xs, ys, zs, w2, nst, L.off, L.counts,
                       L.list, L.nseg, L.seg_len, L.cap, sc);
    e = launch_scan_u32(st, L.nseg, L.seg_off, L.seg_scratch, nst, nullptr, 0, &sc->ds_segments);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ds_seg_count, dim3(seg_blocks), dim3(DS_TPB), 0, st, abort_flag, fc, xs, ys, zs, L.off,
                       L.list, L.seg_off, nst, L.seg_len, w2, L.segcnt);
    hipLaunchKernelGGL(k_ds_meas_tiles, dim3(seg_blocks), dim3(DS_TPB), 0, st, abort_flag, fc, xs, ys, zs, L.off,
                       L.list, L.seg_off, nst, L.seg_len, L.segcnt, w2, meas_occ_stddev, max_range, L.part, meas_occ,
                       meas_free);
    hipLaunchKernelGGL(k_ds_seg_combine, dim3(nst * DS_BLOCKS_PER_ST), dim3(DS_TPB), 0, st, abort_flag, fc,
                       L.seg_off, w2, max_range, L.part, meas_occ, meas_free);
    return hipGetLastError();
}

// Cell update (src/dsphdmib/kernel/mass_update.cu:44-85) of every voxel, the
// predicted occupied mass being the voxel's particle weights (unit sums as in
// k_mass_update).  A voxel above 1 keeps its sum in norm (0 otherwise) for
// k_ds_normalize and continues with 1.  The second mass slot (v.s) holds the
// occupied mass.
__global__ __launch_bounds__(TPB) void k_ds_mass_update(const uint32_t* __restrict__ abort_flag, VoxelArrays v,
                                                        int V, const float* __restrict__ usum,
                                                        const float* __restrict__ meas_occ,
                                                        const float* __restrict__ meas_free, float p_B,
                                                        float* __restrict__ born, float* __restrict__ norm,
                                                        uint32_t* __restrict__ born_parts) {
    if (*abort_flag) return;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * TPB + threadIdx.x;
    // (the voxel's six words in one round trip: unconditional, clamped)
    const int cc = min(c, V - 1);
    const int st0 = v.start[cc], en0 = v.end[cc];
    const float vf = v.f[cc], vs = v.s[cc], mocc = meas_occ[cc], mfree = meas_free[cc];
    int st = -1, nb = 0;
    if (c < V) {
        st = st0;
        if (st != -1) nb = unit_count(st, en0);
    }
    const bool wide = nb > WAVE_CHAIN_UNITS;
    float sum = (st != -1 && !wide) ? chain_sum(usum, st, nb) : 0.0f;
    for (uint64_t hm = __ballot(wide); hm; hm &= hm - 1) {
        const int h = __builtin_ctzll(hm);
        const float s = wave_chain_sum(usum, __builtin_amdgcn_readlane(st, h), __builtin_amdgcn_readlane(nb, h));
        if (lane == h) sum = s;
    }
    float rho = 0.0f;
    if (c < V) {
        float occ_up, free_up, m_occ_pred = 0.0f;
        if (st != -1) {
            m_occ_pred = sum;
            norm[c] = sum > 1.0f ? sum : 0.0f;
            if (sum > 1.0f) m_occ_pred = 1.0f;
            const float m_free_pred = k3ds_predict_free(vf, m_occ_pred);
            k3ds_update_masses(m_occ_pred, m_free_pred, mocc, mfree, &occ_up, &free_up);
            rho = k3ds_newborn_part(m_occ_pred, occ_up, p_B);
        } else {
            const float m_occ = vs;
            const float m_free = k3ds_predict_free(vf, m_occ);
            k3ds_update_masses(m_occ, m_free, mocc, mfree, &occ_up, &free_up);
        }
        born[c] = rho;
        v.pers[c] = occ_up - rho;
        v.f[c] = free_up;
        v.s[c] = occ_up;
        v.pred[c] = m_occ_pred;
    }
    block_max_store(rho, born_parts + blockIdx.x);
}

// Division of the weights of voxels whose sum exceeded 1 (normalize_weights,
// src/dsphdmib/kernel/mass_update.cu:37-42), on the sorted weights that the
// persistent update reads.
__global__ __launch_bounds__(TPB) void k_ds_normalize(const uint32_t* __restrict__ abort_flag,
                                                      const int32_t* __restrict__ keys,
                                                      const float* __restrict__ norm, float* __restrict__ w, int n) {
    if (*abort_flag) return;
    const int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= n) return;
    const float s = norm[keys[j]];
    if (s > 0.0f) w[j] = w[j] / s;
}

hipError_t launch_ds_mass_update(hipStream_t st, const uint32_t* abort_flag, VoxelArrays v, int V, float* weights,
                                 const int32_t* keys, int n, float* usum, const float* meas_occ,
                                 const float* meas_free, float p_B, float* born, float* norm,
                                 uint32_t* born_parts) {
    if (n > 0)
        hipLaunchKernelGGL(k_unit_weight_sums, dim3(unit_blocks(n)), dim3(UNIT_TPB), 0, st, abort_flag, weights,
                           keys, v.start, v.end, n, usum);
    hipLaunchKernelGGL(k_ds_mass_update, dim3(blocks_for(V)), dim3(TPB), 0, st, abort_flag, v, V, usum, meas_occ,
                       meas_free, p_B, born, norm, born_parts);
    if (n > 0)
        hipLaunchKernelGGL(k_ds_normalize, dim3(blocks_for(n)), dim3(TPB), 0, st, abort_flag, keys, norm, weights, n);
    return hipGetLastError();
}

// First frame (src/dsphdmib/dom.cu:55-82, copyMassesKernel at
// measurement_grid.cu:55-62): the born masses are the measurement grid's
// occupied masses.
__global__ __launch_bounds__(TPB) void k_ds_born_from_meas(const float* __restrict__ meas_occ, float* __restrict__ born,
                                                           int V, uint32_t* __restrict__ born_parts) {
    const int c = blockIdx.x * TPB + threadIdx.x;
    float b = 0.0f;
    if (c < V) {
        b = meas_occ[c];
        born[c] = b;
    }
    block_max_store(b, born_parts + blockIdx.x);
}
hipError_t launch_ds_born_from_meas(hipStream_t st, const float* meas_occ, float* born, int V, uint32_t* born_parts) {
    hipLaunchKernelGGL(k_ds_born_from_meas, dim3(blocks_for(V)), dim3(TPB), 0, st, meas_occ, born, V, born_parts);
    return hipGetLastError();
}

// One particle's persistent update (src/common/kernel/update_persistent_particles.cu:23-24,52-53
// with likelihood 1 and p_A 0, which K3DOM never changes): returns the new
// weight and, through *pen, the weight after the static penalty
// (src/common/kernel/resampling.cu:24-28).
__device__ __forceinline__ float persist_weight(float w, float m, float vx, float vy, float vz, float min2,
                                                float* pen) {
    const float pA = 0.0f, lik = 1.0f;
    const float post = lik * w;
    const float wn = pA * m * post + (1.0f - pA) * m * w;
    *pen = (vx * vx + vy * vy + vz * vz < min2) ? 0.0f : wn;
    return wn;
}

// mu of a voxel with particles: persistent over predicted dynamic mass
__device__ __forceinline__ float voxel_mu(float rho, float pr) { return pr > 0.0f ? rho / pr : 0.0f; }

// Persistent update of every sorted particle, with mu of its voxel: the
// penalised weight goes to P.w (each position by its owning wave), and the
// unit sums of new weight x velocity to ux, uy, uz.
__global__ __launch_bounds__(UNIT_TPB) void k_unit_persist(const uint32_t* __restrict__ abort_flag,
                                                           ParticleArrays P, const float* __restrict__ wsrc,
                                                           VoxelArrays v, int n, float min_vel,
                                                           float* __restrict__ ux, float* __restrict__ uy,
                                                           float* __restrict__ uz, uint32_t* __restrict__ joint_parts) {
    __shared__ float lx_all[UNIT_WAVES][UNIT_SPAN];
    __shared__ float ly_all[UNIT_WAVES][UNIT_SPAN];
    __shared__ float lz_all[UNIT_WAVES][UNIT_SPAN];
    __shared__ int32_t lk_all[UNIT_WAVES][UNIT_SPAN];
    if (*abort_flag) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * UNIT_WAVES + wv) * UNIT_WAVE;
    float* lx = lx_all[wv];
    float* ly = ly_all[wv];
    float* lz = lz_all[wv];
    int32_t* lk = lk_all[wv];
    const int m = max(0, min(n - r0, UNIT_SPAN));
    const float min2 = min_vel * min_vel;
    float wmax = 0.0f;
    // All loads first, in two rounds (the particles, then their voxels),
    // unconditional with clamped indices, instead of two dependent round
    // trips per 64 positions.
    // The range of the voxel at this lane's first unit position (r0 + q0) is
    // loaded in the second round too: loaded after the P.w stores, its wait
    // would also wait for them.
    constexpr int NK = UNIT_SPAN / 64;
    int kk[NK];
    float px_[NK], py_[NK], pz_[NK], pw_[NK], pe_[NK], pr_[NK];
    int q_start = 0, q_end = 0;
    if (n > 0) {
        int kq;
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const int i = min(r0 + k * 64 + lane, n - 1);
            kk[k] = P.key[i];
            px_[k] = P.vx[i]; py_[k] = P.vy[i]; pz_[k] = P.vz[i];
            pw_[k] = wsrc[i];
        }
        kq = P.key[min(r0 + UNIT_LANE * lane, n - 1)];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            pe_[k] = v.pers[kk[k]];
            pr_[k] = v.pred[kk[k]];
        }
        q_start = v.start[kq];
        q_end = v.end[kq];
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int j = k * 64 + lane;
        if (j < m) {
            const int i = r0 + j;
            const int kj = kk[k];
            const float x = px_[k], y = py_[k], z = pz_[k];
            const float mu = voxel_mu(pe_[k], pr_[k]);
            float pen;
            const float wn = persist_weight(pw_[k], mu, x, y, z, min2, &pen);
            lx[j] = wn * x;
            ly[j] = wn * y;
            lz[j] = wn * z;
            lk[j] = kj;
            if (j < UNIT_WAVE) {
                P.w[i] = pen;
                if (pen > wmax && pen - pen == 0.0f) wmax = pen;
            }
        }
    }
    block_max_store(wmax, joint_parts + blockIdx.x);
    const int q0 = UNIT_LANE * lane, q1 = min(q0 + UNIT_LANE, m);
    if (q0 >= q1) return;
    for (int u = first_unit_start_in(r0 + q0, q_start, q_end) - r0; u < q1;) {
        const int e = unit_end(lk, u, m);
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        int i = u;
        for (; i + 2 <= e; i += 2) {
            const float x0 = lx[i], y0 = ly[i], z0 = lz[i];
            const float x1 = lx[i + 1], y1 = ly[i + 1], z1 = lz[i + 1];
            sx += x0; sy += y0; sz += z0;
            sx += x1; sy += y1; sz += z1;
        }
        if (i < e) {
            sx += lx[i]; sy += ly[i]; sz += lz[i];
        }
        ux[r0 + u] = sx;
        uy[r0 + u] = sy;
        uz[r0 + u] = sz;
        u = e;
    }
}

// Per voxel: mu and the mean velocity from the updated weights
// (src/common/kernel/update_persistent_particles.cu:17-55).
__global__ __launch_bounds__(TPB) void k_voxel_particles(const uint32_t* __restrict__ abort_flag, VoxelArrays v,
                                                         int V, const float* __restrict__ ux,
                                                         const float* __restrict__ uy,
                                                         const float* __restrict__ uz) {
    if (*abort_flag) return;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * TPB + threadIdx.x;
    // (the voxel's four words in one round trip: unconditional, clamped)
    const int cc = min(c, V - 1);
    const int st0 = v.start[cc], en0 = v.end[cc];
    const float rho = v.pers[cc], pred = v.pred[cc];
    int st = -1, nb = 0;
    if (c < V) {
        st = st0;
        if (st != -1) nb = unit_count(st, en0);
    }
    const bool wide = nb > WAVE_CHAIN_UNITS;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (st != -1 && !wide) {
        sx = chain_sum(ux, st, nb);
        sy = chain_sum(uy, st, nb);
        sz = chain_sum(uz, st, nb);
    }
    for (uint64_t hm = __ballot(wide); hm; hm &= hm - 1) {
        const int h = __builtin_ctzll(hm);
        const int hs = __builtin_amdgcn_readlane(st, h), hn = __builtin_amdgcn_readlane(nb, h);
        const float tx = wave_chain_sum(ux, hs, hn);
        const float ty = wave_chain_sum(uy, hs, hn);
        const float tz = wave_chain_sum(uz, hs, hn);
        if (lane == h) {
            sx = tx;
            sy = ty;
            sz = tz;
        }
    }
    if (c >= V) return;
    if (st == -1) {
        v.mvx[c] = 0.0f;
        v.mvy[c] = 0.0f;
        v.mvz[c] = 0.0f;
    } else {
        v.mu[c] = voxel_mu(rho, pred);
        v.mvx[c] = rho > 0.0f ? sx / rho : 0.0f;
        v.mvy[c] = rho > 0.0f ? sy / rho : 0.0f;
        v.mvz[c] = rho > 0.0f ? sz / rho : 0.0f;
    }
}

int voxel_particles_parts(int n) { return n > 0 ? (int)unit_blocks(n) : 0; }

hipError_t launch_voxel_particles(hipStream_t st, const uint32_t* abort_flag, VoxelArrays v, int V,
                                  ParticleArrays P, const float* wsrc, int n, float min_vel, float* usum3,
                                  uint32_t* joint_parts) {
    float* ux = usum3;
    float* uy = usum3 + n;
    float* uz = usum3 + 2 * (size_t)n;
    if (n > 0)
        hipLaunchKernelGGL(k_unit_persist, dim3(unit_blocks(n)), dim3(UNIT_TPB), 0, st, abort_flag, P, wsrc, v, n,
                           min_vel, ux, uy, uz, joint_parts);
    hipLaunchKernelGGL(k_voxel_particles, dim3(blocks_for(V)), dim3(TPB), 0, st, abort_flag, v, V, ux, uy, uz);
    return hipGetLastError();
}

// ============================================================== fixed-point CDF
// Position i of a[0, na) followed by b[0, nb) (na + nb > 0), 0 past the end:
// one unconditional load from a selected, clamped address, so that a thread's
// loads are all in flight together.
__device__ __forceinline__ float cat_load(const float* a, int na, const float* b, int nb, int i) {
    const int n = na + nb;
    const int c = min(i, n - 1);
    const float* p = c < na ? a + c : b + (c - na);
    const float x = *p;
    return i < n ? x : 0.0f;
}

// Exponent of the fixed-point CDF from the largest of nparts per-block maxima
// (float bits of non-negative values, so their unsigned order is the float order).
__global__ __launch_bounds__(1024) void k_fixed_exponent(const uint32_t* __restrict__ parts, int nparts, int count,
                                                         int32_t* E) {
    __shared__ uint32_t wm[16];
    uint32_t m = 0;
    for (int i = threadIdx.x; i < nparts; i += 1024) m = max(m, parts[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 16; ++i) m = max(m, wm[i]);
        *E = k3_fixed_exponent(k3_u2f(m), (uint32_t)count);
    }
}

// inclusive block scan of one u64 per thread (256 threads); returns the block total
__device__ uint64_t block_scan_incl_u64(uint64_t v, uint64_t* lds4, uint64_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    if (lane == 63) lds4[w] = v;
    __syncthreads();
    uint64_t add = 0;
    for (int ww = 0; ww < w; ++ww) add += lds4[ww];
    *total = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return v + add;
}

// Inclusive u64 scan of the fixed-point values of [a ++ b] in three phases:
// tile sums (u64 sums are exact, so consecutive lanes can take consecutive
// elements), a one-block scan of the tile sums, then per tile a scan with the
// tile staged in LDS (padded by one word in 16 against bank conflicts), each
// thread scanning 16 consecutive elements, and stored with consecutive lanes
// on consecutive elements.
constexpr int FL_ITEMS = 16;
constexpr int FL_TILE = TPB * FL_ITEMS;
__device__ __forceinline__ int fl_pad(int i) { return i + (i >> 4); }

static inline int fl_tiles(long long n) { return (int)std::max(1ll, (n + FL_TILE - 1) / FL_TILE); }
long long fixed_scan_scratch_words(long long n) { return fl_tiles(n); }

__global__ __launch_bounds__(TPB) void k_fixed_reduce(const float* __restrict__ a, int na,
                                                      const float* __restrict__ b, int nb,
                                                      const int32_t* __restrict__ Ep,
                                                      uint64_t* __restrict__ tile_sums) {
    __shared__ uint64_t lds4[4];
    const int E = *Ep;
    const int n = na + nb;
    const int base = blockIdx.x * FL_TILE + threadIdx.x;
    uint64_t s = 0;
    float x[FL_ITEMS] = {};
    if (n > 0) {
#pragma unroll
        for (int k = 0; k < FL_ITEMS; ++k) x[k] = cat_load(a, na, b, nb, base + k * TPB);  // (0 past the end)
    }
#pragma unroll
    for (int k = 0; k < FL_ITEMS; ++k) s += k3_to_fixed(x[k], E);
    uint64_t tot;
    block_scan_incl_u64(s, lds4, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_top_u64(uint64_t* __restrict__ sums, int n,
                                                       uint64_t* __restrict__ total) {
    __shared__ uint64_t part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    uint64_t s = 0;
    for (int i = b; i < e; ++i) s += sums[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint64_t v = (t >= off) ? part[t - off] : 0ull;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = part[t] - s;
    for (int i = b; i < e; ++i) {
        uint64_t v = sums[i];
        sums[i] = run;
        run += v;
    }
    if (t == 1023) *total = part[1023];
}

__global__ __launch_bounds__(TPB) void k_fixed_down(const float* __restrict__ a, int na,
                                                    const float* __restrict__ b, int nb,
                                                    const int32_t* __restrict__ Ep,
                                                    const uint64_t* __restrict__ tile_excl,
                                                    uint64_t* __restrict__ cdf) {
    __shared__ uint64_t tile_v[FL_TILE + FL_TILE / 16];
    __shared__ uint64_t lds4[4];
    const int E = *Ep;
    const int n = na + nb;
    const int t = threadIdx.x;
    const int base = blockIdx.x * FL_TILE;
    float x[FL_ITEMS] = {};
    if (n > 0) {
#pragma unroll
        for (int k = 0; k < FL_ITEMS; ++k) x[k] = cat_load(a, na, b, nb, base + k * TPB + t);  // (0 past the end)
    }
#pragma unroll
    for (int k = 0; k < FL_ITEMS; ++k) tile_v[fl_pad(k * TPB + t)] = k3_to_fixed(x[k], E);
    __syncthreads();
    uint64_t q[FL_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < FL_ITEMS; ++k) {
        q[k] = tile_v[fl_pad(t * FL_ITEMS + k)];
        s += q[k];
    }
    uint64_t tot;
    uint64_t run = tile_excl[blockIdx.x] + block_scan_incl_u64(s, lds4, &tot) - s;
#pragma unroll
    for (int k = 0; k < FL_ITEMS; ++k) {
        run += q[k];
        tile_v[fl_pad(t * FL_ITEMS + k)] = run;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < FL_ITEMS; ++k) {
        const int e = base + k * TPB + t;
        if (e < n) cdf[e] = tile_v[fl_pad(k * TPB + t)];
    }
}

hipError_t launch_fixed_scan(hipStream_t st, const float* a, int na, const float* b, int nb,
                             const uint32_t* max_parts, int nparts, int32_t* E_out, uint64_t* cdf,
                             uint64_t* scratch, uint64_t* total) {
    const int n = na + nb;
    const int nblk = fl_tiles(n);
    hipLaunchKernelGGL(k_fixed_exponent, dim3(1), dim3(1024), 0, st, max_parts, nparts, n, E_out);
    hipLaunchKernelGGL(k_fixed_reduce, dim3(nblk), dim3(TPB), 0, st, a, na, b, nb, E_out, scratch);
    hipLaunchKernelGGL(k_scan_top_u64, dim3(1), dim3(1024), 0, st, scratch, nblk, total);
    hipLaunchKernelGGL(k_fixed_down, dim3(nblk), dim3(TPB), 0, st, a, na, b, nb, E_out, scratch, cdf);
    return hipGetLastError();
}

__device__ __forceinline__ uint32_t slot_bound(uint64_t c, uint64_t tot, uint32_t cnt, double scale) {
    if (c == tot) return cnt;
    uint64_t b = (uint64_t)((double)c * scale);
    return b < (uint64_t)cnt ? (uint32_t)b : cnt;
}

// Slot boundaries of a fixed-point CDF split into cnt particles, and the
// unassociated newborn weight of each cell (src/common/kernel/init_new_particles.cu:119-134).
__global__ __launch_bounds__(TPB) void k_bounds(const uint32_t* __restrict__ abort_flag, const uint64_t* __restrict__ cdf,
                                                const uint64_t* __restrict__ total, int V, uint32_t cnt,
                                                const float* __restrict__ born, uint32_t* __restrict__ bnd,
                                                float* __restrict__ wua) {
    if (*abort_flag) return;
    int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= V) return;
    // (every load issued before the first store, unconditionally)
    const uint64_t tot = *total;
    const uint64_t cj = cdf[j], cp = cdf[max(j - 1, 0)];
    const float bornj = wua ? born[j] : 0.0f;
    uint32_t bj, bp;
    if (tot == 0) {
        bj = (uint32_t)(((uint64_t)(j + 1) * cnt) / (uint64_t)V);
        bp = (uint32_t)(((uint64_t)j * cnt) / (uint64_t)V);
    } else {
        const double scale = (double)cnt / (double)tot;
        bj = slot_bound(cj, tot, cnt, scale);
        bp = j ? slot_bound(cp, tot, cnt, scale) : 0u;
    }
    bnd[j] = bj;
    if (wua) {
        const float pA = 0.0f;
        int nn = (int)(bj - bp);
        wua[j] = nn > 0 ? (float)(((1.0 - (double)pA) * (double)bornj) / (double)nn) : 0.0f;
    }
}
hipError_t launch_bounds(hipStream_t st, const uint32_t* abort_flag, const uint64_t* cdf, const uint64_t* total, int V, uint32_t cnt,
                         const float* born, uint32_t* bnd, float* wua) {
    hipLaunchKernelGGL(k_bounds, dim3(blocks_for(V)), dim3(TPB), 0, st, abort_flag, cdf, total, V, cnt, born, bnd, wua);
    return hipGetLastError();
}

__device__ __forceinline__ int slot_owner(const uint32_t* __restrict__ bnd, int V, uint32_t k) {
    int lo = 0, hi = V;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (bnd[mid] <= k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(TPB) void k_newborn(const uint32_t* __restrict__ abort_flag, const uint32_t* __restrict__ bnd, int V, int B, int gs,
                                                 const float* __restrict__ wua, ParticleArrays NB, float vmax,
                                                 float min_vel, const FrameConsts* __restrict__ fc, uint32_t k0,
                                                 uint32_t k1, uint32_t* __restrict__ joint_parts) {
    if (*abort_flag) return;
    const uint32_t frame = fc->frame;
    int k = blockIdx.x * TPB + threadIdx.x;
    float wmax = 0.0f;
    if (k < B) {
        int c = slot_owner(bnd, V, (uint32_t)k);
        uint32_t r[4];
        k3_philox4x32_10((uint32_t)k, frame, K3_STREAM_NEWBORN, 0u, k0, k1, r);
        float vx, vy, vz;
        k3_sample_velocity(r[0], r[1], r[2], vmax, min_vel, &vx, &vy, &vz);
        float w = wua[c];
        if (vx * vx + vy * vy + vz * vz < min_vel * min_vel) w = 0.0f;
        NB.px[k] = (float)(c % gs) + 0.5f;
        NB.py[k] = (float)((c % (gs * gs)) / gs) + 0.5f;
        NB.pz[k] = (float)(c / (gs * gs)) + 0.5f;
        NB.vx[k] = vx; NB.vy[k] = vy; NB.vz[k] = vz;
        NB.w[k] = w;
        NB.key[k] = c;
        wmax = w;
    }
    block_max_store(wmax, joint_parts + blockIdx.x);
}
int newborn_parts(int B) { return B > 0 ? (int)blocks_for(B) : 0; }
hipError_t launch_newborn(hipStream_t st, const uint32_t* abort_flag, const uint32_t* bnd, int V, int B, int gs, const float* wua,
                          ParticleArrays NB, float vmax, float min_vel, const FrameConsts* fc, uint32_t k0,
                          uint32_t k1, uint32_t* joint_parts) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_newborn, dim3(blocks_for(B)), dim3(TPB), 0, st, abort_flag, bnd, V, B, gs, wua, NB, vmax, min_vel,
                       fc, k0, k1, joint_parts);
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_init_particles(const uint32_t* __restrict__ abort_flag, const uint32_t* __restrict__ bnd, int V, int N, int gs,
                                                        ParticleArrays P, float vmax, float min_vel,
                                                        const FrameConsts* __restrict__ fc, uint32_t k0, uint32_t k1) {
    if (*abort_flag) return;
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= N) return;
    const uint32_t frame = fc->frame;
    int c = slot_owner(bnd, V, (uint32_t)i);
    uint32_t r[4];
    k3_philox4x32_10((uint32_t)i, frame, K3_STREAM_INIT, 0u, k0, k1, r);
    float vx, vy, vz;
    k3_sample_velocity(r[0], r[1], r[2], vmax, min_vel, &vx, &vy, &vz);
    P.px[i] = (float)(c % gs) + 0.5f;
    P.py[i] = (float)((c % (gs * gs)) / gs) + 0.5f;
    P.pz[i] = (float)(c / (gs * gs)) + 0.5f;
    P.vx[i] = vx; P.vy[i] = vy; P.vz[i] = vz;
    P.w[i] = 1.0f / (float)N;
    P.key[i] = c;
}
hipError_t launch_init_particles(hipStream_t st, const uint32_t* abort_flag, const uint32_t* bnd, int V, int N, int gs, ParticleArrays P,
                                 float vmax, float min_vel, const FrameConsts* fc, uint32_t k0, uint32_t k1) {
    hipLaunchKernelGGL(k_init_particles, dim3(blocks_for(N)), dim3(TPB), 0, st, abort_flag, bnd, V, N, gs, P, vmax, min_vel,
                       fc, k0, k1);
    return hipGetLastError();
}

// Systematic resampling on the u64 CDF of [persistent ++ newborn] weights:
// positions T_i = trunc((i + u) * Q / N) (IEEE double), index = first CDF
// entry > T_i; then the gather into the next particle set with weight W/N.
__global__ __launch_bounds__(TPB) void k_resample(const uint32_t* __restrict__ abort_flag, const uint64_t* __restrict__ cdf, int L,
                                                  const uint64_t* __restrict__ total,
                                                  const int32_t* __restrict__ Ep, ParticleArrays S,
                                                  ParticleArrays NB, ParticleArrays O, int N,
                                                  const FrameConsts* __restrict__ fc, uint32_t k0, uint32_t k1,
                                                  int32_t* __restrict__ idx_out) {
    if (*abort_flag) return;
    const uint32_t frame = fc->frame;
    const int i = blockIdx.x * TPB + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // a wavefront past the last output has nothing to do; the others keep
    // every lane for the wavefront-wide steps (lanes past N repeat output N - 1)
    if (i - lane >= N) return;
    const int ic = min(i, N - 1);
    const uint64_t Q = *total;
    const int E = *Ep;
    const float W = (float)((double)Q * k3_pow2d(-E));
    const float nw = W / (float)N;
    int idx;
    if (Q == 0) {
        idx = ic;
    } else {
        uint32_t r[4];
        k3_philox4x32_10(0u, frame, K3_STREAM_RESAMPLE, 0u, k0, k1, r);
        const double u = (double)r[0] * 0x1p-32;
        const double step = (double)Q / (double)N;
        uint64_t T = (uint64_t)(((double)ic + u) * step);
        if (T > Q - 1) T = Q - 1;
        // The ancestor is the number of CDF entries <= T (the first entry
        // above T).  T grows with the output, so lane 0's target is the
        // wavefront's smallest: its count a0 comes from a 64-way sampled
        // search (each round narrows [lo, hi] to one sample interval: four
        // rounds for 2.2M entries instead of 21 dependent loads), and every
        // lane's count from windows of 64 entries read from a0 on, each lane
        // searching the window's sorted values by shuffles.
        // (readfirstlane returns int: each word goes through uint32_t, not sign-extended)
        const uint64_t T0 = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(T >> 32)) << 32) |
                            (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)T);
        uint32_t lo = 0, hi = (uint32_t)L;
        while (hi - lo > 64u) {
            const uint32_t st = (hi - lo + 63u) >> 6;
            const uint32_t j = lo + (uint32_t)lane * st;
            const uint64_t b = __ballot(j < hi && cdf[j] <= T0);
            if (!b) {  // the entry at lo is already above T0
                hi = lo;
                break;
            }
            const uint32_t jk = lo + (uint32_t)(63 - __builtin_clzll(b)) * st;
            lo = jk + 1u;
            hi = min(jk + st, hi);
        }
        if (hi > lo) {
            const uint32_t j = lo + (uint32_t)lane;
            lo += (uint32_t)__popcll(__ballot(j < hi && cdf[j] <= T0));
        }
        uint32_t base = lo;
        bool done = false;
        uint32_t ans = 0;
        for (int win = 0; win < 8; ++win, base += 64u) {
            const uint32_t j = base + (uint32_t)lane;
            const uint64_t v = j < (uint32_t)L ? cdf[j] : ~0ull;
            const int vlo = (int)(uint32_t)v, vhi = (int)(uint32_t)(v >> 32);
            uint32_t c = 0;  // window entries <= T (they increase with the lane)
            for (uint32_t sz = 64; sz > 0; sz >>= 1) {
                const uint32_t at = c + sz - 1u;
                const int src = (int)min(at, 63u);
                const uint64_t x = ((uint64_t)(uint32_t)__shfl(vhi, src) << 32) | (uint64_t)(uint32_t)__shfl(vlo, src);
                if (at < 64u && x <= T) c += sz;
            }
            if (!done && c < 64u) {
                ans = base + c;
                done = true;
            }
            if (!__ballot(!done)) break;
        }
        if (!done) {  // a wavefront whose targets span more than 8 windows
            uint32_t a = base, b = (uint32_t)L;
            while (a < b) {
                const uint32_t mid = (a + b) >> 1;
                if (cdf[mid] <= T) a = mid + 1u;
                else b = mid;
            }
            ans = a;
        }
        idx = (int)ans;
    }
    if (i >= N) return;
    idx_out[i] = idx;
    // the ancestor's seven words from the sorted or the newborn set, all
    // loaded (through selected pointers) before the first store
    const bool pers = idx < N;
    const int j = pers ? idx : idx - N;
    const float px = (pers ? S.px : NB.px)[j], py = (pers ? S.py : NB.py)[j], pz = (pers ? S.pz : NB.pz)[j];
    const float vx = (pers ? S.vx : NB.vx)[j], vy = (pers ? S.vy : NB.vy)[j], vz = (pers ? S.vz : NB.vz)[j];
    const int32_t key = (pers ? S.key : NB.key)[j];
    O.px[i] = px; O.py[i] = py; O.pz[i] = pz;
    O.vx[i] = vx; O.vy[i] = vy; O.vz[i] = vz;
    O.key[i] = key;
    O.w[i] = nw;
}
hipError_t launch_resample(hipStream_t st, const uint32_t* abort_flag, const uint64_t* cdf, int total_len, const uint64_t* total,
                           const int32_t* E, ParticleArrays sorted, ParticleArrays nb, ParticleArrays out, int N,
                           const FrameConsts* fc, uint32_t k0, uint32_t k1, int32_t* idx_out) {
    hipLaunchKernelGGL(k_resample, dim3(blocks_for(N)), dim3(TPB), 0, st, abort_flag, cdf, total_len, total, E, sorted, nb, out,
                       N, fc, k0, k1, idx_out);
    return hipGetLastError();
}

__global__ void k_reset_scalars(DevScalars* sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        DevScalars z = {};
        *sc = z;
    }
}
hipError_t launch_reset_scalars(hipStream_t st, DevScalars* sc) {
    hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(64), 0, st, sc);
    return hipGetLastError();
}

// Clears the overflow state before a re-run of the
// kernel-inference stage (the other scalars are untouched by the aborted tail).
__global__ void k_clear_overflow(DevScalars* sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->overflow = 0;
        sc->overflow_mask = 0;
    }
}
hipError_t launch_clear_overflow(hipStream_t st, DevScalars* sc) {
    hipLaunchKernelGGL(k_clear_overflow, dim3(1), dim3(64), 0, st, sc);
    return hipGetLastError();
}

}  // namespace k3
