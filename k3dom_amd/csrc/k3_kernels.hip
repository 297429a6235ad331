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
//   k_mark_ranges        src/common/kernel/particle_to_grid.cu:13-39
//   k_point_prep, k_list_level, k_bgki, k_bgki_combine
//                        src/k3dom/kernel/mass_update.cu:32-77 (BGKI)
//   k_mass_update        src/k3dom/kernel/mass_update.cu:152-183
//   k_voxel_particles    src/common/kernel/update_persistent_particles.cu:17-55,
//                        src/common/kernel/statistical_moments.cu:22-63,
//                        src/common/kernel/resampling.cu:13-36 (persistent part)
//   k_fixed_*, k_bounds  src/common/dom.cu:213-219 + init_new_particles.cu:44-52 (CDF)
//   k_newborn            src/common/kernel/init_new_particles.cu:106-183
//   k_init_particles     src/common/kernel/init_new_particles.cu:54-104
//   k_resample           src/common/kernel/resampling.cu:38-89
//   k_shift_*            src/common/kernel/ego_motion_compensation.cu:12-41
#include "k3_kernels.h"

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
__global__ __launch_bounds__(TPB) void k_predict(K3PredictConst c, ParticleArrays P, int n) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    float px = P.px[i], py = P.py[i], pz = P.pz[i];
    float vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
    float w = P.w[i];
    int key;
    k3_predict_particle((uint32_t)i, c, &px, &py, &pz, &vx, &vy, &vz, &w, &key);
    P.px[i] = px; P.py[i] = py; P.pz[i] = pz;
    P.vx[i] = vx; P.vy[i] = vy; P.vz[i] = vz;
    P.w[i] = w;
    P.key[i] = key;
}
hipError_t launch_predict(hipStream_t st, const K3PredictConst& c, ParticleArrays P, int n) {
    hipLaunchKernelGGL(k_predict, dim3(blocks_for(n)), dim3(TPB), 0, st, c, P, n);
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
                                                      int my, int mz, float pf, float ps, float pd) {
    int V = gs * gs * gsz;
    int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= V) return;
    int x = c % gs, y = (c % (gs * gs)) / gs, z = c / (gs * gs);
    int nx = x + mx, ny = y + my, nz = z + mz;
    bool in = nx >= 0 && nx < gs && ny >= 0 && ny < gs && nz >= 0 && nz < gsz;
    int sc = in ? nx + gs * ny + gs * gs * nz : c;
    dst.f[c] = in ? src.f[sc] : pf;
    dst.s[c] = in ? src.s[sc] : ps;
    dst.d[c] = in ? src.d[sc] : pd;
    dst.start[c] = in ? src.start[sc] : -1;
    dst.end[c] = in ? src.end[sc] : -1;
    dst.pers[c] = src.pers[sc];
    dst.pred[c] = src.pred[sc];
    dst.mu[c] = src.mu[sc];
    dst.wua[c] = src.wua[sc];
    dst.mvx[c] = src.mvx[sc];
    dst.mvy[c] = src.mvy[sc];
    dst.mvz[c] = src.mvz[sc];
}
hipError_t launch_shift_voxels(hipStream_t st, VoxelArrays dst, VoxelArrays src, int gs, int gsz, int mx, int my,
                               int mz, float pf, float ps, float pd) {
    hipLaunchKernelGGL(k_shift_voxels, dim3(blocks_for((long long)gs * gs * gsz)), dim3(TPB), 0, st, dst, src, gs,
                       gsz, mx, my, mz, pf, ps, pd);
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

// ============================================================== radix sort
// Stable LSD radix sort, 8-bit digits.  Per pass: per-tile LDS histograms
// (digit-major in global memory), one exclusive scan, then a scatter that
// ranks equal digits in input order with wavefront ballots.
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = TPB * RS_ITEMS;

int sort_blocks(int n) { return (n + RS_TILE - 1) / RS_TILE; }

__global__ __launch_bounds__(TPB) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                    uint32_t* __restrict__ hist, int nblk) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; ++r) {
        int e = base + r * TPB + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(TPB) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ iin, int n, int shift,
                                                       const uint32_t* __restrict__ offs, int nblk,
                                                       uint32_t* __restrict__ kout, uint32_t* __restrict__ iout) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[4][256];
    const int t = threadIdx.x, w = t >> 6;
    run[t] = offs[t * nblk + blockIdx.x];
    const int base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        wc[0][t] = 0; wc[1][t] = 0; wc[2][t] = 0; wc[3][t] = 0;
        __syncthreads();
        int e = base + r * TPB + t;
        bool valid = e < n;
        uint32_t k = valid ? kin[e] : 0u;
        uint32_t id = valid ? (iin ? iin[e] : (uint32_t)e) : 0u;
        uint32_t dg = (k >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (dg >> b) & 1u;
            uint64_t bal = __ballot(valid && bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t rank = lanes_below(peers);
        if (valid && rank == 0) wc[w][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[dg] + rank;
            for (int ww = 0; ww < w; ++ww) pos += wc[ww][dg];
            kout[pos] = k;
            iout[pos] = id;
        }
        __syncthreads();
        run[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
        __syncthreads();
    }
}

// Single-workgroup exclusive scan of u32 (n up to a few million).
__global__ __launch_bounds__(1024) void k_scan_u32_single(const uint32_t* __restrict__ in,
                                                          uint32_t* __restrict__ out, int n,
                                                          uint32_t* __restrict__ total) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = b; i < e; ++i) {
        uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
    if (t == 1023 && total) *total = part[1023];
}
hipError_t launch_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out_excl, int n, uint32_t* total) {
    hipLaunchKernelGGL(k_scan_u32_single, dim3(1), dim3(1024), 0, st, in, out_excl, n, total);
    return hipGetLastError();
}

hipError_t launch_radix_sort(hipStream_t st, const int32_t* in_keys, int n, int key_bits, SortScratch s,
                             const uint32_t** out_keys, const uint32_t** out_idx) {
    const int nblk = sort_blocks(n);
    const int passes = key_bits <= 8 ? 1 : (key_bits + 7) / 8;
    const uint32_t* kin = (const uint32_t*)in_keys;
    const uint32_t* iin = nullptr;
    uint32_t* kb[2] = {s.k0, s.k1};
    uint32_t* ib[2] = {s.i0, s.i1};
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(TPB), 0, st, kin, n, shift, s.hist, nblk);
        hipLaunchKernelGGL(k_scan_u32_single, dim3(1), dim3(1024), 0, st, s.hist, s.hist, 256 * nblk,
                           (uint32_t*)nullptr);
        uint32_t* ko = kb[p & 1];
        uint32_t* io = ib[p & 1];
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(TPB), 0, st, kin, iin, n, shift, s.hist, nblk, ko,
                           io);
        kin = ko;
        iin = io;
    }
    *out_keys = kin;
    *out_idx = iin;
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_gather(ParticleArrays src, ParticleArrays dst,
                                                const uint32_t* __restrict__ perm,
                                                const uint32_t* __restrict__ skeys, int n) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    uint32_t j = perm[i];
    dst.px[i] = src.px[j]; dst.py[i] = src.py[j]; dst.pz[i] = src.pz[j];
    dst.vx[i] = src.vx[j]; dst.vy[i] = src.vy[j]; dst.vz[i] = src.vz[j];
    dst.w[i] = src.w[j];
    dst.key[i] = (int32_t)skeys[i];
}
hipError_t launch_gather_sorted(hipStream_t st, ParticleArrays src, ParticleArrays dst, const uint32_t* perm,
                                const uint32_t* sorted_keys, int n) {
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(n)), dim3(TPB), 0, st, src, dst, perm, sorted_keys, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_reset_ranges(int32_t* start, int32_t* end, int V) {
    int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= V) return;
    start[c] = -1;
    end[c] = -1;
}
__global__ __launch_bounds__(TPB) void k_mark_ranges(const int32_t* __restrict__ key, int n, int32_t* start,
                                                     int32_t* end) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    int k = key[i];
    if (i == 0 || key[i - 1] != k) start[k] = i;
    if (i == n - 1 || key[i + 1] != k) end[k] = i;
}
hipError_t launch_ranges(hipStream_t st, const int32_t* sorted_keys, int n, int32_t* start, int32_t* end, int V) {
    hipLaunchKernelGGL(k_reset_ranges, dim3(blocks_for(V)), dim3(TPB), 0, st, start, end, V);
    hipLaunchKernelGGL(k_mark_ranges, dim3(blocks_for(n)), dim3(TPB), 0, st, sorted_keys, n, start, end);
    return hipGetLastError();
}

// ============================================================== BGKI
__global__ __launch_bounds__(TPB) void k_point_prep(const float* __restrict__ xs, const float* __restrict__ ys,
                                                    const float* __restrict__ zs, int m, float res,
                                                    K3PointRec* __restrict__ rec, float4* __restrict__ pxyz) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= m) return;
    K3PointRec r = k3_point_prep(xs[i], ys[i], zs[i], res);
    rec[i] = r;
    pxyz[i] = make_float4(r.x, r.y, r.z, r.valid != 0.0f ? r.dm : -1.0f);
}
hipError_t launch_point_prep(hipStream_t st, const float* xs, const float* ys, const float* zs, int m, float res,
                             K3PointRec* rec, float4* pxyz) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_point_prep, dim3(blocks_for(m)), dim3(TPB), 0, st, xs, ys, zs, m, res, rec, pxyz);
    return hipGetLastError();
}

// Sensor-relative coordinate of the centre of voxel index i along an axis of
// n voxels (the per-axis part of k3_voxel_center).
__device__ __forceinline__ float axis_center(int i, int n, float res, float s) {
    return (((float)i + 0.5f) - (float)n / 2.0f) * res - s;
}

struct BoxSphere {
    float cx, cy, cz, R;
};
// Bounding sphere of the voxel centres of tiles [t0, t1) per axis.
__device__ BoxSphere box_sphere(const GridGeom& g, int tx0, int ty0, int tz0, int tx1, int ty1, int tz1) {
    int i0 = tx0 * 4, i1 = min(tx1 * 4, g.gs) - 1;
    int j0 = ty0 * 4, j1 = min(ty1 * 4, g.gs) - 1;
    int k0 = tz0 * 4, k1 = min(tz1 * 4, g.gsz) - 1;
    float x0 = axis_center(i0, g.gs, g.res, g.sx), x1 = axis_center(i1, g.gs, g.res, g.sx);
    float y0 = axis_center(j0, g.gs, g.res, g.sy), y1 = axis_center(j1, g.gs, g.res, g.sy);
    float z0 = axis_center(k0, g.gsz, g.res, g.sz), z1 = axis_center(k1, g.gsz, g.res, g.sz);
    BoxSphere b;
    b.cx = 0.5f * (x0 + x1);
    b.cy = 0.5f * (y0 + y1);
    b.cz = 0.5f * (z0 + z1);
    float hx = 0.5f * (x1 - x0), hy = 0.5f * (y1 - y0), hz = 0.5f * (z1 - z0);
    b.R = sqrtf(hx * hx + hy * hy + hz * hz);
    return b;
}

// True when some voxel centre of the box may lie within ls of the segment
// [0, p]: distance(box centre, segment) < ls + R + slack.  Every (voxel, point)
// pair with a non-zero term passes (the slack covers FP32 rounding of both
// this test and the exact per-pair distances).
__device__ __forceinline__ bool capsule_hits(const BoxSphere& b, float4 p, float ls) {
    if (!(p.w >= 0.0f)) return false;
    float mm = p.x * p.x + p.y * p.y + p.z * p.z;
    float t = 0.0f;
    if (mm > 0.0f) t = fminf(fmaxf((b.cx * p.x + b.cy * p.y + b.cz * p.z) / mm, 0.0f), 1.0f);
    float dx = b.cx - t * p.x, dy = b.cy - t * p.y, dz = b.cz - t * p.z;
    float d2 = dx * dx + dy * dy + dz * dz;
    float cn = fabsf(b.cx) + fabsf(b.cy) + fabsf(b.cz);
    float thr = ls + b.R + 2e-3f + 1e-5f * (cn + p.w);
    return d2 < thr * thr;
}

// One workgroup per box of this level: scans the parent's candidate list (all
// points at the root) in order and counts (count_mode) or writes, in the same
// ascending order, the candidates whose capsule reaches the box.
__global__ __launch_bounds__(TPB) void k_list_level(int count_mode, GridGeom g, LevelGeom L, int ntx, int nty,
                                                   int ntz, float ls, const float4* __restrict__ pxyz, int m,
                                                   const uint32_t* __restrict__ parent_off,
                                                   const uint32_t* __restrict__ parent_len,
                                                   const uint32_t* __restrict__ parent_list,
                                                   uint32_t* __restrict__ len, const uint32_t* __restrict__ off,
                                                   uint32_t* __restrict__ list) {
    __shared__ uint32_t wcnt[4];
    const int box = blockIdx.x;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int bxi = box % L.nbx, byi = (box / L.nbx) % L.nby, bzi = box / (L.nbx * L.nby);
    const int tx0 = bxi * L.bx, ty0 = byi * L.by, tz0 = bzi * L.bz;
    const int tx1 = min(tx0 + L.bx, ntx), ty1 = min(ty0 + L.by, nty), tz1 = min(tz0 + L.bz, ntz);
    uint32_t poff = 0, plen = (uint32_t)m;
    if (L.pbx > 0) {
        int pb = (tx0 / L.pbx) + L.pnbx * ((ty0 / L.pby) + L.pnby * (tz0 / L.pbz));
        poff = parent_off[pb];
        plen = parent_len[pb];
    }
    const BoxSphere bs = box_sphere(g, tx0, ty0, tz0, tx1, ty1, tz1);
    uint32_t running = 0;
    const uint32_t obase = count_mode ? 0u : off[box];
    for (uint32_t base = 0; base < plen; base += TPB) {
        uint32_t e = base + t;
        bool pass = false;
        uint32_t idx = 0;
        if (e < plen) {
            idx = parent_list ? parent_list[poff + e] : e;
            pass = capsule_hits(bs, pxyz[idx], ls);
        }
        uint64_t mask = __ballot(pass);
        if (lane == 0) wcnt[w] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t pre = 0;
        for (int ww = 0; ww < w; ++ww) pre += wcnt[ww];
        uint32_t tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        if (!count_mode && pass) list[obase + running + pre + lanes_below(mask)] = idx;
        running += tot;
        __syncthreads();
    }
    if (count_mode && t == 0) len[box] = running;
}
hipError_t launch_list_level(hipStream_t st, int count_mode, const GridGeom& g, const LevelGeom& L, int nboxes,
                             float ls, const float4* pxyz, int m, const uint32_t* parent_off,
                             const uint32_t* parent_len, const uint32_t* parent_list, uint32_t* len,
                             const uint32_t* off, uint32_t* list) {
    int ntx = (g.gs + 3) / 4, nty = ntx, ntz = (g.gsz + 3) / 4;
    hipLaunchKernelGGL(k_list_level, dim3(nboxes), dim3(TPB), 0, st, count_mode, g, L, ntx, nty, ntz, ls, pxyz, m,
                       parent_off, parent_len, parent_list, len, off, list);
    return hipGetLastError();
}

// Work items of the BGKI gather.  A tile whose candidate list is short is one
// item; a long list is split into one item per 2^K3_BGKI_CHUNK_LOG2-point chunk
// (each producing that chunk's partial sums), combined afterwards in chunk
// order.  Both give the same two-level association.
constexpr uint32_t BG_HEAVY_LEN = 1u << K3_BGKI_CHUNK_LOG2;
constexpr uint32_t BG_DIRECT = 0xffffffffu;

__global__ __launch_bounds__(TPB) void k_count_items(int ntiles, const uint32_t* __restrict__ toff,
                                                     const uint32_t* __restrict__ tlen,
                                                     const uint32_t* __restrict__ tlist,
                                                     uint32_t* __restrict__ heavy_cnt,
                                                     uint32_t* __restrict__ light_cnt) {
    int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= ntiles) return;
    uint32_t len = tlen[t];
    if (len > BG_HEAVY_LEN) {
        uint32_t c0 = tlist[toff[t]] >> K3_BGKI_CHUNK_LOG2;
        uint32_t c1 = tlist[toff[t] + len - 1] >> K3_BGKI_CHUNK_LOG2;
        heavy_cnt[t] = c1 - c0 + 1;
        light_cnt[t] = 0;
    } else {
        heavy_cnt[t] = 0;
        light_cnt[t] = 1;
    }
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t v) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(TPB) void k_emit_items(int ntiles, const uint32_t* __restrict__ toff,
                                                    const uint32_t* __restrict__ tlen,
                                                    const uint32_t* __restrict__ tlist,
                                                    const uint32_t* __restrict__ heavy_cnt,
                                                    const uint32_t* __restrict__ heavy_off,
                                                    const uint32_t* __restrict__ light_off,
                                                    const uint32_t* __restrict__ n_heavy, uint4* __restrict__ items,
                                                    uint32_t* __restrict__ tile_partial) {
    int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= ntiles) return;
    uint32_t len = tlen[t], o = toff[t];
    uint32_t hc = heavy_cnt[t];
    if (hc > 0) {
        uint32_t c0 = tlist[o] >> K3_BGKI_CHUNK_LOG2;
        uint32_t first = heavy_off[t];
        for (uint32_t k = 0; k < hc; ++k) {
            uint32_t c = c0 + k;
            uint32_t b = lower_bound_u32(tlist, o, o + len, c << K3_BGKI_CHUNK_LOG2);
            uint32_t e = lower_bound_u32(tlist, o, o + len, (c + 1) << K3_BGKI_CHUNK_LOG2);
            items[first + k] = make_uint4((uint32_t)t, b, e, first + k);
        }
        tile_partial[2 * t] = first;
        tile_partial[2 * t + 1] = hc;
    } else {
        items[*n_heavy + light_off[t]] = make_uint4((uint32_t)t, o, o + len, BG_DIRECT);
        tile_partial[2 * t] = 0;
        tile_partial[2 * t + 1] = 0;
    }
}

hipError_t launch_count_items(hipStream_t st, int ntiles, const uint32_t* tile_off, const uint32_t* tile_len,
                              const uint32_t* tile_list, uint32_t* heavy_cnt, uint32_t* light_cnt,
                              uint32_t* heavy_off, uint32_t* light_off, DevScalars* sc) {
    hipLaunchKernelGGL(k_count_items, dim3(blocks_for(ntiles)), dim3(TPB), 0, st, ntiles, tile_off, tile_len,
                       tile_list, heavy_cnt, light_cnt);
    hipLaunchKernelGGL(k_scan_u32_single, dim3(1), dim3(1024), 0, st, heavy_cnt, heavy_off, ntiles,
                       &sc->n_heavy_items);
    hipLaunchKernelGGL(k_scan_u32_single, dim3(1), dim3(1024), 0, st, light_cnt, light_off, ntiles, &sc->n_items);
    return hipGetLastError();
}
hipError_t launch_emit_items(hipStream_t st, int ntiles, const uint32_t* tile_off, const uint32_t* tile_len,
                             const uint32_t* tile_list, const uint32_t* heavy_cnt, const uint32_t* heavy_off,
                             const uint32_t* light_off, uint4* items, uint32_t* tile_partial, DevScalars* sc) {
    hipLaunchKernelGGL(k_emit_items, dim3(blocks_for(ntiles)), dim3(TPB), 0, st, ntiles, tile_off, tile_len,
                       tile_list, heavy_cnt, heavy_off, light_off, &sc->n_heavy_items, items, tile_partial);
    return hipGetLastError();
}

// Kernel-inference gather: one wavefront per work item, lane = voxel of the
// 4x4x4 tile.  Candidates are staged 64 at a time in LDS and visited in
// ascending point order by all lanes (uniform control flow); each lane adds
// its occupied / free terms with the chunked association.
constexpr int BG_WAVES = TPB / 64;

__global__ __launch_bounds__(TPB) void k_bgki(GridGeom g, int ntx, int nty, float ls, float ls2g, float inv_ls,
                                              float sigma, const uint4* __restrict__ items,
                                              const uint32_t* __restrict__ tlist,
                                              const K3PointRec* __restrict__ rec, float* __restrict__ del_occ,
                                              float* __restrict__ del_free, float* __restrict__ partial,
                                              DevScalars* sc) {
    __shared__ K3PointRec srec[BG_WAVES][64];
    __shared__ uint32_t sidx[BG_WAVES][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t n_items = sc->n_items + sc->n_heavy_items;
    while (true) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(&sc->item_counter, 1u);
        it = __shfl(it, 0);
        if (it >= n_items) break;
        const uint4 item = items[it];
        const int tile = (int)item.x;
        const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
        const int i = tx * 4 + (lane & 3), j = ty * 4 + ((lane >> 2) & 3), k = tz * 4 + (lane >> 4);
        const bool vvalid = i < g.gs && j < g.gs && k < g.gsz;
        const float vx = axis_center(i, g.gs, g.res, g.sx);
        const float vy = axis_center(j, g.gs, g.res, g.sy);
        const float vz = axis_center(k, g.gsz, g.res, g.sz);
        const float dc = sqrtf(fmaf(vz, vz, fmaf(vy, vy, vx * vx)));
        const float dc2 = dc * dc;
        float tot_o = 0.0f, tot_f = 0.0f, cur_o = 0.0f, cur_f = 0.0f;
        uint32_t cur_chunk = 0xffffffffu;
        for (uint32_t b = item.y; b < item.z; b += 64) {
            uint32_t e = b + lane;
            if (e < item.z) {
                uint32_t p = tlist[e];
                sidx[w][lane] = p;
                srec[w][lane] = rec[p];
            }
            wave_sync();
            const int cnt = (int)min(64u, item.z - b);
            for (int q = 0; q < cnt; ++q) {
                const uint32_t p = sidx[w][q];
                const uint32_t chunk = p >> K3_BGKI_CHUNK_LOG2;
                if (chunk != cur_chunk) {
                    tot_o += cur_o;
                    tot_f += cur_f;
                    cur_o = 0.0f;
                    cur_f = 0.0f;
                    cur_chunk = chunk;
                }
                const K3PointRec r = srec[w][q];
                float d_o = k3_occ_distance(vx, vy, vz, r, ls, ls2g);
                if (d_o >= 0.0f) cur_o += k3_kernel_value(d_o, inv_ls, sigma);
                float d_f = k3_free_distance(vx, vy, vz, dc, dc2, r, ls, ls2g);
                if (d_f >= 0.0f) cur_f += k3_kernel_value(d_f, inv_ls, sigma);
            }
            wave_sync();
        }
        tot_o += cur_o;
        tot_f += cur_f;
        if (item.w == BG_DIRECT) {
            if (vvalid) {
                int cell = i + g.gs * j + g.gs * g.gs * k;
                del_occ[cell] = tot_o;
                del_free[cell] = tot_f;
            }
        } else {
            partial[(size_t)item.w * 128 + lane] = tot_o;
            partial[(size_t)item.w * 128 + 64 + lane] = tot_f;
        }
    }
}
hipError_t launch_bgki(hipStream_t st, int grid_blocks, const GridGeom& g, int ntx, int nty, float ls, float sigma,
                       const uint4* items, const uint32_t* tile_list, const K3PointRec* rec, float* del_occ,
                       float* del_free, float* partial, DevScalars* sc) {
    float ls2g = k3_ls2_guard(ls);
    float inv_ls = 1.0f / ls;
    hipLaunchKernelGGL(k_bgki, dim3(grid_blocks), dim3(TPB), 0, st, g, ntx, nty, ls, ls2g, inv_ls, sigma, items,
                       tile_list, rec, del_occ, del_free, partial, sc);
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_bgki_combine(int ntiles, GridGeom g, int ntx, int nty,
                                                      const uint32_t* __restrict__ tile_partial,
                                                      const float* __restrict__ partial, float* del_occ,
                                                      float* del_free) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tile = blockIdx.x * BG_WAVES + w;
    if (tile >= ntiles) return;
    const uint32_t first = tile_partial[2 * tile], cnt = tile_partial[2 * tile + 1];
    if (cnt == 0) return;
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int i = tx * 4 + (lane & 3), j = ty * 4 + ((lane >> 2) & 3), k = tz * 4 + (lane >> 4);
    if (!(i < g.gs && j < g.gs && k < g.gsz)) return;
    float ro = 0.0f, rf = 0.0f;
    for (uint32_t c = 0; c < cnt; ++c) {
        ro += partial[(size_t)(first + c) * 128 + lane];
        rf += partial[(size_t)(first + c) * 128 + 64 + lane];
    }
    int cell = i + g.gs * j + g.gs * g.gs * k;
    del_occ[cell] = ro;
    del_free[cell] = rf;
}
hipError_t launch_bgki_combine(hipStream_t st, int ntiles, const GridGeom& g, int ntx, int nty,
                               const uint32_t* tile_partial, const float* partial, float* del_occ, float* del_free) {
    hipLaunchKernelGGL(k_bgki_combine, dim3(blocks_for(ntiles, BG_WAVES)), dim3(TPB), 0, st, ntiles, g, ntx, nty,
                       tile_partial, partial, del_occ, del_free);
    return hipGetLastError();
}

// ============================================================== mass update
__device__ __forceinline__ void atomic_max_nonneg(uint32_t* dst, float v) {
    if (v > 0.0f && v - v == 0.0f) atomicMax(dst, k3_f2u(v));
}

__global__ __launch_bounds__(TPB) void k_mass_update(VoxelArrays v, int V, const float* __restrict__ weights,
                                                     const float* __restrict__ del_occ,
                                                     const float* __restrict__ del_free, float decay,
                                                     float mass_scale, float p_B, float* __restrict__ born,
                                                     DevScalars* sc) {
    int c = blockIdx.x * TPB + threadIdx.x;
    float bmax = 0.0f;
    if (c < V) {
        float f = v.f[c], s = v.s[c], d = v.d[c];
        int st = v.start[c], en = v.end[c];
        float pred_total = (f + d) * decay;
        float dyn_pred = 0.0f;
        if (st != -1)
            for (int j = st; j < en + 1; ++j) dyn_pred += weights[j];
        k3_predict_mass(&f, &s, &d, dyn_pred, pred_total, decay);
        k3_update_mass(&f, &s, &d, del_occ[c], del_free[c], mass_scale);
        float b = k3_newborn_part(dyn_pred, pred_total, d, p_B);
        v.f[c] = f;
        v.s[c] = s;
        v.d[c] = d;
        born[c] = b;
        v.pers[c] = d - b;
        v.pred[c] = dyn_pred;
        if (b > 0.0f && b - b == 0.0f) bmax = b;
    }
    // wave-level max before one atomic per wave
    for (int o = 32; o > 0; o >>= 1) bmax = fmaxf(bmax, __shfl_xor(bmax, o));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&sc->born_max_bits, bmax);
}
hipError_t launch_mass_update(hipStream_t st, VoxelArrays v, int V, const float* weights, const float* del_occ,
                              const float* del_free, float decay, float mass_scale, float p_B, float* born,
                              DevScalars* sc) {
    hipLaunchKernelGGL(k_mass_update, dim3(blocks_for(V)), dim3(TPB), 0, st, v, V, weights, del_occ, del_free,
                       decay, mass_scale, p_B, born, sc);
    return hipGetLastError();
}

// Per voxel with particles: persistent weight update (mu = pers/pred), mean
// velocity from the updated weights, then the static-particle penalty.
__global__ __launch_bounds__(TPB) void k_voxel_particles(VoxelArrays v, int V, ParticleArrays P, float min_vel,
                                                         DevScalars* sc) {
    int c = blockIdx.x * TPB + threadIdx.x;
    float wmax = 0.0f;
    if (c < V) {
        int st = v.start[c];
        if (st == -1) {
            v.mvx[c] = 0.0f;
            v.mvy[c] = 0.0f;
            v.mvz[c] = 0.0f;
        } else {
            int en = v.end[c];
            float rho = v.pers[c], pr = v.pred[c];
            float m = pr > 0.0f ? rho / pr : 0.0f;
            v.mu[c] = m;
            const float pA = 0.0f, lik = 1.0f;
            const float min2 = min_vel * min_vel;
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            for (int j = st; j <= en; ++j) {
                float w = P.w[j];
                float post = lik * w;
                float wn = pA * m * post + (1.0f - pA) * m * w;
                float vx = P.vx[j], vy = P.vy[j], vz = P.vz[j];
                sx += wn * vx;
                sy += wn * vy;
                sz += wn * vz;
                if (vx * vx + vy * vy + vz * vz < min2) wn = 0.0f;
                P.w[j] = wn;
                if (wn > wmax && wn - wn == 0.0f) wmax = wn;
            }
            v.mvx[c] = rho > 0.0f ? sx / rho : 0.0f;
            v.mvy[c] = rho > 0.0f ? sy / rho : 0.0f;
            v.mvz[c] = rho > 0.0f ? sz / rho : 0.0f;
        }
    }
    for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&sc->joint_max_bits, wmax);
}
hipError_t launch_voxel_particles(hipStream_t st, VoxelArrays v, int V, ParticleArrays P, float min_vel,
                                  DevScalars* sc) {
    hipLaunchKernelGGL(k_voxel_particles, dim3(blocks_for(V)), dim3(TPB), 0, st, v, V, P, min_vel, sc);
    return hipGetLastError();
}

// ============================================================== fixed-point CDF
constexpr int FS_ITEMS = 8;
constexpr int FS_TILE = TPB * FS_ITEMS;
int fixed_scan_blocks(int n) { return n <= 0 ? 1 : (n + FS_TILE - 1) / FS_TILE; }

__device__ __forceinline__ float cat_load(const float* a, int na, const float* b, int nb, int i) {
    return i < na ? a[i] : (i < na + nb ? b[i - na] : 0.0f);
}

__global__ void k_fixed_exponent(const uint32_t* max_bits, int count, int32_t* E) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *E = k3_fixed_exponent(k3_u2f(*max_bits), (uint32_t)count);
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

__global__ __launch_bounds__(TPB) void k_fixed_reduce(const float* __restrict__ a, int na,
                                                      const float* __restrict__ b, int nb,
                                                      const int32_t* __restrict__ Ep,
                                                      uint64_t* __restrict__ blocksums) {
    __shared__ uint64_t lds4[4];
    const int E = *Ep;
    const int base = blockIdx.x * FS_TILE + threadIdx.x * FS_ITEMS;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < FS_ITEMS; ++k) s += k3_to_fixed(cat_load(a, na, b, nb, base + k), E);
    uint64_t tot;
    block_scan_incl_u64(s, lds4, &tot);
    if (threadIdx.x == 0) blocksums[blockIdx.x] = tot;
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
                                                    const uint64_t* __restrict__ blocksums,
                                                    uint64_t* __restrict__ cdf) {
    __shared__ uint64_t lds4[4];
    const int E = *Ep;
    const int n = na + nb;
    const int base = blockIdx.x * FS_TILE + threadIdx.x * FS_ITEMS;
    uint64_t q[FS_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < FS_ITEMS; ++k) {
        q[k] = k3_to_fixed(cat_load(a, na, b, nb, base + k), E);
        s += q[k];
    }
    uint64_t tot;
    uint64_t incl = block_scan_incl_u64(s, lds4, &tot);
    uint64_t run = blocksums[blockIdx.x] + incl - s;
#pragma unroll
    for (int k = 0; k < FS_ITEMS; ++k) {
        run += q[k];
        if (base + k < n) cdf[base + k] = run;
    }
}

hipError_t launch_fixed_scan(hipStream_t st, const float* a, int na, const float* b, int nb,
                             const uint32_t* max_bits, int32_t* E_out, uint64_t* cdf, uint64_t* blocksums,
                             uint64_t* total) {
    const int n = na + nb;
    const int nblk = fixed_scan_blocks(n);
    hipLaunchKernelGGL(k_fixed_exponent, dim3(1), dim3(64), 0, st, max_bits, n, E_out);
    hipLaunchKernelGGL(k_fixed_reduce, dim3(nblk), dim3(TPB), 0, st, a, na, b, nb, E_out, blocksums);
    hipLaunchKernelGGL(k_scan_top_u64, dim3(1), dim3(1024), 0, st, blocksums, nblk, total);
    hipLaunchKernelGGL(k_fixed_down, dim3(nblk), dim3(TPB), 0, st, a, na, b, nb, E_out, blocksums, cdf);
    return hipGetLastError();
}

__device__ __forceinline__ uint32_t slot_bound(uint64_t c, uint64_t tot, uint32_t cnt, double scale) {
    if (c == tot) return cnt;
    uint64_t b = (uint64_t)((double)c * scale);
    return b < (uint64_t)cnt ? (uint32_t)b : cnt;
}

// Slot boundaries of a fixed-point CDF split into cnt particles, and the
// unassociated newborn weight of each cell (src/common/kernel/init_new_particles.cu:119-134).
__global__ __launch_bounds__(TPB) void k_bounds(const uint64_t* __restrict__ cdf,
                                                const uint64_t* __restrict__ total, int V, uint32_t cnt,
                                                const float* __restrict__ born, uint32_t* __restrict__ bnd,
                                                float* __restrict__ wua) {
    int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= V) return;
    const uint64_t tot = *total;
    uint32_t bj, bp;
    if (tot == 0) {
        bj = (uint32_t)(((uint64_t)(j + 1) * cnt) / (uint64_t)V);
        bp = (uint32_t)(((uint64_t)j * cnt) / (uint64_t)V);
    } else {
        const double scale = (double)cnt / (double)tot;
        bj = slot_bound(cdf[j], tot, cnt, scale);
        bp = j ? slot_bound(cdf[j - 1], tot, cnt, scale) : 0u;
    }
    bnd[j] = bj;
    if (wua) {
        const float pA = 0.0f;
        int nn = (int)(bj - bp);
        wua[j] = nn > 0 ? (float)(((1.0 - (double)pA) * (double)born[j]) / (double)nn) : 0.0f;
    }
}
hipError_t launch_bounds(hipStream_t st, const uint64_t* cdf, const uint64_t* total, int V, uint32_t cnt,
                         const float* born, uint32_t* bnd, float* wua) {
    hipLaunchKernelGGL(k_bounds, dim3(blocks_for(V)), dim3(TPB), 0, st, cdf, total, V, cnt, born, bnd, wua);
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

__global__ __launch_bounds__(TPB) void k_newborn(const uint32_t* __restrict__ bnd, int V, int B, int gs,
                                                 const float* __restrict__ wua, ParticleArrays NB, float vmax,
                                                 float min_vel, uint32_t frame, uint32_t k0, uint32_t k1,
                                                 DevScalars* sc) {
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
        if (w > 0.0f && w - w == 0.0f) wmax = w;
    }
    for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&sc->joint_max_bits, wmax);
}
hipError_t launch_newborn(hipStream_t st, const uint32_t* bnd, int V, int B, int gs, const float* wua,
                          ParticleArrays NB, float vmax, float min_vel, uint32_t frame, uint32_t k0, uint32_t k1,
                          DevScalars* sc) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_newborn, dim3(blocks_for(B)), dim3(TPB), 0, st, bnd, V, B, gs, wua, NB, vmax, min_vel,
                       frame, k0, k1, sc);
    return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_init_particles(const uint32_t* __restrict__ bnd, int V, int N, int gs,
                                                        ParticleArrays P, float vmax, float min_vel,
                                                        uint32_t frame, uint32_t k0, uint32_t k1) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= N) return;
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
hipError_t launch_init_particles(hipStream_t st, const uint32_t* bnd, int V, int N, int gs, ParticleArrays P,
                                 float vmax, float min_vel, uint32_t frame, uint32_t k0, uint32_t k1) {
    hipLaunchKernelGGL(k_init_particles, dim3(blocks_for(N)), dim3(TPB), 0, st, bnd, V, N, gs, P, vmax, min_vel,
                       frame, k0, k1);
    return hipGetLastError();
}

// Systematic resampling on the u64 CDF of [persistent ++ newborn] weights:
// positions T_i = trunc((i + u) * Q / N) (IEEE double), index = first CDF
// entry > T_i; then the gather into the next particle set with weight W/N.
__global__ __launch_bounds__(TPB) void k_resample(const uint64_t* __restrict__ cdf, int L,
                                                  const uint64_t* __restrict__ total,
                                                  const int32_t* __restrict__ Ep, ParticleArrays S,
                                                  ParticleArrays NB, ParticleArrays O, int N, uint32_t frame,
                                                  uint32_t k0, uint32_t k1, int32_t* __restrict__ idx_out) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= N) return;
    const uint64_t Q = *total;
    const int E = *Ep;
    const float W = (float)((double)Q * k3_pow2d(-E));
    const float nw = W / (float)N;
    int idx;
    if (Q == 0) {
        idx = i;
    } else {
        uint32_t r[4];
        k3_philox4x32_10(0u, frame, K3_STREAM_RESAMPLE, 0u, k0, k1, r);
        const double u = (double)r[0] * 0x1p-32;
        const double step = (double)Q / (double)N;
        uint64_t T = (uint64_t)(((double)i + u) * step);
        if (T > Q - 1) T = Q - 1;
        int lo = 0, hi = L;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (cdf[mid] <= T) lo = mid + 1;
            else hi = mid;
        }
        idx = lo;
    }
    idx_out[i] = idx;
    if (idx < N) {
        O.px[i] = S.px[idx]; O.py[i] = S.py[idx]; O.pz[i] = S.pz[idx];
        O.vx[i] = S.vx[idx]; O.vy[i] = S.vy[idx]; O.vz[i] = S.vz[idx];
        O.key[i] = S.key[idx];
    } else {
        int b = idx - N;
        O.px[i] = NB.px[b]; O.py[i] = NB.py[b]; O.pz[i] = NB.pz[b];
        O.vx[i] = NB.vx[b]; O.vy[i] = NB.vy[b]; O.vz[i] = NB.vz[b];
        O.key[i] = NB.key[b];
    }
    O.w[i] = nw;
}
hipError_t launch_resample(hipStream_t st, const uint64_t* cdf, int total_len, const uint64_t* total,
                           const int32_t* E, ParticleArrays sorted, ParticleArrays nb, ParticleArrays out, int N,
                           uint32_t frame, uint32_t k0, uint32_t k1, int32_t* idx_out) {
    hipLaunchKernelGGL(k_resample, dim3(blocks_for(N)), dim3(TPB), 0, st, cdf, total_len, total, E, sorted, nb, out,
                       N, frame, k0, k1, idx_out);
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

}  // namespace k3
