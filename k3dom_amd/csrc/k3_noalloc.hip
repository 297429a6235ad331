// Allocation-free MapPoint export (k3dom_export_mappoints_noalloc): the
// records of k3dom_export_mappoints written to a caller's device buffer, with
// the visible-voxel compaction done in one pass over a fixed device scratch
// array instead of the flag, offset and scan buffers a handle allocates.
//
// Compaction: one block per tile of NA_TILE voxels, tiles taken in the order
// blocks start (a ticket), so a block waits only for tiles of blocks already
// running.  Each block tests its voxels (voxel base + k * 256 + thread, k <
// 16: coalesced loads), counts them per (k, wavefront) and publishes its tile
// count; wavefront 0 then walks back over the preceding tiles' status words,
// 64 at a time, to the nearest published inclusive prefix (decoupled
// look-back), publishes its own and every visible voxel writes its record at
// prefix + its rank in index order.  The last tile writes the total.
#include "k3_internal.h"
#include "k3_mappoint.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

using namespace k3;

namespace {

constexpr int NA_TPB = 256;
constexpr int NA_ITEMS = 16;
constexpr int NA_TILE = NA_TPB * NA_ITEMS;
constexpr long long NA_MAX_TILES = K3DOM_NOALLOC_MAX_VOXELS / NA_TILE;
constexpr uint64_t NA_SUM = 1ull << 32, NA_PREFIX = 2ull << 32;
constexpr int NA_MAX_DEVICES = 64;
static_assert(K3DOM_NOALLOC_MAX_VOXELS % NA_TILE == 0, "whole tiles");

}  // namespace

// The library's scratch of this export: [0] ticket (low word), [1] record
// count, [2 + t] status of tile t ((flag << 32) | value; flag 1 tile count,
// 2 inclusive prefix, 0 not yet written).  One copy per GPU, at its own
// address; host code uses the address from hipGetSymbolAddress, never this name.
static __device__ __align__(256) uint64_t scratch[2 + NA_MAX_TILES];

__global__ __launch_bounds__(NA_TPB) void k_na_clear(uint64_t* __restrict__ scr, uint32_t words) {
    for (uint32_t i = blockIdx.x * NA_TPB + threadIdx.x; i < words; i += gridDim.x * NA_TPB) scr[i] = 0ull;
}

__global__ __launch_bounds__(NA_TPB) void k_na_compact(VoxelArrays v, int V, MapPointArgs a, uint64_t* scr,
                                                     float4* __restrict__ out, long long cap) {
    __shared__ uint32_t s_tile, s_excl, s_tot;
    __shared__ uint32_t wcnt[NA_ITEMS][NA_TPB / 64];
    __shared__ uint32_t wpre[NA_ITEMS][NA_TPB / 64];
    uint32_t* ticket = reinterpret_cast<uint32_t*>(scr);
    uint64_t* status = scr + 2;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t ntiles = gridDim.x;
    const long long base = (long long)tile * NA_TILE;
    uint32_t vis = 0;  // bit k: voxel base + k * NA_TPB + t is visible
#pragma unroll
    for (int k = 0; k < NA_ITEMS; ++k) {
        const long long c = base + (long long)k * NA_TPB + t;
        const bool in = c < V && mp_visible(mp_cell(v, (int)c, a), a);
        vis |= (uint32_t)in << k;
        const uint64_t b = __ballot(in);
        if (lane == 0) wcnt[k][w] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (int k = 0; k < NA_ITEMS; ++k)
            for (int ww = 0; ww < NA_TPB / 64; ++ww) {
                wpre[k][ww] = run;
                run += wcnt[k][ww];
            }
        s_tot = run;
        __hip_atomic_store(&status[tile], (tile == 0 ? NA_PREFIX : NA_SUM) | run, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t tot = s_tot;
    if (t < 64) {
        uint32_t excl = 0;
        for (int j = (int)tile - 1; j >= 0; j -= 64) {
            const int jj = j - lane;
            uint64_t st = NA_PREFIX;  // before tile 0: an empty prefix
            if (jj >= 0) {
                st = __hip_atomic_load(&status[jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                // tile jj was taken before this one, so its count is on its way
                while ((st >> 32) == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    st = __hip_atomic_load(&status[jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            const uint64_t pm = __ballot((st >> 32) == 2);
            const int stop = pm ? __builtin_ctzll(pm) : 64;
            uint32_t add = lane <= stop ? (uint32_t)st : 0u;
            for (int o = 32; o > 0; o >>= 1) add += __shfl_xor(add, o);
            excl += add;
            if (pm) break;
        }
        if (t == 0) {
            if (tile > 0)
                __hip_atomic_store(&status[tile], NA_PREFIX | (excl + tot), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            s_excl = excl;
            if (tile == ntiles - 1) scr[1] = (uint64_t)(excl + tot);
        }
    }
    __syncthreads();
    const uint32_t excl = s_excl;
#pragma unroll
    for (int k = 0; k < NA_ITEMS; ++k) {
        const bool in = (vis >> k) & 1u;
        const uint64_t b = __ballot(in);
        if (in) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
            const long long slot = (long long)excl + wpre[k][w] + below;
            const int c = (int)(base + (long long)k * NA_TPB + t);
            if (slot < cap) mp_store(out, (size_t)slot, c, mp_cell(v, c, a), a);
        }
    }
}

__global__ __launch_bounds__(NA_TPB) void k_na_fill_all(VoxelArrays v, int V, MapPointArgs a,
                                                      float4* __restrict__ out) {
    const int c = blockIdx.x * NA_TPB + threadIdx.x;
    if (c < V) mp_store(out, (size_t)c, c, mp_cell(v, c, a), a);
}

namespace {

std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[NA_MAX_DEVICES] = {};

// Device address of the scratch array on a GPU, looked up once per device
// with that device current (the lookup loads the code object that holds the
// array, which then stays resident until the process ends).
hipError_t scratch_lookup(int device, uint64_t** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= NA_MAX_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = static_cast<uint64_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

namespace k3api {
// Set-up at map creation, so that the export itself never triggers the load.
bool noalloc_prepare(int device) {
    uint64_t* p = nullptr;
    return scratch_lookup(device, &p) == hipSuccess;
}
}  // namespace k3api

extern "C" k3dom_status k3dom_export_mappoints_noalloc(k3dom_handle* h, int32_t mode, float min_occ,
                                                      float min_mass, k3dom_mappoint* d_out, int64_t capacity,
                                                      int64_t* count) {
    if (!h || !count || (mode != K3DOM_MAPPOINTS_ALL && mode != K3DOM_MAPPOINTS_VISIBLE) || capacity < 0 ||
        (!d_out && capacity > 0))
        return K3DOM_ERR_INVALID_ARG;
    if ((long long)h->V > K3DOM_NOALLOC_MAX_VOXELS)
        return k3api::fail(h, K3DOM_ERR_INVALID_ARG,
                           "export_mappoints_noalloc: the map has " + std::to_string(h->V) +
                               " voxels, above K3DOM_NOALLOC_MAX_VOXELS (" +
                               std::to_string(K3DOM_NOALLOC_MAX_VOXELS) + ") of the fixed scratch");
    if (d_out) {
        // records are written by kernels: d_out must be device-accessible memory
        hipPointerAttribute_t at{};
        if (hipPointerGetAttributes(&at, d_out) != hipSuccess || at.devicePointer == nullptr) {
            (void)hipGetLastError();
            return k3api::fail(h, K3DOM_ERR_INVALID_ARG, "export_mappoints_noalloc: d_out is not device memory");
        }
    }
    {
        k3dom_status st = k3api::settle_shift(h);
        if (st != K3DOM_OK) return st;
    }
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
    float4* out = reinterpret_cast<float4*>(d_out);
    static_assert(sizeof(k3dom_mappoint) == 3 * sizeof(float4), "MapPoint record layout");
    const unsigned vblocks = (unsigned)((h->V + NA_TPB - 1) / NA_TPB);
    if (mode == K3DOM_MAPPOINTS_ALL) {
        *count = h->V;
        if (!d_out) return K3DOM_OK;
        if (capacity < h->V)
            return k3api::fail(h, K3DOM_ERR_INVALID_ARG, "export_mappoints_noalloc: capacity below the record count");
        hipLaunchKernelGGL(k_na_fill_all, dim3(vblocks), dim3(NA_TPB), 0, h->stream, h->vox[h->vcur], h->V, a, out);
        HIP_TRY(h, hipGetLastError());
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        return K3DOM_OK;
    }
    // (looked up when the map was created; again here if that failed)
    uint64_t* scr = nullptr;
    HIP_TRY(h, scratch_lookup(h->device, &scr));
    const uint32_t ntiles = (uint32_t)std::max<long long>(1, ((long long)h->V + NA_TILE - 1) / NA_TILE);
    const uint32_t words = 2 + ntiles;
    // the ticket, the count and this call's tile statuses start at zero
    hipLaunchKernelGGL(k_na_clear, dim3(std::min<uint32_t>(1024u, (words + NA_TPB - 1) / NA_TPB)), dim3(NA_TPB), 0,
                       h->stream, scr, words);
    HIP_TRY(h, hipGetLastError());
    hipLaunchKernelGGL(k_na_compact, dim3(ntiles), dim3(NA_TPB), 0, h->stream, h->vox[h->vcur], h->V, a, scr, out,
                       (long long)(d_out ? capacity : 0));
    HIP_TRY(h, hipGetLastError());
    // the count (below 2^32: V is limited above) into the handle's pinned scalars
    HIP_TRY(h, hipMemcpyAsync(&h->h_sc->mp_total, scr + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    const int64_t n = (int64_t)h->h_sc->mp_total;
    *count = n;
    if (d_out && capacity < n)
        return k3api::fail(h, K3DOM_ERR_INVALID_ARG,
                           "export_mappoints_noalloc: capacity below the record count (the first capacity records "
                           "were written)");
    return K3DOM_OK;
}
