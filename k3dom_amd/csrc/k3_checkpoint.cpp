// Checkpoint and resume of a map's filter state (no counterpart in the
// reference, whose state lives only in device memory for the life of the
// process).
//
// The state after a frame is: the parameters and method, the frame counter
// (the counter word of every random draw, so the random streams continue),
// the time of the last frame, the pose (sensor and map centre), the voxel
// arrays and the current particle set.  Everything else a frame reads is
// rebuilt by the frame itself.  A pending map shift is applied before the
// state is written, so a checkpoint never holds one.
//
// Layout (little-endian), version 1:
//   header (K3CkHeader below, 128 bytes): magic "K3DOMCK1", version, header
//     size, sizeof(k3dom_params), method, V, N, frame, first_pose,
//     updated_time, sensor[3], center[3], payload size, payload checksum
//   k3dom_params (sizeof(k3dom_params) bytes)
//   payload: voxel arrays f, s, d, pers, pred, mu, wua, mvx, mvy, mvz
//     (V float each), start, end (V int32 each); particles px, py, pz, vx, vy,
//     vz, w (N float each), key (N int32)
// The checksum is FNV-1a over the payload's 64-bit little-endian words
// (payload sizes are multiples of 8 bytes after padding to 8).
#include <cstring>

#include "k3_internal.h"

using namespace k3;
using namespace k3api;

namespace {

constexpr char kMagic[8] = {'K', '3', 'D', 'O', 'M', 'C', 'K', '1'};
constexpr uint32_t kVersion = 1;

struct K3CkHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint32_t params_bytes;
    int32_t method;
    int64_t V, N;
    uint32_t frame;
    uint32_t first_pose;
    double updated_time;
    float sensor[3], center[3];
    uint64_t payload_bytes;
    uint64_t checksum;
    uint8_t reserved[128 - 96];
};
static_assert(sizeof(K3CkHeader) == 128, "checkpoint header layout");

constexpr int kVoxelFloats = 10, kVoxelInts = 2, kParticleFloats = 7, kParticleInts = 1;

uint64_t payload_bytes(int64_t V, int64_t N) {
    const uint64_t raw = 4ull * ((uint64_t)V * (kVoxelFloats + kVoxelInts) + (uint64_t)N * (kParticleFloats + kParticleInts));
    return (raw + 7) & ~7ull;
}

uint64_t fnv1a_words(const uint8_t* p, uint64_t bytes) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint64_t i = 0; i < bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h ^= w;
        h *= 0x100000001b3ull;
    }
    return h;
}

// Device arrays of the state in payload order.
void state_arrays(k3dom_handle* h, int vbuf, void* ptrs[20]) {
    VoxelArrays& v = h->vox[vbuf];
    ParticleArrays& P = h->P[0];
    void* a[20] = {v.f, v.s, v.d, v.pers, v.pred, v.mu, v.wua, v.mvx, v.mvy, v.mvz, v.start, v.end,
                   P.px, P.py, P.pz, P.vx, P.vy, P.vz, P.w, P.key};
    std::memcpy(ptrs, a, sizeof(a));
}

}  // namespace

k3dom_status k3dom_checkpoint_size(k3dom_handle* h, int64_t* bytes) {
    if (!h || !bytes) return K3DOM_ERR_INVALID_ARG;
    *bytes = (int64_t)(sizeof(K3CkHeader) + sizeof(k3dom_params) + payload_bytes(h->V, h->N));
    return K3DOM_OK;
}

k3dom_status k3dom_checkpoint_save(k3dom_handle* h, void* dst, int64_t capacity, int64_t* written) {
    if (!h || !dst || !written) return K3DOM_ERR_INVALID_ARG;
    int64_t need = 0;
    k3dom_checkpoint_size(h, &need);
    if (capacity < need) return fail(h, K3DOM_ERR_INVALID_ARG, "checkpoint_save: capacity below checkpoint_size");
    {
        k3dom_status st = settle_shift(h);
        if (st == K3DOM_OK) st = settle_particles(h);
        if (st != K3DOM_OK) return st;
    }
    uint8_t* out = static_cast<uint8_t*>(dst);
    uint8_t* payload = out + sizeof(K3CkHeader) + sizeof(k3dom_params);
    void* arrays[20];
    state_arrays(h, h->vcur, arrays);
    uint64_t off = 0;
    for (int k = 0; k < 20; ++k) {
        const size_t n = (size_t)(k < 12 ? h->V : h->N) * 4;
        HIP_TRY(h, hipMemcpyAsync(payload + off, arrays[k], n, hipMemcpyDeviceToHost, h->stream));
        off += n;
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    const uint64_t pb = payload_bytes(h->V, h->N);
    std::memset(payload + off, 0, pb - off);
    K3CkHeader hd{};
    std::memcpy(hd.magic, kMagic, 8);
    hd.version = kVersion;
    hd.header_bytes = sizeof(K3CkHeader);
    hd.params_bytes = sizeof(k3dom_params);
    hd.method = h->method;
    hd.V = h->V;
    hd.N = h->N;
    hd.frame = h->frame;
    hd.first_pose = h->first_pose ? 1u : 0u;
    hd.updated_time = h->updated_time;
    for (int i = 0; i < 3; ++i) {
        hd.sensor[i] = h->sensor[i];
        hd.center[i] = h->center[i];
    }
    hd.payload_bytes = pb;
    hd.checksum = fnv1a_words(payload, pb);
    std::memcpy(out, &hd, sizeof(hd));
    std::memcpy(out + sizeof(hd), &h->p, sizeof(k3dom_params));
    *written = need;
    return K3DOM_OK;
}

k3dom_status k3dom_checkpoint_load(const void* src, int64_t size, k3dom_handle** out) {
    if (!src || !out) return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: NULL argument");
    *out = nullptr;
    const uint8_t* in = static_cast<const uint8_t*>(src);
    K3CkHeader hd;
    if (size < (int64_t)(sizeof(hd) + sizeof(k3dom_params)))
        return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: buffer shorter than a header");
    std::memcpy(&hd, in, sizeof(hd));
    if (std::memcmp(hd.magic, kMagic, 8) != 0) return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: bad magic");
    if (hd.version != kVersion || hd.header_bytes != sizeof(hd) || hd.params_bytes != sizeof(k3dom_params))
        return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: unsupported version or layout");
    k3dom_params p;
    std::memcpy(&p, in + sizeof(hd), sizeof(p));
    const uint64_t pb = hd.V > 0 && hd.N > 0 ? payload_bytes(hd.V, hd.N) : 0;
    if (pb == 0 || hd.payload_bytes != pb || (uint64_t)size != sizeof(hd) + sizeof(p) + pb)
        return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: size does not match the header");
    const uint8_t* payload = in + sizeof(hd) + sizeof(p);
    if (fnv1a_words(payload, pb) != hd.checksum)
        return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: checksum mismatch");
    k3dom_handle* h = nullptr;
    k3dom_status st = k3dom_create_method(&p, hd.method, &h);
    if (st != K3DOM_OK) return st;
    if ((int64_t)h->V != hd.V || (int64_t)h->N != hd.N) {
        k3dom_destroy(h);
        return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: sizes do not follow from the parameters");
    }
    void* arrays[20];
    state_arrays(h, 0, arrays);
    uint64_t off = 0;
    hipError_t e = hipSuccess;
    for (int k = 0; k < 20 && e == hipSuccess; ++k) {
        const size_t n = (size_t)(k < 12 ? h->V : h->N) * 4;
        e = hipMemcpyAsync(arrays[k], payload + off, n, hipMemcpyHostToDevice, h->stream);
        off += n;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        k3dom_destroy(h);
        return fail(nullptr, K3DOM_ERR_HIP, std::string("checkpoint_load: ") + hipGetErrorString(e));
    }
    h->vcur = 0;
    h->shift_pending = false;
    h->pshift_pending = false;
    h->frame = hd.frame;
    h->first_pose = hd.first_pose != 0;
    h->updated_time = hd.updated_time;
    for (int i = 0; i < 3; ++i) {
        h->sensor[i] = hd.sensor[i];
        h->center[i] = hd.center[i];
    }
    *out = h;
    return K3DOM_OK;
}
