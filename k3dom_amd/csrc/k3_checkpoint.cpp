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
// Record layout and validation: k3_ckformat.h / k3_ckformat.cpp.
#include <cstring>

#include "k3_ckformat.h"
#include "k3_internal.h"

using namespace k3;
using namespace k3api;

namespace {

using k3ck::Header;
using k3ck::payload_bytes;
using k3ck::fnv1a_words;

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
    *bytes = (int64_t)(sizeof(Header) + sizeof(k3dom_params) + payload_bytes(h->V, h->N));
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
    uint8_t* payload = out + sizeof(Header) + sizeof(k3dom_params);
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
    const Header hd = k3ck::make_header(h->method, h->V, h->N, h->frame, h->first_pose, h->updated_time,
                                        h->sensor, h->center, payload);
    std::memcpy(out, &hd, sizeof(hd));
    std::memcpy(out + sizeof(hd), &h->p, sizeof(k3dom_params));
    *written = need;
    return K3DOM_OK;
}

k3dom_status k3dom_checkpoint_load(const void* src, int64_t size, k3dom_handle** out) {
    if (!src || !out) return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: NULL argument");
    *out = nullptr;
    Header hd;
    k3dom_params p;
    const uint8_t* payload = nullptr;
    std::string why;
    if (!k3ck::check(src, size, &hd, &p, &payload, &why))
        return fail(nullptr, K3DOM_ERR_INVALID_ARG, "checkpoint_load: " + why);
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
