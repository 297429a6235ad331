// Checkpoint record, version 1 (little-endian):
//   header (k3ck::Header, 128 bytes): magic "K3DOMCK1", version, header
//     size, sizeof(k3dom_params), method, V, N, frame, first_pose,
//     updated_time, sensor[3], center[3], payload size, payload checksum
//   k3dom_params (sizeof(k3dom_params) bytes)
//   payload: voxel arrays f, s, d, pers, pred, mu, wua, mvx, mvy, mvz
//     (V float each), start, end (V int32 each); particles px, py, pz, vx, vy,
//     vz, w (N float each), key (N int32); zero padding to a multiple of 8
// The checksum is FNV-1a over the payload's 64-bit little-endian words.
#include "k3_ckformat.h"

#include <cstring>

namespace k3ck {

namespace {
constexpr char kMagic[8] = {'K', '3', 'D', 'O', 'M', 'C', 'K', '1'};
constexpr uint32_t kVersion = 1;
}  // namespace

uint64_t payload_bytes(int64_t V, int64_t N) {
    const uint64_t raw = 4ull * ((uint64_t)V * (kVoxelFloats + kVoxelInts) + (uint64_t)N * (kParticleFloats + kParticleInts));
    return (raw + 7) & ~7ull;
}

uint64_t fnv1a_words(const uint8_t* p, uint64_t bytes) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint64_t i = 0; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h ^= w;
        h *= 0x100000001b3ull;
    }
    return h;
}

Header make_header(int32_t method, int64_t V, int64_t N, uint32_t frame, bool first_pose, double updated_time,
                   const float sensor[3], const float center[3], const uint8_t* payload) {
    Header hd{};
    std::memcpy(hd.magic, kMagic, 8);
    hd.version = kVersion;
    hd.header_bytes = sizeof(Header);
    hd.params_bytes = sizeof(k3dom_params);
    hd.method = method;
    hd.V = V;
    hd.N = N;
    hd.frame = frame;
    hd.first_pose = first_pose ? 1u : 0u;
    hd.updated_time = updated_time;
    for (int i = 0; i < 3; ++i) {
        hd.sensor[i] = sensor[i];
        hd.center[i] = center[i];
    }
    hd.payload_bytes = payload_bytes(V, N);
    hd.checksum = fnv1a_words(payload, hd.payload_bytes);
    return hd;
}

bool check(const void* src, int64_t size, Header* hd, k3dom_params* params, const uint8_t** payload,
           std::string* why) {
    const uint8_t* in = static_cast<const uint8_t*>(src);
    if (!src || size < (int64_t)(sizeof(Header) + sizeof(k3dom_params))) {
        *why = "buffer shorter than a header";
        return false;
    }
    std::memcpy(hd, in, sizeof(Header));
    if (std::memcmp(hd->magic, kMagic, 8) != 0) {
        *why = "bad magic";
        return false;
    }
    if (hd->version != kVersion || hd->header_bytes != sizeof(Header) || hd->params_bytes != sizeof(k3dom_params)) {
        *why = "unsupported version or layout";
        return false;
    }
    std::memcpy(params, in + sizeof(Header), sizeof(k3dom_params));
    // sizes a map can have (the C ABI rejects larger grids), so the payload
    // size below cannot wrap around
    const bool sizes_ok = hd->V > 0 && hd->N > 0 && hd->V <= 0x7fffffffll && hd->N <= 0x7fffffffll;
    const uint64_t pb = sizes_ok ? payload_bytes(hd->V, hd->N) : 0;
    const uint64_t rest = (uint64_t)size - sizeof(Header) - sizeof(k3dom_params);
    if (pb == 0 || hd->payload_bytes != pb || rest != pb) {
        *why = "size does not match the header";
        return false;
    }
    *payload = in + sizeof(Header) + sizeof(k3dom_params);
    if (fnv1a_words(*payload, pb) != hd->checksum) {
        *why = "checksum mismatch";
        return false;
    }
    return true;
}

}  // namespace k3ck
