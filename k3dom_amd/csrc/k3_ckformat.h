// Checkpoint record format and its validation: host-only C++ (no HIP), so
// that the checks on untrusted input are built and fuzzed under the
// sanitizers on their own.  Layout in k3_ckformat.cpp.
#pragma once

#include "../../include/k3dom.h"

#include <cstdint>
#include <string>

namespace k3ck {

struct Header {
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
static_assert(sizeof(Header) == 128, "checkpoint header layout");

// Voxel arrays (10 float, 2 int32 per voxel) and particle arrays (7 float,
// 1 int32 per particle) of the payload, in this order.
constexpr int kVoxelFloats = 10, kVoxelInts = 2, kParticleFloats = 7, kParticleInts = 1;
constexpr int kArrays = kVoxelFloats + kVoxelInts + kParticleFloats + kParticleInts;

// Payload bytes for V voxels and N particles (padded to 8).
uint64_t payload_bytes(int64_t V, int64_t N);
// FNV-1a over the payload's 64-bit little-endian words.
uint64_t fnv1a_words(const uint8_t* p, uint64_t bytes);
// A header for a state of V voxels and N particles (checksum of the payload).
Header make_header(int32_t method, int64_t V, int64_t N, uint32_t frame, bool first_pose, double updated_time,
                   const float sensor[3], const float center[3], const uint8_t* payload);
// Checks that [src, src + size) is a whole, intact checkpoint of this
// version: header, layout sizes, payload size and checksum.  On success
// fills *hd, *params and *payload (pointing into src); else sets *why.
bool check(const void* src, int64_t size, Header* hd, k3dom_params* params, const uint8_t** payload,
           std::string* why);

}  // namespace k3ck
