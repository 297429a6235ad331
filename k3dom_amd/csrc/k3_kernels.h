// Launch wrappers of the gfx950 kernels of the per-frame update.  Each wrapper
// enqueues on the given stream and returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "k3_math.h"

namespace k3 {

// Device-side scalars of one map, written and read by kernels within a frame.
struct DevScalars {
    uint32_t born_max_bits;   // max born mass (float bits; masses are >= 0)
    uint32_t joint_max_bits;  // max particle weight after the static penalty
    int32_t E_born, E_joint;  // fixed-point exponents
    uint64_t total_born;      // u64 fixed-point totals
    uint64_t total_joint;
    uint32_t item_counter;    // BGKI work queue head
    uint32_t n_items;         // BGKI items
    uint32_t n_heavy_items;
    uint32_t partial_used;    // partial-sum slots handed out
    uint32_t level_total[4];  // candidate-list totals per level
    uint32_t pad[4];
};

// Voxel state arrays (one buffer set).
struct VoxelArrays {
    float *f, *s, *d, *pers, *pred, *mu, *wua, *mvx, *mvy, *mvz;
    int32_t *start, *end;
};

// Particle set (structure of arrays).
struct ParticleArrays {
    float *px, *py, *pz, *vx, *vy, *vz, *w;
    int32_t* key;
};

// Geometry of one level of the BGKI candidate-list hierarchy: boxes of
// (bx, by, bz) tiles of 4x4x4 voxels.
struct LevelGeom {
    int nbx, nby, nbz;      // boxes per axis
    int bx, by, bz;         // box size in tiles
    int pbx, pby, pbz;      // parent box size in tiles (0 at the root level)
    int pnbx, pnby;         // parent boxes per axis
};

struct GridGeom {
    int gs, gsz;
    float res;
    float sx, sy, sz;  // sensor position relative to the map centre
};

hipError_t launch_predict(hipStream_t st, const K3PredictConst& c, ParticleArrays P, int n);
hipError_t launch_shift_particles(hipStream_t st, ParticleArrays P, int n, int mx, int my, int mz);
hipError_t launch_shift_voxels(hipStream_t st, VoxelArrays dst, VoxelArrays src, int gs, int gsz, int mx, int my,
                               int mz, float prior_f, float prior_s, float prior_d);
hipError_t launch_fill_voxels(hipStream_t st, VoxelArrays v, int V, float prior_f, float prior_s, float prior_d);

// Stable LSD radix sort of keys (values in [0, 2^key_bits)).  in_keys is not
// modified.  On return *out_keys / *out_idx point into the scratch buffers.
struct SortScratch {
    uint32_t *k0, *k1, *i0, *i1;
    uint32_t* hist;  // 256 * nblocks
};
int sort_blocks(int n);
hipError_t launch_radix_sort(hipStream_t st, const int32_t* in_keys, int n, int key_bits, SortScratch s,
                             const uint32_t** out_keys, const uint32_t** out_idx);
hipError_t launch_gather_sorted(hipStream_t st, ParticleArrays src, ParticleArrays dst, const uint32_t* perm,
                                const uint32_t* sorted_keys, int n);
hipError_t launch_ranges(hipStream_t st, const int32_t* sorted_keys, int n, int32_t* start, int32_t* end, int V);

// BGKI
hipError_t launch_point_prep(hipStream_t st, const float* xs, const float* ys, const float* zs, int m, float res,
                             K3PointRec* rec, float4* pxyz);
hipError_t launch_list_level(hipStream_t st, int count_mode, const GridGeom& g, const LevelGeom& L, int nboxes,
                             float ls, const float4* pxyz, int m, const uint32_t* parent_off,
                             const uint32_t* parent_len, const uint32_t* parent_list, uint32_t* len,
                             const uint32_t* off, uint32_t* list);
hipError_t launch_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out_excl, int n, uint32_t* total);
hipError_t launch_count_items(hipStream_t st, int ntiles, const uint32_t* tile_off, const uint32_t* tile_len,
                              const uint32_t* tile_list, uint32_t* heavy_cnt, uint32_t* light_cnt,
                              uint32_t* heavy_off, uint32_t* light_off, DevScalars* sc);
hipError_t launch_emit_items(hipStream_t st, int ntiles, const uint32_t* tile_off, const uint32_t* tile_len,
                             const uint32_t* tile_list, const uint32_t* heavy_cnt, const uint32_t* heavy_off,
                             const uint32_t* light_off, uint4* items, uint32_t* tile_partial, DevScalars* sc);
hipError_t launch_bgki(hipStream_t st, int grid_blocks, const GridGeom& g, int ntx, int nty, float ls, float sigma,
                       const uint4* items, const uint32_t* tile_list, const K3PointRec* rec, float* del_occ,
                       float* del_free, float* partial, DevScalars* sc);
hipError_t launch_bgki_combine(hipStream_t st, int ntiles, const GridGeom& g, int ntx, int nty,
                               const uint32_t* tile_partial, const float* partial, float* del_occ, float* del_free);

// Mass update and particle passes
hipError_t launch_mass_update(hipStream_t st, VoxelArrays v, int V, const float* weights, const float* del_occ,
                              const float* del_free, float decay, float mass_scale, float p_B, float* born,
                              DevScalars* sc);
hipError_t launch_voxel_particles(hipStream_t st, VoxelArrays v, int V, ParticleArrays P, float min_vel,
                                  DevScalars* sc);
hipError_t launch_fixed_scan(hipStream_t st, const float* a, int na, const float* b, int nb,
                             const uint32_t* max_bits, int32_t* E_out, uint64_t* cdf, uint64_t* blocksums,
                             uint64_t* total);
int fixed_scan_blocks(int n);
hipError_t launch_bounds(hipStream_t st, const uint64_t* cdf, const uint64_t* total, int V, uint32_t cnt,
                         const float* born, uint32_t* bnd, float* wua);
hipError_t launch_newborn(hipStream_t st, const uint32_t* bnd, int V, int B, int gs, const float* wua,
                          ParticleArrays NB, float vmax, float min_vel, uint32_t frame, uint32_t k0, uint32_t k1,
                          DevScalars* sc);
hipError_t launch_init_particles(hipStream_t st, const uint32_t* bnd, int V, int N, int gs, ParticleArrays P,
                                 float vmax, float min_vel, uint32_t frame, uint32_t k0, uint32_t k1);
hipError_t launch_resample(hipStream_t st, const uint64_t* cdf, int total_len, const uint64_t* total,
                           const int32_t* E, ParticleArrays sorted, ParticleArrays nb, ParticleArrays out, int N,
                           uint32_t frame, uint32_t k0, uint32_t k1, int32_t* idx_out);
hipError_t launch_reset_scalars(hipStream_t st, DevScalars* sc);

}  // namespace k3
