// Launch wrappers of the gfx950 kernels of the per-frame update.  Each wrapper
// enqueues on the given stream and returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "k3_math.h"

namespace k3 {

// Device-side scalars of one map, written and read by kernels within a frame.
struct DevScalars {
    int32_t E_born, E_joint;  // fixed-point exponents
    uint64_t total_born;      // u64 fixed-point totals
    uint64_t total_joint;
    uint32_t n_light;         // items of tiles handled whole
    uint32_t n_heavy;         // items of tiles split into point chunks
    uint32_t overflow;        // non-zero: a capacity was exceeded, the frame tail is skipped
    uint32_t overflow_mask;   // bit l: candidate level l, bit 4: split items, bit 5: DS-PHD/MIB lists
    uint32_t level_total[3];  // candidate-list entries per level
    uint32_t level_items[3];  // work items per level
    uint32_t ds_total;        // DS-PHD/MIB measurement-grid candidate entries
    uint32_t ds_segments;     // and their segments
    uint32_t mp_total;        // records of the last MapPoint export
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

struct GridGeom {
    int gs, gsz;
    float res;
    float sx, sy, sz;  // sensor position relative to the map centre
};

// Values of one frame that kernels read from device memory (so that a
// captured frame graph replays with each frame's values): written by one
// host-to-device copy at the start of the frame.
struct FrameConsts {
    GridGeom g;          // sensor position relative to the map centre this frame
    K3PredictConst pc;   // prediction constants (dt, frame, noise, keys)
    float decay;         // gamma^dt
    uint32_t frame;      // frame counter: counter word of every random draw
    int32_t m;           // points of the frame
    int32_t pad;
    float pshift[3];     // map shift (voxels) still due on the particle positions, applied by predict
};

// ------------------------------------------------------------------ LiDAR ingest
// dst = rot * src for n points (k3_lidar_rotate; src may equal dst).
hipError_t launch_lidar_rotate(hipStream_t st, const float* rot9, const float* sx, const float* sy, const float* sz,
                               float* dx, float* dy, float* dz, int n);
// Evidence of a further LiDAR added to the running sums (acc += add, per
// voxel); returns at once when the abort flag is set.
hipError_t launch_add_evidence(hipStream_t st, const uint32_t* abort_flag, float* acc_occ, float* acc_free,
                               const float* add_occ, const float* add_free, int V);

// ------------------------------------------------------------------ MapPoint export
// Geometry and field rules of one export (k3dom_export_mappoints).
struct MapPointArgs {
    int gs, gsz;
    float res, size, size_z;
    float cx, cy, cz;   // map centre (global frame)
    int ds;             // 1: DS-PHD/MIB fields (occupied mass in the s slot)
    float prior_s, prior_d, prior_occ;
    float min_occ, min_mass;
};
// mode 0: every voxel to out[c]; mode 1: visible voxels compacted in index
// order (flags, off: V and V + 1 words, scratch: scan_scratch_words(V),
// zeroed once; the count to *total).  out: 3 float4 per record.
hipError_t launch_export_mappoints(hipStream_t st, const VoxelArrays& v, int V, const MapPointArgs& a, int mode,
                                   uint32_t* flags, uint32_t* off, uint32_t* scratch, uint32_t* total,
                                   float4* out);

// ------------------------------------------------------------------ scans
// u32 words of scan scratch for up to n_max elements; the scratch must be
// zeroed once after allocation (every scan leaves it zeroed again).
long long scan_scratch_words(long long n_max);
// Exclusive scan of n = n_dev ? (*n_dev) * n_mul : n_max elements in one
// launch; writes the total to out[n] and *total.  scratch: see above.
hipError_t launch_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t* scratch, long long n_max,
                           const uint32_t* n_dev, uint32_t n_mul, uint32_t* total);

// ------------------------------------------------------------------ particles / voxels
// Predicted state of P to prec (2 float4 per particle: px, py, pz, vx | vy, vz,
// w, 0) and the voxel keys to P.key.
hipError_t launch_predict(hipStream_t st, const FrameConsts* fc, ParticleArrays P, float4* prec, int n);
hipError_t launch_shift_particles(hipStream_t st, ParticleArrays P, int n, int mx, int my, int mz);
// Field sets of a map shift: SHIFT_KEPT are carried from frame to frame (the
// masses and mu, which a frame rewrites only where particles are);
// SHIFT_REWRITTEN are rewritten for every voxel by the next full frame before
// anything reads them (ranges, persistent and predicted mass, unassociated
// newborn weight, mean velocity), so their shift can be deferred.
constexpr int SHIFT_KEPT = 1, SHIFT_REWRITTEN = 2, SHIFT_ALL = 3;
hipError_t launch_shift_voxels(hipStream_t st, VoxelArrays dst, VoxelArrays src, int gs, int gsz, int mx, int my,
                               int mz, float prior_f, float prior_s, float prior_d, int fields);
hipError_t launch_fill_voxels(hipStream_t st, VoxelArrays v, int V, float prior_f, float prior_s, float prior_d);

// Stable LSD radix sort of keys (values in [0, 2^key_bits)).  in_keys is not
// modified.  On return *out_keys / *out_idx point into the scratch buffers.
struct SortScratch {
    uint32_t *k0, *k1, *i0, *i1;
    uint32_t* hist;  // sort_hist_words(n): digit counts per tile and per group of tiles
};
int sort_blocks(int n);
long long sort_hist_words(int n);
// Stable sort of in_keys (key_bits bits); also sets fill_a and fill_b[0, nfill)
// to -1 (the voxel ranges, before launch_gather_sorted marks them).
hipError_t launch_radix_sort(hipStream_t st, const int32_t* in_keys, int n, int key_bits, SortScratch s,
                             const uint32_t** out_keys, const uint32_t** out_idx, int32_t* fill_a, int32_t* fill_b,
                             int nfill);
// Sorted particles from the predicted records to dst, except the weights,
// which go to wsorted; and the voxels' particle ranges (start / end, all -1
// on entry: launch_radix_sort's fill).
hipError_t launch_gather_sorted(hipStream_t st, const float4* prec, ParticleArrays dst, float* wsorted,
                                const uint32_t* perm, const uint32_t* sorted_keys, int n, int32_t* start,
                                int32_t* end);

// ------------------------------------------------------------------ kernel inference (k3_bgki.hip)
// Candidate lists are built over a hierarchy of boxes of 4x4x4-voxel tiles:
// level 0 boxes take their candidates from all points, level l+1 boxes from
// their level-l parent.  Every list holds point indices in ascending order.
struct LevelDesc {
    int nbx, nby, nbz;   // boxes per axis
    int bx, by, bz;      // box size in tiles
    int nparents;        // parent boxes (1 at level 0: all points)
    int pnbx, pnby;      // parent boxes per axis (level >= 1)
    int ncx, ncy, ncz;   // children per parent along each axis (level >= 1)
    int nch;             // children per parent
    int seg;             // parent candidates per work item
};
struct LevelBufs {
    uint32_t *nseg, *item_base;    // nparents, nparents + 1
    uint32_t *counts, *slot_off;   // slot_cap, slot_cap + 1
    uint32_t *box_off, *box_len;   // nboxes
    float4* ctab;                  // levels 1, 2: nparents * nch children's test constants (two float4 each)
    uint32_t* list;                // list_cap
    uint32_t* bsum;                // scan scratch
    uint8_t* mask;                 // per child group and parent-list position: the children a candidate passes
    uint4* ictx;                   // items_cap: each work item's context, from the count kernel to the fill
    long long items_cap, slot_cap, list_cap;
    long long mask_stride;         // parent-list capacity (one mask byte per position per child group)
};
struct TileGeom {
    int ntx, nty, ntz;
};
constexpr int MAX_CHILDREN = 1024;

// m_cap: capacity of the point arrays (the launch covers it; fc->m points are live)
hipError_t launch_point_prep(hipStream_t st, const FrameConsts* fc, const float* xs, const float* ys,
                             const float* zs, int m_cap, float res, float ls, K3PointRec* rec, float4* pxyz);
// Wavefronts share a box's children in groups; masks need
// level_child_groups(L) bytes per parent-list position.
int level_child_groups(const LevelDesc& L);

// Work queues of the gather: BG_NQ heads, one per BG_QSTRIDE words (128 bytes).
constexpr uint32_t BG_NQ = 64;
constexpr uint32_t BG_QSTRIDE = 32;
constexpr size_t BG_QWORDS = (size_t)BG_NQ * BG_QSTRIDE;

struct ItemBufs {
    // per tile: split-item and whole-tile counts, one array of 2 ntiles
    // (light_cnt = heavy_cnt + ntiles), and their exclusive scan (2 ntiles + 1,
    // light_off = heavy_off + ntiles)
    uint32_t *heavy_cnt, *light_cnt, *heavy_off, *light_off;
    uint32_t* bsum;
    uint4* items;             // ntiles + heavy_cap
    uint32_t* tile_partial;   // 2 * ntiles
    float* partial;           // heavy_cap * partial_stride
    uint32_t* qheads;         // BG_QWORDS: the gather's work-queue heads
    int partial_stride = 128; // floats per split item (256 with length-scale derivatives)
    long long heavy_cap;
};
// The candidate lists of the three levels (L, B: three each), and the
// tiles' gather-item counts into I.heavy_cnt / I.light_cnt.
hipError_t launch_build_levels(hipStream_t st, const FrameConsts* fc, const TileGeom& tg, const LevelDesc* L,
                               LevelBufs* B, const ItemBufs& I, float ls, const float4* pxyz, DevScalars* sc);
// The gather's work items from the tiles' lists and item counts; also
// zeroes the gather's work-queue heads.
hipError_t launch_build_items(hipStream_t st, int ntiles, const LevelBufs& tiles, const ItemBufs& I,
                              DevScalars* sc);
// Arguments of the kernel-inference gather over one map.
struct BgArgs {
    const FrameConsts* fc;
    TileGeom tg;
    float ls, kscale, sk, inv_sk2, inv_ls, sigma_o, sigma_f;
    const uint4* items;
    uint32_t* qheads;  // the work-queue heads of the single-map gather
    const uint32_t* tlist;
    const K3PointRec* rec;
    float *del_occ, *del_free, *dls_occ, *dls_free, *partial;
    DevScalars* sc;
};
BgArgs make_bg_args(const FrameConsts* fc, const TileGeom& tg, float ls, float sigma_o, float sigma_f,
                    const ItemBufs& I, const LevelBufs& T, const K3PointRec* rec, float* del_occ, float* del_free,
                    float* dls_occ, float* dls_free, DevScalars* sc);
// The gather over the items of several maps in one persistent launch
// (maps: device array of nmaps argument sets; prefix: nmaps + 1 entries;
// qheads: BG_QWORDS words of work-queue heads, zeroed by the launch).
hipError_t launch_bgki_batch(hipStream_t st, int grid_blocks, const BgArgs* maps, int nmaps, uint32_t* prefix,
                             uint32_t* qheads, int waves = 8);
// Group 0's scalars after a joint gather of ng LiDAR groups (gsc[1 .. ng - 1]):
// overflow flags and bits, largest per-level needs, summed items.
hipError_t launch_merge_group_scalars(hipStream_t st, DevScalars* sc, const DevScalars* gsc, int ng);

// sigma_o / sigma_f weight the occupied / free terms (both the reference's
// single sigma in a plain frame).  Split items need I.partial with 128 floats
// per item (256 for the _grad variants).
// waves: wavefronts per block of the gather (4, 8 or 16)
// queues_zeroed: the work-queue heads are already zero (launch_build_items
// was told to zero them and nothing has claimed since).
hipError_t launch_bgki(hipStream_t st, int grid_blocks, const FrameConsts* fc, const TileGeom& tg, float ls,
                       float sigma_o, float sigma_f, const ItemBufs& I, const LevelBufs& tiles, const K3PointRec* rec,
                       float* del_occ, float* del_free, DevScalars* sc, int waves = 8, bool queues_zeroed = false);
// Evidence of the split tiles: chunk sums added in chunk order, times the
// sigma factors (direct items are written by the gather itself).
hipError_t launch_bgki_combine(hipStream_t st, int ntiles, const GridGeom& g, const TileGeom& tg, const ItemBufs& I,
                               float sigma_o, float sigma_f, float* del_occ, float* del_free, DevScalars* sc);
// The same evidence plus its derivative in the length scale (dls_*).
hipError_t launch_bgki_grad(hipStream_t st, int grid_blocks, const FrameConsts* fc, const TileGeom& tg, float ls,
                            float sigma_o, float sigma_f, const ItemBufs& I, const LevelBufs& tiles,
                            const K3PointRec* rec, float* del_occ, float* del_free, float* dls_occ, float* dls_free,
                            DevScalars* sc, int waves = 4);
hipError_t launch_bgki_combine_grad(hipStream_t st, int ntiles, const GridGeom& g, const TileGeom& tg,
                                    const ItemBufs& I, float ls, float sigma_o, float sigma_f, float* del_occ,
                                    float* del_free, float* dls_occ, float* dls_free, DevScalars* sc);
hipError_t launch_bgki_count(hipStream_t st, int grid_blocks, const GridGeom& g, const TileGeom& tg, float ls,
                             const ItemBufs& I, uint32_t n_items, const LevelBufs& tiles, const K3PointRec* rec,
                             unsigned long long* counts);

// ------------------------------------------------------------------ frame tail
// Every tail kernel returns at once when *abort_flag is set (a capacity of
// the kernel-inference stage overflowed; the host grows it and re-runs).
// The largest born mass and the largest joint particle weight, which fix the
// exponents of the fixed-point CDFs, are kept as one maximum per block of
// their producers (float bits) and reduced by launch_fixed_scan: the born
// maxima come from launch_mass_update (mass_update_parts(V) slots), the joint
// ones from launch_voxel_particles (voxel_particles_parts(N) slots) followed by
// launch_newborn (newborn_parts(B) slots).
int mass_update_parts(int V);
int voxel_particles_parts(int n);
int newborn_parts(int B);
// decay: fc->decay (1 in the initialisation frame when fc is null)
// weights, keys: the n sorted particles (n = 0: no particles, every voxel
// start is -1); usum: n floats of scratch for per-unit sums
hipError_t launch_mass_update(hipStream_t st, const uint32_t* abort_flag, VoxelArrays v, int V,
                              const float* weights, const int32_t* keys, int n, float* usum, const float* del_occ,
                              const float* del_free, const FrameConsts* fc, float mass_scale, float p_B, float* born,
                              uint32_t* born_parts);
// DS-PHD/MIB method: measurement grid of the frame's points into
// meas_occ / meas_free; cell update (born masses into born, block maxima into
// born_parts as launch_mass_update) that also divides the sorted weights of
// voxels whose sum exceeds 1 (norm: V floats of scratch); born masses of the
// first frame from the measurement grid.
// Candidate lists of the culled measurement grid.  counts, nseg: one word per
// supertile (ds_supertiles(gs, gsz)); off, seg_off: that many + 1; scratch,
// seg_scratch: scan_scratch_words of that many (zeroed once); list: cap point
// indices, processed in segments of seg_len; segcnt:
// 64 * 32 * ds_segments_max(gs, gsz, cap, seg_len) bytes; part: as many uint2.
// A frame needing more than cap raises the overflow flag (mask bit 5, total
// in ds_total) and writes no masses.
struct DsLists {
    uint32_t *counts, *off, *scratch, *nseg, *seg_off, *seg_scratch;
    int32_t* list;
    uint8_t* segcnt;
    uint2* part;
    long long cap;
    int seg_len;
};
int ds_supertiles(int gs, int gsz);
long long ds_segments_max(int gs, int gsz, long long cap, int seg_len);
hipError_t launch_ds_meas_grid(hipStream_t st, const uint32_t* abort_flag, const FrameConsts* fc, int gs, int gsz,
                               const float* xs, const float* ys, const float* zs, float resolution,
                               float meas_occ_stddev, float max_range, const DsLists& L, DevScalars* sc,
                               float* meas_occ, float* meas_free);
hipError_t launch_ds_mass_update(hipStream_t st, const uint32_t* abort_flag, VoxelArrays v, int V, float* weights,
                                 const int32_t* keys, int n, float* usum, const float* meas_occ,
                                 const float* meas_free, float p_B, float* born, float* norm,
                                 uint32_t* born_parts);
hipError_t launch_ds_born_from_meas(hipStream_t st, const float* meas_occ, float* born, int V, uint32_t* born_parts);
// P: the n sorted particles, whose weights P.w are written (penalised) from
// wsrc; usum3: 3n floats of scratch
hipError_t launch_voxel_particles(hipStream_t st, const uint32_t* abort_flag, VoxelArrays v, int V,
                                  ParticleArrays P, const float* wsrc, int n, float min_vel, float* usum3,
                                  uint32_t* joint_parts);
// Inclusive fixed-point CDF of [a ++ b].  max_parts: nparts per-block maxima
// of the values; scratch: fixed_scan_scratch_words(na + nb) words.
long long fixed_scan_scratch_words(long long n);
hipError_t launch_fixed_scan(hipStream_t st, const float* a, int na, const float* b, int nb,
                             const uint32_t* max_parts, int nparts, int32_t* E_out, uint64_t* cdf,
                             uint64_t* scratch, uint64_t* total);
hipError_t launch_bounds(hipStream_t st, const uint32_t* abort_flag, const uint64_t* cdf, const uint64_t* total,
                         int V, uint32_t cnt, const float* born, uint32_t* bnd, float* wua);
hipError_t launch_newborn(hipStream_t st, const uint32_t* abort_flag, const uint32_t* bnd, int V, int B, int gs,
                          const float* wua, ParticleArrays NB, float vmax, float min_vel, const FrameConsts* fc,
                          uint32_t k0, uint32_t k1, uint32_t* joint_parts);
hipError_t launch_init_particles(hipStream_t st, const uint32_t* abort_flag, const uint32_t* bnd, int V, int N,
                                 int gs, ParticleArrays P, float vmax, float min_vel, const FrameConsts* fc,
                                 uint32_t k0, uint32_t k1);
hipError_t launch_resample(hipStream_t st, const uint32_t* abort_flag, const uint64_t* cdf, int total_len,
                           const uint64_t* total, const int32_t* E, ParticleArrays sorted, ParticleArrays nb,
                           ParticleArrays out, int N, const FrameConsts* fc, uint32_t k0, uint32_t k1,
                           int32_t* idx_out);
// ------------------------------------------------------------------ differentiated window (k3_adjoint.hip)
// Record of one forward frame: masses after the frame, predicted dynamic mass,
// and the frozen counts per voxel (particles inside the grid, of them not
// penalised; newborns, of them not penalised).
struct WindowFrame {
    float *f, *s, *d, *dyn;
    int32_t *n, *m, *nb, *mb;
};
struct AdjointFrameArgs {
    int V, gs, gsz;
    WindowFrame post;          // this frame
    WindowFrame prev;          // previous frame (prev.f == nullptr: the pre-state is the priors)
    int move[3];               // map shift applied before this frame
    int move_next[3];          // map shift applied before the next frame
    const double *lam_next_f, *lam_next_s, *lam_next_d;  // adjoint of the next frame's pre-state (or null)
    double *lam_f, *lam_s, *lam_d;                        // out: adjoint of this frame's pre-state
    const float *del_occ, *del_free, *dls_occ, *dls_free;
    const uint8_t* gt;
    const uint8_t* mask;       // null: every voxel
    double loss_scale;         // d loss / d (per-voxel BCE term)
    float prior_f, prior_s, prior_d;
    float decay, mass_scale, p_B, sigma_occ, sigma_free;
    float w_coef;              // p_S / N
    int has_w, has_w_prev;     // W of this / of the previous frame depends on the parameters
};
int window_blocks(int V);
// Ps: the N key-sorted persistent particles, NB: the B newborn ones (in voxel order)
hipError_t launch_window_record(hipStream_t st, int V, int gs, int gsz, bool normal, float min_vel,
                                const VoxelArrays& v, const ParticleArrays& Ps, int N, const ParticleArrays& NB,
                                int B, const WindowFrame& rec, const uint8_t* gt, const uint8_t* mask,
                                double* loss_partial);
// acc[0]: lambda_W (in: this frame's, out: the previous frame's); acc[1..6] += gradients
hipError_t launch_window_adjoint(hipStream_t st, const AdjointFrameArgs& a, double* acc, double* partial);
hipError_t launch_window_clear(hipStream_t st, double* acc, int n);
// *sticky |= (*flag != 0), and *sticky = 0 (one-thread kernels, capture-safe)
hipError_t launch_flag_or(hipStream_t st, uint32_t* sticky, const uint32_t* flag);
hipError_t launch_flag_clear(hipStream_t st, uint32_t* sticky);

hipError_t launch_reset_scalars(hipStream_t st, DevScalars* sc);
hipError_t launch_clear_overflow(hipStream_t st, DevScalars* sc);

}  // namespace k3
