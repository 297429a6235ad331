/* C ABI of the MI355X-native K3DOM per-frame dynamic occupancy update.
 *
 * One handle owns one map: its device buffers (structure-of-arrays voxel
 * masses and particles), one HIP stream and the per-frame launch plan.  Every
 * call is synchronous with respect to the host unless its name ends in
 * _async.  Handles are not thread-safe; distinct handles may be used from
 * distinct threads.  Every function returns a k3dom_status; the text of the
 * last error on a handle is available from k3dom_last_error().
 *
 * Each entry point replaces one member of the reference's C++ filter class
 * (youngjae-min/k3dom, namespace dom, classes DOM_c and DOM).  The reference
 * has no C ABI of its own: its ROS node links the .cu sources directly.
 */
#ifndef K3DOM_H
#define K3DOM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum k3dom_status {
    K3DOM_OK = 0,
    K3DOM_ERR_INVALID_ARG = 1,   /* bad pointer, size or parameter value      */
    K3DOM_ERR_HIP = 2,           /* a HIP runtime call failed                  */
    K3DOM_ERR_OUT_OF_MEMORY = 3, /* device or host allocation failed           */
    K3DOM_ERR_STATE = 4          /* call not valid in the handle's state       */
} k3dom_status;

/* Filter parameters: field-for-field the reference's dom::DOM_c::Params
 * (include/common/dom.h:19-52), in the units DOM_c expects (the ROS node
 * divides the stddevs and velocities by the resolution and converts
 * map_shift_thresh to voxels, src/k3dom/k3dom_node.cpp:344-368;
 * k3dom_params_from_yaml() performs that conversion).  `seed` keys the
 * counter-based random generator (the reference seeds cuRAND with 123456,
 * src/common/dom.cu:96). */
typedef struct k3dom_params {
    float size;
    float size_z;
    float mass_scale;
    float resolution;
    int32_t particle_count;
    int32_t new_born_particle_count;
    float persistence_prob;
    float stddev_process_noise_position;
    float stddev_process_noise_velocity;
    float birth_prob;
    float stddev_velocity;
    float init_max_velocity;
    float particle_min_vel;
    float sensor_off_x;
    float sensor_off_y;
    float sensor_off_z;
    int32_t map_shift_thresh;
    float prior_free;
    float prior_static;
    float prior_dynamic;
    float prior_occ;
    float sigma;
    float ls;
    float gamma;
    float max_range;
    float meas_occ_stddev;
    uint64_t seed;
} k3dom_params;

/* Readback record with the field order, types and size (68 bytes) of the
 * reference's dom::GridCell (include/common/dom_types.h:12-45). */
typedef struct k3dom_cell {
    int32_t start_idx;
    int32_t end_idx;
    float mu_A;
    float mu_UA;
    float w_A;
    float w_UA;
    float mean_x_vel;
    float mean_y_vel;
    float mean_z_vel;
    float likelihood;
    float p_A;
    float free_mass;
    float static_mass;
    float dynamic_mass;
    float occ_mass;
    float pers_mass;
    float pred_mass;
} k3dom_cell;

/* Geometry and pose of a map, as returned by the reference's getters
 * (include/common/dom.h:59-83). */
typedef struct k3dom_info {
    int32_t grid_size;
    int32_t grid_size_z;
    int32_t grid_cell_count;
    int32_t particle_count;
    int32_t new_born_particle_count;
    int32_t iteration;          /* frames processed by k3dom_update_grid      */
    float resolution;
    float center_pos[3];
    float sensor_pos[3];
    double updated_time;        /* time of the last frame, -1 before the first */
} k3dom_info;

typedef struct k3dom_handle k3dom_handle;

/* Fills *out with the reference node's defaults (src/k3dom/k3dom_node.cpp:215-258)
 * after unit conversion, seed 123456. */
k3dom_status k3dom_params_default(k3dom_params* out);

/* Reads a method YAML and then a dataset YAML (either path may be NULL),
 * flat `key: value` pairs as in config/k3dom.yaml and config/complex.yaml, with
 * the node's defaults for missing keys and its unit conversion
 * (src/k3dom/k3dom_node.cpp:215-368; rosparam load order of
 * launch/k3dom_demo.launch:33-36).  Writes the error text into err (capacity
 * err_len) on failure. */
k3dom_status k3dom_params_from_yaml(const char* method_yaml, const char* dataset_yaml,
                                    k3dom_params* out, char* err, int32_t err_len);

/* Replaces dom::DOM::DOM / DOM_c::DOM_c + initialize (src/k3dom/dom.cu:20-23,
 * src/common/dom.cu:29-66,90-109): allocates the map on the current HIP device
 * and sets every voxel to the priors.  The map runs the K3DOM method. */
k3dom_status k3dom_create(const k3dom_params* params, k3dom_handle** out);

/* The reference's two filters share DOM_c and its particle stages and differ
 * in the mass update: K3DOM's kernel inference (src/k3dom/dom.cu) or the
 * DS-PHD/MIB measurement grid (src/dsphdmib/dom.cu:20-118). */
#define K3DOM_METHOD_K3DOM 0
#define K3DOM_METHOD_DSPHDMIB 1

/* Replaces the constructor of either filter (dom::DOM of src/k3dom/dom.cu:20-23
 * or of src/dsphdmib/dom.cu:20-30): as k3dom_create, running `method`.  A
 * DS-PHD/MIB map uses resolution, max_range and meas_occ_stddev for its
 * measurement grid and ignores sigma, ls, gamma and mass_scale; it keeps
 * its occupied mass in GridCell::occ_mass, and static/dynamic masses stay at
 * their priors.  Batches and differentiated windows build K3DOM maps only. */
k3dom_status k3dom_create_method(const k3dom_params* params, int32_t method, k3dom_handle** out);

/* Replaces DOM_c::~DOM_c (src/common/dom.cu:68-88).  NULL is accepted. */
k3dom_status k3dom_destroy(k3dom_handle* h);

/* Replaces DOM_c::updatePose (src/common/dom.cu:111-158): sets the sensor
 * position (global frame, metres) and shifts the map by whole voxels when the
 * sensor has moved map_shift_thresh voxels or more. */
k3dom_status k3dom_update_pose(k3dom_handle* h, float x, float y, float z);

/* Replaces DOM::updateGrid (src/k3dom/dom.cu:29-50): one filter frame at time t
 * (seconds) with n points given relative to the sensor in map-aligned axes
 * (host arrays).  The first frame (and any frame while the stored time is
 * <= 0, as in the reference) initialises the particles instead. */
k3dom_status k3dom_update_grid(k3dom_handle* h, double t, const float* xs, const float* ys,
                               const float* zs, int32_t n);

/* As k3dom_update_grid with the three point arrays already in device memory
 * (no host copy of the points). */
k3dom_status k3dom_update_grid_device(k3dom_handle* h, double t, const float* d_xs,
                                      const float* d_ys, const float* d_zs, int32_t n);

/* The reference's public stage methods of DOM_c (include/common/dom.h:89-94)
 * and DOM (include/k3dom/dom.h:24-28), for callers that drive a frame stage by
 * stage.  A normal frame calls them in the order of DOM::updateGrid
 * (src/k3dom/dom.cu:33-44):
 *   k3dom_particle_prediction(h, dt)        particlePrediction(dt)
 *   k3dom_particle_assignment(h)            particleAssignment()
 *   k3dom_grid_cell_occupancy_update(h, dt, xs, ys, zs, n)
 *                                           gridCellOccupancyUpdate(dt, points)
 *   k3dom_update_persistent_particles(h)    updatePersistentParticles()
 *   k3dom_initialize_new_particles(h)       initializeNewParticles()
 *   k3dom_statistical_moments(h)            statisticalMoments()
 *   k3dom_resampling(h)                     resampling() and the copy-back
 * and a first frame calls k3dom_initialize_particles (initializeParticles,
 * src/k3dom/dom.cu:52-77).  Each runs its part of the frame's kernels and
 * returns when it is done; the frame's results equal those of the same
 * frame given to k3dom_update_grid (dt = t - the previous frame's time), bit
 * for bit.  As in the reference, the stages leave the stored time of the last
 * frame to updateGrid.  The persistent update also computes the moments (the
 * same weights and sums), so statisticalMoments may come before or after
 * initializeNewParticles and launches nothing.  A stage called out of this
 * order, a whole-frame call in the middle of a staged frame, or a
 * DS-PHD/MIB map returns K3DOM_ERR_STATE; a stage that fails abandons the
 * frame.  Points are host arrays relative to the sensor in map-aligned axes. */
k3dom_status k3dom_particle_prediction(k3dom_handle* h, float dt);
k3dom_status k3dom_particle_assignment(k3dom_handle* h);
k3dom_status k3dom_grid_cell_occupancy_update(k3dom_handle* h, float dt, const float* xs, const float* ys,
                                              const float* zs, int32_t n);
k3dom_status k3dom_update_persistent_particles(k3dom_handle* h);
k3dom_status k3dom_initialize_new_particles(k3dom_handle* h);
k3dom_status k3dom_statistical_moments(k3dom_handle* h);
k3dom_status k3dom_resampling(k3dom_handle* h);
k3dom_status k3dom_initialize_particles(k3dom_handle* h, const float* xs, const float* ys, const float* zs,
                                        int32_t n);

/* Geometry, pose, time and frame count (include/common/dom.h:59-83). */
k3dom_status k3dom_get_info(const k3dom_handle* h, k3dom_info* out);

/* Copies every voxel into a host array of grid_cell_count GridCell-compatible
 * records, in the reference's index order i = x + gs*y + gs*gs*z.  This is the
 * drop-in for the caller-side cudaMemcpy of grid_cell_array
 * (src/k3dom/k3dom_node.cpp:131-132,162-163). */
k3dom_status k3dom_read_cells(k3dom_handle* h, k3dom_cell* out);

/* Structure-of-arrays readback of the voxel masses and mean velocity (any
 * pointer may be NULL).  Velocities are in voxels per second as in GridCell. */
k3dom_status k3dom_read_masses(k3dom_handle* h, float* free_mass, float* static_mass,
                               float* dynamic_mass, float* mean_vx, float* mean_vy, float* mean_vz);

/* Several LiDARs (the node's PCL transform, src/k3dom/k3dom_node.cpp:73-83,
 * done on the GPU and generalised; no counterpart in the reference, whose
 * node serves one LiDAR).  Each LiDAR gives its points in its own frame, the
 * rotation from its axes to the map's (row-major, rot[0..2] the first row)
 * and its origin relative to the position given to k3dom_update_pose
 * (metres, map axes).  A point becomes rot * p, relative to its LiDAR: each
 * LiDAR's evidence is computed as the reference computes its one sensor's
 * with the sensor at that LiDAR's origin, and the LiDARs' evidence is added
 * in LiDAR order.  One LiDAR with rot = identity and origin 0 is exactly
 * k3dom_update_grid.  DS-PHD/MIB maps take one LiDAR (any rot and origin). */
#define K3DOM_MAX_LIDARS 16
typedef struct k3dom_lidar {
    float rot[9];
    float origin[3];
    const float* xs;  /* points in the LiDAR's frame: host arrays, or device */
    const float* ys;  /* arrays when points_on_device is non-zero            */
    const float* zs;
    int32_t n;
    int32_t pad;
} k3dom_lidar;

/* One filter frame at time t (seconds) from n_lidars (1..K3DOM_MAX_LIDARS)
 * LiDARs, whose points are transformed and kept by LiDAR on the GPU. */
k3dom_status k3dom_update_grid_lidars(k3dom_handle* h, double t, const k3dom_lidar* lidars, int32_t n_lidars,
                                      int32_t points_on_device);

/* One voxel of the evaluation export, MapPoint of include/common/mappoint.h:5-28
 * (PCL_ADD_POINT4D x, y, z + padding, then seven floats, 16-byte aligned:
 * 48 bytes).  Values as the node's publish_mapdata:
 *   K3DOM (src/k3dom/k3dom_node.cpp:161-184): v = mean velocity * resolution
 *     (m/s), occ_val = occ_mean(), dyn_val = dyn_mean(), eval_aug =
 *     total_mass(), dyn_aug = 1.1;
 *   DS-PHD/MIB (src/dsphdmib/dsphdmib_node.cpp:153-177): occ_val =
 *     occ_prob(), dyn_val = mean_vel_sq() * resolution^2, eval_aug =
 *     -(1 - occ_mass - free_mass), dyn_aug = occ_mass.
 * x, y, z = indexToCoordX/Y/Z(i) (voxel centre, global frame); w = 1, pad = 0. */
typedef struct k3dom_mappoint {
    float x, y, z, w;
    float v_x, v_y, v_z, occ_val;
    float dyn_val, eval_aug, dyn_aug, pad;
} k3dom_mappoint;

enum {
    K3DOM_MAPPOINTS_ALL = 0,     /* every voxel, index order (publish_mapdata)            */
    K3DOM_MAPPOINTS_VISIBLE = 1  /* voxels passing the visualisation filter of publish_map,
                                    compacted on the GPU in index order: K3DOM
                                    occ_mean() > min_occ && total_mass() > min_mass
                                    (src/k3dom/k3dom_node.cpp:137-138); DS-PHD/MIB
                                    occ_prob() > min_occ && occ_mass + free_mass > min_mass
                                    (src/dsphdmib/dsphdmib_node.cpp:134-135)              */
};

/* Builds the MapPoint records of the current map on the GPU and copies them to
 * out.  *count receives the number of records; with out == NULL and
 * capacity == 0 only the count is returned.  A capacity below the count is
 * K3DOM_ERR_INVALID_ARG (count still written). */
k3dom_status k3dom_export_mappoints(k3dom_handle* h, int32_t mode, float min_occ, float min_mass,
                                   k3dom_mappoint* out, int64_t capacity, int64_t* count);

/* Allocation-free variant of k3dom_export_mappoints.  No counterpart in the
 * reference: it serves callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The records go to DEVICE memory d_out (capacity records; d_out == NULL with
 * capacity == 0 returns only the count), and the visible-voxel compaction
 * runs in one pass over the library's fixed device scratch (about 0.5 MiB per
 * GPU, held until the process ends) instead of buffers the handle allocates:
 * the call allocates nothing.  Maps of more than K3DOM_NOALLOC_MAX_VOXELS
 * voxels are rejected with K3DOM_ERR_INVALID_ARG.  With capacity below the
 * count, the first capacity records are written, *count receives the count
 * and K3DOM_ERR_INVALID_ARG is returned.  Every call on one GPU shares that
 * scratch: calls that overlap on one GPU (handles used from several threads
 * or streams at once) race. */
#define K3DOM_NOALLOC_MAX_VOXELS (65536ll * 4096ll)
k3dom_status k3dom_export_mappoints_noalloc(k3dom_handle* h, int32_t mode, float min_occ, float min_mass,
                                           k3dom_mappoint* d_out, int64_t capacity, int64_t* count);

/* Checkpoint and resume, no counterpart in the reference (whose filter state
 * exists only in device memory for the life of the process).  A checkpoint
 * holds the parameters and method, the frame counter (which keys every random
 * draw, so the random streams continue), the time of the last frame, the
 * pose, every voxel array and the current particle set: a map restored from
 * it and given the same later frames produces the same results bit for bit
 * as the map it was taken from.  A pending map shift is applied before the
 * state is written.  The format is versioned and checksummed (FNV-1a over the
 * payload's 64-bit words); a 128-byte header is followed by the k3dom_params
 * record, then the arrays (layout in k3dom_amd/csrc/k3_checkpoint.cpp); the
 * per-frame arrays of k3dom_read_array describe frames the handle processed
 * itself, so a restored handle has them only after its first frame.
 *
 * k3dom_checkpoint_size: bytes of a checkpoint of h.
 * k3dom_checkpoint_save: writes it to host memory dst (capacity bytes).
 * k3dom_checkpoint_load: creates a handle from a checkpoint in host memory;
 *   K3DOM_ERR_INVALID_ARG for a buffer that is not a whole, intact
 *   checkpoint of this version (checked before any device work). */
k3dom_status k3dom_checkpoint_size(k3dom_handle* h, int64_t* bytes);
k3dom_status k3dom_checkpoint_save(k3dom_handle* h, void* dst, int64_t capacity, int64_t* written);
k3dom_status k3dom_checkpoint_load(const void* src, int64_t size, k3dom_handle** out);

/* Internal arrays of the last frame, for parity checks and diagnostics. */
typedef enum k3dom_array {
    K3DOM_ARR_PARTICLE_POS = 0,  /* 3*N float: x[N], y[N], z[N] (current set)     */
    K3DOM_ARR_PARTICLE_VEL = 1,  /* 3*N float                                      */
    K3DOM_ARR_PARTICLE_WEIGHT = 2, /* N float                                      */
    K3DOM_ARR_PARTICLE_KEY = 3,  /* N int32 voxel keys (copied with the particle)  */
    K3DOM_ARR_SORT_PERM = 4,     /* N int32: sorted position -> predicted index    */
    K3DOM_ARR_SORTED_KEY = 5,    /* N int32 keys after the stable sort             */
    K3DOM_ARR_RESAMPLE_IDX = 6,  /* N int32 indices into [persistent ++ newborn]   */
    K3DOM_ARR_NEWBORN_POS = 7,   /* 3*B float                                      */
    K3DOM_ARR_NEWBORN_VEL = 8,   /* 3*B float                                      */
    K3DOM_ARR_NEWBORN_WEIGHT = 9,/* B float (after the static penalty)             */
    K3DOM_ARR_NEWBORN_KEY = 10,  /* B int32                                        */
    K3DOM_ARR_BORN_MASS = 11,    /* V float                                        */
    K3DOM_ARR_DEL_OCC = 12,      /* V float occupied evidence of the last frame    */
    K3DOM_ARR_DEL_FREE = 13,     /* V float free evidence of the last frame        */
    K3DOM_ARR_SORTED_WEIGHT = 14 /* N float weights of the sorted set after the    */
                                 /* persistent update and static penalty            */
} k3dom_array;

/* Copies internal array `which` into host memory dst of capacity bytes. */
k3dom_status k3dom_read_array(k3dom_handle* h, int32_t which, void* dst, int64_t capacity);

/* Milliseconds spent in each stage of the last frame, measured with HIP
 * events on the handle's stream: [0] predict, [1] sort and ranges,
 * [2] point preparation and candidate lists, [3] kernel inference (BGKI),
 * [4] mass update, [5] persistent update, [6] newborn, [7] moments,
 * [8] resampling, [9] whole frame, [10] the kernel-inference gather kernel
 * alone.  Timing is recorded only after k3dom_set_timing(h, 1); with
 * k3dom_set_timing(h, 2) only the two events around the gather kernel are
 * recorded (entry [10]; the others read 0), which leaves the frame's launches
 * otherwise undisturbed.  0 turns timing off. */
k3dom_status k3dom_set_timing(k3dom_handle* h, int32_t enable);
k3dom_status k3dom_stage_times(k3dom_handle* h, float* ms, int32_t count);

/* Launch mode, no counterpart in the reference: with enable != 0 every normal
 * frame (after the first) replays one captured HIP graph from the frame
 * constants upload to the readback of the frame's scalars; the graph is
 * captured again when a buffer is reallocated.  Results are identical to the
 * direct launches.  Stage timing (k3dom_set_timing) forces direct launches. */
k3dom_status k3dom_set_graph(k3dom_handle* h, int32_t enable);

/* Diagnostic with no counterpart in the reference: number of (voxel, point)
 * pairs of the last frame whose occupied (out[0]) and free (out[1]) kernel
 * term is evaluated (distance < ls), i.e. the algorithmic work of the
 * kernel-inference gather.  Re-evaluates the last frame's candidate lists. */
k3dom_status k3dom_count_bgki_pairs(k3dom_handle* h, uint64_t* out2);

/* Work statistics of the last frame (diagnostic, no reference counterpart). */
typedef struct k3dom_stats {
    uint64_t points;             /* points of the frame                          */
    uint64_t level_entries[3];   /* candidate-list entries per level (coarse..tile);
                                    DS-PHD/MIB: [0] measurement-grid candidates */
    uint64_t gather_items;       /* work items of the kernel-inference gather    */
    uint64_t gather_split_items; /* items of tiles split into point chunks       */
    uint64_t capacity_regrowths; /* frames re-run after growing a list capacity (all frames) */
} k3dom_stats;
k3dom_status k3dom_get_stats(const k3dom_handle* h, k3dom_stats* out);

/* Launch-shape tuning, no counterpart in the reference: after a normal
 * single-sensor frame, times (1) the candidate-list build over that frame's
 * point records for several list segment lengths (parent-list entries per
 * work item: 128-512 at the top level, 256-2048 below), then (2) the
 * kernel-inference gather over that frame's work items for several block
 * sizes (4, 8 or 16 wavefronts) and grid sizes (1 to 8 blocks per compute
 * unit, at most 32 wavefronts per unit), and keeps the fastest of each.
 * Writes the chosen blocks per compute unit to *blocks_per_cu when not NULL.
 * Results do not depend on the choices. */
k3dom_status k3dom_autotune(k3dom_handle* h, int32_t* blocks_per_cu);

/* The launch shapes in use: shape[0] gather blocks per compute unit,
 * shape[1] wavefronts per gather block, shape[2] and shape[3] the list
 * segment lengths of the top level and of the two finer levels.  No
 * counterpart in the reference. */
k3dom_status k3dom_launch_shape(const k3dom_handle* h, int32_t shape[4]);
/* Sets the launch shapes (the layout of k3dom_launch_shape: blocks per
 * compute unit 1-16, wavefronts per gather block 4, 8 or 16, list segment
 * lengths 32-8192), e.g. those an earlier k3dom_autotune chose.  Between
 * frames only.  Results do not depend on the shapes. */
k3dom_status k3dom_set_launch_shape(k3dom_handle* h, const int32_t shape[4]);

/* ---------------------------------------------------------------------------
 * Batched maps.  No counterpart in the reference (one map per process): many
 * independent maps ("tiles", each with its own parameters, grid size, particle
 * count and point cloud) updated together.  Every map's prediction, sort and
 * candidate lists run concurrently on the maps' streams, the kernel-inference
 * gather of all maps runs in one persistent launch over a joint work queue,
 * then every map's tail runs concurrently.  Each map's results are identical
 * to those of the same map updated alone.
 * ------------------------------------------------------------------------- */
typedef struct k3dom_batch k3dom_batch;

/* params: array of n_maps parameter sets. */
k3dom_status k3dom_batch_create(const k3dom_params* params, int32_t n_maps, k3dom_batch** out);
k3dom_status k3dom_batch_destroy(k3dom_batch* b);

/* One frame of every map at time t: poses = 3 * n_maps sensor positions;
 * xs[i], ys[i], zs[i] the n[i] points of map i (host arrays). */
k3dom_status k3dom_batch_update(k3dom_batch* b, double t, const float* poses, const float* const* xs,
                                const float* const* ys, const float* const* zs, const int32_t* n);

/* Handle of map i, for readback (k3dom_read_cells, ...); owned by the batch. */
k3dom_status k3dom_batch_map(k3dom_batch* b, int32_t i, k3dom_handle** out);

/* Replay each normal batch frame as one captured HIP graph (fork and join
 * over the map streams); captured again when a buffer is reallocated. */
k3dom_status k3dom_batch_set_graph(k3dom_batch* b, int32_t enable);

/* As k3dom_autotune's gather sweep (block and grid sizes) for the joint
 * gather of the batch. */
k3dom_status k3dom_batch_autotune(k3dom_batch* b, int32_t* blocks_per_cu);

/* Host wall time of the last k3dom_batch_update, milliseconds. */
k3dom_status k3dom_batch_last_ms(const k3dom_batch* b, float* ms);
const char* k3dom_batch_last_error(const k3dom_batch* b);

/* ---------------------------------------------------------------------------
 * Differentiated window.  No counterpart in the reference (whose mass update is
 * forward-only, src/k3dom/kernel/mass_update.cu:152-183): a fixed sequence of
 * frames replayed on a fresh map, with the gradient of an occupancy loss with
 * respect to six parameters
 *     theta = {ls, sigma_occ, sigma_free, prior_free, prior_static, prior_dynamic}
 * (sigma_occ / sigma_free weight the occupied / free kernel terms; the
 * reference uses one sigma for both, src/k3dom/kernel/mass_update.cu:51,70).
 * loss = mean over frames of the mean over masked voxels of
 *     BCE(occ_mean, gt),  occ_mean = (s + d) / (f + s + d) clamped to [1e-6, 1 - 1e-6].
 * Sort order, counts, penalties and resampled indices of the forward frames
 * are frozen; the backward pass recomputes each frame's kernel-inference
 * evidence (and its length-scale derivative) instead of storing it.
 * ------------------------------------------------------------------------- */
typedef struct k3dom_window k3dom_window;

/* A window of `frames` frames over a map with these parameters (theta
 * overrides ls, sigma and the priors at each forward). */
k3dom_status k3dom_window_create(const k3dom_params* params, int32_t frames, k3dom_window** out);
k3dom_status k3dom_window_destroy(k3dom_window* w);

/* Input of frame f: time (s), sensor position (global frame, m), n points
 * relative to the sensor (host arrays), and the 0/1 ground-truth occupancy of
 * every voxel (grid_cell_count bytes, index x + gs*y + gs*gs*z). */
k3dom_status k3dom_window_set_frame(k3dom_window* w, int32_t f, double t, const float pose[3], const float* xs,
                                    const float* ys, const float* zs, int32_t n, const uint8_t* gt_occ);

/* Voxels entering the loss (grid_cell_count bytes, non-zero = in); NULL
 * selects every voxel. */
k3dom_status k3dom_window_set_mask(k3dom_window* w, const uint8_t* mask);

/* Frame f of the window from several LiDARs (host points in each LiDAR's
 * frame, as k3dom_update_grid_lidars): the frame's evidence and its
 * length-scale derivative are the LiDARs' (each with the sensor at its
 * origin), added in LiDAR order. */
k3dom_status k3dom_window_set_frame_lidars(k3dom_window* w, int32_t f, double t, const float pose[3],
                                          const k3dom_lidar* lidars, int32_t n_lidars, const uint8_t* gt_occ);

/* Runs the frames from a fresh map with parameters theta[6]; writes the loss. */
k3dom_status k3dom_window_forward(k3dom_window* w, const float theta[6], double* loss);

/* Number of forward passes run on w so far (a backward pass belongs to the
 * last one: a caller that keeps several losses checks this number). */
k3dom_status k3dom_window_forward_id(const k3dom_window* w, uint64_t* id);

/* Gradient of dloss * loss with respect to theta after a forward.  When
 * grad_prior_voxels is not NULL it receives 3 * grid_cell_count floats, the
 * gradient with respect to each voxel's initial (free, static, dynamic) mass. */
k3dom_status k3dom_window_backward(k3dom_window* w, double dloss, double grad[6], float* grad_prior_voxels);

/* Record of forward frame f: which = 0 masses after the frame (3 * V float:
 * free, static, dynamic), 1 predicted dynamic mass (V float), 2 frozen counts
 * (4 * V int32: particles inside the grid, of them not penalised, newborns,
 * of them not penalised), 3 the frame's map shift (3 int32). */
k3dom_status k3dom_window_read_record(k3dom_window* w, int32_t f, int32_t which, void* dst, int64_t capacity);

/* Milliseconds of the last forward and the last backward (host wall clock
 * around the whole pass, device synchronised). */
k3dom_status k3dom_window_times(const k3dom_window* w, float ms[2]);

const char* k3dom_window_last_error(const k3dom_window* w);

/* Text of the last error on h (or of the last failed create when h is NULL). */
const char* k3dom_last_error(const k3dom_handle* h);

#ifdef __cplusplus
}
#endif

#endif /* K3DOM_H */
