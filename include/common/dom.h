// Common part of the C++ drop-in for the reference's filter classes
// (youngjae-min/k3dom, include/common/dom.h:16-148,
// include/common/dom_types.h:12-65), header-only, over the C ABI of k3dom.h.
// The method classes dom::DOM live in k3dom/dom.h (K3DOM) and dsphdmib/dom.h
// (DS-PHD/MIB); a program includes one of them, as with the reference.
//
// Same namespace, class names, member function names, argument meaning and
// units as the reference: a caller written against dom::DOM compiles against
// this header and links libk3dom.so instead of the reference's .cu sources.
// Differences a caller sees:
//   - grid_cell_array (a device array of GridCell in the reference) does not
//     exist: voxel state lives in device memory as structure-of-arrays.  Call
//     copyCellsToHost() where the reference's caller did
//     cudaMemcpy(grid_cell_array_host, grid_cell_array, ...)
//     (src/k3dom/k3dom_node.cpp:131-132,162-163); it fills
//     grid_cell_array_host with GridCell records of the reference's layout.
//   - Params carries one more field, `seed`, of the counter-based generator
//     (the reference seeds cuRAND with 123456, src/common/dom.cu:96); the
//     default member initialiser keeps aggregate initialisation working.
//   - updateGrid also accepts a double time stamp with plain host arrays.
// Errors follow the reference's convention (include/common/cuda_utils.h:9-25):
// a failed call prints the error to stderr and the program continues;
// lastStatus() returns the k3dom_status of the last call for callers that
// want to check.
#pragma once

#include <algorithm>
#include <cassert>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../k3dom.h"
#include "mappoint.h"

namespace dom {

// include/common/dom_types.h:12-58 (68 bytes, same field order).
struct GridCell {
    int start_idx;
    int end_idx;
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

    inline float total_mass() const { return free_mass + static_mass + dynamic_mass; }
    inline float occ_mean() const { return (static_mass + dynamic_mass) / total_mass(); }
    inline float dyn_mean() const { return dynamic_mass / total_mass(); }
    float calcVar() const {
        const float mass_sum = total_mass();
        if (mass_sum == 0.0f) return 1e6f;
        const float mean = std::max(std::max(free_mass, static_mass), dynamic_mass) / mass_sum;
        return mean * (1.0f - mean) / (mass_sum + 1.0f);
    }
    inline float occ_prob() const { return occ_mass + (1.0f - occ_mass - free_mass) / 2.0f; }
    inline float mean_vel_sq() const {
        return mean_x_vel * mean_x_vel + mean_y_vel * mean_y_vel + mean_z_vel * mean_z_vel;
    }
};
static_assert(sizeof(GridCell) == sizeof(k3dom_cell), "GridCell must match k3dom_cell");

// include/common/dom_types.h:67-74: one voxel of the DS-PHD/MIB measurement grid.
struct MeasurementCell {
    float free_mass;
    float occ_mass;
};

class DOM_c {
public:
    // include/common/dom.h:19-52, same fields, order and units.
    struct Params {
        float size;
        float size_z;
        float mass_scale;
        float resolution;
        int particle_count;
        int new_born_particle_count;
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
        int map_shift_thresh;
        float prior_free;
        float prior_static;
        float prior_dynamic;
        float prior_occ;
        float sigma;
        float ls;
        float gamma;
        float max_range;
        float meas_occ_stddev;
        unsigned long long seed = 123456ull;
    };

    explicit DOM_c(const Params& p) : DOM_c(p, K3DOM_METHOD_K3DOM) {}
    // The mass-update method (K3DOM_METHOD_*) is chosen once, at creation.
    DOM_c(const Params& p, int method) : params(p) {
        k3dom_params c = toC(p);
        check(k3dom_create_method(&c, method, &handle_), "DOM_c::DOM_c");
        k3dom_info info{};
        if (handle_ && k3dom_get_info(handle_, &info) == K3DOM_OK) {
            grid_size = info.grid_size;
            grid_size_z = info.grid_size_z;
            grid_cell_count = info.grid_cell_count;
            particle_count = info.particle_count;
            new_born_particle_count = info.new_born_particle_count;
            grid_cell_array_host = new GridCell[(size_t)grid_cell_count];
        }
    }
    ~DOM_c() {
        delete[] grid_cell_array_host;
        k3dom_destroy(handle_);
    }
    DOM_c(const DOM_c&) = delete;
    DOM_c& operator=(const DOM_c&) = delete;

    // src/common/dom.cu:111-158
    void updatePose(float new_x, float new_y, float new_z) {
        check(k3dom_update_pose(handle_, new_x, new_y, new_z), "updatePose");
    }

    int getGridSize() const { return grid_size; }
    float getResolution() const { return params.resolution; }
    float getCenterPositionX() const { return info().center_pos[0]; }
    float getCenterPositionY() const { return info().center_pos[1]; }
    float getCenterPositionZ() const { return info().center_pos[2]; }
    float getSensorPositionX() const { return info().sensor_pos[0]; }
    float getSensorPositionY() const { return info().sensor_pos[1]; }
    float getSensorPositionZ() const { return info().sensor_pos[2]; }

    // include/common/dom.h:70-81 (same float expressions)
    float indexToCoordX(int i) {
        return (i % grid_size + 0.5f) * params.resolution - params.size / 2.0f + getCenterPositionX();
    }
    float indexToCoordY(int i) {
        return ((i % (grid_size * grid_size)) / grid_size + 0.5f) * params.resolution - params.size / 2.0f +
               getCenterPositionY();
    }
    float indexToCoordZ(int i) {
        return (i / (grid_size * grid_size) + 0.5f) * params.resolution - params.size_z / 2.0f +
               getCenterPositionZ();
    }

    // The reference's stage methods (include/common/dom.h:89-94), called in
    // the order of DOM::updateGrid (k3dom.h: k3dom_particle_prediction ...).
    void particlePrediction(float dt) { check(k3dom_particle_prediction(handle_, dt), "particlePrediction"); }
    void particleAssignment() { check(k3dom_particle_assignment(handle_), "particleAssignment"); }
    void updatePersistentParticles() {
        check(k3dom_update_persistent_particles(handle_), "updatePersistentParticles");
    }
    void initializeNewParticles() { check(k3dom_initialize_new_particles(handle_), "initializeNewParticles"); }
    void statisticalMoments() { check(k3dom_statistical_moments(handle_), "statisticalMoments"); }
    void resampling() { check(k3dom_resampling(handle_), "resampling"); }

    // The reference never initialises or advances its iteration counter
    // (src/common/dom.cu:29-34); this one counts the frames processed.
    int getIteration() const { return info().iteration; }

    // Fills grid_cell_array_host with every voxel (index x + gs*y + gs*gs*z).
    void copyCellsToHost() {
        check(k3dom_read_cells(handle_, reinterpret_cast<k3dom_cell*>(grid_cell_array_host)), "copyCellsToHost");
    }

    // One frame from several LiDARs (k3dom_update_grid_lidars): each with its
    // points in its own frame, its rotation to the map's axes and its origin
    // relative to the pose of updatePose; the transform runs on the GPU.
    void updateGridLidars(double t, const std::vector<k3dom_lidar>& lidars, bool points_on_device = false) {
        check(k3dom_update_grid_lidars(handle_, t, lidars.data(), (int32_t)lidars.size(), points_on_device ? 1 : 0),
              "updateGridLidars");
    }

    // The node's /map_data records (src/k3dom/k3dom_node.cpp:161-184,
    // src/dsphdmib/dsphdmib_node.cpp:153-177) built on the GPU: every voxel
    // in index order, or with visible = true only the voxels passing the
    // visualisation filter (min_occ, min_mass) of publish_map, in index order.
    void copyMapPointsToHost(std::vector<MapPoint>& out, bool visible = false, float min_occ = 0.0f,
                             float min_mass = 0.0f) {
        out.resize((size_t)grid_cell_count);
        int64_t n = 0;
        check(k3dom_export_mappoints(handle_, visible ? K3DOM_MAPPOINTS_VISIBLE : K3DOM_MAPPOINTS_ALL, min_occ,
                                     min_mass, reinterpret_cast<k3dom_mappoint*>(out.data()), (int64_t)out.size(),
                                     &n),
              "copyMapPointsToHost");
        out.resize(status_ == K3DOM_OK ? (size_t)n : 0);
    }

    // Checkpoint of the filter state (k3dom_checkpoint_save): parameters,
    // frame counter, time, pose, voxels and particles, versioned and checksummed.
    bool saveCheckpoint(std::vector<unsigned char>& out) {
        int64_t n = 0, written = 0;
        check(k3dom_checkpoint_size(handle_, &n), "saveCheckpoint");
        if (status_ != K3DOM_OK) return false;
        out.resize((size_t)n);
        check(k3dom_checkpoint_save(handle_, out.data(), n, &written), "saveCheckpoint");
        return status_ == K3DOM_OK;
    }
    // Replaces this map with the one a checkpoint holds (its parameters and
    // method included); later frames continue it bit for bit.
    bool loadCheckpoint(const void* data, size_t size) {
        k3dom_handle* nh = nullptr;
        status_ = k3dom_checkpoint_load(data, (int64_t)size, &nh);
        if (status_ != K3DOM_OK) {
            std::fprintf(stderr, "k3dom: loadCheckpoint failed: %s\n", k3dom_last_error(nullptr));
            return false;
        }
        k3dom_destroy(handle_);
        handle_ = nh;
        k3dom_params c{};
        std::memcpy(&c, static_cast<const unsigned char*>(data) + 128, sizeof(c));
        params = fromC(c);
        k3dom_info i = info();
        grid_size = i.grid_size;
        grid_size_z = i.grid_size_z;
        particle_count = i.particle_count;
        new_born_particle_count = i.new_born_particle_count;
        if (i.grid_cell_count != grid_cell_count) {
            delete[] grid_cell_array_host;
            grid_cell_count = i.grid_cell_count;
            grid_cell_array_host = new GridCell[(size_t)grid_cell_count];
        }
        return true;
    }

    k3dom_status lastStatus() const { return status_; }
    k3dom_handle* handle() const { return handle_; }

    Params params;
    GridCell* grid_cell_array_host = nullptr;
    int grid_size = 0;
    int grid_size_z = 0;
    int grid_cell_count = 0;
    int particle_count = 0;
    int new_born_particle_count = 0;

protected:
    void check(k3dom_status s, const char* where) {
        status_ = s;
        if (s != K3DOM_OK) std::fprintf(stderr, "k3dom: %s failed: %s\n", where, k3dom_last_error(handle_));
    }
    k3dom_info info() const {
        k3dom_info i{};
        if (handle_) k3dom_get_info(handle_, &i);
        return i;
    }
    static Params fromC(const k3dom_params& c) {
        Params p{};
        p.size = c.size;
        p.size_z = c.size_z;
        p.mass_scale = c.mass_scale;
        p.resolution = c.resolution;
        p.particle_count = c.particle_count;
        p.new_born_particle_count = c.new_born_particle_count;
        p.persistence_prob = c.persistence_prob;
        p.stddev_process_noise_position = c.stddev_process_noise_position;
        p.stddev_process_noise_velocity = c.stddev_process_noise_velocity;
        p.birth_prob = c.birth_prob;
        p.stddev_velocity = c.stddev_velocity;
        p.init_max_velocity = c.init_max_velocity;
        p.particle_min_vel = c.particle_min_vel;
        p.sensor_off_x = c.sensor_off_x;
        p.sensor_off_y = c.sensor_off_y;
        p.sensor_off_z = c.sensor_off_z;
        p.map_shift_thresh = c.map_shift_thresh;
        p.prior_free = c.prior_free;
        p.prior_static = c.prior_static;
        p.prior_dynamic = c.prior_dynamic;
        p.prior_occ = c.prior_occ;
        p.sigma = c.sigma;
        p.ls = c.ls;
        p.gamma = c.gamma;
        p.max_range = c.max_range;
        p.meas_occ_stddev = c.meas_occ_stddev;
        p.seed = c.seed;
        return p;
    }
    static k3dom_params toC(const Params& p) {
        k3dom_params c{};
        c.size = p.size;
        c.size_z = p.size_z;
        c.mass_scale = p.mass_scale;
        c.resolution = p.resolution;
        c.particle_count = p.particle_count;
        c.new_born_particle_count = p.new_born_particle_count;
        c.persistence_prob = p.persistence_prob;
        c.stddev_process_noise_position = p.stddev_process_noise_position;
        c.stddev_process_noise_velocity = p.stddev_process_noise_velocity;
        c.birth_prob = p.birth_prob;
        c.stddev_velocity = p.stddev_velocity;
        c.init_max_velocity = p.init_max_velocity;
        c.particle_min_vel = p.particle_min_vel;
        c.sensor_off_x = p.sensor_off_x;
        c.sensor_off_y = p.sensor_off_y;
        c.sensor_off_z = p.sensor_off_z;
        c.map_shift_thresh = p.map_shift_thresh;
        c.prior_free = p.prior_free;
        c.prior_static = p.prior_static;
        c.prior_dynamic = p.prior_dynamic;
        c.prior_occ = p.prior_occ;
        c.sigma = p.sigma;
        c.ls = p.ls;
        c.gamma = p.gamma;
        c.max_range = p.max_range;
        c.meas_occ_stddev = p.meas_occ_stddev;
        c.seed = p.seed;
        return c;
    }

    k3dom_handle* handle_ = nullptr;
    k3dom_status status_ = K3DOM_OK;
};

}  // namespace dom
