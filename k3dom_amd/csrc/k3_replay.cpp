// Frame-replay driver over the C++ drop-in headers: the per-frame sequence
// of the reference's ROS node callbacks (src/k3dom/k3dom_node.cpp:64-97,
// src/dsphdmib/dsphdmib_node.cpp:61-94): updatePose with the sensor position,
// then updateGrid with the sensor-relative points, then copyCellsToHost.
// Built twice: k3dom_replay over include/k3dom/dom.h and, with
// K3_REPLAY_DSPHDMIB defined, k3dom_replay_dsphdmib over include/dsphdmib/dom.h.
//
//   k3dom_replay PARAMS FRAMES OUT [stages]
//     stages  (K3DOM only) drive each frame through the stage methods in the
//             order DOM::updateGrid calls them (src/k3dom/dom.cu:29-50):
//             initializeParticles for a first frame, else particlePrediction,
//             particleAssignment, gridCellOccupancyUpdate,
//             updatePersistentParticles, initializeNewParticles,
//             statisticalMoments, resampling
//     PARAMS  raw k3dom_params record (sizeof(k3dom_params) bytes)
//     FRAMES  "K3FR", int32 frame count, then per frame: double t,
//             float pose[3], int32 n, float x[n], float y[n], float z[n]
//     OUT     int32 grid_cell_count, then that many 68-byte GridCell records
//             after the last frame (DS-PHD/MIB: then as many MeasurementCell
//             records of the last frame)
// Exit status 0 on success; any failed call is reported and gives status 1.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef K3_REPLAY_DSPHDMIB
#include "dsphdmib/dom.h"
#else
#include "k3dom/dom.h"
#endif

namespace {
bool read_exact(FILE* f, void* dst, size_t n) { return std::fread(dst, 1, n, f) == n; }
}  // namespace

int main(int argc, char** argv) {
    if (argc != 4 && argc != 5) {
        std::fprintf(stderr, "usage: %s PARAMS FRAMES OUT [stages]\n", argv[0]);
        return 2;
    }
#ifdef K3_REPLAY_DSPHDMIB
    const bool stages = false;
#else
    const bool stages = argc == 5 && std::strcmp(argv[4], "stages") == 0;
#endif
    double prev_t = -1.0;  // the updateGrid time bookkeeping of src/k3dom/dom.cu:31,46
    k3dom_params cp;
    FILE* pf = std::fopen(argv[1], "rb");
    if (!pf || !read_exact(pf, &cp, sizeof(cp))) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::fclose(pf);
    dom::DOM_c::Params p{};
    p.size = cp.size;
    p.size_z = cp.size_z;
    p.mass_scale = cp.mass_scale;
    p.resolution = cp.resolution;
    p.particle_count = cp.particle_count;
    p.new_born_particle_count = cp.new_born_particle_count;
    p.persistence_prob = cp.persistence_prob;
    p.stddev_process_noise_position = cp.stddev_process_noise_position;
    p.stddev_process_noise_velocity = cp.stddev_process_noise_velocity;
    p.birth_prob = cp.birth_prob;
    p.stddev_velocity = cp.stddev_velocity;
    p.init_max_velocity = cp.init_max_velocity;
    p.particle_min_vel = cp.particle_min_vel;
    p.sensor_off_x = cp.sensor_off_x;
    p.sensor_off_y = cp.sensor_off_y;
    p.sensor_off_z = cp.sensor_off_z;
    p.map_shift_thresh = cp.map_shift_thresh;
    p.prior_free = cp.prior_free;
    p.prior_static = cp.prior_static;
    p.prior_dynamic = cp.prior_dynamic;
    p.prior_occ = cp.prior_occ;
    p.sigma = cp.sigma;
    p.ls = cp.ls;
    p.gamma = cp.gamma;
    p.max_range = cp.max_range;
    p.meas_occ_stddev = cp.meas_occ_stddev;
    p.seed = cp.seed;

    FILE* ff = std::fopen(argv[2], "rb");
    char magic[4];
    int32_t nframes = 0;
    if (!ff || !read_exact(ff, magic, 4) || std::memcmp(magic, "K3FR", 4) != 0 || !read_exact(ff, &nframes, 4)) {
        std::fprintf(stderr, "cannot read %s\n", argv[2]);
        return 2;
    }
    dom::DOM grid_map(p);
    if (grid_map.lastStatus() != K3DOM_OK) return 1;
    std::vector<float> xs, ys, zs;
    for (int32_t f = 0; f < nframes; ++f) {
        double t;
        float pose[3];
        int32_t n;
        if (!read_exact(ff, &t, 8) || !read_exact(ff, pose, 12) || !read_exact(ff, &n, 4) || n < 0) {
            std::fprintf(stderr, "truncated frame %d\n", f);
            return 2;
        }
        xs.resize((size_t)n);
        ys.resize((size_t)n);
        zs.resize((size_t)n);
        if (!read_exact(ff, xs.data(), 4 * (size_t)n) || !read_exact(ff, ys.data(), 4 * (size_t)n) ||
            !read_exact(ff, zs.data(), 4 * (size_t)n)) {
            std::fprintf(stderr, "truncated frame %d\n", f);
            return 2;
        }
        grid_map.updatePose(pose[0], pose[1], pose[2]);
        if (grid_map.lastStatus() != K3DOM_OK) return 1;
        if (!stages) {
            grid_map.updateGrid(t, xs.data(), ys.data(), zs.data(), n);
            if (grid_map.lastStatus() != K3DOM_OK) return 1;
            continue;
        }
#ifndef K3_REPLAY_DSPHDMIB
        if (prev_t > 0) {
            const float dt = (float)(t - prev_t);
            grid_map.particlePrediction(dt);
            grid_map.particleAssignment();
            grid_map.gridCellOccupancyUpdate(dt, xs, ys, zs);
            grid_map.updatePersistentParticles();
            grid_map.initializeNewParticles();
            grid_map.statisticalMoments();
            grid_map.resampling();
        } else {
            grid_map.initializeParticles(xs, ys, zs);
        }
        if (grid_map.lastStatus() != K3DOM_OK) return 1;
        prev_t = t;
#endif
    }
    std::fclose(ff);
    grid_map.copyCellsToHost();
    if (grid_map.lastStatus() != K3DOM_OK) return 1;
    FILE* of = std::fopen(argv[3], "wb");
    const int32_t V = grid_map.grid_cell_count;
    if (!of || std::fwrite(&V, 4, 1, of) != 1 ||
        std::fwrite(grid_map.grid_cell_array_host, sizeof(dom::GridCell), (size_t)V, of) != (size_t)V) {
        std::fprintf(stderr, "cannot write %s\n", argv[3]);
        return 2;
    }
#ifdef K3_REPLAY_DSPHDMIB
    std::vector<dom::MeasurementCell> meas;
    grid_map.copyMeasGridToHost(meas);
    if (grid_map.lastStatus() != K3DOM_OK) return 1;
    if (std::fwrite(meas.data(), sizeof(dom::MeasurementCell), meas.size(), of) != meas.size()) {
        std::fprintf(stderr, "cannot write %s\n", argv[3]);
        return 2;
    }
#endif
    std::fclose(of);
    std::printf("replayed %d frames, %d cells\n", nframes, V);
    return 0;
}
