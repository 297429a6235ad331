// Synthetic box scene with a ray-cast LiDAR (k3_scene.cpp): the descriptor
// and the C entry points used by the benchmark, the tests and the sanitizer
// build.
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
typedef struct k3scene_desc {
    uint64_t seed;
    int32_t n_static, n_moving;
    float side_min, side_max;      // box edge lengths ~ U[side_min, side_max] per axis
    float speed_min, speed_max;    // moving boxes: speed ~ U[.,.], heading ~ U[0, 2pi)
    float xmin, xmax, ymin, ymax;  // box centres ~ U over this rectangle (global frame)
    float ground_z;                // boxes stand on this plane
    float clearance;               // box footprints keep this distance from the ego path
    float ego_x0, ego_y0, ego_z0;  // ego (sensor) position at frame 0
    float ego_vx, ego_vy;          // ego velocity (m/s)
    int32_t beams, az_bins;
    float elev_min_deg, elev_max_deg;
    float max_range, range_noise;
    float frame_dt;                // seconds between frames
    double t0;                     // time stamp of frame 0
    int32_t ground;                // 1: the plane z = ground_z also returns points
    float walls;                   // > 0: every ray also hits the walls of a square of this
                                   // half-size (metres) centred on the sensor, unbounded in z
} k3scene_desc;
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
void* k3scene_create(const k3scene_desc* d);
void k3scene_destroy(void* h);
int k3scene_frame(void* h, int f, float* xs, float* ys, float* zs, int cap, double* pose_out);
int k3scene_frame_lidar(void* h, int f, const double* mount, double yaw, float* xs, float* ys, float* zs, int cap,
                        double* pose_out);
#ifdef __cplusplus
}
#endif
