// Synthetic workload for the K3DOM update: axis-aligned boxes (static and
// moving at constant velocity) observed by a ray-cast spinning LiDAR carried by
// an ego vehicle.  The reference ships no data (its rosbags are absent), so
// benchmarks and parity tests drive the filter with this seeded scene.
//
// Output convention matches what the reference node hands to DOM::updateGrid
// (src/k3dom/k3dom_node.cpp:73-85): points relative to the sensor origin in
// map-aligned axes, plus the sensor position in the global frame for
// DOM_c::updatePose.  Points are emitted column by column (azimuth outer, beam
// inner), the firing order of a spinning LiDAR.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "k3_scene.h"

namespace {

// SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) for the scene layout and the range noise.
static inline uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline double u01(uint64_t h) { return (double)(h >> 11) * 0x1p-53; }

struct Box {
    double cx, cy, hx, hy, hz;  // centre (xy at t=0), half extents
    double vx, vy;
};

struct Scene {
    k3scene_desc d;
    std::vector<Box> boxes;

    explicit Scene(const k3scene_desc& desc) : d(desc) {
        uint64_t s = d.seed * 0x100000001B3ull + 17;
        auto next = [&]() { s = splitmix(s); return u01(s); };
        const int total = d.n_static + d.n_moving;
        for (int b = 0; b < total; ++b) {
            Box bx;
            for (int tries = 0; tries < 1000; ++tries) {
                bx.hx = 0.5 * (d.side_min + (d.side_max - d.side_min) * next());
                bx.hy = 0.5 * (d.side_min + (d.side_max - d.side_min) * next());
                bx.hz = 0.5 * (d.side_min + (d.side_max - d.side_min) * next());
                bx.cx = d.xmin + (d.xmax - d.xmin) * next();
                bx.cy = d.ymin + (d.ymax - d.ymin) * next();
                bx.vx = bx.vy = 0.0;
                if (b >= d.n_static) {
                    double sp = d.speed_min + (d.speed_max - d.speed_min) * next();
                    double hd = 2.0 * M_PI * next();
                    bx.vx = sp * std::cos(hd);
                    bx.vy = sp * std::sin(hd);
                }
                if (clear_of_path(bx)) break;
            }
            boxes.push_back(bx);
        }
    }

    // distance from the ego path (a segment over the run) to the box footprint,
    // checked at t = 0 only for static boxes and over the run for moving ones
    bool clear_of_path(const Box& b) const {
        const int steps = 64;
        const double T = 200.0 * d.frame_dt;
        for (int k = 0; k <= steps; ++k) {
            double t = T * k / steps;
            double ex = d.ego_x0 + d.ego_vx * t, ey = d.ego_y0 + d.ego_vy * t;
            double bx = b.cx + b.vx * t, by = b.cy + b.vy * t;
            double dx = std::max(0.0, std::fabs(ex - bx) - b.hx);
            double dy = std::max(0.0, std::fabs(ey - by) - b.hy);
            if (dx * dx + dy * dy < (double)d.clearance * d.clearance) return false;
            if (b.vx == 0.0 && b.vy == 0.0 && d.ego_vx == 0.0 && d.ego_vy == 0.0) break;
        }
        return true;
    }

    void pose(int frame, double* x, double* y, double* z, double* t) const {
        double tr = frame * (double)d.frame_dt;
        *x = d.ego_x0 + d.ego_vx * tr;
        *y = d.ego_y0 + d.ego_vy * tr;
        *z = d.ego_z0;
        *t = d.t0 + tr;
    }

    // ray/box slab test: nearest entry distance in (0, tmax], or -1
    static inline double hit(double ox, double oy, double oz, double dx, double dy, double dz,
                             double lx, double ly, double lz, double ux, double uy, double uz, double tmax) {
        double t0 = 0.0, t1 = tmax;
        const double o[3] = {ox, oy, oz}, dd[3] = {dx, dy, dz}, lo[3] = {lx, ly, lz}, hi[3] = {ux, uy, uz};
        for (int a = 0; a < 3; ++a) {
            if (std::fabs(dd[a]) < 1e-12) {
                if (o[a] < lo[a] || o[a] > hi[a]) return -1.0;
                continue;
            }
            double inv = 1.0 / dd[a];
            double ta = (lo[a] - o[a]) * inv, tb = (hi[a] - o[a]) * inv;
            if (ta > tb) std::swap(ta, tb);
            t0 = std::max(t0, ta);
            t1 = std::min(t1, tb);
            if (t0 > t1) return -1.0;
        }
        return t0 > 0.0 ? t0 : -1.0;
    }

    int frame(int f, float* xs, float* ys, float* zs, int cap, double* pose_out) const {
        const double mount[3] = {0.0, 0.0, 0.0};
        return frame_from(f, mount, 0.0, xs, ys, zs, cap, pose_out);
    }

    // A LiDAR mounted at `mount` (metres, world axes, from the ego position)
    // and turned by yaw (radians) about z: its points in its own frame
    // (azimuth 0 along its x axis); pose_out = the ego position and time.
    int frame_from(int f, const double* mount, double yaw, float* xs, float* ys, float* zs, int cap,
                   double* pose_out) const {
        double px, py, pz, t;
        pose(f, &px, &py, &pz, &t);
        pose_out[0] = px; pose_out[1] = py; pose_out[2] = pz; pose_out[3] = t;
        const double ex = px + mount[0], ey = py + mount[1], ez = pz + mount[2];
        const double cyaw = std::cos(yaw), syaw = std::sin(yaw);
        const double tr = f * (double)d.frame_dt;
        // boxes at this time, with the azimuth interval each one covers
        struct Live { double lx, ly, lz, ux, uy, uz; int a0, a1; bool all; };
        std::vector<Live> live;
        const double R = d.max_range;
        const int A = d.az_bins;
        for (const Box& b : boxes) {
            double cx = b.cx + b.vx * tr, cy = b.cy + b.vy * tr;
            Live L;
            L.lx = cx - b.hx; L.ux = cx + b.hx;
            L.ly = cy - b.hy; L.uy = cy + b.hy;
            L.lz = d.ground_z; L.uz = d.ground_z + 2.0 * b.hz;
            double nx = std::max(L.lx - ex, std::max(0.0, ex - L.ux));
            double ny = std::max(L.ly - ey, std::max(0.0, ey - L.uy));
            if (nx * nx + ny * ny > R * R) continue;
            L.all = (ex >= L.lx && ex <= L.ux && ey >= L.ly && ey <= L.uy);
            L.a0 = 0; L.a1 = A - 1;
            if (!L.all) {
                // angular interval of the four footprint corners (it spans < pi)
                double ang[4] = {std::atan2(L.ly - ey, L.lx - ex) - yaw, std::atan2(L.ly - ey, L.ux - ex) - yaw,
                                 std::atan2(L.uy - ey, L.lx - ex) - yaw, std::atan2(L.uy - ey, L.ux - ex) - yaw};
                double ref = ang[0], lo = 0.0, hi = 0.0;
                for (int k = 1; k < 4; ++k) {
                    double dlt = std::remainder(ang[k] - ref, 2.0 * M_PI);
                    lo = std::min(lo, dlt);
                    hi = std::max(hi, dlt);
                }
                double a_lo = (ref + lo) / (2.0 * M_PI) * A, a_hi = (ref + hi) / (2.0 * M_PI) * A;
                L.a0 = (int)std::floor(a_lo) - 1;
                L.a1 = (int)std::ceil(a_hi) + 1;
            }
            live.push_back(L);
        }
        // per-azimuth lists of candidate boxes
        std::vector<std::vector<int>> cols(A);
        for (int i = 0; i < (int)live.size(); ++i) {
            const Live& L = live[i];
            if (L.all || L.a1 - L.a0 + 1 >= A) {
                for (int a = 0; a < A; ++a) cols[a].push_back(i);
            } else {
                for (int a = L.a0; a <= L.a1; ++a) cols[((a % A) + A) % A].push_back(i);
            }
        }
        const int NB = d.beams;
        std::vector<float> rng((size_t)A * NB, -1.0f);
        std::vector<double> ce(NB), se(NB);
        for (int b = 0; b < NB; ++b) {
            double e = (NB == 1) ? d.elev_min_deg
                                 : d.elev_min_deg + (d.elev_max_deg - d.elev_min_deg) * b / (NB - 1.0);
            e *= M_PI / 180.0;
            ce[b] = std::cos(e);
            se[b] = std::sin(e);
        }
#pragma omp parallel for schedule(static)
        for (int a = 0; a < A; ++a) {
            double az = 2.0 * M_PI * a / A, ca = std::cos(az), sa = std::sin(az);
            // world direction of the LiDAR's azimuth a (a yaw of 0 leaves it unchanged)
            const double wca = yaw == 0.0 ? ca : cyaw * ca - syaw * sa, wsa = yaw == 0.0 ? sa : syaw * ca + cyaw * sa;
            for (int b = 0; b < NB; ++b) {
                double dx = ce[b] * wca, dy = ce[b] * wsa, dz = se[b];
                double best = -1.0;
                for (int i : cols[a]) {
                    const Live& L = live[i];
                    double th = hit(ex, ey, ez, dx, dy, dz, L.lx, L.ly, L.lz, L.ux, L.uy, L.uz, R);
                    if (th > 0.0 && (best < 0.0 || th < best)) best = th;
                }
                if (d.ground && dz < -1e-9) {
                    double tg = ((double)d.ground_z - ez) / dz;
                    if (tg > 0.0 && tg <= R && (best < 0.0 || tg < best)) best = tg;
                }
                if (d.walls > 0.0f) {
                    const double hx = std::max(std::fabs(dx), std::fabs(dy));
                    if (hx > 1e-12) {
                        const double tw = (double)d.walls / hx;
                        if (tw <= R && (best < 0.0 || tw < best)) best = tw;
                    }
                }
                if (best > 0.0) {
                    uint64_t h = splitmix(d.seed ^ splitmix(((uint64_t)f << 32) ^ (uint64_t)(a * NB + b)));
                    double u1 = (double)((h >> 40) + 1) * 0x1p-24, u2 = (double)((h >> 16) & 0xffffff) * 0x1p-24;
                    double n = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
                    double r = best + d.range_noise * n;
                    if (r > 1e-3) rng[(size_t)a * NB + b] = (float)r;
                }
            }
        }
        int n = 0;
        for (int a = 0; a < A; ++a) {
            double az = 2.0 * M_PI * a / A, ca = std::cos(az), sa = std::sin(az);
            for (int b = 0; b < NB; ++b) {
                float r = rng[(size_t)a * NB + b];
                if (r < 0.0f) continue;
                if (n < cap) {
                    xs[n] = (float)(r * ce[b] * ca);
                    ys[n] = (float)(r * ce[b] * sa);
                    zs[n] = (float)(r * se[b]);
                }
                ++n;
            }
        }
        return n;
    }

    // 1 where the voxel centre (global frame) lies inside a box at frame f
    void occupancy(int f, int gs, int gsz, float res, float cxm, float cym, float czm, float size, float size_z,
                   uint8_t* out) const {
        const double tr = f * (double)d.frame_dt;
        const size_t V = (size_t)gs * gs * gsz;
        std::memset(out, 0, V);
        for (const Box& b : boxes) {
            double cx = b.cx + b.vx * tr, cy = b.cy + b.vy * tr;
            double lo[3] = {cx - b.hx, cy - b.hy, d.ground_z}, hi[3] = {cx + b.hx, cy + b.hy, d.ground_z + 2 * b.hz};
            for (int z = 0; z < gsz; ++z) {
                double wz = (z + 0.5) * res - size_z / 2.0 + czm;
                if (wz < lo[2] || wz > hi[2]) continue;
                for (int y = 0; y < gs; ++y) {
                    double wy = (y + 0.5) * res - size / 2.0 + cym;
                    if (wy < lo[1] || wy > hi[1]) continue;
                    for (int x = 0; x < gs; ++x) {
                        double wx = (x + 0.5) * res - size / 2.0 + cxm;
                        if (wx >= lo[0] && wx <= hi[0]) out[(size_t)x + (size_t)gs * y + (size_t)gs * gs * z] = 1;
                    }
                }
            }
        }
    }
};

}  // namespace

extern "C" {
void* k3scene_create(const k3scene_desc* d) { return d ? new Scene(*d) : nullptr; }
void k3scene_destroy(void* h) { delete (Scene*)h; }
// Writes up to cap points of frame f; returns the number of points of the
// frame (which may exceed cap).  pose_out = {x, y, z, t}.
int k3scene_frame(void* h, int f, float* xs, float* ys, float* zs, int cap, double* pose_out) {
    return ((Scene*)h)->frame(f, xs, ys, zs, cap, pose_out);
}
// The same for a LiDAR mounted at mount[3] (from the ego, world axes) and
// turned by yaw: points in the LiDAR's own frame.
int k3scene_frame_lidar(void* h, int f, const double* mount, double yaw, float* xs, float* ys, float* zs, int cap,
                        double* pose_out) {
    return ((Scene*)h)->frame_from(f, mount, yaw, xs, ys, zs, cap, pose_out);
}
void k3scene_occupancy(void* h, int f, int gs, int gsz, float res, float cx, float cy, float cz, float size,
                       float size_z, uint8_t* out) {
    ((Scene*)h)->occupancy(f, gs, gsz, res, cx, cy, cz, size, size_z, out);
}
int k3scene_box_count(void* h) { return (int)((Scene*)h)->boxes.size(); }
}
