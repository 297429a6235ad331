// C ABI of the K3DOM update on MI355X: handle lifetime, device buffers, the
// per-frame launch sequence and readback.  Declarations and the reference
// interface each entry point replaces are in include/k3dom.h.
#include "../../include/k3dom.h"
#include "k3_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

using namespace k3;

namespace {

thread_local std::string g_create_error;

struct Level {
    LevelGeom g;
    int nboxes = 0;
    uint32_t *len = nullptr, *off = nullptr, *list = nullptr;
    size_t cap = 0;  // entries of list
};

enum { ST_PREDICT = 0, ST_SORT, ST_LISTS, ST_BGKI, ST_MASS, ST_PERSIST, ST_NEWBORN, ST_MOMENTS, ST_RESAMPLE,
       ST_FRAME, ST_COUNT };

}  // namespace

struct k3dom_handle {
    k3dom_params p{};
    int gs = 0, gsz = 0, V = 0, N = 0, B = 0, key_bits = 1;
    int ntx = 0, nty = 0, ntz = 0, ntiles = 0;
    int device = 0, num_cus = 256;
    hipStream_t stream = nullptr;

    std::vector<void*> allocs;
    VoxelArrays vox[2]{};
    int vcur = 0;
    float *born = nullptr, *del_occ = nullptr, *del_free = nullptr;
    ParticleArrays P[2]{};  // 0: current set, 1: sorted set
    ParticleArrays NB{};
    SortScratch sort{};
    const uint32_t* last_perm = nullptr;
    const uint32_t* last_skeys = nullptr;
    uint64_t *cdf_v = nullptr, *cdf_j = nullptr, *blocksums = nullptr;
    uint32_t* bnd = nullptr;
    int32_t* ridx = nullptr;
    DevScalars* sc = nullptr;

    int m_cap = 0;
    float* d_pts = nullptr;
    K3PointRec* rec = nullptr;
    float4* pxyz = nullptr;
    float* h_pts = nullptr;

    Level lev[3];
    uint32_t *heavy_cnt = nullptr, *light_cnt = nullptr, *heavy_off = nullptr, *light_off = nullptr;
    uint32_t* tile_partial = nullptr;
    uint4* items = nullptr;
    size_t item_cap = 0;
    float* partial = nullptr;
    size_t partial_cap = 0;  // items
    DevScalars* h_sc = nullptr;  // pinned mirror

    uint32_t frame = 0;
    bool first_pose = false;
    float sensor[3] = {0, 0, 0}, center[3] = {0, 0, 0};
    double updated_time = -1.0;
    std::string err;

    bool timing = false;
    hipEvent_t ev[ST_COUNT + 1]{};
    float stage_ms[ST_COUNT]{};
};

namespace {

k3dom_status fail(k3dom_handle* h, k3dom_status s, const std::string& msg) {
    if (h) h->err = msg;
    else g_create_error = msg;
    return s;
}

#define HIP_TRY(h, expr)                                                                          \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return fail((h), _e == hipErrorOutOfMemory ? K3DOM_ERR_OUT_OF_MEMORY : K3DOM_ERR_HIP, \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                       \
    } while (0)

template <class T>
hipError_t dalloc(k3dom_handle* h, T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return e;
    h->allocs.push_back(q);
    *p = (T*)q;
    return hipSuccess;
}

void dfree(k3dom_handle* h, void* p) {
    if (!p) return;
    auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
    if (it != h->allocs.end()) h->allocs.erase(it);
    (void)hipFree(p);
}

k3dom_status validate(const k3dom_params* p, std::string* why) {
    if (!p) { *why = "params is NULL"; return K3DOM_ERR_INVALID_ARG; }
    if (!(p->resolution > 0.0f)) { *why = "resolution must be > 0"; return K3DOM_ERR_INVALID_ARG; }
    int gs = (int)(p->size / p->resolution), gsz = (int)(p->size_z / p->resolution);
    if (gs < 1 || gsz < 1) { *why = "size and size_z must be >= resolution"; return K3DOM_ERR_INVALID_ARG; }
    if ((double)gs * gs * gsz > 2147483647.0 / 2) { *why = "grid too large"; return K3DOM_ERR_INVALID_ARG; }
    if (p->particle_count <= 0) { *why = "particle_count must be > 0"; return K3DOM_ERR_INVALID_ARG; }
    if (p->new_born_particle_count < 0) { *why = "new_born_particle_count must be >= 0"; return K3DOM_ERR_INVALID_ARG; }
    if (!(p->ls > 0.0f)) { *why = "ls must be > 0"; return K3DOM_ERR_INVALID_ARG; }
    if (!(p->mass_scale != 0.0f)) { *why = "mass_scale must be non-zero"; return K3DOM_ERR_INVALID_ARG; }
    return K3DOM_OK;
}

void setup_levels(k3dom_handle* h) {
    // Boxes of 8x8x4 tiles (32x32x16 voxels), then 2x2x2 tiles, then single tiles.
    const int sizes[3][3] = {{8, 8, 4}, {2, 2, 2}, {1, 1, 1}};
    for (int l = 0; l < 3; ++l) {
        LevelGeom& g = h->lev[l].g;
        g.bx = sizes[l][0]; g.by = sizes[l][1]; g.bz = sizes[l][2];
        g.nbx = (h->ntx + g.bx - 1) / g.bx;
        g.nby = (h->nty + g.by - 1) / g.by;
        g.nbz = (h->ntz + g.bz - 1) / g.bz;
        if (l == 0) {
            g.pbx = g.pby = g.pbz = 0;
            g.pnbx = g.pnby = 0;
        } else {
            const LevelGeom& pg = h->lev[l - 1].g;
            g.pbx = pg.bx; g.pby = pg.by; g.pbz = pg.bz;
            g.pnbx = pg.nbx; g.pnby = pg.nby;
        }
        h->lev[l].nboxes = g.nbx * g.nby * g.nbz;
    }
}

k3dom_status ensure_points(k3dom_handle* h, int m) {
    if (m <= h->m_cap) return K3DOM_OK;
    int cap = std::max(m, h->m_cap * 3 / 2);
    dfree(h, h->d_pts); dfree(h, h->rec); dfree(h, h->pxyz);
    if (h->h_pts) { (void)hipHostFree(h->h_pts); h->h_pts = nullptr; }
    h->d_pts = nullptr; h->rec = nullptr; h->pxyz = nullptr;
    h->m_cap = 0;
    HIP_TRY(h, dalloc(h, &h->d_pts, (size_t)3 * cap));
    HIP_TRY(h, dalloc(h, &h->rec, (size_t)cap));
    HIP_TRY(h, dalloc(h, &h->pxyz, (size_t)cap));
    HIP_TRY(h, hipHostMalloc((void**)&h->h_pts, sizeof(float) * 3 * (size_t)cap, hipHostMallocDefault));
    h->m_cap = cap;
    return K3DOM_OK;
}

k3dom_status ensure_level(k3dom_handle* h, Level& L, size_t need) {
    if (need <= L.cap) return K3DOM_OK;
    size_t cap = std::max(need + need / 4 + 1024, L.cap * 3 / 2);
    dfree(h, L.list);
    L.list = nullptr;
    L.cap = 0;
    HIP_TRY(h, dalloc(h, &L.list, cap));
    L.cap = cap;
    return K3DOM_OK;
}

k3dom_status sync_scalars(k3dom_handle* h) {
    HIP_TRY(h, hipMemcpyAsync(h->h_sc, h->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return K3DOM_OK;
}

void mark(k3dom_handle* h, int stage) {
    if (h->timing) (void)hipEventRecord(h->ev[stage], h->stream);
}

GridGeom grid_geom(const k3dom_handle* h) {
    GridGeom g;
    g.gs = h->gs;
    g.gsz = h->gsz;
    g.res = h->p.resolution;
    g.sx = h->sensor[0] - h->center[0];
    g.sy = h->sensor[1] - h->center[1];
    g.sz = h->sensor[2] - h->center[2];
    return g;
}

// Kernel inference for the current point set (already in h->d_pts): writes
// del_occ / del_free for every voxel.
k3dom_status run_bgki(k3dom_handle* h, const float* xs, const float* ys, const float* zs, int m) {
    const GridGeom g = grid_geom(h);
    mark(h, ST_LISTS);
    HIP_TRY(h, launch_point_prep(h->stream, xs, ys, zs, m, h->p.resolution, h->rec, h->pxyz));
    for (int l = 0; l < 3; ++l) {
        Level& L = h->lev[l];
        const Level* P = l ? &h->lev[l - 1] : nullptr;
        HIP_TRY(h, launch_list_level(h->stream, 1, g, L.g, L.nboxes, h->p.ls, h->pxyz, m, P ? P->off : nullptr,
                                     P ? P->len : nullptr, P ? P->list : nullptr, L.len, nullptr, nullptr));
        HIP_TRY(h, launch_scan_u32(h->stream, L.len, L.off, L.nboxes, &h->sc->level_total[l]));
        k3dom_status s = sync_scalars(h);
        if (s != K3DOM_OK) return s;
        s = ensure_level(h, L, h->h_sc->level_total[l]);
        if (s != K3DOM_OK) return s;
        HIP_TRY(h, launch_list_level(h->stream, 0, g, L.g, L.nboxes, h->p.ls, h->pxyz, m, P ? P->off : nullptr,
                                     P ? P->len : nullptr, P ? P->list : nullptr, L.len, L.off, L.list));
    }
    const Level& T = h->lev[2];
    HIP_TRY(h, launch_count_items(h->stream, h->ntiles, T.off, T.len, T.list, h->heavy_cnt, h->light_cnt,
                                  h->heavy_off, h->light_off, h->sc));
    k3dom_status s = sync_scalars(h);
    if (s != K3DOM_OK) return s;
    size_t n_heavy = h->h_sc->n_heavy_items, n_items = n_heavy + h->h_sc->n_items;
    if (n_items > h->item_cap) {
        dfree(h, h->items);
        h->items = nullptr;
        h->item_cap = 0;
        size_t cap = n_items + n_items / 4 + 1024;
        HIP_TRY(h, dalloc(h, &h->items, cap));
        h->item_cap = cap;
    }
    if (n_heavy > h->partial_cap) {
        dfree(h, h->partial);
        h->partial = nullptr;
        h->partial_cap = 0;
        size_t cap = n_heavy + n_heavy / 4 + 256;
        HIP_TRY(h, dalloc(h, &h->partial, cap * 128));
        h->partial_cap = cap;
    }
    HIP_TRY(h, launch_emit_items(h->stream, h->ntiles, T.off, T.len, T.list, h->heavy_cnt, h->heavy_off,
                                 h->light_off, h->items, h->tile_partial, h->sc));
    mark(h, ST_BGKI);
    const int grid_blocks = std::max(1, std::min<int>((int)((n_items + 3) / 4), h->num_cus * 8));
    HIP_TRY(h, launch_bgki(h->stream, grid_blocks, g, h->ntx, h->nty, h->p.ls, h->p.sigma, h->items, T.list, h->rec,
                           h->del_occ, h->del_free, h->partial, h->sc));
    HIP_TRY(h, launch_bgki_combine(h->stream, h->ntiles, g, h->ntx, h->nty, h->tile_partial, h->partial,
                                   h->del_occ, h->del_free));
    return K3DOM_OK;
}

k3dom_status frame(k3dom_handle* h, double t, const float* dx, const float* dy, const float* dz, int m) {
    VoxelArrays& v = h->vox[h->vcur];
    const uint32_t k0 = (uint32_t)h->p.seed, k1 = (uint32_t)(h->p.seed >> 32);
    if (h->timing) for (float& x : h->stage_ms) x = 0.0f;
    mark(h, ST_COUNT);  // frame start
    HIP_TRY(h, launch_reset_scalars(h->stream, h->sc));
    if (h->updated_time > 0) {
        const float dt = (float)(t - h->updated_time);
        const float decay = (float)std::pow((double)h->p.gamma, (double)dt);
        K3PredictConst c;
        c.dt = dt;
        c.p_S = h->p.persistence_prob;
        c.sigma_pos = h->p.stddev_process_noise_position;
        c.sigma_vel = h->p.stddev_process_noise_velocity;
        c.vclamp = 5 * h->p.init_max_velocity;
        c.gs = h->gs;
        c.gsz = h->gsz;
        c.frame = h->frame;
        c.key0 = k0;
        c.key1 = k1;
        mark(h, ST_PREDICT);
        HIP_TRY(h, launch_predict(h->stream, c, h->P[0], h->N));
        mark(h, ST_SORT);
        HIP_TRY(h, launch_radix_sort(h->stream, h->P[0].key, h->N, h->key_bits, h->sort, &h->last_skeys,
                                     &h->last_perm));
        HIP_TRY(h, launch_gather_sorted(h->stream, h->P[0], h->P[1], h->last_perm, h->last_skeys, h->N));
        HIP_TRY(h, launch_ranges(h->stream, h->P[1].key, h->N, v.start, v.end, h->V));
        k3dom_status s = run_bgki(h, dx, dy, dz, m);
        if (s != K3DOM_OK) return s;
        mark(h, ST_MASS);
        HIP_TRY(h, launch_mass_update(h->stream, v, h->V, h->P[1].w, h->del_occ, h->del_free, decay,
                                      h->p.mass_scale, h->p.birth_prob, h->born, h->sc));
        mark(h, ST_PERSIST);
        HIP_TRY(h, launch_voxel_particles(h->stream, v, h->V, h->P[1], h->p.particle_min_vel, h->sc));
        mark(h, ST_NEWBORN);
        HIP_TRY(h, launch_fixed_scan(h->stream, h->born, h->V, nullptr, 0, &h->sc->born_max_bits, &h->sc->E_born,
                                     h->cdf_v, h->blocksums, &h->sc->total_born));
        HIP_TRY(h, launch_bounds(h->stream, h->cdf_v, &h->sc->total_born, h->V, (uint32_t)h->B, h->born, h->bnd,
                                 v.wua));
        HIP_TRY(h, launch_newborn(h->stream, h->bnd, h->V, h->B, h->gs, v.wua, h->NB, h->p.init_max_velocity,
                                  h->p.particle_min_vel, h->frame, k0, k1, h->sc));
        mark(h, ST_MOMENTS);
        mark(h, ST_RESAMPLE);
        HIP_TRY(h, launch_fixed_scan(h->stream, h->P[1].w, h->N, h->NB.w, h->B, &h->sc->joint_max_bits,
                                     &h->sc->E_joint, h->cdf_j, h->blocksums, &h->sc->total_joint));
        HIP_TRY(h, launch_resample(h->stream, h->cdf_j, h->N + h->B, &h->sc->total_joint, &h->sc->E_joint, h->P[1],
                                   h->NB, h->P[0], h->N, h->frame, k0, k1, h->ridx));
    } else {
        k3dom_status s = run_bgki(h, dx, dy, dz, m);
        if (s != K3DOM_OK) return s;
        mark(h, ST_MASS);
        HIP_TRY(h, launch_mass_update(h->stream, v, h->V, h->P[0].w, h->del_occ, h->del_free, 1.0f,
                                      h->p.mass_scale, h->p.birth_prob, h->born, h->sc));
        mark(h, ST_PERSIST);
        mark(h, ST_NEWBORN);
        HIP_TRY(h, launch_fixed_scan(h->stream, h->born, h->V, nullptr, 0, &h->sc->born_max_bits, &h->sc->E_born,
                                     h->cdf_v, h->blocksums, &h->sc->total_born));
        HIP_TRY(h, launch_bounds(h->stream, h->cdf_v, &h->sc->total_born, h->V, (uint32_t)h->N, h->born, h->bnd,
                                 nullptr));
        HIP_TRY(h, launch_init_particles(h->stream, h->bnd, h->V, h->N, h->gs, h->P[0], h->p.init_max_velocity,
                                         h->p.particle_min_vel, h->frame, k0, k1));
        mark(h, ST_MOMENTS);
        mark(h, ST_RESAMPLE);
    }
    mark(h, ST_FRAME);
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (h->timing) {
        // stage boundaries: events ST_PREDICT..ST_FRAME in launch order; the
        // frame start is ev[ST_COUNT]
        const int order[] = {ST_PREDICT, ST_SORT, ST_LISTS, ST_BGKI, ST_MASS, ST_PERSIST, ST_NEWBORN, ST_MOMENTS,
                             ST_RESAMPLE, ST_FRAME};
        bool normal = h->updated_time > 0;
        for (int k = 0; k + 1 < 10; ++k) {
            int a = order[k], b = order[k + 1];
            if (!normal && (a == ST_PREDICT || a == ST_SORT)) continue;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, h->ev[a], h->ev[b]) == hipSuccess) h->stage_ms[a] = ms;
        }
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, h->ev[ST_COUNT], h->ev[ST_FRAME]) == hipSuccess) h->stage_ms[ST_FRAME] = ms;
    }
    h->updated_time = t;
    h->frame++;
    return K3DOM_OK;
}

// ---------------------------------------------------------------- YAML params
std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

bool parse_flat_yaml(const char* path, std::map<std::string, std::string>& kv, std::string* err) {
    std::ifstream in(path);
    if (!in) {
        *err = std::string("cannot open ") + path;
        return false;
    }
    std::string line;
    while (std::getline(in, line)) {
        size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        size_t colon = line.find(':');
        if (colon == std::string::npos) continue;
        std::string k = trim(line.substr(0, colon)), v = trim(line.substr(colon + 1));
        if (k.empty()) continue;
        if (v.size() >= 2 && (v.front() == '"' || v.front() == '\'') && v.back() == v.front())
            v = v.substr(1, v.size() - 2);
        kv[k] = v;
    }
    return true;
}

// nh.param<float> keeps the default when the value is not a number.
void get_float(const std::map<std::string, std::string>& kv, const char* key, float* v) {
    auto it = kv.find(key);
    if (it == kv.end()) return;
    const char* s = it->second.c_str();
    char* end = nullptr;
    double d = std::strtod(s, &end);
    if (end == s || *end != '\0') return;
    *v = (float)d;
}

}  // namespace

// ============================================================== C ABI
extern "C" {

k3dom_status k3dom_params_default(k3dom_params* out) {
    if (!out) return K3DOM_ERR_INVALID_ARG;
    return k3dom_params_from_yaml(nullptr, nullptr, out, nullptr, 0);
}

// src/k3dom/k3dom_node.cpp:215-368
k3dom_status k3dom_params_from_yaml(const char* method_yaml, const char* dataset_yaml, k3dom_params* out, char* err,
                                    int32_t err_len) {
    if (!out) return K3DOM_ERR_INVALID_ARG;
    std::map<std::string, std::string> kv;
    std::string why;
    for (const char* path : {method_yaml, dataset_yaml}) {
        if (path && !parse_flat_yaml(path, kv, &why)) {
            if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", why.c_str());
            return K3DOM_ERR_INVALID_ARG;
        }
    }
    float resolution = 0.1f, size = 10.0f, size_z = 5.0f, mass_scale = 3.0f;
    float sensor_off_x = 0.0f, sensor_off_y = 0.0f, sensor_off_z = 0.0f, map_shift_thresh = 1.0f;
    float particle_count = 3 * static_cast<int>(10e6);
    float new_born_particle_count = 3 * static_cast<int>(10e5);
    float persistence_prob = 0.99f, stddev_process_noise_position = 0.1f, stddev_process_noise_velocity = 1.0f;
    float birth_prob = 0.02f, stddev_velocity = 1.0f, init_max_velocity = 5.0f, particle_min_vel = 1.0f;
    float sigma = 1.0f, ls = 1.0f, gamma = 0.9f;
    float prior_free = 1.0f, prior_static = 1.0f, prior_dynamic = 1.0f, prior_occ = 0.0f;
    get_float(kv, "resolution", &resolution);
    get_float(kv, "size", &size);
    get_float(kv, "size_z", &size_z);
    get_float(kv, "mass_scale", &mass_scale);
    get_float(kv, "sensor_off_x", &sensor_off_x);
    get_float(kv, "sensor_off_y", &sensor_off_y);
    get_float(kv, "sensor_off_z", &sensor_off_z);
    get_float(kv, "map_shift_thresh", &map_shift_thresh);
    get_float(kv, "particle_count", &particle_count);
    get_float(kv, "new_born_particle_count", &new_born_particle_count);
    get_float(kv, "persistence_prob", &persistence_prob);
    get_float(kv, "stddev_process_noise_position", &stddev_process_noise_position);
    get_float(kv, "stddev_process_noise_velocity", &stddev_process_noise_velocity);
    get_float(kv, "birth_prob", &birth_prob);
    get_float(kv, "stddev_velocity", &stddev_velocity);
    get_float(kv, "init_max_velocity", &init_max_velocity);
    get_float(kv, "particle_min_vel", &particle_min_vel);
    get_float(kv, "sigma", &sigma);
    get_float(kv, "ls", &ls);
    get_float(kv, "gamma", &gamma);
    get_float(kv, "prior_free", &prior_free);
    get_float(kv, "prior_static", &prior_static);
    get_float(kv, "prior_dynamic", &prior_dynamic);
    get_float(kv, "prior_occ", &prior_occ);
    k3dom_params p{};
    p.size = size;
    p.size_z = size_z;
    p.mass_scale = mass_scale;
    p.resolution = resolution;
    p.particle_count = (int32_t)particle_count;
    p.new_born_particle_count = (int32_t)new_born_particle_count;
    p.persistence_prob = persistence_prob;
    p.stddev_process_noise_position = stddev_process_noise_position / resolution;
    p.stddev_process_noise_velocity = stddev_process_noise_velocity / resolution;
    p.birth_prob = birth_prob;
    p.stddev_velocity = stddev_velocity / resolution;
    p.init_max_velocity = init_max_velocity / resolution;
    p.particle_min_vel = particle_min_vel / resolution;
    p.sigma = sigma;
    p.ls = ls;
    p.gamma = gamma;
    p.prior_free = prior_free;
    p.prior_static = prior_static;
    p.prior_dynamic = prior_dynamic;
    p.prior_occ = prior_occ;
    p.sensor_off_x = sensor_off_x;
    p.sensor_off_y = sensor_off_y;
    p.sensor_off_z = sensor_off_z;
    p.map_shift_thresh = static_cast<int>(map_shift_thresh / resolution);
    p.max_range = 0.0f;
    p.meas_occ_stddev = 0.0f;
    p.seed = 123456;
    *out = p;
    return K3DOM_OK;
}

k3dom_status k3dom_create(const k3dom_params* params, k3dom_handle** out) {
    if (!out) return fail(nullptr, K3DOM_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    std::string why;
    k3dom_status vs = validate(params, &why);
    if (vs != K3DOM_OK) return fail(nullptr, vs, why);
    k3dom_handle* h = new (std::nothrow) k3dom_handle();
    if (!h) return fail(nullptr, K3DOM_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto bail = [&](k3dom_status s) {
        g_create_error = h->err;
        k3dom_destroy(h);
        return s;
    };
#define CREATE_TRY(expr)                                                                           \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            h->err = std::string(#expr) + ": " + hipGetErrorString(_e);                            \
            return bail(_e == hipErrorOutOfMemory ? K3DOM_ERR_OUT_OF_MEMORY : K3DOM_ERR_HIP);      \
        }                                                                                          \
    } while (0)
    h->p = *params;
    h->gs = (int)(h->p.size / h->p.resolution);
    h->gsz = (int)(h->p.size_z / h->p.resolution);
    h->V = h->gs * h->gs * h->gsz;
    h->N = h->p.particle_count;
    h->B = h->p.new_born_particle_count;
    h->key_bits = 1;
    while ((1ll << h->key_bits) < (long long)h->V) ++h->key_bits;
    h->ntx = (h->gs + 3) / 4;
    h->nty = h->ntx;
    h->ntz = (h->gsz + 3) / 4;
    h->ntiles = h->ntx * h->nty * h->ntz;
    CREATE_TRY(hipGetDevice(&h->device));
    hipDeviceProp_t prop;
    CREATE_TRY(hipGetDeviceProperties(&prop, h->device));
    h->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    CREATE_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    const size_t V = (size_t)h->V, N = (size_t)h->N, B = (size_t)std::max(h->B, 1);
    for (int b = 0; b < 2; ++b) {
        VoxelArrays& v = h->vox[b];
        for (float** f : {&v.f, &v.s, &v.d, &v.pers, &v.pred, &v.mu, &v.wua, &v.mvx, &v.mvy, &v.mvz})
            CREATE_TRY(dalloc(h, f, V));
        CREATE_TRY(dalloc(h, &v.start, V));
        CREATE_TRY(dalloc(h, &v.end, V));
    }
    CREATE_TRY(dalloc(h, &h->born, V));
    CREATE_TRY(dalloc(h, &h->del_occ, V));
    CREATE_TRY(dalloc(h, &h->del_free, V));
    for (int s = 0; s < 2; ++s) {
        ParticleArrays& P = h->P[s];
        for (float** f : {&P.px, &P.py, &P.pz, &P.vx, &P.vy, &P.vz, &P.w}) CREATE_TRY(dalloc(h, f, N));
        CREATE_TRY(dalloc(h, &P.key, N));
    }
    for (float** f : {&h->NB.px, &h->NB.py, &h->NB.pz, &h->NB.vx, &h->NB.vy, &h->NB.vz, &h->NB.w})
        CREATE_TRY(dalloc(h, f, B));
    CREATE_TRY(dalloc(h, &h->NB.key, B));
    CREATE_TRY(dalloc(h, &h->sort.k0, N));
    CREATE_TRY(dalloc(h, &h->sort.k1, N));
    CREATE_TRY(dalloc(h, &h->sort.i0, N));
    CREATE_TRY(dalloc(h, &h->sort.i1, N));
    CREATE_TRY(dalloc(h, &h->sort.hist, (size_t)256 * sort_blocks(h->N)));
    CREATE_TRY(dalloc(h, &h->cdf_v, V));
    CREATE_TRY(dalloc(h, &h->cdf_j, N + B));
    CREATE_TRY(dalloc(h, &h->blocksums, (size_t)fixed_scan_blocks((int)std::max(V, N + B))));
    CREATE_TRY(dalloc(h, &h->bnd, V));
    CREATE_TRY(dalloc(h, &h->ridx, N));
    CREATE_TRY(dalloc(h, &h->sc, 1));
    setup_levels(h);
    for (int l = 0; l < 3; ++l) {
        CREATE_TRY(dalloc(h, &h->lev[l].len, (size_t)h->lev[l].nboxes));
        CREATE_TRY(dalloc(h, &h->lev[l].off, (size_t)h->lev[l].nboxes));
    }
    const size_t T = (size_t)h->ntiles;
    for (uint32_t** a : {&h->heavy_cnt, &h->light_cnt, &h->heavy_off, &h->light_off}) CREATE_TRY(dalloc(h, a, T));
    CREATE_TRY(dalloc(h, &h->tile_partial, 2 * T));
    CREATE_TRY(hipHostMalloc((void**)&h->h_sc, sizeof(DevScalars), hipHostMallocDefault));
    for (int e = 0; e <= ST_COUNT; ++e) CREATE_TRY(hipEventCreate(&h->ev[e]));
    // prior fill and zeroed particle state (src/common/kernel/init.cu:20-35)
    for (int b = 0; b < 2; ++b)
        CREATE_TRY(launch_fill_voxels(h->stream, h->vox[b], h->V, h->p.prior_free, h->p.prior_static,
                                      h->p.prior_dynamic));
    for (int s = 0; s < 2; ++s) {
        ParticleArrays& P = h->P[s];
        for (float* f : {P.px, P.py, P.pz, P.vx, P.vy, P.vz, P.w})
            CREATE_TRY(hipMemsetAsync(f, 0, N * sizeof(float), h->stream));
        CREATE_TRY(hipMemsetAsync(P.key, 0, N * sizeof(int32_t), h->stream));
    }
    CREATE_TRY(hipMemsetAsync(h->born, 0, V * sizeof(float), h->stream));
    CREATE_TRY(hipMemsetAsync(h->del_occ, 0, V * sizeof(float), h->stream));
    CREATE_TRY(hipMemsetAsync(h->del_free, 0, V * sizeof(float), h->stream));
    CREATE_TRY(hipMemsetAsync(h->ridx, 0, N * sizeof(int32_t), h->stream));
    CREATE_TRY(hipStreamSynchronize(h->stream));
#undef CREATE_TRY
    *out = h;
    return K3DOM_OK;
}

k3dom_status k3dom_destroy(k3dom_handle* h) {
    if (!h) return K3DOM_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    if (h->h_pts) (void)hipHostFree(h->h_pts);
    if (h->h_sc) (void)hipHostFree(h->h_sc);
    for (int e = 0; e <= ST_COUNT; ++e)
        if (h->ev[e]) (void)hipEventDestroy(h->ev[e]);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return K3DOM_OK;
}

// src/common/dom.cu:111-158
k3dom_status k3dom_update_pose(k3dom_handle* h, float x, float y, float z) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    h->sensor[0] = x;
    h->sensor[1] = y;
    h->sensor[2] = z;
    if (!h->first_pose) {
        h->first_pose = true;
        h->center[0] = h->sensor[0] - h->p.sensor_off_x;
        h->center[1] = h->sensor[1] - h->p.sensor_off_y;
        h->center[2] = h->sensor[2] - h->p.sensor_off_z;
        return K3DOM_OK;
    }
    const int mx = static_cast<int>((h->sensor[0] - h->center[0] - h->p.sensor_off_x) / h->p.resolution);
    const int my = static_cast<int>((h->sensor[1] - h->center[1] - h->p.sensor_off_y) / h->p.resolution);
    const int mz = static_cast<int>((h->sensor[2] - h->center[2] - h->p.sensor_off_z) / h->p.resolution);
    if (std::abs(mx) >= h->p.map_shift_thresh || std::abs(my) >= h->p.map_shift_thresh ||
        std::abs(mz) >= h->p.map_shift_thresh) {
        const int nxt = 1 - h->vcur;
        HIP_TRY(h, launch_shift_voxels(h->stream, h->vox[nxt], h->vox[h->vcur], h->gs, h->gsz, mx, my, mz,
                                       h->p.prior_free, h->p.prior_static, h->p.prior_dynamic));
        HIP_TRY(h, launch_shift_particles(h->stream, h->P[0], h->N, mx, my, mz));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        h->vcur = nxt;
        h->center[0] += mx * h->p.resolution;
        h->center[1] += my * h->p.resolution;
        h->center[2] += mz * h->p.resolution;
    }
    return K3DOM_OK;
}

k3dom_status k3dom_update_grid(k3dom_handle* h, double t, const float* xs, const float* ys, const float* zs,
                               int32_t n) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && (!xs || !ys || !zs))) return fail(h, K3DOM_ERR_INVALID_ARG, "bad point arrays");
    k3dom_status s = ensure_points(h, std::max(n, 1));
    if (s != K3DOM_OK) return s;
    const size_t cap = (size_t)h->m_cap;
    if (n > 0) {
        std::memcpy(h->h_pts, xs, sizeof(float) * n);
        std::memcpy(h->h_pts + cap, ys, sizeof(float) * n);
        std::memcpy(h->h_pts + 2 * cap, zs, sizeof(float) * n);
        HIP_TRY(h, hipMemcpyAsync(h->d_pts, h->h_pts, sizeof(float) * 3 * cap, hipMemcpyHostToDevice, h->stream));
    }
    return frame(h, t, h->d_pts, h->d_pts + cap, h->d_pts + 2 * cap, n);
}

k3dom_status k3dom_update_grid_device(k3dom_handle* h, double t, const float* d_xs, const float* d_ys,
                                      const float* d_zs, int32_t n) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && (!d_xs || !d_ys || !d_zs))) return fail(h, K3DOM_ERR_INVALID_ARG, "bad point arrays");
    k3dom_status s = ensure_points(h, std::max(n, 1));
    if (s != K3DOM_OK) return s;
    return frame(h, t, d_xs, d_ys, d_zs, n);
}

k3dom_status k3dom_get_info(const k3dom_handle* h, k3dom_info* out) {
    if (!h || !out) return K3DOM_ERR_INVALID_ARG;
    out->grid_size = h->gs;
    out->grid_size_z = h->gsz;
    out->grid_cell_count = h->V;
    out->particle_count = h->N;
    out->new_born_particle_count = h->B;
    out->iteration = (int32_t)h->frame;
    out->resolution = h->p.resolution;
    for (int i = 0; i < 3; ++i) {
        out->center_pos[i] = h->center[i];
        out->sensor_pos[i] = h->sensor[i];
    }
    out->updated_time = h->updated_time;
    return K3DOM_OK;
}

k3dom_status k3dom_read_masses(k3dom_handle* h, float* f, float* s, float* d, float* mvx, float* mvy, float* mvz) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    const VoxelArrays& v = h->vox[h->vcur];
    const size_t bytes = sizeof(float) * (size_t)h->V;
    const float* src[6] = {v.f, v.s, v.d, v.mvx, v.mvy, v.mvz};
    float* dst[6] = {f, s, d, mvx, mvy, mvz};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) HIP_TRY(h, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return K3DOM_OK;
}

k3dom_status k3dom_read_cells(k3dom_handle* h, k3dom_cell* out) {
    if (!h || !out) return K3DOM_ERR_INVALID_ARG;
    const VoxelArrays& v = h->vox[h->vcur];
    const size_t V = (size_t)h->V;
    std::vector<float> buf(V * 10);
    std::vector<int32_t> ib(V * 2);
    const float* src[10] = {v.f, v.s, v.d, v.pers, v.pred, v.mu, v.wua, v.mvx, v.mvy, v.mvz};
    for (int k = 0; k < 10; ++k)
        HIP_TRY(h, hipMemcpyAsync(buf.data() + k * V, src[k], V * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(ib.data(), v.start, V * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(ib.data() + V, v.end, V * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (size_t c = 0; c < V; ++c) {
        k3dom_cell& o = out[c];
        o.start_idx = ib[c];
        o.end_idx = ib[V + c];
        o.mu_A = buf[5 * V + c];
        o.mu_UA = buf[5 * V + c];
        o.w_A = 0.0f;
        o.w_UA = buf[6 * V + c];
        o.mean_x_vel = buf[7 * V + c];
        o.mean_y_vel = buf[8 * V + c];
        o.mean_z_vel = buf[9 * V + c];
        o.likelihood = 1.0f;
        o.p_A = 0.0f;
        o.free_mass = buf[0 * V + c];
        o.static_mass = buf[1 * V + c];
        o.dynamic_mass = buf[2 * V + c];
        o.occ_mass = h->p.prior_occ;
        o.pers_mass = buf[3 * V + c];
        o.pred_mass = buf[4 * V + c];
    }
    return K3DOM_OK;
}

k3dom_status k3dom_read_array(k3dom_handle* h, int32_t which, void* dst, int64_t capacity) {
    if (!h || !dst) return K3DOM_ERR_INVALID_ARG;
    const size_t N = (size_t)h->N, B = (size_t)h->B, V = (size_t)h->V;
    auto cp = [&](const void* src, size_t bytes, size_t at) -> k3dom_status {
        if ((int64_t)(at + bytes) > capacity) return fail(h, K3DOM_ERR_INVALID_ARG, "capacity too small");
        if (!src) return fail(h, K3DOM_ERR_STATE, "array not available before the first normal frame");
        HIP_TRY(h, hipMemcpyAsync((char*)dst + at, src, bytes, hipMemcpyDeviceToHost, h->stream));
        return K3DOM_OK;
    };
    k3dom_status s = K3DOM_OK;
    const ParticleArrays& P = h->P[0];
    switch (which) {
    case K3DOM_ARR_PARTICLE_POS:
        if ((s = cp(P.px, N * 4, 0)) || (s = cp(P.py, N * 4, N * 4)) || (s = cp(P.pz, N * 4, 2 * N * 4))) return s;
        break;
    case K3DOM_ARR_PARTICLE_VEL:
        if ((s = cp(P.vx, N * 4, 0)) || (s = cp(P.vy, N * 4, N * 4)) || (s = cp(P.vz, N * 4, 2 * N * 4))) return s;
        break;
    case K3DOM_ARR_PARTICLE_WEIGHT: s = cp(P.w, N * 4, 0); break;
    case K3DOM_ARR_PARTICLE_KEY: s = cp(P.key, N * 4, 0); break;
    case K3DOM_ARR_SORT_PERM: s = cp(h->last_perm, N * 4, 0); break;
    case K3DOM_ARR_SORTED_KEY: s = cp(h->last_skeys, N * 4, 0); break;
    case K3DOM_ARR_RESAMPLE_IDX: s = cp(h->ridx, N * 4, 0); break;
    case K3DOM_ARR_NEWBORN_POS:
        if ((s = cp(h->NB.px, B * 4, 0)) || (s = cp(h->NB.py, B * 4, B * 4)) || (s = cp(h->NB.pz, B * 4, 2 * B * 4)))
            return s;
        break;
    case K3DOM_ARR_NEWBORN_VEL:
        if ((s = cp(h->NB.vx, B * 4, 0)) || (s = cp(h->NB.vy, B * 4, B * 4)) || (s = cp(h->NB.vz, B * 4, 2 * B * 4)))
            return s;
        break;
    case K3DOM_ARR_NEWBORN_WEIGHT: s = cp(h->NB.w, B * 4, 0); break;
    case K3DOM_ARR_NEWBORN_KEY: s = cp(h->NB.key, B * 4, 0); break;
    case K3DOM_ARR_BORN_MASS: s = cp(h->born, V * 4, 0); break;
    case K3DOM_ARR_DEL_OCC: s = cp(h->del_occ, V * 4, 0); break;
    case K3DOM_ARR_DEL_FREE: s = cp(h->del_free, V * 4, 0); break;
    case K3DOM_ARR_SORTED_WEIGHT: s = cp(h->P[1].w, N * 4, 0); break;
    default: return fail(h, K3DOM_ERR_INVALID_ARG, "unknown array id");
    }
    if (s != K3DOM_OK) return s;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return K3DOM_OK;
}

k3dom_status k3dom_set_timing(k3dom_handle* h, int32_t enable) {
    if (!h) return K3DOM_ERR_INVALID_ARG;
    h->timing = enable != 0;
    return K3DOM_OK;
}

k3dom_status k3dom_stage_times(k3dom_handle* h, float* ms, int32_t count) {
    if (!h || !ms || count < 0) return K3DOM_ERR_INVALID_ARG;
    for (int i = 0; i < count && i < ST_COUNT; ++i) ms[i] = h->stage_ms[i];
    return K3DOM_OK;
}

const char* k3dom_last_error(const k3dom_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

}  // extern "C"
