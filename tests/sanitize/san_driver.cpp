// Driver of the sanitizer build (Makefile beside it): every host path that
// reads untrusted input, fuzzed, and the serial restatement run through the
// frames the tests use, all under AddressSanitizer and UndefinedBehavior-
// Sanitizer.  Any report aborts the process; "sanitize ok" is printed at the end.
#include "k3dom.h"
#include "../../k3dom_amd/csrc/k3_ckformat.h"
#include "../../k3dom_amd/csrc/k3_scene.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

extern "C" {
void* orc_create_method(const k3dom_params* p, int method);
void orc_destroy(void* h);
void orc_set_brute(void* h, int brute);
void orc_set_threads(int n);
void orc_update_pose(void* h, float x, float y, float z);
int orc_update_grid(void* h, double t, const float* xs, const float* ys, const float* zs, int n);
int orc_update_grid_lidars(void* h, double t, int n_lidars, const float* rot, const float* origin, const int* counts,
                           const float* xs, const float* ys, const float* zs);
void orc_read_cells(void* h, k3dom_cell* out);
int64_t orc_read_array(void* h, int which, void* dst, int64_t capacity);
void orc_get_info(void* h, k3dom_info* out);
void orc_bgki64(int gs, int gsz, float res, float sx, float sy, float sz, float ls, float sigma, const float* xs,
                const float* ys, const float* zs, int n, const int32_t* cells, int ncell, double* del_occ,
                double* del_free);
}

namespace {

int g_checks = 0;
void expect(bool ok, const char* what) {
    ++g_checks;
    if (!ok) {
        std::fprintf(stderr, "sanitize driver: check failed: %s\n", what);
        std::exit(1);
    }
}

// ---------------------------------------------------------------- YAML loader
void fuzz_yaml(const std::string& dir) {
    const std::string path = dir + "/p.yaml";
    auto load = [&](const std::string& text, int err_len) {
        FILE* f = std::fopen(path.c_str(), "wb");
        std::fwrite(text.data(), 1, text.size(), f);
        std::fclose(f);
        k3dom_params p;
        char err[64];
        return k3dom_params_from_yaml(path.c_str(), nullptr, &p, err_len > 0 ? err : nullptr, err_len);
    };
    const std::string cases[] = {
        "resolution: 0.2\nsize: 40\nsize_z: 5\nparticle_count: 1000000\n",
        "resolution: 0.1 # comment\n  ls : '0.5'\nsigma: \"0.1\"\n",
        "resolution:\n: 3\n::::\n#only comment\nkey without colon\n",
        "resolution: 1e999\nsize: nan\nsize_z: -inf\nparticle_count: 1e30\nnew_born_particle_count: -5\n",
        "resolution: 0x10\nls: 1.5.5\nsigma: ''\ngamma: '\n",
        "",
        std::string(5000, 'a') + ": " + std::string(5000, '1') + "\n",
        std::string("resolution: 0.2\0size: 3\n", 24),
    };
    for (const std::string& c : cases) load(c, 64);
    std::mt19937_64 rng(1);
    const char alphabet[] = "ab:_# \t\n'\"0123456789.eE+-xn\r";
    for (int it = 0; it < 3000; ++it) {
        std::string s(rng() % 400, ' ');
        for (char& ch : s) ch = (rng() % 20 == 0) ? (char)(rng() & 0xff) : alphabet[rng() % (sizeof(alphabet) - 1)];
        load(s, (int)(rng() % 3) * 32);
    }
    k3dom_params p;
    expect(k3dom_params_from_yaml((dir + "/missing.yaml").c_str(), nullptr, &p, nullptr, 0) == K3DOM_ERR_INVALID_ARG,
           "missing file rejected");
    expect(k3dom_params_default(&p) == K3DOM_OK && p.seed == 123456, "defaults");
}

// ---------------------------------------------------------------- checkpoint record
void fuzz_checkpoint() {
    std::mt19937_64 rng(2);
    const int64_t V = 64, N = 100;
    const uint64_t pb = k3ck::payload_bytes(V, N);
    std::vector<uint8_t> payload(pb);
    for (auto& b : payload) b = (uint8_t)rng();
    const float zero3[3] = {0, 0, 0};
    k3dom_params params;
    k3dom_params_default(&params);
    auto record = [&](const k3ck::Header& hd) {
        std::vector<uint8_t> r(sizeof(hd) + sizeof(params) + payload.size());
        std::memcpy(r.data(), &hd, sizeof(hd));
        std::memcpy(r.data() + sizeof(hd), &params, sizeof(params));
        std::memcpy(r.data() + sizeof(hd) + sizeof(params), payload.data(), payload.size());
        return r;
    };
    const k3ck::Header good = k3ck::make_header(0, V, N, 7, true, 1.5, zero3, zero3, payload.data());
    const std::vector<uint8_t> rec = record(good);
    k3ck::Header hd;
    k3dom_params p;
    const uint8_t* pl = nullptr;
    std::string why;
    expect(k3ck::check(rec.data(), (int64_t)rec.size(), &hd, &p, &pl, &why), "intact record accepted");
    expect(!k3ck::check(nullptr, 0, &hd, &p, &pl, &why), "null rejected");
    for (int it = 0; it < 20000; ++it) {
        std::vector<uint8_t> r = rec;
        int64_t size = (int64_t)r.size();
        switch (it % 4) {
        case 0:  // truncated or extended
            size = (int64_t)(rng() % (r.size() + 64));
            r.resize((size_t)size + 8);
            break;
        case 1: {  // a flipped bit anywhere
            const size_t at = rng() % r.size();
            r[at] ^= (uint8_t)(1u << (rng() % 8));
            break;
        }
        case 2: {  // header sizes edited, checksum recomputed
            k3ck::Header h2 = good;
            const int64_t vals[] = {0, -1, 1, V + 1, 0x7fffffffll, 0x80000000ll, 0x7fffffffffffffffll,
                                    (int64_t)0x8000000000000000ull};
            h2.V = vals[rng() % 8];
            h2.N = vals[rng() % 8];
            h2.payload_bytes = rng() % 2 ? good.payload_bytes : rng();
            r = record(h2);
            size = (int64_t)r.size();
            break;
        }
        default:  // random bytes behind a valid magic
            for (size_t k = 8; k < r.size(); ++k) r[k] = (uint8_t)rng();
            break;
        }
        const bool ok = k3ck::check(r.data(), size, &hd, &p, &pl, &why);
        // an accepted record is the intact one (a flip in a byte no check
        // reads: params or reserved header bytes)
        if (ok) expect(size == (int64_t)rec.size() && hd.V == V && hd.N == N, "only intact records accepted");
    }
}

// ---------------------------------------------------------------- restatement frames
k3dom_params small_params(int method) {
    k3dom_params p;
    k3dom_params_default(&p);
    p.resolution = 0.2f;
    p.size = 3.3f;    // 16 voxels
    p.size_z = 1.7f;  // 8 voxels
    p.particle_count = 4096;
    p.new_born_particle_count = 409;
    p.mass_scale = 2.5f;
    p.stddev_process_noise_position = 0.05f;
    p.stddev_process_noise_velocity = 0.5f;
    p.init_max_velocity = 25.0f;
    p.particle_min_vel = method ? 0.0f : 0.5f;
    p.map_shift_thresh = 5;
    p.prior_free = p.prior_static = p.prior_dynamic = method ? 0.0f : 0.001f;
    p.sigma = 0.1f;
    p.ls = 0.5f;
    p.gamma = 0.99f;
    p.max_range = 50.0f;
    p.meas_occ_stddev = 0.1f;
    return p;
}

k3scene_desc small_scene(uint64_t seed) {
    k3scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.seed = seed;
    d.n_static = 6;
    d.n_moving = 2;
    d.side_min = 0.5f;
    d.side_max = 3.0f;
    d.speed_max = 2.0f;
    d.xmin = d.ymin = -3.2f;
    d.xmax = d.ymax = 3.2f;
    d.ground_z = -0.8f;
    d.clearance = 0.6f;
    d.ego_vx = 4.0f;
    d.beams = 8;
    d.az_bins = 128;
    d.elev_min_deg = -15.0f;
    d.elev_max_deg = 15.0f;
    d.max_range = 4.8f;
    d.range_noise = 0.02f;
    d.frame_dt = 0.1f;
    d.t0 = 1.0;
    d.ground = 1;
    return d;
}

void read_back(void* o, const k3dom_params& p) {
    k3dom_info info;
    orc_get_info(o, &info);
    std::vector<k3dom_cell> cells((size_t)info.grid_cell_count);
    orc_read_cells(o, cells.data());
    std::vector<uint8_t> buf(64 << 20);
    for (int a = 0; a <= K3DOM_ARR_SORTED_WEIGHT; ++a) orc_read_array(o, a, buf.data(), (int64_t)buf.size());
    for (const k3dom_cell& c : cells)
        expect(std::isfinite(c.free_mass) && c.free_mass >= 0.0f, "finite non-negative masses");
    (void)p;
}

void run_frames(int method, bool brute, int threads, int lidars, int frames, uint64_t seed) {
    const k3dom_params p = small_params(method);
    void* o = orc_create_method(&p, method);
    expect(o != nullptr, "restatement created");
    orc_set_brute(o, brute ? 1 : 0);
    orc_set_threads(threads);
    k3scene_desc d = small_scene(seed);
    void* sc = k3scene_create(&d);
    const int cap = d.beams * d.az_bins;
    std::vector<float> xs(cap + 5), ys(cap + 5), zs(cap + 5);
    for (int f = 0; f < frames; ++f) {
        double pose[4];
        int n = k3scene_frame(sc, f, xs.data(), ys.data(), zs.data(), cap, pose);
        if (f == 2) {  // non-finite and zero points
            const float bad[5] = {NAN, INFINITY, 0.0f, -INFINITY, 1e30f};
            for (int k = 0; k < 5; ++k) {
                xs[n + k] = bad[k];
                ys[n + k] = bad[4 - k];
                zs[n + k] = 0.0f;
            }
            n += 5;
        }
        orc_update_pose(o, (float)pose[0], (float)pose[1], (float)pose[2]);
        if (lidars > 0) {
            std::vector<float> rot(9 * lidars), org(3 * lidars);
            std::vector<int> counts(lidars);
            for (int k = 0; k < lidars; ++k) {
                const float a = 2.0f * 3.14159265f * k / lidars, c = std::cos(a), s = std::sin(a);
                const float r[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
                std::memcpy(&rot[9 * k], r, sizeof(r));
                org[3 * k] = 0.3f * c;
                org[3 * k + 1] = 0.3f * s;
                org[3 * k + 2] = 0.05f * k;
                counts[k] = n / lidars + (k < n % lidars ? 1 : 0);
            }
            expect(orc_update_grid_lidars(o, pose[3], lidars, rot.data(), org.data(), counts.data(), xs.data(),
                                          ys.data(), zs.data()) == K3DOM_OK, "lidar frame");
        } else {
            expect(orc_update_grid(o, pose[3], xs.data(), ys.data(), zs.data(), n) == K3DOM_OK, "frame");
        }
        read_back(o, p);
    }
    // FP64 evaluation of a few voxels of the last frame's points
    k3dom_info info;
    orc_get_info(o, &info);
    const int32_t cells[3] = {0, info.grid_cell_count / 2, info.grid_cell_count - 1};
    double o64[3], f64[3];
    orc_bgki64(info.grid_size, info.grid_size_z, p.resolution, 0.1f, -0.1f, 0.05f, p.ls, p.sigma, xs.data(),
               ys.data(), zs.data(), cap, cells, 3, o64, f64);
    k3scene_destroy(sc);
    orc_destroy(o);
    orc_set_threads(1);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    fuzz_yaml(dir);
    fuzz_checkpoint();
    run_frames(K3DOM_METHOD_K3DOM, false, 1, 0, 8, 3);
    run_frames(K3DOM_METHOD_K3DOM, true, 1, 0, 3, 5);
    run_frames(K3DOM_METHOD_K3DOM, false, 4, 0, 4, 7);
    run_frames(K3DOM_METHOD_K3DOM, false, 1, 3, 4, 9);
    run_frames(K3DOM_METHOD_DSPHDMIB, false, 1, 0, 5, 11);
    std::printf("sanitize ok: %d checks\n", g_checks);
    return 0;
}
