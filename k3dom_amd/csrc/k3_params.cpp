// Parameter loading of the C ABI (k3dom_params_default,
// k3dom_params_from_yaml): the reference node's defaults, YAML keys and unit
// conversion (src/k3dom/k3dom_node.cpp:215-368).  Host-only C++, no HIP: the
// library builds it with the rest, and the sanitizer build of the host code
// compiles it on its own.
#include "../../include/k3dom.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>

namespace {

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

}  // extern "C"
