// DS-PHD/MIB filter class of the C++ drop-in (youngjae-min/k3dom,
// include/dsphdmib/dom.h:17-41): dom::DOM whose evidence is the measurement
// grid of rays passing each voxel and whose cell update is Dempster's rule
// (src/dsphdmib/dom.cu, src/dsphdmib/kernel/measurement_grid.cu,
// src/dsphdmib/kernel/mass_update.cu).  GridCell, DOM_c and the differences a
// caller sees are in common/dom.h; the occupied mass of a voxel is
// GridCell::occ_mass, static_mass and dynamic_mass keep their priors.
//
// meas_cell_array (a device array in the reference) does not exist: the
// measurement grid lives in device memory as two float arrays.
// copyMeasGridToHost() returns it as MeasurementCell records.
#pragma once

#include <cassert>
#include <vector>

#include "../common/dom.h"

namespace dom {

class DOM : public DOM_c {
public:
    // src/dsphdmib/dom.cu:21-22
    explicit DOM(const Params& p) : DOM_c(p, K3DOM_METHOD_DSPHDMIB), window_radius(p.resolution * 0.866f) {}

    // src/dsphdmib/dom.cu:33-52: one frame at time t (seconds), points
    // relative to the sensor in map-aligned axes (metres).  Lengths must be equal.
    void updateGrid(float t, std::vector<float>& measurements_x, std::vector<float>& measurements_y,
                    std::vector<float>& measurements_z) {
        assert(measurements_x.size() == measurements_y.size() && measurements_y.size() == measurements_z.size());
        updateGrid((double)t, measurements_x.data(), measurements_y.data(), measurements_z.data(),
                   (int)measurements_x.size());
    }
    // Same with a double time stamp (seconds) and plain host arrays of n points.
    void updateGrid(double t, const float* xs, const float* ys, const float* zs, int n) {
        check(k3dom_update_grid(handle_, t, xs, ys, zs, (int32_t)n), "updateGrid");
    }

    // The measurement grid of the last frame, one record per voxel.
    void copyMeasGridToHost(std::vector<MeasurementCell>& out) {
        std::vector<float> occ((size_t)grid_cell_count), free_m((size_t)grid_cell_count);
        const int64_t bytes = (int64_t)grid_cell_count * 4;
        check(k3dom_read_array(handle_, K3DOM_ARR_DEL_OCC, occ.data(), bytes), "copyMeasGridToHost");
        if (status_ != K3DOM_OK) return;
        check(k3dom_read_array(handle_, K3DOM_ARR_DEL_FREE, free_m.data(), bytes), "copyMeasGridToHost");
        if (status_ != K3DOM_OK) return;
        out.resize((size_t)grid_cell_count);
        for (size_t c = 0; c < out.size(); ++c) out[c] = MeasurementCell{free_m[c], occ[c]};
    }

    float window_radius;  // rays passing within this distance of a voxel centre are its neighbours
};

}  // namespace dom
