// K3DOM filter class of the C++ drop-in (youngjae-min/k3dom,
// include/k3dom/dom.h:17-30): dom::DOM with kernel-inference evidence.
// GridCell, DOM_c and the differences a caller sees are in common/dom.h.
#pragma once

#include <cassert>
#include <vector>

#include "../common/dom.h"

namespace dom {

class DOM : public DOM_c {
public:
    explicit DOM(const Params& p) : DOM_c(p) {}

    // src/k3dom/dom.cu:29-50: one frame at time t (seconds), points relative
    // to the sensor in map-aligned axes (metres).  Lengths must be equal.
    void updateGrid(float t, std::vector<float>& measurements_x, std::vector<float>& measurements_y,
                    std::vector<float>& measurements_z) {
        assert(measurements_x.size() == measurements_y.size() && measurements_y.size() == measurements_z.size());
        updateGrid((double)t, measurements_x.data(), measurements_y.data(), measurements_z.data(),
                   (int)measurements_x.size());
    }
    // src/k3dom/dom.cu:52-77 and :79-106, the first-frame and evidence stages
    // of updateGrid for callers that drive the stages themselves (k3dom.h).
    void initializeParticles(std::vector<float>& measurements_x, std::vector<float>& measurements_y,
                             std::vector<float>& measurements_z) {
        assert(measurements_x.size() == measurements_y.size() && measurements_y.size() == measurements_z.size());
        check(k3dom_initialize_particles(handle_, measurements_x.data(), measurements_y.data(),
                                         measurements_z.data(), (int32_t)measurements_x.size()),
              "initializeParticles");
    }
    void gridCellOccupancyUpdate(float dt, std::vector<float>& measurements_x, std::vector<float>& measurements_y,
                                 std::vector<float>& measurements_z) {
        assert(measurements_x.size() == measurements_y.size() && measurements_y.size() == measurements_z.size());
        check(k3dom_grid_cell_occupancy_update(handle_, dt, measurements_x.data(), measurements_y.data(),
                                               measurements_z.data(), (int32_t)measurements_x.size()),
              "gridCellOccupancyUpdate");
    }
    // Same with a double time stamp (seconds) and plain host arrays of n points.
    void updateGrid(double t, const float* xs, const float* ys, const float* zs, int n) {
        check(k3dom_update_grid(handle_, t, xs, ys, zs, (int32_t)n), "updateGrid");
    }
};

}  // namespace dom
