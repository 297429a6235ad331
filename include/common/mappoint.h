// Evaluation record of the C++ drop-in (youngjae-min/k3dom,
// include/common/mappoint.h:5-28): MapPoint with the same fields, order and
// 16-byte alignment (48 bytes), without the PCL point-type registration.
// DOM_c::copyMapPointsToHost fills it from the GPU (k3dom_export_mappoints).
#pragma once

#include "../k3dom.h"

struct alignas(16) MapPoint {
    float x, y, z;
    float data_w;   // the padding float of PCL_ADD_POINT4D (1)
    float v_x, v_y, v_z;
    float occ_val;
    float dyn_val;
    float eval_aug;
    float dyn_aug;
    float pad_;
};
static_assert(sizeof(MapPoint) == sizeof(k3dom_mappoint), "MapPoint must match k3dom_mappoint");
