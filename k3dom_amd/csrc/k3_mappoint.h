// MapPoint records of the node's evaluation topic (publish_mapdata,
// src/k3dom/k3dom_node.cpp:161-184, src/dsphdmib/dsphdmib_node.cpp:153-177)
// with the GridCell accessors of include/common/dom_types.h:47-64 and
// DOM_c::indexToCoord of include/common/dom.h:70-81, same float
// expressions.  Shared by the export kernels of k3_kernels.hip and the
// allocation-free export of k3_noalloc.hip.
#pragma once

#include "k3_kernels.h"

namespace k3 {

struct MpCell {
    float f, s, d, occ, vx, vy, vz;
};

__device__ __forceinline__ MpCell mp_cell(const VoxelArrays& v, int c, const MapPointArgs& a) {
    MpCell m;
    m.f = v.f[c];
    if (a.ds) {
        m.s = a.prior_s;
        m.d = a.prior_d;
        m.occ = v.s[c];
    } else {
        m.s = v.s[c];
        m.d = v.d[c];
        m.occ = a.prior_occ;
    }
    m.vx = v.mvx[c];
    m.vy = v.mvy[c];
    m.vz = v.mvz[c];
    return m;
}

// The visualisation filter of publish_map (src/k3dom/k3dom_node.cpp:137-138,
// src/dsphdmib/dsphdmib_node.cpp:134-135).
__device__ __forceinline__ bool mp_visible(const MpCell& m, const MapPointArgs& a) {
    if (a.ds) return m.occ + (1.0f - m.occ - m.f) / 2.0f > a.min_occ && m.occ + m.f > a.min_mass;
    const float total = m.f + m.s + m.d;
    return (m.s + m.d) / total > a.min_occ && total > a.min_mass;
}

// Record of voxel c as three float4 stores at out[3 * slot ..].
__device__ __forceinline__ void mp_store(float4* __restrict__ out, size_t slot, int c, const MpCell& m,
                                         const MapPointArgs& a) {
    const int gs2 = a.gs * a.gs;
    const float x = ((float)(c % a.gs) + 0.5f) * a.res - a.size / 2.0f + a.cx;
    const float y = ((float)((c % gs2) / a.gs) + 0.5f) * a.res - a.size / 2.0f + a.cy;
    const float z = ((float)(c / gs2) + 0.5f) * a.res - a.size_z / 2.0f + a.cz;
    float occ_val, dyn_val, eval_aug, dyn_aug;
    if (a.ds) {
        occ_val = m.occ + (1.0f - m.occ - m.f) / 2.0f;
        dyn_val = (m.vx * m.vx + m.vy * m.vy + m.vz * m.vz) * a.res * a.res;
        eval_aug = -(1.0f - m.occ - m.f);
        dyn_aug = m.occ;
    } else {
        const float total = m.f + m.s + m.d;
        occ_val = (m.s + m.d) / total;
        dyn_val = m.d / total;
        eval_aug = total;
        dyn_aug = 1.1f;
    }
    const size_t o = 3 * slot;
    out[o] = make_float4(x, y, z, 1.0f);
    out[o + 1] = make_float4(m.vx * a.res, m.vy * a.res, m.vz * a.res, occ_val);
    out[o + 2] = make_float4(dyn_val, eval_aug, dyn_aug, 0.0f);
}

}  // namespace k3
