// TEST INFRASTRUCTURE ONLY.  Serial C++ restatement of the K3DOM per-frame
// update (youngjae-min/k3dom) used as the parity checker for the HIP path and
// as the timed CPU baseline.  Nothing in the product links, loads or calls
// this file; only tests/, the smoke check and bench.py's CPU-baseline leg do.
//
// Restated from (read as text, never compiled or run):
//   src/k3dom/dom.cu:29-106          updateGrid / initializeParticles / gridCellOccupancyUpdate
//   src/common/dom.cu:111-297        updatePose and the stage drivers
//   src/k3dom/kernel/mass_update.cu  BGKI, predict_mass, update_mass, newborn split
//   src/common/kernel/*.cu           predict, particle_to_grid, update_persistent_particles,
//                                    init_new_particles, statistical_moments, resampling,
//                                    ego_motion_compensation, init
//   include/common/dom_types.h       GridCell fields and accessors
//   src/dsphdmib/dom.cu, src/dsphdmib/kernel/{measurement_grid,mass_update}.cu
//                                    the DS-PHD/MIB method (measurement grid and cell update)
//   src/k3dom/k3dom_node.cpp:73-83   the LiDAR-to-map point transform, generalised to
//                                    several LiDARs (no counterpart in the reference:
//                                    each LiDAR's evidence is computed as the reference
//                                    computes its one sensor's, with the sensor at that
//                                    LiDAR's origin, and the LiDARs' evidence is added
//                                    in LiDAR order)
//
// Deliberate, documented deviations from the reference (each also in the
// repository design notes):
//   * Philox4x32-10 keyed by (seed) with counters (index, frame, stream, word)
//     instead of per-thread cuRAND XORWOW states;
//   * systematic resampling on a u64 fixed-point CDF instead of multinomial
//     resampling on an FP32 scan (src/common/dom.cu:265-289);
//   * per-voxel segment sums (predicted dynamic mass, velocity moments) with
//     the two-level association of K3_SEG_BLOCK in k3_math.h (left-to-right
//     sums of 64-particle blocks, added left to right) instead of
//     differences of a global FP32 inclusive scan (include/common/common.h:24-28);
//   * newborn / initial particle counts from a u64 fixed-point CDF of the born
//     masses (every slot is assigned; a zero total spreads slots uniformly)
//     instead of truncating a normalised FP32 CDF
//     (src/common/kernel/init_new_particles.cu:14-52);
//   * BGKI sums use the two-level chunked association of k3_math.h and skip
//     pairs whose distance is >= ls (their term is exactly zero); the sparse
//     kernel is read as a function of the squared distance from the
//     piecewise-quadratic table of k3_kernel_table.h (<= 8e-8 of sigma from the
//     reference formula, no square root), exactly zero from d^2 = ls^2 on;
//   * gamma^dt is computed once per frame on the host.
//
// Parity pins for this file live in tests/test_oracle.py (brute-force BGKI ==
// culled BGKI bit for bit, closed-form mass rules, invariants: stable sort
// permutation, ranges tile [0,N), sum of newborn counts == B, uniform
// resampled weights), tests/test_oracle_pins.py (motion step, velocity draw,
// moments and map shift against closed forms, statistics and hand-built
// cases), tests/test_invariants.py (frame properties) and
// tests/test_fp64_errors.py (the FP64 stage evaluators and FP64 shadow below).

#include "../include/k3dom.h"
#include "../k3dom_amd/csrc/k3_math.h"
#include "../k3dom_amd/csrc/k3_dual.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;
static double ms_since(clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

// predictKernel, src/common/kernel/predict.cu:20-56: every particle i
// through k3_predict_particle with its Philox counters (i, frame, PREDICT).
static void predict_particles(const K3PredictConst& c, int n, float* px, float* py, float* pz, float* vx,
                              float* vy, float* vz, float* w, int32_t* key) {
    for (int i = 0; i < n; ++i)
        k3_predict_particle((uint32_t)i, c, &px[i], &py[i], &pz[i], &vx[i], &vy[i], &vz[i], &w[i], &key[i]);
}

// statisticalMomentsKernel1/2, src/common/kernel/statistical_moments.cu:12-63:
// per voxel with particles [start, end], sum(w v) / pers when pers > 0, else
// 0; 0 for voxels without particles.  The sums take the two-level
// association (left-to-right sums of K3_SEG_BLOCK-particle blocks from the
// voxel's first particle, added left to right).
static void voxel_moments(int V, const int32_t* start, const int32_t* end, const float* w, const float* vx,
                          const float* vy, const float* vz, const float* pers, float* mvx, float* mvy, float* mvz) {
    for (int c = 0; c < V; ++c) {
        if (start[c] != -1) {
            float sx = 0, sy = 0, sz = 0;
            for (int b = start[c]; b <= end[c]; b += K3_SEG_BLOCK) {
                float bx = 0, by = 0, bz = 0;
                for (int j = b; j <= std::min(end[c], b + K3_SEG_BLOCK - 1); ++j) {
                    bx += w[j] * vx[j];
                    by += w[j] * vy[j];
                    bz += w[j] * vz[j];
                }
                sx += bx;
                sy += by;
                sz += bz;
            }
            const float rho = pers[c];
            mvx[c] = rho > 0.0f ? sx / rho : 0.0f;
            mvy[c] = rho > 0.0f ? sy / rho : 0.0f;
            mvz[c] = rho > 0.0f ? sz / rho : 0.0f;
        } else {
            mvx[c] = 0; mvy[c] = 0; mvz[c] = 0;
        }
    }
}

// FP64 restatement of one voxel's predict_mass, update_mass and newborn split
// (src/k3dom/kernel/mass_update.cu:12-150,161-181) in double with libm: from
// (f, s, d) before prediction to out4 = (f, s, d) after the update and born.
static void mass_update64(double f, double s, double d, double dyn_pred, double decay, double del_occ,
                          double del_free, double mass_scale, double p_B, double* out4) {
    double pred_total = (f + d) * decay;
    s *= decay;
    d = std::min(dyn_pred, std::max(0.0, pred_total - s));
    f = pred_total - d;
    auto credit = [&](double m) { return std::tanh(m / mass_scale); };
    if (del_free > 0 || del_occ > 0) {
        double ds = del_free + del_occ, ps = f + s + d;
        double c = credit(std::sqrt(ds * ps));
        double l_DtoF = c * std::max(0.0, del_free - del_occ) / ds * d / ps;
        double l_StoFD = c * std::max(0.0, del_free - del_occ) / ds * std::max(0.0, s + d - f) / ps;
        double l_FtoD = c * std::max(0.0, del_occ - del_free) / ds * std::max(0.0, f - s - d) / ps;
        double df = l_StoFD / 2 * s + l_DtoF * d - l_FtoD * f;
        double dss = -l_StoFD * s;
        double dd = l_FtoD * f + l_StoFD / 2 * s - l_DtoF * d;
        f += df; s += dss; d += dd;
        c = credit(s + d);
        double ratio = (s <= 0 || d < 0) ? 0.0 : 1.0 / (1.0 + d / s);
        double beta = (1 - c) + c * ratio;
        s = s + beta * del_occ;
        d = d + (1 - beta) * del_occ;
        f = f + del_free;
    }
    double born;
    if (dyn_pred <= 0) born = d;
    else if (pred_total <= dyn_pred) born = 0;
    else born = d * p_B * (pred_total - dyn_pred) / (dyn_pred + p_B * (pred_total - dyn_pred));
    out4[0] = f; out4[1] = s; out4[2] = d; out4[3] = born;
}


struct BgkiStats {
    uint64_t occ_pairs = 0;
    uint64_t free_pairs = 0;
};

// ----------------------------------------------------------------------------
// BGKI over all voxels: del_occ[v], del_free[v] with the two-level chunked
// association.  brute = true visits every (voxel, point) pair (the reference's
// literal double loop, src/k3dom/kernel/mass_update.cu:43-72); brute = false
// visits, for every point in ascending order, only voxels whose centre lies
// within L-infinity distance ls (+slack) of the segment [0, p] — a superset of
// the pairs with d < ls, whose terms are the only non-zero ones.
struct ChunkAcc {
    std::vector<float> tot_o, tot_f, cur_o, cur_f;
    std::vector<int32_t> chunk_o, chunk_f;
    void init(size_t V) {
        tot_o.assign(V, 0.0f); tot_f.assign(V, 0.0f);
        cur_o.assign(V, 0.0f); cur_f.assign(V, 0.0f);
        chunk_o.assign(V, -1); chunk_f.assign(V, -1);
    }
    inline void add(std::vector<float>& tot, std::vector<float>& cur, std::vector<int32_t>& ch,
                    size_t v, int32_t c, float term) {
        if (ch[v] != c) {
            tot[v] += cur[v];
            cur[v] = 0.0f;
            ch[v] = c;
        }
        cur[v] += term;
    }
    // the sums of kernel brackets times the sigma factor (k3_kernel_bracket_s)
    void finish(float* del_occ, float* del_free, float sigma) {
        const float sf = k3_sigma_factor(sigma);
        for (size_t v = 0; v < tot_o.size(); ++v) {
            del_occ[v] = sf * (tot_o[v] + cur_o[v]);
            del_free[v] = sf * (tot_f[v] + cur_f[v]);
        }
    }
};

struct Geom {
    int gs, gsz;
    float res, sx, sy, sz;
    std::vector<float> cx, cy, cz;  // voxel centre coordinate per index along each axis
    void build() {
        cx.resize(gs); cy.resize(gs); cz.resize(gsz);
        for (int i = 0; i < gs; ++i) {
            float x, y, z;
            k3_voxel_center(i, gs, gsz, res, sx, sy, sz, &x, &y, &z);
            cx[i] = x;
            k3_voxel_center(i * gs, gs, gsz, res, sx, sy, sz, &x, &y, &z);
            cy[i] = y;
        }
        for (int k = 0; k < gsz; ++k) {
            float x, y, z;
            k3_voxel_center(k * gs * gs, gs, gsz, res, sx, sy, sz, &x, &y, &z);
            cz[k] = z;
        }
    }
};

// Two-level sum of x[st..en] (K3_SEG_BLOCK): blocks of K3_SEG_BLOCK elements
// from st, each summed left to right, block sums added left to right.
static float seg_sum(const float* x, int st, int en) {
    float s = 0.0f;
    for (int b = st; b <= en; b += K3_SEG_BLOCK) {
        float bs = 0.0f;
        for (int j = b; j <= std::min(en, b + K3_SEG_BLOCK - 1); ++j) bs += x[j];
        s += bs;
    }
    return s;
}

static void bgki_pair(const Geom& g, int i, int j, int k, const K3PointRec& pr, float kscale, int32_t chunk,
                      ChunkAcc& acc, BgkiStats& st) {
    float vx = g.cx[i], vy = g.cy[j], vz = g.cz[k];
    size_t v = (size_t)i + (size_t)g.gs * j + (size_t)g.gs * g.gs * k;
    float dc2s = fmaf(vz, vz, fmaf(vy, vy, vx * vx)) * kscale;
    float s_o = k3_occ_s(vx, vy, vz, pr, kscale);
    if (s_o >= 0.0f) {
        acc.add(acc.tot_o, acc.cur_o, acc.chunk_o, v, chunk, k3_kernel_bracket_s(s_o));
        st.occ_pairs++;
    }
    float s_f = k3_free_s(vx, vy, vz, dc2s, pr, kscale);
    if (s_f >= 0.0f) {
        acc.add(acc.tot_f, acc.cur_f, acc.chunk_f, v, chunk, k3_kernel_bracket_s(s_f));
        st.free_pairs++;
    }
}

// index range [lo, hi] of voxel centres c[idx] = (idx + 0.5 - n/2) * res - s that
// may lie in [a, b] (slack of one voxel on each side; the exact test decides).
static inline bool axis_range(double a, double b, double res, double s, int n, int* lo, int* hi) {
    double fa = (a + s) / res + n / 2.0 - 0.5;
    double fb = (b + s) / res + n / 2.0 - 0.5;
    double l = std::floor(fa) - 1.0, h = std::ceil(fb) + 1.0;
    if (h < 0.0 || l > n - 1.0) return false;
    *lo = (int)std::max(0.0, l);
    *hi = (int)std::min((double)(n - 1), h);
    return *lo <= *hi;
}

// t-interval of the segment t*m (t in [t0,t1]) where |t*m_a - c| <= R
static inline bool narrow(double m, double c, double R, double* t0, double* t1) {
    if (std::fabs(m) < 1e-30) return std::fabs(c) <= R;
    double ta = (c - R) / m, tb = (c + R) / m;
    if (ta > tb) std::swap(ta, tb);
    *t0 = std::max(*t0, ta);
    *t1 = std::min(*t1, tb);
    return *t0 <= *t1;
}

// Threads of the culled kernel inference (orc_set_threads; 1 = serial).
static int g_bgki_threads = 1;

// Culled kernel inference over the voxels with x index in [xlo, xhi): every
// point in order, so each voxel's terms arrive in point order whatever the
// slab, and the sums do not depend on how the x axis is split.
static void bgki_culled_slab(const Geom& g, const float* xs, const float* ys, const float* zs, int n, float ls,
                             int xlo, int xhi, ChunkAcc& acc, BgkiStats& st) {
    const float kscale = k3_kernel_scale(ls);
    const float sk = k3_kernel_sk(ls);
    const double R = (double)ls * (1.0 + 1e-5) + 1e-4;
    for (int p = 0; p < n; ++p) {
        K3PointRec pr = k3_point_prep(xs[p], ys[p], zs[p], g.res, sk);
        if (pr.valid == 0.0f) continue;
        int32_t chunk = p >> K3_BGKI_CHUNK_LOG2;
        double mx = pr.x, my = pr.y, mz = pr.z;
        int i0, i1;
        if (!axis_range(std::min(0.0, mx) - R, std::max(0.0, mx) + R, g.res, g.sx, g.gs, &i0, &i1)) continue;
        i0 = std::max(i0, xlo);
        i1 = std::min(i1, xhi - 1);
        for (int i = i0; i <= i1; ++i) {
            double ta = 0.0, tb = 1.0;
            if (!narrow(mx, g.cx[i], R, &ta, &tb)) continue;
            int j0, j1;
            double ylo = std::min(ta * my, tb * my) - R, yhi = std::max(ta * my, tb * my) + R;
            if (!axis_range(ylo, yhi, g.res, g.sy, g.gs, &j0, &j1)) continue;
            for (int j = j0; j <= j1; ++j) {
                double tc = ta, td = tb;
                if (!narrow(my, g.cy[j], R, &tc, &td)) continue;
                int k0, k1;
                double zlo = std::min(tc * mz, td * mz) - R, zhi = std::max(tc * mz, td * mz) + R;
                if (!axis_range(zlo, zhi, g.res, g.sz, g.gsz, &k0, &k1)) continue;
                for (int k = k0; k <= k1; ++k)
                    bgki_pair(g, i, j, k, pr, kscale, chunk, acc, st);
            }
        }
    }
}

static BgkiStats bgki_all(const Geom& g, const float* xs, const float* ys, const float* zs, int n,
                          float ls, float sigma, bool brute, float* del_occ, float* del_free) {
    const size_t V = (size_t)g.gs * g.gs * g.gsz;
    ChunkAcc acc;
    acc.init(V);
    BgkiStats st;
    if (brute) {
        const float kscale = k3_kernel_scale(ls);
        const float sk = k3_kernel_sk(ls);
        for (int p = 0; p < n; ++p) {
            K3PointRec pr = k3_point_prep(xs[p], ys[p], zs[p], g.res, sk);
            if (pr.valid == 0.0f) continue;
            int32_t chunk = p >> K3_BGKI_CHUNK_LOG2;
            for (int k = 0; k < g.gsz; ++k)
                for (int j = 0; j < g.gs; ++j)
                    for (int i = 0; i < g.gs; ++i)
                        bgki_pair(g, i, j, k, pr, kscale, chunk, acc, st);
        }
    } else if (g_bgki_threads <= 1) {
        bgki_culled_slab(g, xs, ys, zs, n, ls, 0, g.gs, acc, st);
    } else {
        // x slabs of at most 4 voxels, taken dynamically: voxels of different
        // slabs are disjoint, so the threads never write the same element
        const int slabs = (g.gs + 3) / 4;
        std::vector<BgkiStats> sst(slabs);
#pragma omp parallel for num_threads(g_bgki_threads) schedule(dynamic, 1)
        for (int s = 0; s < slabs; ++s)
            bgki_culled_slab(g, xs, ys, zs, n, ls, 4 * s, std::min(g.gs, 4 * s + 4), acc, sst[s]);
        for (const BgkiStats& x : sst) {
            st.occ_pairs += x.occ_pairs;
            st.free_pairs += x.free_pairs;
        }
    }
    acc.finish(del_occ, del_free, sigma);
    return st;
}

// Boundaries of a fixed-point CDF split into `cnt` slots: bnd[j] = number of
// slots owned by cells 0..j.  q is the u64 fixed-point mass per cell.
static void cdf_bounds(const std::vector<uint64_t>& q, uint32_t cnt, std::vector<uint32_t>& bnd) {
    const size_t V = q.size();
    bnd.resize(V);
    uint64_t tot = 0;
    for (size_t j = 0; j < V; ++j) tot += q[j];
    if (tot == 0) {
        for (size_t j = 0; j < V; ++j) bnd[j] = (uint32_t)(((uint64_t)(j + 1) * cnt) / V);
        return;
    }
    double scale = (double)cnt / (double)tot;
    uint64_t c = 0;
    for (size_t j = 0; j < V; ++j) {
        c += q[j];
        if (c == tot) {
            bnd[j] = cnt;
        } else {
            uint64_t b = (uint64_t)((double)c * scale);
            bnd[j] = (uint32_t)std::min<uint64_t>(b, cnt);
        }
    }
}

static void to_fixed(const float* w, size_t n, std::vector<uint64_t>& q, int* E_out) {
    float wmax = 0.0f;
    for (size_t i = 0; i < n; ++i)
        if (w[i] > wmax && w[i] - w[i] == 0.0f) wmax = w[i];
    int E = k3_fixed_exponent(wmax, (uint32_t)n);
    q.resize(n);
    for (size_t i = 0; i < n; ++i) q[i] = k3_to_fixed(w[i], E);
    *E_out = E;
}

struct Oracle {
    k3dom_params p;
    int gs, gsz, V, N, B;
    // voxel state (GridCell fields as arrays)
    std::vector<int32_t> start, end;
    std::vector<float> muA, muUA, wA, wUA, mvx, mvy, mvz, lik, pA, f, s, d, occ, pers, pred;
    std::vector<float> born;
    // particles (current set) and newborn set
    std::vector<float> px, py, pz, vx, vy, vz, w;
    std::vector<int32_t> key;
    std::vector<float> bpx, bpy, bpz, bvx, bvy, bvz, bw;
    std::vector<int32_t> bkey;
    // diagnostics of the last frame
    std::vector<int32_t> perm, sorted_key, resample_idx;
    std::vector<float> sorted_w, del_occ, del_free;
    // frozen counts of the last frame per voxel (differentiated window): sorted
    // particles inside the grid, of them not penalised; newborns, of them not penalised
    std::vector<int32_t> cnt_n, cnt_m, cnt_nb, cnt_mb;
    BgkiStats bgki_stats;
    double stage_ms[10];
    uint32_t frame = 0;
    bool first_pose = false;
    float sensor[3] = {0, 0, 0}, center[3] = {0, 0, 0};
    double updated_time = -1.0;
    bool brute = false;
    int method = K3DOM_METHOD_K3DOM;
    // LiDAR groups of the current frame (empty: one sensor at the pose):
    // points [start, start + n) with their ray origin o (relative to the
    // pose, map-aligned axes)
    struct Group {
        int start, n;
        float o[3];
    };
    std::vector<Group> groups;

    explicit Oracle(const k3dom_params& prm) : p(prm) {
        gs = (int)(p.size / p.resolution);
        gsz = (int)(p.size_z / p.resolution);
        V = gs * gs * gsz;
        N = p.particle_count;
        B = p.new_born_particle_count;
        start.assign(V, -1); end.assign(V, -1);
        muA.assign(V, 0); muUA.assign(V, 0); wA.assign(V, 0); wUA.assign(V, 0);
        mvx.assign(V, 0); mvy.assign(V, 0); mvz.assign(V, 0);
        lik.assign(V, 1.0f); pA.assign(V, 0.0f);
        f.assign(V, p.prior_free); s.assign(V, p.prior_static); d.assign(V, p.prior_dynamic);
        occ.assign(V, p.prior_occ); pers.assign(V, 0); pred.assign(V, 0);
        born.assign(V, 0);
        px.assign(N, 0); py.assign(N, 0); pz.assign(N, 0);
        vx.assign(N, 0); vy.assign(N, 0); vz.assign(N, 0); w.assign(N, 0); key.assign(N, 0);
        bpx.assign(B, 0); bpy.assign(B, 0); bpz.assign(B, 0);
        bvx.assign(B, 0); bvy.assign(B, 0); bvz.assign(B, 0); bw.assign(B, 0); bkey.assign(B, 0);
        for (double& x : stage_ms) x = 0.0;
    }

    uint32_t key0() const { return (uint32_t)p.seed; }
    uint32_t key1() const { return (uint32_t)(p.seed >> 32); }

    // FP64 shadow of the last frame's floating-point stages (keep64): the masses
    // before the mass update, then in double from the frame's own FP32 inputs:
    // the predicted dynamic mass (segment sum of the sorted weights), the
    // mass rules (f, s, d, born) and the mean velocities.
    bool keep64 = false;
    enum { E_PRE_F, E_PRE_S, E_PRE_D, E_DYN, E_F, E_S, E_D, E_BORN, E_MVX, E_MVY, E_MVZ, E_COUNT };
    std::vector<double> e64[E_COUNT];

    // DOM_c::updatePose, src/common/dom.cu:111-158
    void update_pose(float nx, float ny, float nz) {
        sensor[0] = nx; sensor[1] = ny; sensor[2] = nz;
        if (!first_pose) {
            first_pose = true;
            center[0] = sensor[0] - p.sensor_off_x;
            center[1] = sensor[1] - p.sensor_off_y;
            center[2] = sensor[2] - p.sensor_off_z;
            return;
        }
        const int mx = (int)((sensor[0] - center[0] - p.sensor_off_x) / p.resolution);
        const int my = (int)((sensor[1] - center[1] - p.sensor_off_y) / p.resolution);
        const int mz = (int)((sensor[2] - center[2] - p.sensor_off_z) / p.resolution);
        if (std::abs(mx) >= p.map_shift_thresh || std::abs(my) >= p.map_shift_thresh ||
            std::abs(mz) >= p.map_shift_thresh) {
            shift_map(mx, my, mz);
            center[0] += mx * p.resolution;
            center[1] += my * p.resolution;
            center[2] += mz * p.resolution;
        }
    }

    // initGridCellsKernel + moveParticlesKernel + moveMapKernel
    // (src/common/kernel/init.cu:20-35, src/common/kernel/ego_motion_compensation.cu:12-41)
    void shift_map(int mx, int my, int mz) {
        for (int i = 0; i < N; ++i) {
            px[i] -= mx;
            py[i] -= my;
            pz[i] -= mz;
        }
        auto shift = [&](std::vector<float>& a, bool reset, float prior) {
            std::vector<float> old = a;
            for (int c = 0; c < V; ++c) {
                int x = c % gs, y = (c % (gs * gs)) / gs, z = c / (gs * gs);
                int nx = x + mx, ny = y + my, nz = z + mz;
                if (nx >= 0 && nx < gs && ny >= 0 && ny < gs && nz >= 0 && nz < gsz)
                    a[c] = old[nx + gs * ny + gs * gs * nz];
                else if (reset)
                    a[c] = prior;
            }
        };
        auto shifti = [&](std::vector<int32_t>& a) {
            std::vector<int32_t> old = a;
            for (int c = 0; c < V; ++c) {
                int x = c % gs, y = (c % (gs * gs)) / gs, z = c / (gs * gs);
                int nx = x + mx, ny = y + my, nz = z + mz;
                if (nx >= 0 && nx < gs && ny >= 0 && ny < gs && nz >= 0 && nz < gsz)
                    a[c] = old[nx + gs * ny + gs * gs * nz];
                else
                    a[c] = -1;
            }
        };
        shift(f, true, p.prior_free);
        shift(s, true, p.prior_static);
        shift(d, true, p.prior_dynamic);
        shift(occ, true, p.prior_occ);
        shift(lik, true, 1.0f);
        shift(pA, true, 0.0f);
        shifti(start);
        shifti(end);
        shift(muA, false, 0); shift(muUA, false, 0); shift(wA, false, 0); shift(wUA, false, 0);
        shift(mvx, false, 0); shift(mvy, false, 0); shift(mvz, false, 0);
        shift(pers, false, 0); shift(pred, false, 0);
    }

    // gridCellPredictionUpdateKernel, src/k3dom/kernel/mass_update.cu:152-183
    void mass_update(float decay, const float* xs, const float* ys, const float* zs, int n,
                     const std::vector<float>& wsorted) {
        auto t0 = clk::now();
        if (groups.empty()) {
            Geom g{gs, gsz, p.resolution, sensor[0] - center[0], sensor[1] - center[1], sensor[2] - center[2]};
            g.build();
            del_occ.assign(V, 0.0f);
            del_free.assign(V, 0.0f);
            bgki_stats = bgki_all(g, xs, ys, zs, n, p.ls, p.sigma, brute, del_occ.data(), del_free.data());
        } else {
            // one BGKI per LiDAR with the sensor at its origin, added in LiDAR order
            std::vector<float> go(V), gf(V);
            for (size_t k = 0; k < groups.size(); ++k) {
                const Group& G = groups[k];
                Geom g{gs, gsz, p.resolution, (sensor[0] + G.o[0]) - center[0], (sensor[1] + G.o[1]) - center[1],
                       (sensor[2] + G.o[2]) - center[2]};
                g.build();
                float* to = k == 0 ? (del_occ.assign(V, 0.0f), del_occ.data()) : go.data();
                float* tf = k == 0 ? (del_free.assign(V, 0.0f), del_free.data()) : gf.data();
                bgki_all(g, xs + G.start, ys + G.start, zs + G.start, G.n, p.ls, p.sigma, brute, to, tf);
                if (k > 0)
                    for (int i = 0; i < V; ++i) {
                        del_occ[i] = del_occ[i] + go[i];
                        del_free[i] = del_free[i] + gf[i];
                    }
            }
        }
        stage_ms[3] = ms_since(t0);
        t0 = clk::now();
        if (keep64)
            for (auto& a : e64) a.assign(V, 0.0);
        for (int i = 0; i < V; ++i) {
            if (keep64) {
                double dyn64 = 0.0;
                if (start[i] != -1)
                    for (int j = start[i]; j <= end[i]; ++j) dyn64 += (double)wsorted[j];
                double o4[4];
                const float dyn32 = start[i] != -1 ? seg_sum(wsorted.data(), start[i], end[i]) : 0.0f;
                mass_update64(f[i], s[i], d[i], dyn32, decay, del_occ[i], del_free[i], p.mass_scale, p.birth_prob,
                              o4);
                e64[E_PRE_F][i] = f[i]; e64[E_PRE_S][i] = s[i]; e64[E_PRE_D][i] = d[i];
                e64[E_DYN][i] = dyn64;
                e64[E_F][i] = o4[0]; e64[E_S][i] = o4[1]; e64[E_D][i] = o4[2]; e64[E_BORN][i] = o4[3];
            }
            float pred_total = (f[i] + d[i]) * decay;
            const float dyn_pred = start[i] != -1 ? seg_sum(wsorted.data(), start[i], end[i]) : 0.0f;
            k3_predict_mass(&f[i], &s[i], &d[i], dyn_pred, pred_total, decay);
            k3_update_mass(&f[i], &s[i], &d[i], del_occ[i], del_free[i], p.mass_scale);
            born[i] = k3_newborn_part(dyn_pred, pred_total, d[i], p.birth_prob);
            pers[i] = d[i] - born[i];
            pred[i] = dyn_pred;
        }
        stage_ms[4] = ms_since(t0);
    }

    // DS-PHD/MIB measurement grid, src/dsphdmib/kernel/measurement_grid.cu:64-160:
    // every voxel scans the points in index order and keeps the first
    // K3DS_MAX_NEIGHBORS whose half-line from the sensor passes within
    // 0.866 resolution of its centre (src/dsphdmib/dom.cu:21-22).  The masses go
    // to del_occ / del_free, the arrays that hold the evidence in the K3DOM path.
    void ds_meas_grid(const float* xs, const float* ys, const float* zs, int n) {
        auto t0 = clk::now();
        // one LiDAR at most: with the sensor at its origin
        const float o[3] = {groups.empty() ? 0.0f : groups[0].o[0], groups.empty() ? 0.0f : groups[0].o[1],
                            groups.empty() ? 0.0f : groups[0].o[2]};
        Geom g{gs, gsz, p.resolution, (sensor[0] + o[0]) - center[0], (sensor[1] + o[1]) - center[1],
               (sensor[2] + o[2]) - center[2]};
        g.build();
        const float wr = p.resolution * 0.866f;
        const float w2 = wr * wr;
        std::vector<float> m2(n);
        for (int q = 0; q < n; ++q) m2[q] = xs[q] * xs[q] + ys[q] * ys[q] + zs[q] * zs[q];
        del_occ.assign(V, 0.0f);
        del_free.assign(V, 0.0f);
        for (int c = 0; c < V; ++c) {
            const float x = g.cx[c % gs], y = g.cy[(c % (gs * gs)) / gs], z = g.cz[c / (gs * gs)];
            const float c2 = x * x + y * y + z * z;
            int nb = 0;
            bool all_farther = true;
            float max_occ = 0.0f;
            for (int q = 0; q < n && nb < K3DS_MAX_NEIGHBORS; ++q) {
                if (!k3ds_ray_near(x, y, z, c2, xs[q], ys[q], zs[q], m2[q], w2)) continue;
                if (all_farther && c2 > m2[q]) all_farther = false;
                const float o = k3ds_pocc(x, y, z, xs[q], ys[q], zs[q], p.meas_occ_stddev);
                if (o > max_occ) max_occ = o;
                ++nb;
            }
            k3ds_meas_masses(nb, all_farther, max_occ, x, y, z, p.max_range, &del_occ[c], &del_free[c]);
        }
        stage_ms[3] = ms_since(t0);
    }

    // DS-PHD/MIB cell update, gridCellPredictionUpdateKernel
    // (src/dsphdmib/kernel/mass_update.cu:44-85): predicted occupied mass =
    // the voxel's particle weights (two-level sum); above 1 the weights are
    // divided by it and it becomes 1.  The sorted weights are updated in place.
    void ds_mass_update(const float* xs, const float* ys, const float* zs, int n) {
        ds_meas_grid(xs, ys, zs, n);
        auto t0 = clk::now();
        for (int c = 0; c < V; ++c) {
            float occ_up, free_up, rho = 0.0f, m_occ_pred = 0.0f;
            if (start[c] != -1) {
                m_occ_pred = seg_sum(w.data(), start[c], end[c]);
                if (m_occ_pred > 1.0f) {
                    for (int j = start[c]; j <= end[c]; ++j) w[j] /= m_occ_pred;
                    m_occ_pred = 1.0f;
                }
                const float m_free_pred = k3ds_predict_free(f[c], m_occ_pred);
                k3ds_update_masses(m_occ_pred, m_free_pred, del_occ[c], del_free[c], &occ_up, &free_up);
                rho = k3ds_newborn_part(m_occ_pred, occ_up, p.birth_prob);
            } else {
                const float m_occ = occ[c];
                const float m_free = k3ds_predict_free(f[c], m_occ);
                k3ds_update_masses(m_occ, m_free, del_occ[c], del_free[c], &occ_up, &free_up);
            }
            born[c] = rho;
            pers[c] = occ_up - rho;
            f[c] = free_up;
            occ[c] = occ_up;
            pred[c] = m_occ_pred;
        }
        stage_ms[4] = ms_since(t0);
    }

    // particle slots of a fixed-point CDF over the born masses
    void slots_from_born(uint32_t cnt, std::vector<uint32_t>& bnd) {
        std::vector<uint64_t> q;
        int E;
        to_fixed(born.data(), (size_t)V, q, &E);
        cdf_bounds(q, cnt, bnd);
    }

    static int owner(const std::vector<uint32_t>& bnd, uint32_t k) {
        return (int)(std::upper_bound(bnd.begin(), bnd.end(), k) - bnd.begin());
    }

    // DOM::initializeParticles, src/k3dom/dom.cu:52-77 (DS-PHD/MIB:
    // src/dsphdmib/dom.cu:55-82, born masses = the measurement grid's occupied
    // masses, cells untouched) and initParticlesKernel1/2,
    // src/common/kernel/init_new_particles.cu:54-104
    void initialize_particles(const float* xs, const float* ys, const float* zs, int n) {
        if (method == K3DOM_METHOD_DSPHDMIB) {
            ds_meas_grid(xs, ys, zs, n);
            born = del_occ;
        } else {
            mass_update(1.0f, xs, ys, zs, n, w);
        }
        auto t0 = clk::now();
        std::vector<uint32_t> bnd;
        slots_from_born((uint32_t)N, bnd);
        const float new_weight = 1.0f / N;
        for (int i = 0; i < N; ++i) {
            int c = owner(bnd, (uint32_t)i);
            uint32_t r[4];
            k3_philox4x32_10((uint32_t)i, frame, K3_STREAM_INIT, 0u, key0(), key1(), r);
            px[i] = (float)(c % gs) + 0.5f;
            py[i] = (float)((c % (gs * gs)) / gs) + 0.5f;
            pz[i] = (float)(c / (gs * gs)) + 0.5f;
            k3_sample_velocity(r[0], r[1], r[2], p.init_max_velocity, p.particle_min_vel, &vx[i], &vy[i], &vz[i]);
            w[i] = new_weight;
            key[i] = c;
        }
        stage_ms[6] = ms_since(t0);
    }

    void predict(float dt) {
        K3PredictConst c;
        c.dt = dt; c.p_S = p.persistence_prob; c.sigma_pos = p.stddev_process_noise_position;
        c.sigma_vel = p.stddev_process_noise_velocity; c.vclamp = 5 * p.init_max_velocity;
        c.gs = gs; c.gsz = gsz; c.frame = frame; c.key0 = key0(); c.key1 = key1();
        predict_particles(c, N, px.data(), py.data(), pz.data(), vx.data(), vy.data(), vz.data(), w.data(),
                          key.data());
    }

    // stable counting sort by key + particleToGridKernel
    // (src/common/dom.cu:169-188, src/common/kernel/particle_to_grid.cu:13-39)
    void assign() {
        std::vector<int32_t> cnt(V + 1, 0);
        for (int i = 0; i < N; ++i) cnt[key[i] + 1]++;
        for (int v = 0; v < V; ++v) cnt[v + 1] += cnt[v];
        perm.assign(N, 0);
        for (int i = 0; i < N; ++i) perm[cnt[key[i]]++] = i;
        auto gather = [&](std::vector<float>& a) {
            std::vector<float> o(N);
            for (int i = 0; i < N; ++i) o[i] = a[perm[i]];
            a.swap(o);
        };
        gather(px); gather(py); gather(pz); gather(vx); gather(vy); gather(vz); gather(w);
        std::vector<int32_t> k2(N);
        for (int i = 0; i < N; ++i) k2[i] = key[perm[i]];
        key.swap(k2);
        sorted_key = key;
        std::fill(start.begin(), start.end(), -1);
        std::fill(end.begin(), end.end(), -1);
        for (int i = 0; i < N; ++i) {
            int j = key[i];
            if (i == 0 || key[i] != key[i - 1]) start[j] = i;
            if (i == N - 1 || key[i] != key[i + 1]) end[j] = i;
        }
    }

    // updatePersistentParticlesKernel1-3, src/common/kernel/update_persistent_particles.cu:17-55
    void update_persistent() {
        for (int c = 0; c < V; ++c) {
            if (start[c] != -1) {
                float rho = pers[c];
                float m = pred[c] > 0.0f ? rho / pred[c] : 0.0f;
                muA[c] = m;
                muUA[c] = m;
            }
        }
        for (int i = 0; i < N; ++i) {
            int c = key[i];
            float post = lik[c] * w[i];
            w[i] = pA[c] * muA[c] * post + (1.0f - pA[c]) * muUA[c] * w[i];
        }
    }

    // initializeNewParticles, src/common/dom.cu:213-232,
    // initNewParticlesKernel1/2, src/common/kernel/init_new_particles.cu:106-183
    void newborn() {
        std::vector<uint32_t> bnd;
        slots_from_born((uint32_t)B, bnd);
        for (int c = 0; c < V; ++c) {
            uint32_t lo = c ? bnd[c - 1] : 0;
            int n = (int)(bnd[c] - lo);
            wA[c] = 0.0f;
            wUA[c] = n > 0 ? (float)(((1.0 - (double)pA[c]) * (double)born[c]) / (double)n) : 0.0f;
        }
        for (int k = 0; k < B; ++k) {
            int c = owner(bnd, (uint32_t)k);
            uint32_t r[4];
            k3_philox4x32_10((uint32_t)k, frame, K3_STREAM_NEWBORN, 0u, key0(), key1(), r);
            bpx[k] = (float)(c % gs) + 0.5f;
            bpy[k] = (float)((c % (gs * gs)) / gs) + 0.5f;
            bpz[k] = (float)(c / (gs * gs)) + 0.5f;
            k3_sample_velocity(r[0], r[1], r[2], p.init_max_velocity, p.particle_min_vel, &bvx[k], &bvy[k], &bvz[k]);
            bw[k] = wUA[c];
            bkey[k] = c;
        }
    }

    // statisticalMomentsKernel1/2, src/common/kernel/statistical_moments.cu:22-63
    void moments() {
        voxel_moments(V, start.data(), end.data(), w.data(), vx.data(), vy.data(), vz.data(), pers.data(),
                      mvx.data(), mvy.data(), mvz.data());
        if (keep64 && (int)e64[E_MVX].size() == V)
            for (int c = 0; c < V; ++c) {
                double sx = 0, sy = 0, sz = 0;
                if (start[c] != -1)
                    for (int j = start[c]; j <= end[c]; ++j) {
                        sx += (double)w[j] * vx[j];
                        sy += (double)w[j] * vy[j];
                        sz += (double)w[j] * vz[j];
                    }
                const double rho = pers[c];
                e64[E_MVX][c] = rho > 0 ? sx / rho : 0.0;
                e64[E_MVY][c] = rho > 0 ? sy / rho : 0.0;
                e64[E_MVZ][c] = rho > 0 ? sz / rho : 0.0;
            }
    }

    // resampling, src/common/dom.cu:259-297 and src/common/kernel/resampling.cu:13-89,
    // systematic on a u64 fixed-point CDF
    void resample() {
        const float min2 = p.particle_min_vel * p.particle_min_vel;
        for (int i = 0; i < N; ++i)
            if (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i] < min2) w[i] = 0.0f;
        for (int k = 0; k < B; ++k)
            if (bvx[k] * bvx[k] + bvy[k] * bvy[k] + bvz[k] * bvz[k] < min2) bw[k] = 0.0f;
        sorted_w = w;
        std::vector<float> joint(w);
        joint.insert(joint.end(), bw.begin(), bw.end());
        std::vector<uint64_t> q;
        int E;
        to_fixed(joint.data(), joint.size(), q, &E);
        std::vector<uint64_t> cdf(q.size());
        uint64_t acc = 0;
        for (size_t i = 0; i < q.size(); ++i) { acc += q[i]; cdf[i] = acc; }
        const uint64_t Q = acc;
        const float W = (float)((double)Q * k3_pow2d(-E));
        resample_idx.assign(N, 0);
        if (Q == 0) {
            for (int i = 0; i < N; ++i) resample_idx[i] = i;
        } else {
            uint32_t r[4];
            k3_philox4x32_10(0u, frame, K3_STREAM_RESAMPLE, 0u, key0(), key1(), r);
            const double u = (double)r[0] * 0x1p-32;
            const double step = (double)Q / (double)N;
            for (int i = 0; i < N; ++i) {
                uint64_t T = (uint64_t)(((double)i + u) * step);
                if (T > Q - 1) T = Q - 1;
                resample_idx[i] = (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), T) - cdf.begin());
            }
        }
        const float new_weight = W / N;
        std::vector<float> npx(N), npy(N), npz(N), nvx(N), nvy(N), nvz(N), nw(N);
        std::vector<int32_t> nk(N);
        for (int i = 0; i < N; ++i) {
            int idx = resample_idx[i];
            if (idx < N) {
                npx[i] = px[idx]; npy[i] = py[idx]; npz[i] = pz[idx];
                nvx[i] = vx[idx]; nvy[i] = vy[idx]; nvz[i] = vz[idx]; nk[i] = key[idx];
            } else {
                int b = idx - N;
                npx[i] = bpx[b]; npy[i] = bpy[b]; npz[i] = bpz[b];
                nvx[i] = bvx[b]; nvy[i] = bvy[b]; nvz[i] = bvz[b]; nk[i] = bkey[b];
            }
            nw[i] = new_weight;
        }
        px.swap(npx); py.swap(npy); pz.swap(npz); vx.swap(nvx); vy.swap(nvy); vz.swap(nvz);
        w.swap(nw); key.swap(nk);
    }

    // inside test of k3_predict_particle, penalty test of resample()
    void record_counts() {
        const float min2 = p.particle_min_vel * p.particle_min_vel;
        cnt_n.assign(V, 0); cnt_m.assign(V, 0); cnt_nb.assign(V, 0); cnt_mb.assign(V, 0);
        for (int i = 0; i < N; ++i) {
            const float x = px[i], y = py[i], z = pz[i];
            const bool inb = !((x >= (float)gs || x < 0.0f) || (y >= (float)gs || y < 0.0f) ||
                               (z >= (float)gsz || z < 0.0f));
            const bool pen = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i] < min2;
            cnt_n[key[i]] += inb;
            cnt_m[key[i]] += inb && !pen;
        }
        for (int k = 0; k < B; ++k) {
            cnt_nb[bkey[k]] += 1;
            cnt_mb[bkey[k]] += !(bvx[k] * bvx[k] + bvy[k] * bvy[k] + bvz[k] * bvz[k] < min2);
        }
    }

    // DOM::updateGrid, src/k3dom/dom.cu:29-50
    void update_grid(double t, const float* xs, const float* ys, const float* zs, int n) {
        for (double& x : stage_ms) x = 0.0;
        auto tf = clk::now();
        if (updated_time > 0) {
            float dt = (float)(t - updated_time);
            float decay = (float)std::pow((double)p.gamma, (double)dt);
            auto t0 = clk::now();
            predict(dt);
            stage_ms[0] = ms_since(t0);
            t0 = clk::now();
            assign();
            stage_ms[1] = ms_since(t0);
            if (method == K3DOM_METHOD_DSPHDMIB) ds_mass_update(xs, ys, zs, n);
            else mass_update(decay, xs, ys, zs, n, w);
            t0 = clk::now();
            update_persistent();
            stage_ms[5] = ms_since(t0);
            t0 = clk::now();
            newborn();
            stage_ms[6] = ms_since(t0);
            t0 = clk::now();
            moments();
            stage_ms[7] = ms_since(t0);
            record_counts();
            t0 = clk::now();
            resample();
            stage_ms[8] = ms_since(t0);
        } else {
            initialize_particles(xs, ys, zs, n);
            cnt_n.assign(V, 0); cnt_m.assign(V, 0); cnt_nb.assign(V, 0); cnt_mb.assign(V, 0);
        }
        stage_ms[9] = ms_since(tf);
        updated_time = t;
        frame++;
    }
};

// ---------------------------------------------------------------------------
// FP64 evaluators of the floating-point stages, for error statements.  These
// take the FP32 inputs of a stage and compute its outputs in double with libm
// and exact division, following the reference formulas directly.
static double kern64(double dd, double ls, double sigma) {
    const double PI = 3.14159265358979323846;
    double q = dd / ls;
    return std::max(0.0, sigma * ((2.0 + std::cos(2 * PI * q)) * (1.0 - q) / 3.0 +
                                  std::sin(2 * PI * q) / (2.0 * (double)3.141592f)));
}

static void bgki64_voxel(double x, double y, double z, const float* xs, const float* ys, const float* zs,
                         int n, double res, double ls, double sigma, double* o, double* fr) {
    double occ = 0, fre = 0;
    double dc = std::sqrt(x * x + y * y + z * z);
    for (int i = 0; i < n; ++i) {
        double xm = xs[i], ym = ys[i], zm = zs[i];
        if (!(std::isfinite(xm) && std::isfinite(ym) && std::isfinite(zm))) continue;
        double dd = std::sqrt((x - xm) * (x - xm) + (y - ym) * (y - ym) + (z - zm) * (z - zm));
        if (dd < ls) occ += kern64(dd, ls, sigma);
        double dm = std::sqrt(xm * xm + ym * ym + zm * zm);
        if (!(dm > 0)) continue;
        double dp = (x * xm + y * ym + z * zm) / dm;
        double mg = std::max(0.9 * dm, dm - 2.0 * res);
        if (dp <= 0) dd = dc;
        else if (dp < mg) dd = std::sqrt(std::max(0.0, dc * dc - dp * dp));
        else if (dp < dm) {
            double r = mg / dm;
            dd = std::sqrt((x - r * xm) * (x - r * xm) + (y - r * ym) * (y - r * ym) + (z - r * zm) * (z - r * zm));
        } else continue;
        if (dd < ls) fre += kern64(dd, ls, sigma);
    }
    *o = occ;
    *fr = fre;
}

// ---------------------------------------------------------------------------
// Differentiated window in double precision: the continuous chain of a window
// (masses per voxel and the joint weight W) with the frozen integer decisions
// of a forward run (map shifts, counts), evaluated with X = double for the
// loss and central differences, and with X = K3Dual<double, 6> for the
// gradient by forward-mode differentiation.  theta = (ls, sigma_occ,
// sigma_free, prior_free, prior_static, prior_dynamic).
struct WindowInput {
    const k3dom_params* p;
    int T, gs, gsz;
    const int32_t* npts;      // T
    const float* pts;         // per frame x[n], y[n], z[n], frames back to back
    const float* sensor_rel;  // 3 T: sensor minus map centre (float, as the frame uses)
    const int32_t* move;      // 3 T: map shift before each frame
    const float* decay;       // T
    const int32_t* counts;    // T * 4 * V: n, m, nb, mb
    const uint8_t* gt;        // T * V
    const uint8_t* mask;      // V or null
    // several LiDARs per frame (null: one sensor): group count per frame, the
    // points per group (T * K3DOM_MAX_LIDARS, consecutive within the frame)
    // and each group's sensor position relative to the map centre (3 per group)
    const int32_t* ngroups = nullptr;
    const int32_t* gcount = nullptr;
    const float* gsrel = nullptr;
};

template <class X>
static X kern_x(double dd, const X& ls, const X& sigma) {
    const double PI = 3.14159265358979323846;
    const X q = X(k3d_cast<X>(dd)) / ls;
    const X k = (2.0 + k3g_cos(2.0 * PI * q)) * (1.0 - q) / 3.0 + k3g_sin(2.0 * PI * q) / (2.0 * (double)3.141592f);
    const X v = sigma * k;
    return k3g_val(v) > 0.0 ? v : v * 0.0;
}

template <class X>
static void bgki_x_points(const WindowInput& in, const Geom& g, const float* xs, const float* ys, const float* zs,
                          int n, const X& ls, const X& so, const X& sf, std::vector<X>& dO, std::vector<X>& dF);

// Evidence of frame t (per LiDAR group with its sensor position, added)
template <class X>
static void bgki_x(const WindowInput& in, int t, const X& ls, const X& so, const X& sf, std::vector<X>& dO,
                   std::vector<X>& dF) {
    size_t off = 0;
    for (int f = 0; f < t; ++f) off += 3 * (size_t)in.npts[f];
    const int n = in.npts[t];
    const float *xs = in.pts + off, *ys = xs + n, *zs = ys + n;
    const int K = in.ngroups ? in.ngroups[t] : 0;
    if (K == 0) {
        Geom g{in.gs, in.gsz, in.p->resolution, in.sensor_rel[3 * t], in.sensor_rel[3 * t + 1],
               in.sensor_rel[3 * t + 2]};
        g.build();
        bgki_x_points(in, g, xs, ys, zs, n, ls, so, sf, dO, dF);
        return;
    }
    int at = 0;
    for (int k = 0; k < K; ++k) {
        const float* sr = in.gsrel + 3 * ((size_t)t * K3DOM_MAX_LIDARS + k);
        Geom g{in.gs, in.gsz, in.p->resolution, sr[0], sr[1], sr[2]};
        g.build();
        const int c = in.gcount[(size_t)t * K3DOM_MAX_LIDARS + k];
        bgki_x_points(in, g, xs + at, ys + at, zs + at, c, ls, so, sf, dO, dF);
        at += c;
    }
}

template <class X>
static void bgki_x_points(const WindowInput& in, const Geom& g, const float* xs, const float* ys, const float* zs,
                          int n, const X& ls, const X& so, const X& sf, std::vector<X>& dO, std::vector<X>& dF) {
    const int gs = in.gs, gsz = in.gsz;
    const double lsv = k3g_val(ls);
    const double R = lsv * (1.0 + 1e-5) + 1e-4;
    const double res = in.p->resolution;
    for (int pi = 0; pi < n; ++pi) {
        // only the validity flag of the record is used here
        const K3PointRec pr = k3_point_prep(xs[pi], ys[pi], zs[pi], in.p->resolution, 1.0f);
        if (pr.valid == 0.0f) continue;
        const double xm = xs[pi], ym = ys[pi], zm = zs[pi];
        const double dm = std::sqrt(xm * xm + ym * ym + zm * zm);
        const double mg = std::max(0.9 * dm, dm - 2.0 * res);
        int i0, i1;
        if (!axis_range(std::min(0.0, xm) - R, std::max(0.0, xm) + R, g.res, g.sx, gs, &i0, &i1)) continue;
        for (int i = i0; i <= i1; ++i) {
            double ta = 0.0, tb = 1.0;
            if (!narrow(xm, g.cx[i], R, &ta, &tb)) continue;
            int j0, j1;
            if (!axis_range(std::min(ta * ym, tb * ym) - R, std::max(ta * ym, tb * ym) + R, g.res, g.sy, gs, &j0, &j1))
                continue;
            for (int j = j0; j <= j1; ++j) {
                double tc = ta, td = tb;
                if (!narrow(ym, g.cy[j], R, &tc, &td)) continue;
                int k0, k1;
                if (!axis_range(std::min(tc * zm, td * zm) - R, std::max(tc * zm, td * zm) + R, g.res, g.sz, gsz,
                                &k0, &k1))
                    continue;
                for (int k = k0; k <= k1; ++k) {
                    const double x = g.cx[i], y = g.cy[j], z = g.cz[k];
                    const size_t v = (size_t)i + (size_t)gs * j + (size_t)gs * gs * k;
                    const double dd = std::sqrt((x - xm) * (x - xm) + (y - ym) * (y - ym) + (z - zm) * (z - zm));
                    if (dd < lsv) dO[v] = dO[v] + kern_x(dd, ls, so);
                    if (!(dm > 0)) continue;
                    const double dc = std::sqrt(x * x + y * y + z * z);
                    const double dp = (x * xm + y * ym + z * zm) / dm;
                    double df;
                    if (dp <= 0) df = dc;
                    else if (dp < mg) df = std::sqrt(std::max(0.0, dc * dc - dp * dp));
                    else if (dp < dm) {
                        const double r = mg / dm;
                        df = std::sqrt((x - r * xm) * (x - r * xm) + (y - r * ym) * (y - r * ym) +
                                       (z - r * zm) * (z - r * zm));
                    } else continue;
                    if (df < lsv) dF[v] = dF[v] + kern_x(df, ls, sf);
                }
            }
        }
    }
}

template <class X>
static X window_chain(const WindowInput& in, const X th[6]) {
    const int gs = in.gs, gsz = in.gsz, V = gs * gs * gsz;
    const k3dom_params& p = *in.p;
    std::vector<X> f(V, th[3]), s(V, th[4]), d(V, th[5]);
    std::vector<X> dO(V), dF(V);
    X W = k3d_cast<X>(1.0);
    X loss = k3d_cast<X>(0.0);
    long long cnt_mask = 0;
    for (int v = 0; v < V; ++v) cnt_mask += (!in.mask || in.mask[v]) ? 1 : 0;
    for (int t = 0; t < in.T; ++t) {
        const int* mv = in.move + 3 * t;
        if (t > 0 && (mv[0] || mv[1] || mv[2])) {
            std::vector<X> nf(V), ns(V), nd(V);
            for (int c = 0; c < V; ++c) {
                const int x = c % gs, y = (c / gs) % gs, z = c / (gs * gs);
                const int sx = x + mv[0], sy = y + mv[1], sz = z + mv[2];
                if (sx >= 0 && sx < gs && sy >= 0 && sy < gs && sz >= 0 && sz < gsz) {
                    const int src = sx + gs * sy + gs * gs * sz;
                    nf[c] = f[src]; ns[c] = s[src]; nd[c] = d[src];
                } else {
                    nf[c] = th[3]; ns[c] = th[4]; nd[c] = th[5];
                }
            }
            f.swap(nf); s.swap(ns); d.swap(nd);
        }
        for (int v = 0; v < V; ++v) {
            dO[v] = k3d_cast<X>(0.0);
            dF[v] = k3d_cast<X>(0.0);
        }
        bgki_x(in, t, th[0], th[1], th[2], dO, dF);
        const int32_t* cn = in.counts + (size_t)t * 4 * V;
        X Wn = k3d_cast<X>(0.0);
        const double decay = t == 0 ? 1.0 : (double)in.decay[t];
        for (int v = 0; v < V; ++v) {
            const int n = cn[v], m = cn[V + v], nb = cn[2 * V + v], mb = cn[3 * V + v];
            const X dyn = t == 0 ? k3d_cast<X>(0.0) : W * ((double)n * (double)p.persistence_prob / (double)p.particle_count);
            X fo, so, dout, born;
            k3g_mass_step<X, double>(f[v], s[v], d[v], dyn, dO[v], dF[v], decay, (double)p.mass_scale,
                                     (double)p.birth_prob, &fo, &so, &dout, &born);
            f[v] = fo; s[v] = so; d[v] = dout;
            if (t >= 1) {
                if (n > 0) Wn = Wn + (dout - born) * ((double)m / (double)n);
                if (nb > 0) Wn = Wn + born * ((double)mb / (double)nb);
            }
            if (!in.mask || in.mask[v]) loss = loss + k3g_bce_occ<X, double>(fo, so, dout, in.gt[(size_t)t * V + v]);
        }
        if (t >= 1) W = Wn;
    }
    return loss / ((double)in.T * (double)cnt_mask);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI for the Python tests and the benchmark's CPU-baseline leg.
extern "C" {

void* orc_create(const k3dom_params* p) {
    if (!p || p->resolution <= 0 || p->size < p->resolution || p->size_z < p->resolution ||
        p->particle_count <= 0 || p->new_born_particle_count < 0)
        return nullptr;
    return new Oracle(*p);
}
void* orc_create_method(const k3dom_params* p, int method) {
    if (method != K3DOM_METHOD_K3DOM && method != K3DOM_METHOD_DSPHDMIB) return nullptr;
    Oracle* o = (Oracle*)orc_create(p);
    if (o) o->method = method;
    return o;
}
void orc_destroy(void* h) { delete (Oracle*)h; }
void orc_set_brute(void* h, int brute) { ((Oracle*)h)->brute = brute != 0; }
// Threads of the culled kernel inference of every oracle map (1 = serial;
// the results do not depend on it).
void orc_set_threads(int n) { g_bgki_threads = n < 1 ? 1 : n; }
void orc_update_pose(void* h, float x, float y, float z) { ((Oracle*)h)->update_pose(x, y, z); }
int orc_update_grid(void* h, double t, const float* xs, const float* ys, const float* zs, int n) {
    if (!h || n < 0 || (n > 0 && (!xs || !ys || !zs))) return K3DOM_ERR_INVALID_ARG;
    ((Oracle*)h)->update_grid(t, xs, ys, zs, n);
    return K3DOM_OK;
}

// Several LiDARs: counts[k] points per LiDAR, concatenated in xs/ys/zs in
// their own frames; rot: 9 floats per LiDAR (row-major, LiDAR to map axes);
// origin: 3 floats per LiDAR (relative to the pose, map axes).
int orc_update_grid_lidars(void* hh, double t, int n_lidars, const float* rot, const float* origin,
                           const int* counts, const float* xs, const float* ys, const float* zs) {
    Oracle* h = (Oracle*)hh;
    if (!h || n_lidars < 1 || !rot || !origin || !counts) return K3DOM_ERR_INVALID_ARG;
    if (h->method == K3DOM_METHOD_DSPHDMIB && n_lidars > 1) return K3DOM_ERR_INVALID_ARG;
    int total = 0;
    for (int k = 0; k < n_lidars; ++k) {
        if (counts[k] < 0) return K3DOM_ERR_INVALID_ARG;
        total += counts[k];
    }
    if (total > 0 && (!xs || !ys || !zs)) return K3DOM_ERR_INVALID_ARG;
    std::vector<float> rx(total), ry(total), rz(total);
    h->groups.clear();
    int at = 0;
    for (int k = 0; k < n_lidars; ++k) {
        Oracle::Group G{at, counts[k], {origin[3 * k], origin[3 * k + 1], origin[3 * k + 2]}};
        for (int q = at; q < at + counts[k]; ++q) k3_lidar_rotate(rot + 9 * k, xs[q], ys[q], zs[q], &rx[q], &ry[q], &rz[q]);
        h->groups.push_back(G);
        at += counts[k];
    }
    h->update_grid(t, rx.data(), ry.data(), rz.data(), total);
    h->groups.clear();
    return K3DOM_OK;
}

void orc_get_info(void* hh, k3dom_info* out) {
    Oracle* h = (Oracle*)hh;
    out->grid_size = h->gs;
    out->grid_size_z = h->gsz;
    out->grid_cell_count = h->V;
    out->particle_count = h->N;
    out->new_born_particle_count = h->B;
    out->iteration = (int32_t)h->frame;
    out->resolution = h->p.resolution;
    for (int i = 0; i < 3; ++i) { out->center_pos[i] = h->center[i]; out->sensor_pos[i] = h->sensor[i]; }
    out->updated_time = h->updated_time;
}

void orc_read_cells(void* hh, k3dom_cell* out) {
    Oracle* h = (Oracle*)hh;
    for (int c = 0; c < h->V; ++c) {
        k3dom_cell& o = out[c];
        o.start_idx = h->start[c]; o.end_idx = h->end[c];
        o.mu_A = h->muA[c]; o.mu_UA = h->muUA[c]; o.w_A = h->wA[c]; o.w_UA = h->wUA[c];
        o.mean_x_vel = h->mvx[c]; o.mean_y_vel = h->mvy[c]; o.mean_z_vel = h->mvz[c];
        o.likelihood = h->lik[c]; o.p_A = h->pA[c];
        o.free_mass = h->f[c]; o.static_mass = h->s[c]; o.dynamic_mass = h->d[c];
        o.occ_mass = h->occ[c]; o.pers_mass = h->pers[c]; o.pred_mass = h->pred[c];
    }
}

// Same numbering as k3dom_array in include/k3dom.h.
int64_t orc_read_array(void* hh, int which, void* dst, int64_t capacity) {
    Oracle* h = (Oracle*)hh;
    auto put3 = [&](const std::vector<float>& a, const std::vector<float>& b, const std::vector<float>& c) -> int64_t {
        int64_t bytes = (int64_t)(a.size() * 3 * 4);
        if (capacity < bytes) return -1;
        if (bytes == 0) return 0;  // (no memcpy from an empty vector's null data)
        float* o = (float*)dst;
        memcpy(o, a.data(), a.size() * 4);
        memcpy(o + a.size(), b.data(), b.size() * 4);
        memcpy(o + 2 * a.size(), c.data(), c.size() * 4);
        return bytes;
    };
    auto put = [&](const void* src, size_t bytes) -> int64_t {
        if (capacity < (int64_t)bytes) return -1;
        if (bytes == 0) return 0;
        memcpy(dst, src, bytes);
        return (int64_t)bytes;
    };
    switch (which) {
    case K3DOM_ARR_PARTICLE_POS: return put3(h->px, h->py, h->pz);
    case K3DOM_ARR_PARTICLE_VEL: return put3(h->vx, h->vy, h->vz);
    case K3DOM_ARR_PARTICLE_WEIGHT: return put(h->w.data(), h->w.size() * 4);
    case K3DOM_ARR_PARTICLE_KEY: return put(h->key.data(), h->key.size() * 4);
    case K3DOM_ARR_SORT_PERM: return put(h->perm.data(), h->perm.size() * 4);
    case K3DOM_ARR_SORTED_KEY: return put(h->sorted_key.data(), h->sorted_key.size() * 4);
    case K3DOM_ARR_RESAMPLE_IDX: return put(h->resample_idx.data(), h->resample_idx.size() * 4);
    case K3DOM_ARR_NEWBORN_POS: return put3(h->bpx, h->bpy, h->bpz);
    case K3DOM_ARR_NEWBORN_VEL: return put3(h->bvx, h->bvy, h->bvz);
    case K3DOM_ARR_NEWBORN_WEIGHT: return put(h->bw.data(), h->bw.size() * 4);
    case K3DOM_ARR_NEWBORN_KEY: return put(h->bkey.data(), h->bkey.size() * 4);
    case K3DOM_ARR_BORN_MASS: return put(h->born.data(), h->born.size() * 4);
    case K3DOM_ARR_DEL_OCC: return put(h->del_occ.data(), h->del_occ.size() * 4);
    case K3DOM_ARR_DEL_FREE: return put(h->del_free.data(), h->del_free.size() * 4);
    case K3DOM_ARR_SORTED_WEIGHT: return put(h->sorted_w.data(), h->sorted_w.size() * 4);
    default: return -1;
    }
}

void orc_stage_times(void* hh, double* out10) {
    Oracle* h = (Oracle*)hh;
    for (int i = 0; i < 10; ++i) out10[i] = h->stage_ms[i];
}
void orc_bgki_pairs(void* hh, uint64_t* occ_free) {
    Oracle* h = (Oracle*)hh;
    occ_free[0] = h->bgki_stats.occ_pairs;
    occ_free[1] = h->bgki_stats.free_pairs;
}

// ---- primitives (pinned by known answers and libm comparisons in the tests)
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    k3_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}
float orc_ln(float x) { return k3_ln(x); }
float orc_exp(float x) { return k3_exp(x); }
void orc_sincospi(float t, float* s, float* c) { k3_sincospi(t, s, c); }
float orc_seg_sum(const float* x, int32_t st, int32_t en) { return seg_sum(x, st, en); }

// Frozen counts of the last frame (4 * V int32: n, m, nb, mb).
int64_t orc_read_counts(void* hh, int32_t* dst, int64_t capacity) {
    Oracle* h = (Oracle*)hh;
    const int64_t bytes = 16 * (int64_t)h->V;
    if (capacity < bytes || (int64_t)h->cnt_n.size() != h->V) return -1;
    std::memcpy(dst, h->cnt_n.data(), 4 * (size_t)h->V);
    std::memcpy(dst + h->V, h->cnt_m.data(), 4 * (size_t)h->V);
    std::memcpy(dst + 2 * h->V, h->cnt_nb.data(), 4 * (size_t)h->V);
    std::memcpy(dst + 3 * h->V, h->cnt_mb.data(), 4 * (size_t)h->V);
    return bytes;
}

// Loss of a window in double (grad == NULL), or loss and its gradient by
// forward-mode differentiation (grad[6]).
int orc_window_chain64_groups(const k3dom_params* p, int32_t T, const int32_t* npts, const float* pts,
                              const float* sensor_rel, const int32_t* move, const float* decay, const int32_t* counts,
                              const uint8_t* gt, const uint8_t* mask, const int32_t* ngroups, const int32_t* gcount,
                              const float* gsrel, const double* theta, double* loss, double* grad) {
    if (!p || T <= 0 || !npts || !pts || !sensor_rel || !move || !decay || !counts || !gt || !theta || !loss)
        return K3DOM_ERR_INVALID_ARG;
    if (ngroups && (!gcount || !gsrel)) return K3DOM_ERR_INVALID_ARG;
    WindowInput in{p, T, (int)(p->size / p->resolution), (int)(p->size_z / p->resolution), npts, pts, sensor_rel,
                   move, decay, counts, gt, mask, ngroups, gcount, gsrel};
    if (!grad) {
        double th[6];
        for (int k = 0; k < 6; ++k) th[k] = theta[k];
        *loss = window_chain<double>(in, th);
        return K3DOM_OK;
    }
    typedef K3Dual<double, 6> D;
    D th[6];
    for (int k = 0; k < 6; ++k) th[k] = k3d_var<double, 6>(theta[k], k);
    const D L = window_chain<D>(in, th);
    *loss = L.v;
    for (int k = 0; k < 6; ++k) grad[k] = L.d[k];
    return K3DOM_OK;
}
int orc_window_chain64(const k3dom_params* p, int32_t T, const int32_t* npts, const float* pts,
                       const float* sensor_rel, const int32_t* move, const float* decay, const int32_t* counts,
                       const uint8_t* gt, const uint8_t* mask, const double* theta, double* loss, double* grad) {
    return orc_window_chain64_groups(p, T, npts, pts, sensor_rel, move, decay, counts, gt, mask, nullptr, nullptr,
                                     nullptr, theta, loss, grad);
}
// length-scale derivative of sigma * kbar(d / ls): closed form, and from the table
float orc_kernel_dls(float d, float ls, float sigma) { return k3_kernel_dls(d, ls, 1.0f / ls, sigma); }
float orc_kernel_dls_sq(float a2, float ls, float sigma) {
    return k3_kernel_dls_sq(a2, k3_kernel_scale(ls), 1.0f / ls, sigma);
}
// credit in forward-mode dual numbers: value and derivative by the mass
void orc_dual_credit(double mass, double scale, double* value, double* deriv) {
    const K3Dual<double, 6> c = k3g_credit(k3d_var<double, 6>(mass, 0), scale);
    *value = c.v;
    *deriv = c.d[0];
}
// kernel at distance d (evaluated from the float32 square d * d)
float orc_kernel(float d, float ls, float sigma) { return k3_kernel_value_sq(d * d, k3_kernel_scale(ls), sigma); }
float orc_credit(float m, float scale) { return k3_credit(m, scale); }
float orc_ds_pocc(float x, float y, float z, float mx, float my, float mz, float delta) {
    return k3ds_pocc(x, y, z, mx, my, mz, delta);
}
float orc_ds_pfree(float x, float y, float z, float max_range) { return k3ds_pfree(x, y, z, max_range); }
void orc_ds_update_masses(float occ_pred, float free_pred, float meas_occ, float meas_free, float* out2) {
    k3ds_update_masses(occ_pred, free_pred, meas_occ, meas_free, &out2[0], &out2[1]);
}
float orc_ds_newborn_part(float occ_pred, float occ_up, float p_B) { return k3ds_newborn_part(occ_pred, occ_up, p_B); }
void orc_box_muller(uint32_t a, uint32_t b, float* n0, float* n1) { k3_box_muller(a, b, n0, n1); }
void orc_update_mass(float* fsd, float del_occ, float del_free, float mass_scale) {
    k3_update_mass(&fsd[0], &fsd[1], &fsd[2], del_occ, del_free, mass_scale);
}
void orc_predict_mass(float* fsd, float dyn_pred, float pred_total, float decay) {
    k3_predict_mass(&fsd[0], &fsd[1], &fsd[2], dyn_pred, pred_total, decay);
}
float orc_newborn_part(float dyn_pred, float total_pred, float dyn_up, float p_B) {
    return k3_newborn_part(dyn_pred, total_pred, dyn_up, p_B);
}
int orc_key(float x, float y, float z, int gs, int gsz) { return k3_key(x, y, z, gs, gsz); }

// FP64 shadow of the floating-point stages of later frames (on != 0).
void orc_set_keep64(void* hh, int on) { ((Oracle*)hh)->keep64 = on != 0; }
// Array `which` of the FP64 shadow of the last normal frame (V doubles; order
// of Oracle::E_*: pre-update f, s, d; dynamic mass; f, s, d, born; mean vx,
// vy, vz); -1 when not kept.
int64_t orc_read64(void* hh, int which, double* dst, int64_t capacity) {
    Oracle* h = (Oracle*)hh;
    if (which < 0 || which >= Oracle::E_COUNT || (int64_t)h->e64[which].size() != h->V) return -1;
    const int64_t bytes = 8 * (int64_t)h->V;
    if (capacity < bytes) return -1;
    std::memcpy(dst, h->e64[which].data(), (size_t)bytes);
    return bytes;
}
// The motion step of the frame (predict_particles) on caller arrays.
void orc_predict(int n, float dt, float p_S, float sigma_pos, float sigma_vel, float vclamp, int gs, int gsz,
                 uint32_t frame, uint64_t seed, float* px, float* py, float* pz, float* vx, float* vy, float* vz,
                 float* w, int32_t* key) {
    K3PredictConst c;
    c.dt = dt; c.p_S = p_S; c.sigma_pos = sigma_pos; c.sigma_vel = sigma_vel; c.vclamp = vclamp;
    c.gs = gs; c.gsz = gsz; c.frame = frame; c.key0 = (uint32_t)seed; c.key1 = (uint32_t)(seed >> 32);
    predict_particles(c, n, px, py, pz, vx, vy, vz, w, key);
}
// The velocity draw of new particles (k3_sample_velocity) from three Philox words.
void orc_sample_velocity(uint32_t r0, uint32_t r1, uint32_t r2, float vmax, float min_vel, float* out3) {
    k3_sample_velocity(r0, r1, r2, vmax, min_vel, &out3[0], &out3[1], &out3[2]);
}
// The moments of the frame (voxel_moments) on caller arrays.
void orc_moments(int V, const int32_t* start, const int32_t* end, const float* w, const float* vx, const float* vy,
                 const float* vz, const float* pers, float* mvx, float* mvy, float* mvz) {
    voxel_moments(V, start, end, w, vx, vy, vz, pers, mvx, mvy, mvz);
}
// Overwrites every cell of a map (GridCell records, index order), for
// hand-built states.
void orc_set_cells(void* hh, const k3dom_cell* in) {
    Oracle* h = (Oracle*)hh;
    for (int c = 0; c < h->V; ++c) {
        const k3dom_cell& o = in[c];
        h->start[c] = o.start_idx; h->end[c] = o.end_idx;
        h->muA[c] = o.mu_A; h->muUA[c] = o.mu_UA; h->wA[c] = o.w_A; h->wUA[c] = o.w_UA;
        h->mvx[c] = o.mean_x_vel; h->mvy[c] = o.mean_y_vel; h->mvz[c] = o.mean_z_vel;
        h->lik[c] = o.likelihood; h->pA[c] = o.p_A;
        h->f[c] = o.free_mass; h->s[c] = o.static_mass; h->d[c] = o.dynamic_mass;
        h->occ[c] = o.occ_mass; h->pers[c] = o.pers_mass; h->pred[c] = o.pred_mass;
    }
}
// Overwrites the particle positions (3 * N floats: x[N], y[N], z[N]).
void orc_set_particle_pos(void* hh, const float* pos) {
    Oracle* h = (Oracle*)hh;
    for (int i = 0; i < h->N; ++i) {
        h->px[i] = pos[i];
        h->py[i] = pos[h->N + i];
        h->pz[i] = pos[2 * h->N + i];
    }
}

// BGKI of one frame on a grid, brute force or culled, returning the
// contributing pair counts.
void orc_bgki(int gs, int gsz, float res, float sx, float sy, float sz, float ls, float sigma,
              const float* xs, const float* ys, const float* zs, int n, int brute,
              float* del_occ, float* del_free, uint64_t* pairs) {
    Geom g{gs, gsz, res, sx, sy, sz};
    g.build();
    BgkiStats st = bgki_all(g, xs, ys, zs, n, ls, sigma, brute != 0, del_occ, del_free);
    if (pairs) { pairs[0] = st.occ_pairs; pairs[1] = st.free_pairs; }
}

// FP64 BGKI for a list of voxels (cell indices), from the reference formula.
void orc_bgki64(int gs, int gsz, float res, float sx, float sy, float sz, float ls, float sigma,
                const float* xs, const float* ys, const float* zs, int n, const int32_t* cells, int ncell,
                double* del_occ, double* del_free) {
    for (int t = 0; t < ncell; ++t) {
        float x, y, z;
        k3_voxel_center(cells[t], gs, gsz, res, sx, sy, sz, &x, &y, &z);
        bgki64_voxel(x, y, z, xs, ys, zs, n, res, ls, sigma, &del_occ[t], &del_free[t]);
    }
}

// FP64 mass update of one voxel from FP32 inputs: in fsd = (f, s, d) before
// prediction; out fsd after update, plus born.
void orc_mass_update64(const float* fsd_in, double dyn_pred, double decay, double del_occ, double del_free,
                       double mass_scale, double p_B, double* out4) {
    mass_update64(fsd_in[0], fsd_in[1], fsd_in[2], dyn_pred, decay, del_occ, del_free, mass_scale, p_B, out4);
}

}  // extern "C"
