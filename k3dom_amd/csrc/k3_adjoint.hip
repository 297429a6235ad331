// Differentiated window on gfx950: the record kernel of the forward frames and
// the per-voxel adjoint of the mass update (no counterpart in the reference;
// the forward arithmetic is src/k3dom/kernel/mass_update.cu:79-183).
//
// Continuous state of a window: per voxel (free, static, dynamic) masses, and
// one scalar, the joint particle weight W_t of frame t.  After resampling every
// particle weighs W_t / N, so the predicted dynamic mass of voxel v in frame
// t+1 is n_v * p_S * W_t / N (n_v: particles predicted into v and inside the
// grid), and
//   W_t = sum_v m_v / n_v * pers_v + sum_v mb_v / nb_v * born_v
// (m_v of the n_v escape the static penalty; nb_v newborns were placed in v,
// mb_v of them escape it).  Keys, sort, counts, penalties and resampled
// indices are frozen integer decisions; the initialisation frame's weights
// (1/N) do not depend on the parameters.
#include "k3_kernels.h"
#include "k3_dual.h"

namespace k3 {

namespace {
constexpr int TPB = 256;
inline unsigned blocks_for(long long n) {
    long long b = (n + TPB - 1) / TPB;
    return (unsigned)(b < 1 ? 1 : b);
}

// block-wide sum of K doubles into out[blockIdx.x * K + k]
template <int K>
__device__ void block_sum_out(double (&v)[K], double* out) {
    __shared__ double red[TPB / 64][K];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < K; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        if (lane == 0) red[w][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int ww = 0; ww < TPB / 64; ++ww) s += red[ww][threadIdx.x];
        out[(size_t)blockIdx.x * K + threadIdx.x] = s;
    }
}
}  // namespace

// ------------------------------------------------------------------ record
// Voxel pass: masses, dynamic mass and loss of every voxel; the particle
// counts are zeroed here and added by the particle-parallel counts below.
__global__ __launch_bounds__(TPB) void k_window_record(int V, VoxelArrays v, WindowFrame rec,
                                                       const uint8_t* __restrict__ gt,
                                                       const uint8_t* __restrict__ mask, double* loss_partial) {
    const int c = blockIdx.x * TPB + threadIdx.x;
    double loss[1] = {0.0};
    if (c < V) {
        const float f = v.f[c], s = v.s[c], d = v.d[c];
        rec.f[c] = f;
        rec.s[c] = s;
        rec.d[c] = d;
        rec.dyn[c] = v.pred[c];
        rec.n[c] = 0;
        rec.m[c] = 0;
        rec.nb[c] = 0;
        rec.mb[c] = 0;
        if (!mask || mask[c]) loss[0] = (double)k3g_bce_occ<float, float>(f, s, d, gt[c]);
    }
    block_sum_out<1>(loss, loss_partial);
}

// Particle-parallel counts per voxel over a key-sorted particle set (the
// sorted persistent set, or the newborn set in voxel order): n[key] counts
// the particles inside the grid (persistent set only: the inside test of
// k3_predict_particle), m[key] those of them above the minimum speed (the
// penalty test of the persistent update).  Each wavefront adds, per run of
// equal keys among its 64 particles, the run's counts with one atomic add.
__global__ __launch_bounds__(TPB) void k_window_count(int n, int gs, int gsz, int test_inside, float min2,
                                                      ParticleArrays P, int32_t* cnt_n, int32_t* cnt_m) {
    const int j = blockIdx.x * TPB + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = j < n;
    int key = -1;
    bool in = false, fast = false;
    if (valid) {
        key = P.key[j];
        const float x = P.px[j], y = P.py[j], z = P.pz[j];
        in = !test_inside || !((x >= (float)gs || x < 0.0f) || (y >= (float)gs || y < 0.0f) ||
                               (z >= (float)gsz || z < 0.0f));
        const float vx = P.vx[j], vy = P.vy[j], vz = P.vz[j];
        fast = !(vx * vx + vy * vy + vz * vz < min2);
    }
    // (the permutes run in every lane)
    const int prev = __builtin_amdgcn_ds_bpermute(((lane + 63) & 63) << 2, key);
    const int next = __builtin_amdgcn_ds_bpermute(((lane + 1) & 63) << 2, key);
    const uint64_t heads = __ballot(valid && (lane == 0 || prev != key));
    const uint64_t inm = __ballot(valid && in), okm = __ballot(valid && in && fast);
    const bool last = valid && (lane == 63 || next != key);
    if (last) {
        const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const int start = 63 - __clzll(heads & upto);
        const uint64_t run = upto & ~((1ull << start) - 1ull);
        const int rn = __popcll(inm & run), rm = __popcll(okm & run);
        if (rn) atomicAdd(&cnt_n[key], rn);
        if (rm) atomicAdd(&cnt_m[key], rm);
    }
}

hipError_t launch_window_record(hipStream_t st, int V, int gs, int gsz, bool normal, float min_vel,
                                const VoxelArrays& v, const ParticleArrays& Ps, int N, const ParticleArrays& NB,
                                int B, const WindowFrame& rec, const uint8_t* gt, const uint8_t* mask,
                                double* loss_partial) {
    hipLaunchKernelGGL(k_window_record, dim3(blocks_for(V)), dim3(TPB), 0, st, V, v, rec, gt, mask, loss_partial);
    if (normal) {
        const float min2 = min_vel * min_vel;
        if (N > 0)
            hipLaunchKernelGGL(k_window_count, dim3(blocks_for(N)), dim3(TPB), 0, st, N, gs, gsz, 1, min2, Ps,
                               rec.n, rec.m);
        if (B > 0)
            hipLaunchKernelGGL(k_window_count, dim3(blocks_for(B)), dim3(TPB), 0, st, B, gs, gsz, 0, min2, NB,
                               rec.nb, rec.mb);
    }
    return hipGetLastError();
}
int window_blocks(int V) { return (int)blocks_for(V); }

// ------------------------------------------------------------------ adjoint of one frame
// Partials per block: [0] lambda_W of the previous frame, [1] d/d ls,
// [2] d/d sigma_occ, [3] d/d sigma_free, [4..6] d/d prior (free, static, dynamic).
constexpr int ADJ_K = 7;

__global__ __launch_bounds__(TPB) void k_window_adjoint(AdjointFrameArgs a, const double* __restrict__ acc,
                                                        double* partial) {
    const int c = blockIdx.x * TPB + threadIdx.x;
    double out[ADJ_K] = {0, 0, 0, 0, 0, 0, 0};
    if (c < a.V) {
        const int gs = a.gs, gsz = a.gsz;
        const int i = c % gs, j = (c / gs) % gs, k = c / (gs * gs);
        // adjoint of the masses after this frame: loss term + next frame's pre-state adjoint
        // (the adjoint arithmetic runs in double; the forward values are the frame's floats)
        // (the cell's words, its next-frame adjoint and its previous masses are
        // loaded unconditionally, from a valid cell, before any is used, so
        // that they are all in flight together)
        const bool use = !a.mask || a.mask[c];
        const float post_f = a.post.f[c], post_s = a.post.s[c], post_d = a.post.d[c];
        const auto gt = a.gt[c];
        // new cell c' = c - move of the next frame
        const int ni = i - a.move_next[0], nj = j - a.move_next[1], nk = k - a.move_next[2];
        const bool next_in = a.lam_next_f && ni >= 0 && ni < gs && nj >= 0 && nj < gs && nk >= 0 && nk < gsz;
        const int cn = next_in ? ni + gs * nj + gs * gs * nk : c;
        double nf = 0.0, ns = 0.0, nd = 0.0;
        if (a.lam_next_f) {
            nf = a.lam_next_f[cn];
            ns = a.lam_next_s[cn];
            nd = a.lam_next_d[cn];
        }
        // masses before this frame: the previous frame's, moved by this frame's shift, or priors
        const int si = i + a.move[0], sj = j + a.move[1], sk = k + a.move[2];
        const bool prev_in = a.prev.f && si >= 0 && si < gs && sj >= 0 && sj < gs && sk >= 0 && sk < gsz;
        const int cs = prev_in ? si + gs * sj + gs * gs * sk : c;
        double pf = a.prior_f, ps = a.prior_s, pd = a.prior_d;
        if (a.prev.f) {
            const float xf = a.prev.f[cs], xs = a.prev.s[cs], xd = a.prev.d[cs];
            if (prev_in) {
                pf = xf;
                ps = xs;
                pd = xd;
            }
        }
        const bool prior_cell = !prev_in;
        double lf = 0.0, ls_ = 0.0, ld = 0.0;
        if (use) {
            typedef K3Dual<double, 3> D3;
            const D3 bce = k3g_bce_occ<D3, double>(k3d_var<double, 3>(post_f, 0), k3d_var<double, 3>(post_s, 1),
                                                   k3d_var<double, 3>(post_d, 2), gt);
            lf = a.loss_scale * bce.d[0];
            ls_ = a.loss_scale * bce.d[1];
            ld = a.loss_scale * bce.d[2];
        }
        if (next_in) {
            lf += nf;
            ls_ += ns;
            ld += nd;
        }
        typedef K3Dual<double, 6> D6;
        const double del_o = a.del_occ[c], del_f = a.del_free[c];
        D6 fo, so, dout, born;
        k3g_mass_step<D6, double>(k3d_var<double, 6>(pf, 0), k3d_var<double, 6>(ps, 1), k3d_var<double, 6>(pd, 2),
                                  k3d_var<double, 6>((double)a.post.dyn[c], 3), k3d_var<double, 6>(del_o, 4),
                                  k3d_var<double, 6>(del_f, 5), (double)a.decay, (double)a.mass_scale,
                                  (double)a.p_B, &fo, &so, &dout, &born);
        // W of this frame through pers = d - born and born
        const double lamW = acc[0];
        const int n = a.post.n[c], nb = a.post.nb[c];
        const double ca = n > 0 ? (double)a.post.m[c] / (double)n : 0.0;
        const double cb = nb > 0 ? (double)a.post.mb[c] / (double)nb : 0.0;
        const double l_pers = a.has_w ? lamW * ca : 0.0;
        const double l_born = a.has_w ? lamW * cb - l_pers : 0.0;
        const double l_d = ld + l_pers;
        double lin[6];
        for (int q = 0; q < 6; ++q) lin[q] = lf * fo.d[q] + ls_ * so.d[q] + l_d * dout.d[q] + l_born * born.d[q];
        a.lam_f[c] = lin[0];
        a.lam_s[c] = lin[1];
        a.lam_d[c] = lin[2];
        if (prior_cell) {
            out[4] = lin[0];
            out[5] = lin[1];
            out[6] = lin[2];
        }
        if (a.has_w_prev) out[0] = lin[3] * (double)n * (double)a.w_coef;
        out[1] = lin[4] * (double)a.dls_occ[c] + lin[5] * (double)a.dls_free[c];
        out[2] = a.sigma_occ != 0.0f ? lin[4] * del_o / (double)a.sigma_occ : 0.0;
        out[3] = a.sigma_free != 0.0f ? lin[5] * del_f / (double)a.sigma_free : 0.0;
    }
    block_sum_out<ADJ_K>(out, partial);
}

// Sums the block partials in block order: acc[0] becomes lambda_W of the
// previous frame, acc[1..6] accumulate the parameter gradients.
// Sum of the per-block partials in a fixed order (deterministic): thread t
// adds blocks t, t + RED_T, ... in turn, then a tree over the threads.
constexpr int RED_T = 256;
__global__ __launch_bounds__(RED_T) void k_window_reduce(int nblocks, const double* __restrict__ partial,
                                                         double* acc) {
    __shared__ double sh[ADJ_K][RED_T];
    const int t = threadIdx.x;
    double s[ADJ_K];
#pragma unroll
    for (int k = 0; k < ADJ_K; ++k) s[k] = 0.0;
    for (int b = t; b < nblocks; b += RED_T)
#pragma unroll
        for (int k = 0; k < ADJ_K; ++k) s[k] += partial[(size_t)b * ADJ_K + k];
#pragma unroll
    for (int k = 0; k < ADJ_K; ++k) sh[k][t] = s[k];
    __syncthreads();
    for (int w = RED_T / 2; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int k = 0; k < ADJ_K; ++k) sh[k][t] += sh[k][t + w];
        __syncthreads();
    }
    if (t < ADJ_K) {
        if (t == 0)
            acc[0] = sh[0][0];
        else
            acc[t] += sh[t][0];
    }
}

hipError_t launch_window_adjoint(hipStream_t st, const AdjointFrameArgs& a, double* acc, double* partial) {
    const unsigned nb = blocks_for(a.V);
    hipLaunchKernelGGL(k_window_adjoint, dim3(nb), dim3(TPB), 0, st, a, acc, partial);
    hipLaunchKernelGGL(k_window_reduce, dim3(1), dim3(RED_T), 0, st, (int)nb, partial, acc);
    return hipGetLastError();
}

// Clears the accumulators (a kernel, so the sequence can be captured).
__global__ void k_window_clear(double* acc, int n) {
    if ((int)threadIdx.x < n) acc[threadIdx.x] = 0.0;
}
hipError_t launch_window_clear(hipStream_t st, double* acc, int n) {
    hipLaunchKernelGGL(k_window_clear, dim3(1), dim3(64), 0, st, acc, n);
    return hipGetLastError();
}

// sticky |= *flag (one thread): keeps a per-frame flag that the next frame's
// reset clears
__global__ void k_flag_or(uint32_t* sticky, const uint32_t* flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && *flag) *sticky = 1u;
}
__global__ void k_flag_clear(uint32_t* sticky) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *sticky = 0u;
}
hipError_t launch_flag_or(hipStream_t st, uint32_t* sticky, const uint32_t* flag) {
    hipLaunchKernelGGL(k_flag_or, dim3(1), dim3(64), 0, st, sticky, flag);
    return hipGetLastError();
}
hipError_t launch_flag_clear(hipStream_t st, uint32_t* sticky) {
    hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(64), 0, st, sticky);
    return hipGetLastError();
}

}  // namespace k3
