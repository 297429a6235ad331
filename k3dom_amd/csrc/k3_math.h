// Deterministic FP32 arithmetic shared by the gfx950 kernels and the host-side
// serial restatement of the update step.
//
// Every function here uses only IEEE-754 operations that are correctly rounded
// on both x86-64 (gcc, -ffp-contract=off) and gfx950 (hipcc, -ffp-contract=off,
// correctly rounded f32 divide/sqrt): +, -, *, /, sqrtf, fmaf, rintf, min/max
// and integer bit manipulation.  No libm/ocml transcendental is called, so the
// same inputs give the same bits on the CPU and on the GPU.
//
// Reference formulas restated here (youngjae-min/k3dom):
//   kernel_func / distance_with_margin / BGKI  src/k3dom/kernel/mass_update.cu:24-77
//   calc_credit / positive_diff / calc_ratio    src/k3dom/kernel/mass_update.cu:12-22
//   predict_mass / update_mass                  src/k3dom/kernel/mass_update.cu:79-139
//   separate_newborn_part                       src/k3dom/kernel/mass_update.cu:141-150
//   particle key                                src/common/kernel/predict.cu:51-55
// The reference draws its noise from cuRAND XORWOW; this code uses the
// counter-based Philox4x32-10 generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the Box-Muller
// transform (Box & Muller 1958) with the polynomial ln/sin/cos below.
#pragma once

#include <stdint.h>
#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define K3_HD __host__ __device__ __forceinline__
#else
#define K3_HD static inline
#endif

// ---------------------------------------------------------------- constants
// sin(pi r) = r (S1 + r^2 (S3 + r^2 (S5 + r^2 (S7 + r^2 S9))))  and
// cos(pi r) = 1 + r^2 (C2 + ... + r^2 C10): Taylor coefficients pi^k/k!
// rounded to float.  For |r| <= 1/4 the truncation error is below 2e-9.
#define K3_S1 3.1415927410125732f
#define K3_S3 (-5.167712688446045f)
#define K3_S5 2.550163984298706f
#define K3_S7 (-0.5992645025253296f)
#define K3_S9 0.08214588463306427f
#define K3_C2 (-4.934802055358887f)
#define K3_C4 4.058712005615234f
#define K3_C6 (-1.3352627754211426f)
#define K3_C8 0.23533062636852264f
#define K3_C10 (-0.025806890800595284f)
// Cody-Waite split of ln 2: LN2_HI has 16 significant bits so k*LN2_HI is exact.
#define K3_LN2_HI 0.693145751953125f
#define K3_LN2_LO 1.428606765330187e-06f
#define K3_LOG2E 1.4426950216293335f
#define K3_SQRT2 1.4142135381698608f

// Philox4x32-10 round multipliers and Weyl key increments (Salmon et al. 2011).
#define K3_PHILOX_M0 0xD2511F53u
#define K3_PHILOX_M1 0xCD9E8D57u
#define K3_PHILOX_W0 0x9E3779B9u
#define K3_PHILOX_W1 0xBB67AE85u

// BGKI sums are two-level: terms are added in ascending point index inside
// aligned chunks of 2^K3_BGKI_CHUNK_LOG2 points, and the chunk partial sums are
// added in ascending chunk order.  Both the GPU and the serial restatement use
// this association, so the sum does not depend on how the work is split.
#define K3_BGKI_CHUNK_LOG2 10

// A voxel's particle sums (predicted dynamic mass, velocity moments) are also
// two-level: the voxel's particles, in sorted order, form blocks of
// K3_SEG_BLOCK starting at its first particle; each block is summed left to
// right and the block sums are added left to right.  A voxel with at most
// K3_SEG_BLOCK particles is one plain left-to-right sum.
#define K3_SEG_BLOCK 64

// Philox stream ids (third counter word) of the random draws of one frame.
#define K3_STREAM_PREDICT 1u
#define K3_STREAM_NEWBORN 2u
#define K3_STREAM_INIT 3u
#define K3_STREAM_RESAMPLE 4u

K3_HD uint32_t k3_f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
K3_HD float k3_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
K3_HD double k3_pow2d(int n) { return __builtin_bit_cast(double, (uint64_t)(n + 1023) << 52); }  // n in [-1022,1023]

// ---------------------------------------------------------------- Philox4x32-10
K3_HD void k3_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                            uint32_t k0, uint32_t k1, uint32_t out[4]) {
#ifdef __HIPCC__
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)K3_PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)K3_PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += K3_PHILOX_W0;
        k1 += K3_PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 24-bit uniform in [0, 1)
K3_HD float k3_uniform01(uint32_t r) { return (float)(r >> 8) * 0x1p-24f; }
// 24-bit uniform in (0, 1]
K3_HD float k3_uniform01_open0(uint32_t r) { return (float)((r >> 8) + 1u) * 0x1p-24f; }

// ---------------------------------------------------------------- sin/cos(pi t)
// Exact reduction t = n/2 + r with |r| <= 1/4, then the Taylor polynomials.
K3_HD void k3_sincospi(float t, float* sn, float* cs) {
    float n = rintf(t * 2.0f);
    float r = t - n * 0.5f;
    float r2 = r * r;
    float ps = fmaf(r2, K3_S9, K3_S7);
    ps = fmaf(r2, ps, K3_S5);
    ps = fmaf(r2, ps, K3_S3);
    ps = fmaf(r2, ps, K3_S1);
    ps = ps * r;
    float pc = fmaf(r2, K3_C10, K3_C8);
    pc = fmaf(r2, pc, K3_C6);
    pc = fmaf(r2, pc, K3_C4);
    pc = fmaf(r2, pc, K3_C2);
    pc = fmaf(r2, pc, 1.0f);
    // quadrant q = n mod 4: (sin, cos) = (ps, pc), (pc, -ps), (-ps, -pc), (-pc, ps)
    int q = ((int)n) & 3;
    bool swap = (q & 1) != 0;
    float a = swap ? pc : ps;
    float b = swap ? ps : pc;
    uint32_t sa = (uint32_t)(q & 2) << 30;
    uint32_t sb = (uint32_t)((q + 1) & 2) << 30;
    *sn = k3_u2f(k3_f2u(a) ^ sa);
    *cs = k3_u2f(k3_f2u(b) ^ sb);
}

// ---------------------------------------------------------------- ln(x), x > 0
// x = 2^e m with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1),
// by the atanh series up to s^11 (|s| <= 0.172, truncation < 1e-10).
K3_HD float k3_ln(float x) {
    int e = 0;
    if (x < 0x1p-126f) { x = x * 0x1p24f; e = -24; }
    uint32_t b = k3_f2u(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = k3_u2f((b & 0x7fffffu) | 0x3f800000u);
    if (m > K3_SQRT2) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float s2 = s * s;
    float p = fmaf(s2, 0.09090909361839294f /*1/11*/, 0.1111111119389534f /*1/9*/);
    p = fmaf(s2, p, 0.1428571492433548f /*1/7*/);
    p = fmaf(s2, p, 0.20000000298023224f /*1/5*/);
    p = fmaf(s2, p, 0.3333333432674408f /*1/3*/);
    p = p * s2;
    float two_s = s + s;
    float lnm = fmaf(two_s, p, two_s);
    float fe = (float)e;
    return fmaf(fe, K3_LN2_HI, fmaf(fe, K3_LN2_LO, lnm));
}

// ---------------------------------------------------------------- exp(x)
// Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor polynomial to r^8.
K3_HD float k3_pow2i(int n) { return k3_u2f((uint32_t)(n + 127) << 23); }  // n in [-126,127]
K3_HD float k3_exp(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return k3_u2f(0x7f800000u);
    if (x < -103.97208f) return 0.0f;
    float k = rintf(x * K3_LOG2E);
    float r = fmaf(-k, K3_LN2_HI, x);
    r = fmaf(-k, K3_LN2_LO, r);
    float p = fmaf(r, 2.4801587642822415e-05f /*1/8!*/, 0.0001984127011382952f /*1/7!*/);
    p = fmaf(r, p, 0.0013888889225199819f /*1/6!*/);
    p = fmaf(r, p, 0.008333333767950535f /*1/5!*/);
    p = fmaf(r, p, 0.0416666679084301f /*1/4!*/);
    p = fmaf(r, p, 0.1666666716337204f /*1/3!*/);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    int ki = (int)k;
    int k1 = ki / 2;
    int k2 = ki - k1;
    return (p * k3_pow2i(k1)) * k3_pow2i(k2);
}

// ---------------------------------------------------------------- Box-Muller
K3_HD void k3_box_muller(uint32_t a, uint32_t b, float* n0, float* n1) {
    float u1 = k3_uniform01_open0(a);
    float u2 = k3_uniform01(b);
    float r = sqrtf(-2.0f * k3_ln(u1));
    float s, c;
    k3_sincospi(u2 + u2, &s, &c);
    *n0 = r * c;
    *n1 = r * s;
}

// ---------------------------------------------------------------- kernel inference
// Sparse kernel of src/k3dom/kernel/mass_update.cu:24-26:
//   k(d) = max(0, sigma * ((2 + cospi(2d/ls)) (1 - d/ls) / 3 + sinpi(2d/ls) / (2*3.141592)))
// (with the reference's truncated pi), evaluated as a function of the squared
// distance a2 = d^2 so that no square root is taken: with u = a2 / ls^2 the
// bracket is smooth in u on [0, 1] and is read from the piecewise-quadratic
// table of k3_kernel_table.h (K3_KTAB_N intervals; float32 error <= 8e-8 of the
// bracket, against 4e-7 for a sin/cos evaluation).
#include "k3_kernel_table.h"

// Scale of the table argument: s = a2 * scale = N u.
K3_HD float k3_kernel_scale(float ls) {
    return (float)((double)K3_KTAB_N / ((double)ls * (double)ls));
}
// Its square root, sqrt(N) / ls: lengths multiplied by it are in table units
// (the point record's direction and segment bounds).
K3_HD float k3_kernel_sk(float ls) {
    return (float)(sqrt((double)K3_KTAB_N) / (double)ls);
}

// One table entry (a, b, c): a + f (b + f c), f in [0, 1).  Entries at
// i >= K3_KTAB_N are zero, so s >= N (d >= ls) gives exactly 0.
K3_HD float k3_kernel_poly_raw(float a, float b, float c, float f) { return fmaf(f, fmaf(f, c, b), a); }
K3_HD float k3_kernel_poly(float a, float b, float c, float f, float sigma) {
    const float v = k3_kernel_poly_raw(a, b, c, f);
    return fmaxf(0.0f, sigma * v);
}
// Table index and fraction of the argument s (clamped to [0, N]; a rounded
// difference of squares may be slightly negative and stands for 0).
K3_HD int k3_kernel_index_s(float s, float* f) {
    const float c = fminf(fmaxf(s, 0.0f), (float)K3_KTAB_N);
    const int i = (int)c;
    *f = c - (float)i;
    return i;
}
// The bracket of the kernel (sigma = 1), clamped at 0, for the table argument
// s = d^2 * k3_kernel_scale(ls).  Kernel inference sums these per voxel (in
// the chunked order above) and multiplies the sum by k3_sigma_factor(sigma)
// once: max(0, sigma k) = sigma max(0, k) for sigma >= 0, and the bracket is
// non-negative on [0, ls), so a negative sigma gives no evidence, as the
// reference's max(0, sigma k) does there.
K3_HD float k3_kernel_bracket_s(float s) {
    float f;
    const int i = k3_kernel_index_s(s, &f);
    return fmaxf(0.0f, k3_kernel_poly_raw(k3_ktab[4 * i], k3_ktab[4 * i + 1], k3_ktab[4 * i + 2], f));
}
K3_HD float k3_sigma_factor(float sigma) { return fmaxf(sigma, 0.0f); }
// One term, sigma k(d), for the table argument s.
K3_HD float k3_kernel_value_s(float s, float sigma) { return k3_sigma_factor(sigma) * k3_kernel_bracket_s(s); }
// k(sqrt(a2)) for a2 >= 0 with scale = k3_kernel_scale(ls).
K3_HD float k3_kernel_value_sq(float a2, float scale, float sigma) { return k3_kernel_value_s(a2 * scale, sigma); }

// Derivative in ls of sigma * kbar(d / ls), kbar the bracket of the
// reference's kernel (src/k3dom/kernel/mass_update.cu:24-26) differentiated in
// closed form, q = d / ls < 1:
//   kbar'(q) = -(2 pi / 3) sin(2 pi q) (1 - q) - (2 + cos(2 pi q)) / 3 + cos(2 pi q) pi / 3.141592
//   d/dls [sigma kbar(d / ls)] = sigma kbar'(q) (-q / ls)
// Used only by the differentiated window (gradients, not parity values).
K3_HD float k3_kernel_dls(float d, float ls, float inv_ls, float sigma) {
    (void)ls;
    const float q = d * inv_ls;
    float sn, cs;
    k3_sincospi(2.0f * q, &sn, &cs);
    const float two_pi = 6.28318530717958647692f;
    const float dk = -(two_pi / 3.0f) * sn * (1.0f - q) - (2.0f + cs) / 3.0f + cs * (3.14159265358979323846f / 3.141592f);
    return sigma * dk * (-q * inv_ls);
}

// The same derivative from the second table: sigma G(s / N) / ls with
// G(u) = -q kbar'(q), q = sqrt(u), where the kernel value is positive (the
// gather's derivative evaluation; no square root or sine per pair).
K3_HD float k3_kernel_dls_s(float s, float inv_ls, float sigma) {
    float f;
    const int i = k3_kernel_index_s(s, &f);
    if (!(k3_kernel_poly_raw(k3_ktab[4 * i], k3_ktab[4 * i + 1], k3_ktab[4 * i + 2], f) > 0.0f)) return 0.0f;
    return sigma * k3_kernel_poly_raw(k3_ktab_dls[4 * i], k3_ktab_dls[4 * i + 1], k3_ktab_dls[4 * i + 2], f) * inv_ls;
}
K3_HD float k3_kernel_dls_sq(float a2, float scale, float inv_ls, float sigma) {
    return k3_kernel_dls_s(a2 * scale, inv_ls, sigma);
}

// distance_with_margin, src/k3dom/kernel/mass_update.cu:29-31
K3_HD float k3_margin(float distance, float resolution) {
    return fmaxf(0.9f * distance, distance - 2.0f * resolution);
}

// Per-point values of one measurement that every voxel reuses (one 48-byte
// record per point).  p = (x, y, z) is the point relative to the sensor in
// map-aligned axes; dm = |p|; mg = margin(dm); e = (mg/dm) p is the end of the
// free segment.  Distances along the ray are kept in table units: with
// sk = k3_kernel_sk(ls), u = p (sk / dm) is the ray's direction scaled by sk,
// mgs = mg sk and dms = dm sk, so that for a voxel v the projection v . u and
// the squared perpendicular distance |v|^2 sk^2 - (v . u)^2 come out in the
// units of the table argument.  u is NaN when dm == 0, which disables the
// free term exactly as the reference's 0/0 does.
struct K3PointRec {
    float ux, uy, uz, mgs;
    float x, y, z, dms;
    float ex, ey, ez, valid;
};

// valid = 0 for points with a non-finite coordinate or |p|^2 overflow: such a
// point contributes nothing.
K3_HD K3PointRec k3_point_prep(float x, float y, float z, float resolution, float sk) {
    K3PointRec r;
    float ss = fmaf(z, z, fmaf(y, y, x * x));
    bool ok = (x - x == 0.0f) && (y - y == 0.0f) && (z - z == 0.0f) && (ss - ss == 0.0f);
    float dm = sqrtf(ss);
    float mg = k3_margin(dm, resolution);
    float resize = mg / dm;
    float us = (dm > 0.0f) ? sk / dm : k3_u2f(0x7fc00000u);
    r.ux = x * us; r.uy = y * us; r.uz = z * us;
    r.mgs = mg * sk;
    r.x = x; r.y = y; r.z = z;
    r.dms = dm * sk;
    r.ex = resize * x; r.ey = resize * y; r.ez = resize * z;
    r.valid = ok ? 1.0f : 0.0f;
    return r;
}

// Table argument of the occupied evidence of BGKI (src/k3dom/kernel/mass_update.cu:49),
// s = |v - m|^2 * scale, when its kernel term can be non-zero (s < N), else -1.
K3_HD float k3_occ_s(float vx, float vy, float vz, const K3PointRec& p, float scale) {
    float dx = vx - p.x, dy = vy - p.y, dz = vz - p.z;
    float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) * scale;
    return (s < (float)K3_KTAB_N) ? s : -1.0f;
}

// Table argument of the free evidence of BGKI (src/k3dom/kernel/mass_update.cu:54-68):
// the squared distance from the voxel centre to the free segment [0, e] of
// the ray in table units, with the region beyond the measured point
// excluded; -1 when there is no free term or its kernel term is zero.
// dc2s = |v|^2 * scale.  The perpendicular distance squared is clamped at 0
// (the reference can take sqrtf of a rounding-negative value and drop the
// term).
K3_HD float k3_free_s(float vx, float vy, float vz, float dc2s, const K3PointRec& p, float scale) {
    float dp = fmaf(vz, p.uz, fmaf(vy, p.uy, vx * p.ux));
    float a;
    if (dp <= 0.0f) {
        a = dc2s;
    } else if (dp < p.mgs) {
        a = fmaxf(0.0f, fmaf(-dp, dp, dc2s));
    } else if (dp < p.dms) {
        float dx = vx - p.ex, dy = vy - p.ey, dz = vz - p.ez;
        a = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) * scale;
    } else {
        return -1.0f;
    }
    return (a < (float)K3_KTAB_N) ? a : -1.0f;
}

// Voxel centre relative to the sensor, src/k3dom/kernel/mass_update.cu:92-99.
K3_HD void k3_voxel_center(int cell, int gs, int gsz, float res, float sx, float sy, float sz,
                           float* x, float* y, float* z) {
    float xi = (float)(cell % gs) + 0.5f;
    float yi = (float)((cell % (gs * gs)) / gs) + 0.5f;
    float zi = (float)(cell / (gs * gs)) + 0.5f;
    *x = (xi - (float)gs / 2.0f) * res - sx;
    *y = (yi - (float)gs / 2.0f) * res - sy;
    *z = (zi - (float)gsz / 2.0f) * res - sz;
}

// ---------------------------------------------------------------- LiDAR ingest
// A point of a LiDAR (its own frame) in map-aligned axes relative to that
// LiDAR: R p with R row-major (LiDAR axes to map axes), each row as
// (r0 x + r1 y) + r2 z.  The reference's node does this with PCL on the host
// for its one LiDAR (src/k3dom/k3dom_node.cpp:73-76); several LiDARs keep
// their own origins (each one's evidence is computed with the sensor at its
// origin), so no translation enters the point.
K3_HD void k3_lidar_rotate(const float* R, float x, float y, float z, float* ox, float* oy, float* oz) {
    *ox = R[0] * x + R[1] * y + R[2] * z;
    *oy = R[3] * x + R[4] * y + R[5] * z;
    *oz = R[6] * x + R[7] * y + R[8] * z;
}

// ---------------------------------------------------------------- mass rules
// calc_credit, src/k3dom/kernel/mass_update.cu:12-14 (= tanh(mass/scale))
K3_HD float k3_credit(float mass, float mass_scale) {
    return 1.0f - 2.0f / (1.0f + k3_exp(2.0f * mass / mass_scale));
}
K3_HD float k3_positive_diff(float a, float b) { return fmaxf(0.0f, a - b); }
K3_HD float k3_calc_ratio(float numerator, float rest) {
    return (numerator <= 0.0f || rest < 0.0f) ? 0.0f : (1.0f / (1.0f + rest / numerator));
}

// predict_mass, src/k3dom/kernel/mass_update.cu:79-84; decay = gamma^dt.
K3_HD void k3_predict_mass(float* f, float* s, float* d, float dyn_pred, float pred_total, float decay) {
    *s = *s * decay;
    *d = fminf(dyn_pred, fmaxf(0.0f, pred_total - *s));
    *f = pred_total - *d;
}

// update_mass, src/k3dom/kernel/mass_update.cu:106-137 (the evidence part).
K3_HD void k3_update_mass(float* f, float* s, float* d, float del_occ, float del_free, float mass_scale) {
    if (del_free > 0.0f || del_occ > 0.0f) {
        float m_f = *f, m_s = *s, m_d = *d;
        float del_sigma = del_free + del_occ;
        float prime_sigma = m_f + m_s + m_d;
        float credit = k3_credit(sqrtf(del_sigma * prime_sigma), mass_scale);
        float lamda_DtoF = credit * k3_positive_diff(del_free, del_occ) / del_sigma * m_d / prime_sigma;
        float lamda_StoFD = credit * k3_positive_diff(del_free, del_occ) / del_sigma *
                            k3_positive_diff(m_s + m_d, m_f) / prime_sigma;
        float lamda_FtoD = credit * k3_positive_diff(del_occ, del_free) / del_sigma *
                           k3_positive_diff(m_f, m_s + m_d) / prime_sigma;
        float diff_f = lamda_StoFD / 2.0f * m_s + lamda_DtoF * m_d - lamda_FtoD * m_f;
        float diff_s = -lamda_StoFD * m_s;
        float diff_d = lamda_FtoD * m_f + lamda_StoFD / 2.0f * m_s - lamda_DtoF * m_d;
        m_f += diff_f;
        m_s += diff_s;
        m_d += diff_d;
        credit = k3_credit(m_s + m_d, mass_scale);
        float beta = (1.0f - credit) + credit * k3_calc_ratio(m_s, m_d);
        *s = m_s + beta * del_occ;
        *d = m_d + (1.0f - beta) * del_occ;
        *f = m_f + del_free;
    }
}

// separate_newborn_part, src/k3dom/kernel/mass_update.cu:141-150
K3_HD float k3_newborn_part(float m_dyn_pred, float m_total_pred, float m_dyn_up, float p_B) {
    if (m_dyn_pred <= 0.0f) return m_dyn_up;
    if (m_total_pred <= m_dyn_pred) return 0.0f;
    return (m_dyn_up * p_B * (m_total_pred - m_dyn_pred)) / (m_dyn_pred + p_B * (m_total_pred - m_dyn_pred));
}

// ---------------------------------------------------------------- DS-PHD/MIB method
// The reference's second filter (src/dsphdmib/): a measurement grid of
// occupied/free masses from the rays near each voxel, then a Dempster-Shafer
// update of the cell's occupied/free masses (src/dsphdmib/kernel/
// measurement_grid.cu:12-163, src/dsphdmib/kernel/mass_update.cu:12-87).
// Where the reference mixes double literals into float expressions, the
// restatement computes in double at the same points.
#define K3DS_MAX_NEIGHBORS 100   // rays per voxel (first ones in point order), measurement_grid.cu:82
#define K3DS_OCC_MAX 0.95f       // measurement_grid.cu:37
#define K3DS_P_MIN 0.15f         // measurement_grid.cu:50-51
#define K3DS_P_MAX 1.0f
#define K3DS_EPS 0.00001f        // mass clipping, measurement_grid.cu:146
#define K3DS_ALPHA 0.9f          // free-mass decay of predict_free_mass, mass_update.cu:12

// Half-line of point m passes near the voxel at c (|c|^2 = c2): positive
// projection and squared distance to the line below w2 (measurement_grid.cu:86-93).
K3_HD bool k3ds_ray_near(float x, float y, float z, float c2, float mx, float my, float mz, float m2, float w2) {
    const float inner = x * mx + y * my + z * mz;
    return inner > 0.0f && c2 - inner * inner / m2 < w2;
}

// pOcc, measurement_grid.cu:31-41 (own exp)
K3_HD float k3ds_pocc(float x, float y, float z, float mx, float my, float mz, float delta) {
    const float delta_sqr = delta * delta;
    const float r_sqr = (x - mx) * (x - mx) + (y - my) * (y - my) + (z - mz) * (z - mz);
    return K3DS_OCC_MAX * k3_exp(-0.5f * r_sqr / delta_sqr);
}

// pFree, measurement_grid.cu:44-52
K3_HD float k3ds_pfree(float x, float y, float z, float max_range) {
    const float r = sqrtf(x * x + y * y + z * z);
    return K3DS_P_MIN + r / max_range * (K3DS_P_MAX - K3DS_P_MIN);
}

// Measurement masses of a voxel from its near rays: n of them (at most
// K3DS_MAX_NEIGHBORS), all_farther = none of their points is nearer the
// sensor than the voxel, max_occ = their largest pOcc (measurement_grid.cu:100-151).
K3_HD void k3ds_meas_masses(int n, bool all_farther, float max_occ, float x, float y, float z, float max_range,
                            float* m_occ, float* m_free) {
    if (n == 0) {
        *m_occ = 0.0f;
        *m_free = 0.0f;
        return;
    }
    float mo = 0.0f, mf = 0.0f;
    if (all_farther) {
        const float uncertain = k3ds_pfree(x, y, z, max_range);
        if (max_occ > uncertain) {
            mo = max_occ;
        } else {
            mf = 1.0f - uncertain;
        }
    } else {
        mo = max_occ > 0.5f ? max_occ : 0.0f;
    }
    *m_occ = fmaxf(K3DS_EPS, fminf(1.0f - K3DS_EPS, mo));
    *m_free = fmaxf(K3DS_EPS, fminf(1.0f - K3DS_EPS, mf));
}

// |m|^2 of a point, the reference's expression (measurement_grid.cu:87)
K3_HD float k3ds_m2(float x, float y, float z) { return x * x + y * y + z * z; }

// ---- culling bound of the GPU measurement grid (no counterpart in the
// reference, whose kernel tests every point against every voxel)
//
// Voxel-centre coordinate along one axis, the expression of k3_voxel_center
// (non-decreasing in i, so the centres of voxels i0..i1 lie in [c(i0), c(i1)]).
K3_HD float k3ds_axis(int i, int n, float res, float s) { return (((float)i + 0.5f) - (float)n / 2.0f) * res - s; }

// Box of the voxel centres [i0, i1] x [j0, j1] x [k0, k1] (sensor at the
// origin) and the squared distance from its centre below which a point's line
// may pass within the window (w2 = window^2) of one of them.  A voxel centre
// c lies within R of the box centre C.  When 1e-20 <= m2 <= 1e30 and
// |c| <= 1e4, the float test c2 - inner^2/m2 < w2 is off the true squared
// distance d(c)^2 by at most about 15 ulps of c2, so it accepts only if
// d(c)^2 < w2 + 4e-6 c2; then d(C) <= d(c) + R.  The extra terms cover the
// rounding of C, R and of the cross product in k3ds_may_pass.  Beyond
// |c| = 1e4 every point is kept.
struct K3DsBox {
    float cx, cy, cz, thr;
};
K3_HD K3DsBox k3ds_box(int gs, int gsz, float res, float sx, float sy, float sz, int i0, int i1, int j0, int j1,
                       int k0, int k1, float w2) {
    const float lx = k3ds_axis(i0, gs, res, sx), hx = k3ds_axis(i1, gs, res, sx);
    const float ly = k3ds_axis(j0, gs, res, sy), hy = k3ds_axis(j1, gs, res, sy);
    const float lz = k3ds_axis(k0, gsz, res, sz), hz = k3ds_axis(k1, gsz, res, sz);
    K3DsBox b;
    b.cx = 0.5f * (lx + hx);
    b.cy = 0.5f * (ly + hy);
    b.cz = 0.5f * (lz + hz);
    const float ex = 0.5f * (hx - lx), ey = 0.5f * (hy - ly), ez = 0.5f * (hz - lz);
    const float R = sqrtf(ex * ex + ey * ey + ez * ez);
    const float D = sqrtf(b.cx * b.cx + b.cy * b.cy + b.cz * b.cz) + R;
    const float T = sqrtf(w2 + 4e-6f * D * D) + R + 2e-6f * D + 1e-6f;
    b.thr = D <= 1e4f ? T * T * 1.00001f : INFINITY;
    return b;
}

// Whether the line of point m (|m|^2 = m2) may pass near a voxel of box b:
// |C x m|^2 <= thr |m|^2 (the rounding of both sides is within thr's margin);
// true for every point the bound does not cover (m2 outside [1e-20, 1e30], NaN).
K3_HD bool k3ds_may_pass(const K3DsBox& b, float mx, float my, float mz, float m2) {
    if (!(m2 >= 1e-20f && m2 <= 1e30f)) return true;
    const float qx = b.cy * mz - b.cz * my, qy = b.cz * mx - b.cx * mz, qz = b.cx * my - b.cy * mx;
    return !(qx * qx + qy * qy + qz * qz > b.thr * m2);
}

// k3ds_ray_near decided from r, an approximation of 1/m2 within 1 ulp,
// where that is safe: the value c2 - inner^2 r is within 4.8e-7 c2 of the
// test's c2 - inner^2/m2 (1e-20 <= m2 <= 1e30, c2 <= 1e8), so when it is
// farther than 1e-6 c2 + 1e-12 from w2 it falls on the same side; otherwise,
// and outside those ranges, the division decides.
K3_HD bool k3ds_ray_near_rcp(float x, float y, float z, float c2, float mx, float my, float mz, float m2, float r,
                             float w2) {
    const float inner = x * mx + y * my + z * mz;
    if (!(inner > 0.0f)) return false;
    const float q = inner * inner;
    if (m2 >= 1e-20f && m2 <= 1e30f && c2 <= 1e8f) {
        const float a = c2 - q * r;
        const float tol = 1e-6f * c2 + 1e-12f;
        if (a > w2 + tol) return false;
        if (a < w2 - tol) return true;
    }
    return c2 - q / m2 < w2;
}

// predict_free_mass, mass_update.cu:12-15
K3_HD float k3ds_predict_free(float free_mass, float m_occ_pred) {
    return fminf(K3DS_ALPHA * free_mass, 1.0f - m_occ_pred);
}

// update_masses, mass_update.cu:17-30 (Dempster's rule with the measurement;
// the 1.0 literals make those terms double)
K3_HD void k3ds_update_masses(float m_occ_pred, float m_free_pred, float meas_occ, float meas_free, float* occ_up,
                              float* free_up) {
    const float unknown_pred = (float)(1.0 - (double)m_occ_pred - (double)m_free_pred);
    const float meas_unknown = (float)(1.0 - (double)meas_free - (double)meas_occ);
    const float K = m_free_pred * meas_occ + m_occ_pred * meas_free;
    const float no = m_occ_pred * meas_unknown + unknown_pred * meas_occ + m_occ_pred * meas_occ;
    const float nf = m_free_pred * meas_unknown + unknown_pred * meas_free + m_free_pred * meas_free;
    *occ_up = (float)((double)no / (1.0 - (double)K));
    *free_up = (float)((double)nf / (1.0 - (double)K));
}

// separate_newborn_part, mass_update.cu:32-35 (double where the literal makes it)
K3_HD float k3ds_newborn_part(float m_occ_pred, float m_occ_up, float p_B) {
    const double rest = 1.0 - (double)m_occ_pred;
    return (float)(((double)(m_occ_up * p_B) * rest) / ((double)m_occ_pred + (double)p_B * rest));
}

// ---------------------------------------------------------------- particles
// Voxel key of a position in voxel-index units, src/common/kernel/predict.cu:51-55.
// Clamping in float before truncation equals the reference's truncate-then-clamp
// for every finite input; NaN maps to index 0.
K3_HD int k3_key(float x, float y, float z, int gs, int gsz) {
    int ix = (int)fminf(fmaxf(x, 0.0f), (float)(gs - 1));
    int iy = (int)fminf(fmaxf(y, 0.0f), (float)(gs - 1));
    int iz = (int)fminf(fmaxf(z, 0.0f), (float)(gsz - 1));
    return ix + gs * iy + gs * gs * iz;
}

K3_HD float k3_clampf(float a, float lo, float hi) { return fmaxf(fminf(a, hi), lo); }

// Velocity sampled U[-vmax, vmax) per component with the minimum-speed rescale of
// src/common/kernel/init_new_particles.cu:168-176 (and :87-95).
K3_HD void k3_sample_velocity(uint32_t r0, uint32_t r1, uint32_t r2, float vmax, float min_vel,
                              float* vx, float* vy, float* vz) {
    float lo = -vmax, span = vmax - (-vmax);
    float x = lo + span * k3_uniform01(r0);
    float y = lo + span * k3_uniform01(r1);
    float z = lo + span * k3_uniform01(r2);
    float vel_sq = x * x + y * y + z * z;
    if (vel_sq < min_vel * min_vel && vel_sq != 0.0f) {
        float rate = min_vel / sqrtf(vel_sq);
        x *= rate; y *= rate; z *= rate;
    }
    *vx = x; *vy = y; *vz = z;
}

// One particle's prediction step, src/common/kernel/predict.cu:20-56, with the
// six N(0, sigma) draws taken from Philox counters (i, frame, PREDICT, 0..1).
struct K3PredictConst {
    float dt, p_S, sigma_pos, sigma_vel, vclamp;
    int gs, gsz;
    uint32_t frame, key0, key1;
};

K3_HD void k3_predict_particle(uint32_t i, const K3PredictConst& c, float* px, float* py, float* pz,
                               float* vx, float* vy, float* vz, float* w, int* key) {
    uint32_t r[8];
    k3_philox4x32_10(i, c.frame, K3_STREAM_PREDICT, 0u, c.key0, c.key1, r);
    k3_philox4x32_10(i, c.frame, K3_STREAM_PREDICT, 1u, c.key0, c.key1, r + 4);
    float n0, n1, n2, n3, n4, n5;
    k3_box_muller(r[0], r[1], &n0, &n1);
    k3_box_muller(r[2], r[3], &n2, &n3);
    k3_box_muller(r[4], r[5], &n4, &n5);
    float npx = n0 * c.sigma_pos + 0.0f, npy = n1 * c.sigma_pos + 0.0f, npz = n2 * c.sigma_pos + 0.0f;
    float nvx = n3 * c.sigma_vel + 0.0f, nvy = n4 * c.sigma_vel + 0.0f, nvz = n5 * c.sigma_vel + 0.0f;
    float x = *px + (*vx * c.dt + npx);
    float y = *py + (*vy * c.dt + npy);
    float z = *pz + (*vz * c.dt + npz);
    float ux = *vx + nvx, uy = *vy + nvy, uz = *vz + nvz;
    float ww = *w * c.p_S;
    ux = k3_clampf(ux, -c.vclamp, c.vclamp);
    uy = k3_clampf(uy, -c.vclamp, c.vclamp);
    uz = k3_clampf(uz, -c.vclamp, c.vclamp);
    if ((x >= (float)c.gs || x < 0.0f) || (y >= (float)c.gs || y < 0.0f) || (z >= (float)c.gsz || z < 0.0f))
        ww = 0.0f;
    *px = x; *py = y; *pz = z;
    *vx = ux; *vy = uy; *vz = uz;
    *w = ww;
    *key = k3_key(x, y, z, c.gs, c.gsz);
}

// ---------------------------------------------------------------- fixed-point CDFs
// Weights become u64 fixed point q = trunc(w * 2^E) (computed exactly in double)
// with E chosen from the largest weight so that the total stays below 2^62.
// Integer prefix sums do not depend on the summation order.
K3_HD int k3_fixed_exponent(float wmax, uint32_t count) {
    if (!(wmax > 0.0f)) return 0;
    int lg = 0;
    while (lg < 32 && ((uint64_t)1 << lg) < (uint64_t)count) ++lg;
    uint32_t b = k3_f2u(wmax);
    int e = (int)((b >> 23) & 0xffu) - 127;
    if (((b >> 23) & 0xffu) == 0) e = -126;  // subnormal max
    return 61 - lg - e;
}
K3_HD uint64_t k3_to_fixed(float w, int E) {
    if (!(w > 0.0f) || !(w - w == 0.0f)) return 0;
    double v = (double)w * k3_pow2d(E);
    return (uint64_t)v;
}
