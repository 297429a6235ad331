// Forward-mode dual numbers and the per-voxel mass update written once for
// any scalar type.  Used by the differentiated window: the GPU backward pass
// takes the local Jacobian of one voxel's update (6 inputs) with
// K3Dual<float, 6>, and the double-precision reference gradient of the serial
// restatement propagates K3Dual<double, 6> through a whole window.
//
// The operations follow the float functions of k3_math.h step for step
// (src/k3dom/kernel/mass_update.cu:12-22,79-150); a branch is decided on the
// value part, so the derivative is that of the branch taken (the integer and
// branch decisions of the forward frame are frozen).
#pragma once

#include "k3_math.h"

#ifdef __HIPCC__
#define K3_HDT __host__ __device__ __forceinline__
#else
#define K3_HDT inline
#include <cmath>
#endif

template <class T, int N>
struct K3Dual {
    T v;
    T d[N];
};

template <class T, int N>
K3_HDT K3Dual<T, N> k3d_const(T v) {
    K3Dual<T, N> r;
    r.v = v;
    for (int i = 0; i < N; ++i) r.d[i] = T(0);
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> k3d_var(T v, int which) {
    K3Dual<T, N> r = k3d_const<T, N>(v);
    r.d[which] = T(1);
    return r;
}

template <class T, int N>
K3_HDT K3Dual<T, N> operator+(const K3Dual<T, N>& a, const K3Dual<T, N>& b) {
    K3Dual<T, N> r;
    r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> operator-(const K3Dual<T, N>& a, const K3Dual<T, N>& b) {
    K3Dual<T, N> r;
    r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> operator-(const K3Dual<T, N>& a) {
    K3Dual<T, N> r;
    r.v = -a.v;
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> operator*(const K3Dual<T, N>& a, const K3Dual<T, N>& b) {
    K3Dual<T, N> r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> operator/(const K3Dual<T, N>& a, const K3Dual<T, N>& b) {
    K3Dual<T, N> r;
    r.v = a.v / b.v;
    const T inv = T(1) / b.v;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
// with plain scalars
template <class T, int N>
K3_HDT K3Dual<T, N> operator+(const K3Dual<T, N>& a, T b) { return a + k3d_const<T, N>(b); }
template <class T, int N>
K3_HDT K3Dual<T, N> operator+(T a, const K3Dual<T, N>& b) { return k3d_const<T, N>(a) + b; }
template <class T, int N>
K3_HDT K3Dual<T, N> operator-(const K3Dual<T, N>& a, T b) { return a - k3d_const<T, N>(b); }
template <class T, int N>
K3_HDT K3Dual<T, N> operator-(T a, const K3Dual<T, N>& b) { return k3d_const<T, N>(a) - b; }
template <class T, int N>
K3_HDT K3Dual<T, N> operator*(const K3Dual<T, N>& a, T b) {
    K3Dual<T, N> r;
    r.v = a.v * b;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> operator*(T a, const K3Dual<T, N>& b) { return b * a; }
template <class T, int N>
K3_HDT K3Dual<T, N> operator/(const K3Dual<T, N>& a, T b) { return a / k3d_const<T, N>(b); }
template <class T, int N>
K3_HDT K3Dual<T, N> operator/(T a, const K3Dual<T, N>& b) { return k3d_const<T, N>(a) / b; }

// a constant of type X (plain scalar or dual number)
template <class X>
struct K3CastT {
    static K3_HDT X of(double v) { return X(v); }
};
template <class T, int N>
struct K3CastT<K3Dual<T, N>> {
    static K3_HDT K3Dual<T, N> of(double v) { return k3d_const<T, N>((T)v); }
};
template <class X>
K3_HDT X k3d_cast(double v) { return K3CastT<X>::of(v); }

// value of a plain scalar or a dual number (for branch decisions)
K3_HDT float k3g_val(float x) { return x; }
K3_HDT double k3g_val(double x) { return x; }
template <class T, int N>
K3_HDT T k3g_val(const K3Dual<T, N>& x) { return x.v; }

// elementary functions
K3_HDT float k3g_exp(float x) { return k3_exp(x); }
K3_HDT float k3g_sqrt(float x) { return sqrtf(x); }
K3_HDT float k3g_sin(float x) { return sinf(x); }
K3_HDT float k3g_cos(float x) { return cosf(x); }
K3_HDT double k3g_exp(double x) { return exp(x); }
K3_HDT double k3g_sqrt(double x) { return sqrt(x); }
K3_HDT double k3g_sin(double x) { return sin(x); }
K3_HDT double k3g_cos(double x) { return cos(x); }
K3_HDT float k3g_log(float x) { return logf(x); }
K3_HDT double k3g_log(double x) { return log(x); }
template <class T, int N>
K3_HDT K3Dual<T, N> k3g_log(const K3Dual<T, N>& a) {
    K3Dual<T, N> r;
    r.v = k3g_log(a.v);
    const T inv = T(1) / a.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * inv;
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> k3g_exp(const K3Dual<T, N>& a) {
    K3Dual<T, N> r;
    r.v = k3g_exp(a.v);
    for (int i = 0; i < N; ++i) r.d[i] = r.v * a.d[i];
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> k3g_sqrt(const K3Dual<T, N>& a) {
    K3Dual<T, N> r;
    r.v = k3g_sqrt(a.v);
    const T h = r.v > T(0) ? T(0.5) / r.v : T(0);
    for (int i = 0; i < N; ++i) r.d[i] = h * a.d[i];
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> k3g_sin(const K3Dual<T, N>& a) {
    K3Dual<T, N> r;
    r.v = k3g_sin(a.v);
    const T c = k3g_cos(a.v);
    for (int i = 0; i < N; ++i) r.d[i] = c * a.d[i];
    return r;
}
template <class T, int N>
K3_HDT K3Dual<T, N> k3g_cos(const K3Dual<T, N>& a) {
    K3Dual<T, N> r;
    r.v = k3g_cos(a.v);
    const T s = k3g_sin(a.v);
    for (int i = 0; i < N; ++i) r.d[i] = -s * a.d[i];
    return r;
}
// max / min pick an operand by value (ties: the first operand)
template <class X>
K3_HDT X k3g_max(const X& a, const X& b) { return k3g_val(b) > k3g_val(a) ? b : a; }
template <class X>
K3_HDT X k3g_min(const X& a, const X& b) { return k3g_val(b) < k3g_val(a) ? b : a; }

// ------------------------------------------------------------------ mass update
// calc_credit, positive_diff, calc_ratio (src/k3dom/kernel/mass_update.cu:12-22)
template <class X, class T>
K3_HDT X k3g_credit(const X& mass, T mass_scale) {
    return T(1) - T(2) / (T(1) + k3g_exp(T(2) * mass / mass_scale));
}
// Dual numbers: the same value; the derivative in closed form,
// d/dm (1 - 2 / (1 + e^(2m/s))) = (1 - c^2) / s with c = tanh(m/s) the value,
// which stays finite where e^(2m/s) overflows (the quotient rule would give
// inf / inf there).
template <class T, int N, class S>
K3_HDT K3Dual<T, N> k3g_credit(const K3Dual<T, N>& mass, S mass_scale) {
    K3Dual<T, N> r;
    r.v = T(1) - T(2) / (T(1) + k3g_exp(T(2) * mass.v / T(mass_scale)));
    const T dc = (T(1) - r.v * r.v) / T(mass_scale);
    for (int i = 0; i < N; ++i) r.d[i] = dc * mass.d[i];
    return r;
}
template <class X>
K3_HDT X k3g_positive_diff(const X& a, const X& b) {
    const X d = a - b;
    return k3g_val(d) > 0 ? d : d * decltype(k3g_val(d))(0);
}
template <class X, class T>
K3_HDT X k3g_calc_ratio(const X& num, const X& rest) {
    if (k3g_val(num) <= 0 || k3g_val(rest) < 0) return num * T(0);
    return T(1) / (T(1) + rest / num);
}

// One voxel's predict_mass + update_mass + separate_newborn_part
// (src/k3dom/kernel/mass_update.cu:79-150,161-181): in (f, s, d) are the
// masses before the frame, out the masses after it and the born mass.
template <class X, class T>
K3_HDT void k3g_mass_step(X f, X s, X d, const X& dyn_pred, const X& del_occ, const X& del_free, T decay,
                          T mass_scale, T p_B, X* f_out, X* s_out, X* d_out, X* born) {
    const X pred_total = (f + d) * decay;
    s = s * decay;
    {
        const X zero = s * T(0);
        d = k3g_min(dyn_pred, k3g_max(zero, pred_total - s));
        f = pred_total - d;
    }
    if (k3g_val(del_free) > 0 || k3g_val(del_occ) > 0) {
        const X del_sigma = del_free + del_occ;
        const X prime_sigma = f + s + d;
        X credit = k3g_credit(k3g_sqrt(del_sigma * prime_sigma), mass_scale);
        const X fo = k3g_positive_diff(del_free, del_occ), of = k3g_positive_diff(del_occ, del_free);
        const X l_dtof = credit * fo / del_sigma * d / prime_sigma;
        const X l_stofd = credit * fo / del_sigma * k3g_positive_diff(s + d, f) / prime_sigma;
        const X l_ftod = credit * of / del_sigma * k3g_positive_diff(f, s + d) / prime_sigma;
        const X diff_f = l_stofd / T(2) * s + l_dtof * d - l_ftod * f;
        const X diff_s = -(l_stofd * s);
        const X diff_d = l_ftod * f + l_stofd / T(2) * s - l_dtof * d;
        f = f + diff_f;
        s = s + diff_s;
        d = d + diff_d;
        credit = k3g_credit(s + d, mass_scale);
        const X beta = (T(1) - credit) + credit * k3g_calc_ratio<X, T>(s, d);
        s = s + beta * del_occ;
        d = d + (T(1) - beta) * del_occ;
        f = f + del_free;
    }
    // separate_newborn_part
    X b;
    if (k3g_val(dyn_pred) <= 0) b = d;
    else if (k3g_val(pred_total) <= k3g_val(dyn_pred)) b = d * T(0);
    else b = (d * p_B * (pred_total - dyn_pred)) / (dyn_pred + p_B * (pred_total - dyn_pred));
    *f_out = f;
    *s_out = s;
    *d_out = d;
    *born = b;
}

// Binary cross-entropy of the occupancy mean (s + d) / (f + s + d) against a
// 0/1 target, the mean clamped to [eps, 1 - eps].
#define K3_BCE_EPS 1e-6
template <class X, class T>
K3_HDT X k3g_bce_occ(const X& f, const X& s, const X& d, int target) {
    X p = (s + d) / (f + s + d);
    const T lo = T(K3_BCE_EPS), hi = T(1) - T(K3_BCE_EPS);
    if (k3g_val(p) < lo) p = p * T(0) + lo;
    if (k3g_val(p) > hi) p = p * T(0) + hi;
    return target ? -k3g_log(p) : -k3g_log(T(1) - p);
}
