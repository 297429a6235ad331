"""Generates k3_kernel_table.h: the sparse kernel as a function of the squared
distance, tabulated for piecewise-quadratic interpolation.

The reference evaluates, for q = d / ls in [0, 1) (src/k3dom/kernel/mass_update.cu:24-26),
    kbar(q) = (2 + cos(2 pi q)) (1 - q) / 3 + sin(2 pi q) / (2 * 3.141592)
(a truncated pi in the last divisor) and multiplies by sigma.  In u = q^2 =
d^2 / ls^2 the function kbar(sqrt(u)) is smooth on [0, 1], so the kernel needs
no square root: u in [0, 1) is split into N equal intervals, and on interval i
with f = N u - i in [0, 1)
    p_i(f) = a_i + f (b_i + f c_i)
interpolates kbar(sqrt(u)) at f = 0, 1/2, 1 (values in double, coefficients
rounded to float32).  Entry N is zero, so every u >= 1 evaluates to exactly 0.
The float32 evaluation (two fmaf) stays within 8e-8 of kbar over [0, 1).
A second table holds G(u) = -q kbar'(q) (the length-scale derivative factor of
the differentiated window) with the same intervals and interpolation.
Run: python3 k3dom_amd/csrc/gen_kernel_table.py > k3dom_amd/csrc/k3_kernel_table.h
"""
import numpy as np

N = 512
PI_T = float(np.float32(3.141592))


def kbar(q):
    return (2 + np.cos(2 * np.pi * q)) * (1 - q) / 3 + np.sin(2 * np.pi * q) / (2 * PI_T)


def table(n=N):
    h = 1.0 / n
    u0 = np.arange(n) * h
    y0 = kbar(np.sqrt(u0))
    ym = kbar(np.sqrt(u0 + h / 2))
    y1 = kbar(np.sqrt(u0 + h))
    y1[-1] = 0.0  # kbar(1) = 0
    c = 2 * y0 - 4 * ym + 2 * y1
    b = y1 - y0 - c
    t = np.zeros((n + 1, 4), np.float32)
    t[:n, 0] = y0
    t[:n, 1] = b
    t[:n, 2] = c
    return t


def dkbar(q):
    """kbar'(q) in closed form."""
    return (-(2 * np.pi / 3) * np.sin(2 * np.pi * q) * (1 - q) - (2 + np.cos(2 * np.pi * q)) / 3
            + np.cos(2 * np.pi * q) * np.pi / PI_T)


def gls(u):
    """G(u) = -q kbar'(q), q = sqrt(u): d/dls [sigma kbar(d / ls)] = sigma G(d^2 / ls^2) / ls."""
    q = np.sqrt(u)
    return -q * dkbar(q)


def table_dls(n=N):
    """G(u) tabulated like kbar: same intervals, same interpolation, entry n zero."""
    h = 1.0 / n
    u0 = np.arange(n) * h
    y0, ym, y1 = gls(u0), gls(u0 + h / 2), gls(u0 + h)
    c = 2 * y0 - 4 * ym + 2 * y1
    b = y1 - y0 - c
    t = np.zeros((n + 1, 4), np.float32)
    t[:n, 0] = y0
    t[:n, 1] = b
    t[:n, 2] = c
    return t


def evaluate(t, u):
    """float32 evaluation exactly as k3_kernel_value_sq does (sigma = 1, before the max)."""
    n = t.shape[0] - 1
    s = np.minimum((np.asarray(u, np.float32) * np.float32(n)).astype(np.float32), np.float32(n))
    i = s.astype(np.int64)
    f = (s - i.astype(np.float32)).astype(np.float32)
    inner = (f.astype(np.float64) * t[i, 2] + t[i, 1]).astype(np.float32)
    return (f.astype(np.float64) * inner + t[i, 0]).astype(np.float32)


if __name__ == "__main__":
    t = table()
    u = np.linspace(0, 1, 2_000_001)[:-1].astype(np.float32)
    err = np.abs(evaluate(t, u).astype(np.float64) - kbar(np.sqrt(u.astype(np.float64))))
    print("// Generated by gen_kernel_table.py (beside this file); do not edit.")
    print(f"// Sparse kernel bracket of src/k3dom/kernel/mass_update.cu:24-26 as a function of")
    print(f"// u = d^2 / ls^2: {N} intervals of piecewise-quadratic interpolation, entry {N} zero.")
    print(f"// Max |float32 evaluation - kbar| over [0, 1): {err.max():.2e}")
    print("#pragma once")
    print(f"#define K3_KTAB_N {N}")
    print(f"static constexpr float k3_ktab[{(N + 1) * 4}] = {{")
    for row in t:
        print("    " + ", ".join(f"{float(v)!r}f" for v in row) + ",")
    print("};")
    g = table_dls()
    gerr = np.abs(evaluate(g, u).astype(np.float64) - gls(u.astype(np.float64)))
    print(f"// Length-scale derivative factor G(u) = -q kbar'(q), q = sqrt(u), tabulated the")
    print(f"// same way (d/dls [sigma kbar(d/ls)] = sigma G(u) / ls).  Max |float32 evaluation - G|: {gerr.max():.2e}")
    print(f"static constexpr float k3_ktab_dls[{(N + 1) * 4}] = {{")
    for row in g:
        print("    " + ", ".join(f"{float(v)!r}f" for v in row) + ",")
    print("};")
