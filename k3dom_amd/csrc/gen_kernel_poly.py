"""Generates the polynomial form of the sparse kernel used by k3_math.h.

The reference evaluates, for q = d / ls in [0, 1) (src/k3dom/kernel/mass_update.cu:24-26),
    k(q) = (2 + cos(2 pi q)) (1 - q) / 3 + sin(2 pi q) / (2 * 3.141592)
(a truncated pi in the last divisor).  This script interpolates k at the
Chebyshev points of degree DEG in x = 2q - 1, converts to monomial
coefficients in x, rounds them to float32 and prints them with the maximum
error of the float32 Horner evaluation (fmaf) against float64 k on a dense
grid.  The constant coefficient is then set to minus the float32 Horner
value of the others at x = 1, so that the evaluation at x = 1 (q = 1, where
k vanishes) is exactly zero: the kernel evaluator clamps x to at most 1, which
makes every term with d >= ls an exact zero.
Run: python3 k3dom_amd/csrc/gen_kernel_poly.py
"""
import numpy as np

DEG = 14
PI_T = float(np.float32(3.141592))


def k(q):
    return (2 + np.cos(2 * np.pi * q)) * (1 - q) / 3 + np.sin(2 * np.pi * q) / (2 * PI_T)


def coefficients(deg=DEG):
    n = deg + 1
    xs = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    cheb = np.polynomial.chebyshev.chebfit(xs, k((xs + 1) / 2), deg)
    mono = np.polynomial.chebyshev.cheb2poly(cheb).astype(np.float32)
    acc = mono[-1]
    for a in mono[-2:0:-1]:
        acc = np.float32(np.float64(acc) + np.float64(a))  # fmaf(acc, 1, a)
    mono[0] = -acc
    return mono


def horner32(c, q):
    q = np.asarray(q, np.float32)
    x = (np.float32(2) * q - np.float32(1)).astype(np.float32)
    acc = np.full_like(x, c[-1])
    for a in c[-2::-1]:
        acc = (acc.astype(np.float64) * x + np.float64(a)).astype(np.float32)  # fmaf: one rounding
    return acc


if __name__ == "__main__":
    c = coefficients()
    q = np.linspace(0, 1, 2000001, dtype=np.float64).astype(np.float32)
    err = np.abs(horner32(c, q).astype(np.float64) - k(q.astype(np.float64)))
    print(f"degree {DEG}: max |float32 Horner - k| = {err.max():.3e} at q = {q[err.argmax()]:.6f}")
    print(f"value at q = 1: {float(horner32(c, np.float32(1.0))):.3e}")
    for i, a in enumerate(c):
        print(f"#define K3_KP{i} {float(a)!r}f")
