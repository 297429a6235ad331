"""Pins of the deterministic math shared by the kernels and the serial
restatement: published Philox4x32-10 known answers, and the polynomial
ln / exp / sin / cos / kernel against libm in double precision."""
import ctypes as C
import math

import numpy as np
import pytest


def _philox(L, ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox(c, k, o)
    return [o[i] for i in range(4)]


# Known-answer vectors of Philox4x32-10 published with Random123 (Salmon et al. 2011).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(oracle_lib, ctr, key, expect):
    assert _philox(oracle_lib, ctr, key) == list(expect)


def test_philox_output_is_uniform(oracle_lib):
    # 4096 counters x 4 words: byte histogram chi-square and bit balance
    words = np.array([_philox(oracle_lib, (i, 7, 1, 0), (123456, 0)) for i in range(4096)], np.uint32).ravel()
    bits = np.unpackbits(words.view(np.uint8))
    assert abs(bits.mean() - 0.5) < 0.005
    hist = np.bincount(words.view(np.uint8), minlength=256)
    exp = len(words) * 4 / 256
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 350  # 255 dof, p ~ 1e-4


def _ulp(x):
    return np.spacing(np.float32(abs(x))).astype(np.float64)


def test_ln_matches_libm(oracle_lib):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(2 ** -24, 1.0, 20000), np.geomspace(1e-30, 1e30, 2000),
                         [2 ** -24, 0.5, 1.0, 2.0, math.sqrt(2), 1 - 2 ** -24]]).astype(np.float32)
    worst = 0.0
    for x in xs:
        got = oracle_lib.orc_ln(float(x))
        ref = math.log(float(x))
        err = abs(got - ref)
        tol = 2 * _ulp(ref) + 1e-45
        worst = max(worst, err / max(_ulp(ref), 1e-45))
        assert err <= tol, (x, got, ref)
    assert oracle_lib.orc_ln(1.0) == 0.0


def test_exp_matches_libm(oracle_lib):
    xs = np.concatenate([np.linspace(-80, 88.7, 20001), np.linspace(-1, 1, 2001)]).astype(np.float32)
    for x in xs:
        got = oracle_lib.orc_exp(float(x))
        ref = math.exp(float(x))
        assert abs(got - ref) <= 2 * _ulp(ref), (x, got, ref)
    assert oracle_lib.orc_exp(0.0) == 1.0
    assert math.isinf(oracle_lib.orc_exp(89.0))
    assert oracle_lib.orc_exp(-200.0) == 0.0


def test_sincospi_matches_libm(oracle_lib):
    s, c = C.c_float(), C.c_float()
    for t in np.linspace(0, 2, 40001).astype(np.float32):
        oracle_lib.orc_sincospi(float(t), C.byref(s), C.byref(c))
        assert abs(s.value - math.sin(math.pi * float(t))) <= 2.5e-7
        assert abs(c.value - math.cos(math.pi * float(t))) <= 2.5e-7
    # exact special points
    for t, es, ec in [(0.0, 0.0, 1.0), (0.5, 1.0, 0.0), (1.0, 0.0, -1.0), (1.5, -1.0, 0.0)]:
        oracle_lib.orc_sincospi(t, C.byref(s), C.byref(c))
        assert abs(s.value - es) == 0.0 and abs(c.value - ec) == 0.0


def _kern64(d, ls, sigma):
    q = d / ls
    return max(0.0, sigma * ((2 + math.cos(2 * math.pi * q)) * (1 - q) / 3 +
                             math.sin(2 * math.pi * q) / (2 * float(np.float32(3.141592)))))


def test_kernel_closed_form(oracle_lib):
    ls, sigma = np.float32(0.5), np.float32(0.1)
    inv = ls  # the evaluator takes the length scale itself
    # k(0) = sigma exactly; k(d) -> 0 as d -> ls; monotone non-increasing
    assert oracle_lib.orc_kernel(0.0, float(inv), float(sigma)) == float(sigma)
    ds = np.linspace(0, float(ls), 5001, endpoint=False).astype(np.float32)
    vals = np.array([oracle_lib.orc_kernel(float(d), float(inv), float(sigma)) for d in ds])
    ref = np.array([_kern64(float(d), float(ls), float(sigma)) for d in ds])
    assert np.max(np.abs(vals - ref)) < 4e-8
    assert np.all(np.diff(vals) <= 2 * np.spacing(np.float32(sigma)))
    assert vals[-1] < 1e-9
    # the sparse kernel is flat at both ends: k'(0) = 0 and k'(ls) = 0 (a dropped
    # sin or cos term makes one of these slopes non-zero)
    h = np.float32(1e-3) * ls
    k0 = oracle_lib.orc_kernel(0.0, float(inv), float(sigma))
    kh = oracle_lib.orc_kernel(float(h), float(inv), float(sigma))
    k1 = oracle_lib.orc_kernel(float(ls - h), float(inv), float(sigma))
    assert abs((kh - k0) / float(h)) < 1e-3 * float(sigma) / float(ls) * 10
    assert abs(k1 / float(h)) < 1e-3 * float(sigma) / float(ls) * 10
    # and its value at half the length scale: (2 + cos pi)(1/2)/3 + sin(pi)/(2 pi) = 1/6
    assert abs(oracle_lib.orc_kernel(float(ls / 2), float(inv), 1.0) - 1.0 / 6.0) < 1e-7


def test_credit_is_tanh(oracle_lib):
    for scale in [2.5, 3.0, 0.7]:
        for m in np.concatenate([np.linspace(0, 20, 2001), [1e3, 1e7, 1e30]]):
            got = oracle_lib.orc_credit(float(np.float32(m)), float(np.float32(scale)))
            ref = math.tanh(float(np.float32(m)) / float(np.float32(scale)))
            assert abs(got - ref) <= 3e-7, (m, scale, got, ref)


def test_box_muller_gaussian(oracle_lib):
    a, b = C.c_float(), C.c_float()
    out = []
    for i in range(20000):
        r = _philox(oracle_lib, (i, 0, 1, 0), (5, 6))
        oracle_lib.orc_box_muller(r[0], r[1], C.byref(a), C.byref(b))
        u1 = ((r[0] >> 8) + 1) * 2.0 ** -24
        u2 = (r[1] >> 8) * 2.0 ** -24
        rr = math.sqrt(-2 * math.log(u1))
        assert abs(a.value - rr * math.cos(2 * math.pi * u2)) < 2e-6
        assert abs(b.value - rr * math.sin(2 * math.pi * u2)) < 2e-6
        out += [a.value, b.value]
    out = np.array(out)
    assert abs(out.mean()) < 0.02 and abs(out.var() - 1) < 0.03
    # fraction beyond 2 sigma ~ 4.55 %
    assert abs((np.abs(out) > 2).mean() - 0.0455) < 0.005


def test_key_matches_truncate_then_clamp(oracle_lib):
    rng = np.random.default_rng(3)
    gs, gsz = 64, 16
    vals = np.concatenate([rng.uniform(-5, 70, 3000), [-1.0, -0.999, -0.0, 0.0, 63.999, 64.0, 1e9, -1e9]])
    for x, y, z in zip(vals, vals[::-1], rng.uniform(-3, 20, len(vals))):
        x, y, z = np.float32(x), np.float32(y), np.float32(z)

        def ref_idx(v, n):
            i = int(np.trunc(v)) if abs(v) < 2 ** 31 else (n if v > 0 else -1)
            return min(max(i, 0), n - 1)
        exp = ref_idx(x, gs) + gs * ref_idx(y, gs) + gs * gs * ref_idx(z, gsz)
        assert oracle_lib.orc_key(float(x), float(y), float(z), gs, gsz) == exp
    assert oracle_lib.orc_key(float("nan"), 1.0, 1.0, gs, gsz) == 0 + gs * 1 + gs * gs * 1


def test_kernel_table_matches_generator_and_formula(oracle_lib):
    """The header's kernel table is what gen_kernel_table.py prints, and the
    float32 evaluation through it stays within 1.5e-7 (relative to sigma) of
    the reference formula in double over the whole support."""
    import importlib.util
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("gkt", os.path.join(root, "k3dom_amd", "csrc", "gen_kernel_table.py"))
    gkt = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gkt)
    header = open(os.path.join(root, "k3dom_amd", "csrc", "k3_kernel_table.h")).read()
    body = header[header.index("{") + 1:header.index("}")]
    got = np.array([np.float32(float(v)) for v in re.findall(r"([-0-9.e]+)f", body)], np.float32)
    want = gkt.table().reshape(-1)
    assert got.shape == want.shape and np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.all(want[-4:] == 0)  # beyond the support
    ls, sigma = 0.5, 1.0
    ds = np.linspace(0, ls, 20001, endpoint=False).astype(np.float32)
    vals = np.array([oracle_lib.orc_kernel(float(d), ls, sigma) for d in ds])
    ref = np.maximum(0.0, gkt.kbar(ds.astype(np.float64) / ls))
    assert np.max(np.abs(vals - ref)) < 1.5e-7
    # a dropped coefficient or a shifted interval is far outside that bound
    assert gkt.evaluate(gkt.table(), np.float32(0.3)) == gkt.evaluate(want.reshape(-1, 4), np.float32(0.3))


def test_kernel_table_never_exceeds_one():
    """The gather clamps each table value to [0, 1] (one clamp bit on its last
    FMA) where the restatement takes max(0, value): the same only if no
    interval's quadratic a + f (b + f c) exceeds 1 for f in [0, 1].  Checked
    exactly at the interval ends and the vertex, and through the float32 FMA
    evaluation on a dense grid of the first intervals (the largest values)."""
    import importlib.util
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("gkt", os.path.join(root, "k3dom_amd", "csrc", "gen_kernel_table.py"))
    gkt = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gkt)
    t = gkt.table().astype(np.float64)
    a, b, c = t[:, 0], t[:, 1], t[:, 2]
    top = np.maximum(a, a + b + c)
    inner = (c != 0) & (-b / np.where(c != 0, 2 * c, 1) > 0) & (-b / np.where(c != 0, 2 * c, 1) < 1)
    fv = np.where(inner, -b / np.where(c != 0, 2 * c, 1), 0.0)
    top = np.maximum(top, np.where(inner, a + fv * (b + fv * c), -np.inf))
    assert np.max(top) <= 1.0
    assert np.max(top[1:]) < 1.0 - 1e-3  # only entry 0 reaches 1, at f = 0
    f = np.linspace(0.0, 1.0, 100001).astype(np.float32)
    t32 = gkt.table()
    for i in range(4):
        # float32 FMA: the product is exact in double, one rounding to float32
        inner32 = (f.astype(np.float64) * np.float64(t32[i, 2]) + np.float64(t32[i, 1])).astype(np.float32)
        v = (f.astype(np.float64) * inner32.astype(np.float64) + np.float64(t32[i, 0])).astype(np.float32)
        assert np.max(v) <= np.float32(1.0)


@pytest.mark.parametrize("ls", [0.5, 0.3, 0.1, 0.7, 1.3, 0.05, 2.0, 0.123456])
def test_kernel_is_exactly_zero_from_the_length_scale_on(oracle_lib, ls):
    """Every d beyond ls (past the rounding of d^2 / ls^2) gives an exact zero,
    so a term outside the support adds nothing and culling such pairs cannot
    change a sum, for either sign of sigma; just below ls the value is tiny and
    non-negative."""
    ls32 = np.float32(ls)
    above = [ls32 * np.float32(1 + 3e-7), ls32 * np.float32(1.0001), ls32 * 2, np.float32(1e30)]
    for d in above:
        for sigma in (0.1, 1.0, -0.1):
            assert oracle_lib.orc_kernel(float(d), float(ls32), sigma) == 0.0, (ls, d, sigma)
    below = np.nextafter(ls32, np.float32(0))
    v = oracle_lib.orc_kernel(float(below), float(ls32), 1.0)
    assert 0.0 <= v < 1e-6


def test_kernel_length_scale_derivative_table(oracle_lib):
    """The gather's length-scale derivative comes from a second table,
    G(u) = -q kbar'(q) with u = q^2 (k3_ktab_dls): it is what the generator
    prints, the closed form kbar'(q) matches central differences of kbar in
    double, and the table evaluation sigma G(d^2/ls^2) / ls stays within 1e-6
    (relative to sigma / ls) of the closed form over the support, with exact
    zeros from ls on."""
    import importlib.util
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("gkt", os.path.join(root, "k3dom_amd", "csrc", "gen_kernel_table.py"))
    gkt = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gkt)
    header = open(os.path.join(root, "k3dom_amd", "csrc", "k3_kernel_table.h")).read()
    body = header[header.index("k3_ktab_dls"):]
    body = body[body.index("{") + 1:body.index("}")]
    got = np.array([np.float32(float(v)) for v in re.findall(r"([-0-9.e]+)f", body)], np.float32)
    want = gkt.table_dls().reshape(-1)
    assert got.shape == want.shape and np.array_equal(got.view(np.uint32), want.view(np.uint32))
    q = np.linspace(0.01, 0.99, 99)
    h = 1e-6
    fd = (gkt.kbar(q + h) - gkt.kbar(q - h)) / (2 * h)
    assert np.max(np.abs(fd - gkt.dkbar(q))) < 1e-8
    for ls, sigma in ((0.5, 1.0), (0.3, 0.1), (1.7, 2.5)):
        ds = np.linspace(0, ls, 4001, endpoint=False).astype(np.float32)
        tab = np.array([oracle_lib.orc_kernel_dls_sq(float(d) * float(d), ls, sigma) for d in ds])
        ref = np.array([oracle_lib.orc_kernel_dls(float(d), ls, sigma) for d in ds])
        assert np.max(np.abs(tab - ref)) <= 1e-6 * sigma / ls
        assert oracle_lib.orc_kernel_dls_sq(float(np.float32(ls) * np.float32(1.0001)) ** 2, ls, sigma) == 0.0
