"""The multi-limb fixed-point oracle of the Strict-Metric exhaustive fallback
(math_core/cuda_kernels.py:2237-2467; math_core/high_precision.py:581-656)
and the host-side conversion of coordinates to limbs.

Pins of the oracle:
  * limb arithmetic against Python integers: add/sub/cmp exactly; the
    truncated product against the closed form of the reference's column
    sums (sum over i+j <= L-1 of a_i b_j 2^(32(L-1-i-j)) plus the high words
    of the column i+j = L) and against the exact product, which it may fall
    below by at most L+1 units of the last limb;
  * the samples on closed forms: integer orbits on the real axis, bounded
    points, the Julia set of c = 0 where |z_k| = |z_0|^(2^k) lands on the
    bailout exactly (the 34-limb tier escapes on >= of the integer limbs, the
    arbitrary-precision tier on |z|^2 > 65536);
  * both semantics against the mpmath subsamples of
    tests/golden/mpmath_subsamples.json at the widths whose tier they serve,
    and type 4 (with exponents whose powers leave the fixed-point range)
    against an mpmath direct loop.
"""
import json
import os
import random
from decimal import Decimal, getcontext, localcontext
from fractions import Fraction

import mpmath
import numpy as np
import pytest

from fractura_amd import _abi, exhaustive
from fractura_amd.config import INV_LN2_F32, LN_BAILOUT_LOG
from oracle import oracle

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "mpmath_subsamples.json")
GPU34, HIGHPREC = _abi.AP_LIMBS_GPU34, _abi.AP_LIMBS_HIGHPREC


def to_int(limbs):
    x = 0
    for v in limbs:
        x = (x << 32) | int(v)
    return x


def from_int(x, L):
    out = np.empty(L, np.uint32)
    for i in range(L - 1, -1, -1):
        out[i] = x & 0xFFFFFFFF
        x >>= 32
    return out


def rand_limbs(rng, L, int_max=2 ** 32):
    a = np.array([rng.randrange(2 ** 32) for _ in range(L)], np.uint32)
    a[0] = rng.randrange(int_max)
    return a


@pytest.mark.parametrize("L", [2, 3, 34, 64, 65, 129, 256])
def test_mul_matches_closed_form_and_bounds(L):
    rng = random.Random(L)
    for trial in range(12):
        a, b = rand_limbs(rng, L), rand_limbs(rng, L)
        if trial == 0:
            a[:] = 0xFFFFFFFF
            b[:] = 0xFFFFFFFF
        _, out = oracle.limb_op(3, a, b)
        A, B = [int(x) for x in a], [int(x) for x in b]
        S = 0
        for i in range(L):
            for j in range(L - i):
                S += (A[i] * B[j]) << (32 * (L - 1 - i - j))
        S += sum((A[i] * B[L - i]) >> 32 for i in range(1, L))
        mod = 1 << (32 * L)
        assert to_int(out) == S % mod
        # Against the exact product: never above, at most L+1 units below.
        exact = (to_int(a) * to_int(b)) >> (32 * (L - 1))
        assert 0 <= exact - S <= L + 1


@pytest.mark.parametrize("L", [2, 34, 100])
def test_add_sub_cmp_exact(L):
    rng = random.Random(1000 + L)
    mod = 1 << (32 * L)
    for trial in range(40):
        a, b = rand_limbs(rng, L), rand_limbs(rng, L)
        if trial % 5 == 0:
            b = a.copy()
            b[-1] ^= 1 if trial % 10 == 0 else 0
        A, B = to_int(a), to_int(b)
        c, _ = oracle.limb_op(0, a, b)
        assert c == (A > B) - (A < B)
        carry, s = oracle.limb_op(1, a, b)
        assert to_int(s) == (A + B) % mod and carry == (A + B) >> (32 * L)
        hi, lo = (a, b) if A >= B else (b, a)
        _, d = oracle.limb_op(2, hi, lo)
        assert to_int(d) == abs(A - B)
        # Signed forms: value identity, and a zero result takes a's sign.
        for sa in (0, 1):
            for sb in (0, 1):
                sign, r = oracle.limb_op(4, a, b, sa, sb)
                val = (-1) ** sign * to_int(r)
                want = (-1) ** sa * A + (-1) ** sb * B
                if abs(want) < mod:
                    assert val == want
                if A == B and sa != sb:
                    assert to_int(r) == 0 and sign == sa
                sign2, r2 = oracle.limb_op(5, a, b, sa, sb)
                sign3, r3 = oracle.limb_op(4, a, b, sa, 1 - sb)
                assert sign2 == sign3 and np.array_equal(r2, r3)


def ref_decimal_to_limbs(d, L):
    """The reference's digit loop (math_core/cuda_kernels.py:2262-2274) at
    its 500-digit Decimal precision."""
    with localcontext() as ctx:
        ctx.prec = 500
        d = Decimal(str(d))
        sign = 1 if d < 0 else 0
        val = abs(d)
        ip = int(val)
        limbs = [ip]
        frac = val - Decimal(ip)
        for _ in range(1, L):
            frac = frac * Decimal(4294967296)
            lv = int(frac)
            limbs.append(lv & 0xFFFFFFFF)
            frac = frac - Decimal(lv)
    return sign, np.array(limbs, np.uint32)


def test_decimal_to_limbs_exact_and_as_reference():
    rng = random.Random(7)
    cases = ["0", "-0", "1", "-2", "0.5", "-1.75", "0.1", "3.999999999999999999999999",
             "-1.7548776662466927600495088963585286918946066177727931", "1e-300", "-2.5e-320"]
    for _ in range(30):
        digits = "".join(rng.choice("0123456789") for _ in range(rng.randrange(1, 200)))
        cases.append(("-" if rng.random() < 0.5 else "") + str(rng.randrange(4)) + "." + digits)
    for s in cases:
        for L in (3, 34, 109):
            sign, limbs = exhaustive.decimal_to_limbs(Decimal(s), L)
            q = abs(Fraction(s))
            assert to_int(limbs) == int(q * (1 << (32 * (L - 1))))
            assert sign == (1 if Decimal(s) < 0 else 0)
            rs, rl = ref_decimal_to_limbs(s, L)
            assert rs == sign and np.array_equal(rl, limbs), s


def limb_view(ftype=0, max_iter=1000, n=3, kr=0.25, ki=1.0, W=64, H=1):
    v = _abi.ApView()
    v.width, v.height, v.factor, v.max_iter = W, H, 1, max_iter
    v.fractal_type, v.gen_n, v.gen_kr, v.gen_ki = ftype, n, kr, ki
    v.inv_ln2, v.ln_bailout_log = INV_LN2_F32, LN_BAILOUT_LOG
    return v


def run_points(points, semantics, L, ftype=0, max_iter=1000, julia=(0.0, 0.0), n=3, kr=0.25,
               ki=1.0, lut=None, scheme=0):
    """Evaluate a list of complex points (Decimal strings or numbers) with the
    limb oracle; returns Maps of a 1 x len(points) frame."""
    v = limb_view(ftype, max_iter, n, kr, ki, W=len(points))
    v.scheme_id = scheme
    coords = np.empty((len(points), 2, L), np.uint32)
    signs = np.empty((len(points), 2), np.int32)
    for i, (x, y) in enumerate(points):
        signs[i, 0], coords[i, 0] = exhaustive.decimal_to_limbs(Decimal(str(x)), L)
        signs[i, 1], coords[i, 1] = exhaustive.decimal_to_limbs(Decimal(str(y)), L)
    params, psigns = exhaustive._param_limbs(julia[0], julia[1], kr, ki, L)
    return oracle.render_limbs(v, semantics, L, np.arange(len(points)), coords, signs, params,
                               psigns, lut=lut)


def test_real_axis_integer_orbits():
    # c = 1: 0, 1, 2, 5, 26, 677 -> |z_5|^2 = 458329 escapes at k = 5, both
    # semantics; the float32 value equals the FP64 loop's (same |z|^2).
    v64 = limb_view(0, 1000)
    k64, s64, sm64 = oracle.fp64_sample(v64, 1.0, 0.0)
    for sem, L in ((GPU34, 34), (HIGHPREC, 5)):
        m = run_points([(1, 0), (0.5, 0), (2, 0)], sem, L)
        assert m.counts[0, 0] == 5 == k64
        assert m.iter_map[0, 0] == np.float32(s64)
        assert abs(m.smooth[0, 0] - sm64) < 1e-12
        # c = 1/2: 0, .5, .75, 1.0625, ... escapes at k = 8 (exact rationals).
        z, k = Fraction(0), 0
        while z * z <= 65536:
            z = z * z + Fraction(1, 2)
            k += 1
        assert m.counts[0, 1] == k
        assert m.counts[0, 2] == 4  # 0, 2, 6, 38, 1446: |z_4|^2 > 65536


@pytest.mark.parametrize("sem,L", [(GPU34, 34), (HIGHPREC, 12)])
def test_bounded_points_are_interior(sem, L):
    m = run_points([(-2, 0), (0, 1), (-1, 0), (0, 0), (-0.75, 0)], sem, L, max_iter=300)
    assert np.all(m.counts[0] == 300)
    assert np.all(m.smooth[0] == -1.0)
    want = -1.0 if sem == GPU34 else 300.0  # the CPU tier stores float(max_iter)
    assert np.all(m.iter_map[0] == np.float32(want))


@pytest.mark.parametrize("ftype", [2, 3])
def test_julia_zero_bailout_boundary(ftype):
    # c = 0: |z_k| = 2^(2^k); |z_3|^2 = 65536 exactly.  The integer-limb test
    # (>= 65536) escapes at k = 3, the strict |z|^2 > 65536 at k = 4.
    pts = [(2, 0), (0, -2), (Decimal(2).sqrt(), Decimal(2).sqrt())]
    g = run_points(pts, GPU34, 34, ftype=ftype, julia=(0.0, 0.0))
    h = run_points(pts[:2], HIGHPREC, 8, ftype=ftype, julia=(0.0, 0.0))
    assert g.counts[0, 0] == 3 and g.counts[0, 1] == 3
    assert h.counts[0, 0] == 4 and h.counts[0, 1] == 4
    # sqrt 2 (1 + i) is inexact: |z_3|^2 lies just below 65536 on both.
    assert g.counts[0, 2] == 4


def test_gpu34_does_not_evaluate_type4():
    m = run_points([(0.1, 0.2)], GPU34, 34, ftype=4)
    assert m.counts[0, 0] == -1 and m.iter_map[0, 0] == -1.0


def mp_escape(c, ftype, max_iter, n=3, kr=0.25, ki=1.0, julia=(0, 0), dps=60):
    with mpmath.workdps(dps):
        cx, cy = mpmath.mpf(str(c[0])), mpmath.mpf(str(c[1]))
        k_r, k_i = mpmath.mpf(repr(float(kr))), mpmath.mpf(repr(float(ki)))
        zx, zy = mpmath.mpf(0), mpmath.mpf(0)
        for k in range(max_iter):
            m2 = zx * zx + zy * zy
            if m2 > 65536:
                lnz = mpmath.log(m2) / 2
                return k, (k + 1) - float(mpmath.log(lnz / mpmath.log(256))) / float(mpmath.log(n if ftype == 4 else 2))
            if ftype == 4:
                pr, pi = zx, zy
                for _ in range(n - 1):
                    pr, pi = pr * zx - pi * zy, pr * zy + pi * zx
                zx, zy = pr + k_r * zx - k_i * zy + cx, pi + k_r * zy + k_i * zx + cy
            else:
                zx, zy = zx * zx - zy * zy + cx, 2 * zx * zy + cy
        return max_iter, -1.0


@pytest.mark.parametrize("n,kr,ki", [(2, 0.0, 0.0), (3, 0.25, 1.0), (5, -0.2, 0.1), (8, 0.05, -0.3),
                                     (17, 0.0, 0.5)])
def test_highprec_type4_against_mpmath(n, kr, ki):
    # Exponents 5, 8 and 17 take the float64 path once |z^m| reaches 2^22.
    rng = np.random.Generator(np.random.PCG64(n))
    pts = [(Decimal(f"{x:.12f}"), Decimal(f"{y:.12f}")) for x, y in rng.uniform(-1.3, 1.3, (48, 2))]
    m = run_points(pts, HIGHPREC, 6, ftype=4, max_iter=150, n=n, kr=kr, ki=ki)
    agree, errs = 0, []
    for i, p in enumerate(pts):
        k, sm = mp_escape(p, 4, 150, n, kr, ki)
        if k == m.counts[0, i]:
            agree += 1
            if k < 150:
                errs.append(abs(sm - m.smooth[0, i]))
    assert agree == len(pts)
    assert max(errs) < 1e-9


def fixture(name):
    with open(FIXTURE) as f:
        return {v["name"]: v for v in json.load(f)["views"]}[name]


def fixture_points(view, count):
    """Limbs of the fixture view's first `count` sample pixels."""
    xs, ys = np.array(view["x"][:count]), np.array(view["y"][:count])
    return xs, ys


def run_fixture(name, semantics, L=None, count=128, threads=8):
    view = fixture(name)
    W, H = view["W"], view["H"]
    w = Decimal(view["width"])
    with localcontext() as ctx:
        ctx.prec = 1200
        h = w * Decimal(H) / Decimal(W)
    L = L or (exhaustive.ML_LIMBS if semantics == GPU34 else exhaustive.highprec_limbs(w))
    xs, ys = fixture_points(view, count)
    coords, signs = exhaustive.pixel_coordinate_limbs(ys, xs, W, H, Decimal(view["cx"]),
                                                      Decimal(view["cy"]), w, h, L)
    n, kr, ki = view.get("gen", (3, 0.25, 1.0))
    params, psigns = exhaustive._param_limbs(0.0, 0.0, kr, ki, L)
    v = limb_view(view["type"], view["max_iter"], n, kr, ki, W=W, H=H)
    m = oracle.render_limbs(v, semantics, L, ys * W + xs, coords, signs, params, psigns,
                            threads=threads)
    got_c = m.counts[ys, xs]
    got_s = m.smooth[ys, xs]
    ref_c = np.array(view["counts"][:count])
    ref_s = np.array(view["smooth"][:count])
    return got_c, got_s, ref_c, ref_s, m, (xs, ys)


@pytest.mark.parametrize("name,semantics", [
    ("tier2_misiurewicz_i_1e-250", GPU34), ("tier2_bship_tip_1e-250", GPU34),
    ("tier2_misiurewicz_i_1e-250", HIGHPREC), ("config4", GPU34),
    ("deep_misiurewicz_i_1e-320", HIGHPREC), ("config5_frame12", HIGHPREC),
    ("config5_frame31", HIGHPREC)])
def test_limb_oracle_against_mpmath(name, semantics):
    got_c, got_s, ref_c, ref_s, _, _ = run_fixture(name, semantics)
    assert np.array_equal(got_c, ref_c)
    esc = ref_c < fixture(name)["max_iter"]
    assert np.max(np.abs(got_s[esc] - ref_s[esc])) < 1e-9


def test_coordinate_limbs_are_the_exact_grid():
    view = fixture("tier2_misiurewicz_i_1e-250")
    W, H, L = view["W"], view["H"], 34
    w = Decimal(view["width"])
    xs, ys = np.array([0, 17, W - 1]), np.array([H - 1, 5, 0])
    with localcontext() as ctx:
        ctx.prec = 1200
        h = w * Decimal(H) / Decimal(W)
        coords, signs = exhaustive.pixel_coordinate_limbs(ys, xs, W, H, Decimal(0), Decimal(1), w, h, L)
    for i, (x, y) in enumerate(zip(xs, ys)):
        ex = Fraction(0) - Fraction(w) / 2 + x * Fraction(w) / (W - 1)
        ey = Fraction(1) + Fraction(w) * H / W / 2 - y * Fraction(w) * H / W / (H - 1)
        for comp, e in ((0, ex), (1, ey)):
            got = Fraction(to_int(coords[i, comp]), 1 << (32 * (L - 1))) * (-1 if signs[i, comp] else 1)
            assert abs(got - e) <= Fraction(1, 1 << (32 * (L - 1)))
