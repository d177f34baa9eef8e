"""The host reference orbit: ap_reference_orbit, the library's C++ routine
behind compute_single_highprec_orbit (math_core/high_precision.py:706-875).

Pins:
  * the native routine equals the Python fixed-point restatement
    (oracle.highprec_orbit) bit for bit (stored doubles, ref_len, the periodic
    fill) on seeded points of every type, at 320 to 3513 fraction bits, and at
    2500 to 33300 bits where the products split (Karatsuba),
    including values that round to subnormal doubles and orbits that escape
    on the first steps;
  * the restatement, and with it the routine, against mpmath at the same
    binary precision: escape index and stored doubles (within one ulp) of
    seeded orbits;
  * period detection on exact cycles: c = -1 (period 2), c = i (preperiodic,
    then period 2), c = -2 (fixed point 2 after two steps), c = 0, and the
    fill of the rest of the arrays from the cycle;
  * the argument checks of the C entry point.
"""
import ctypes as C
from decimal import Decimal

import mpmath as mp
import numpy as np
import pytest

from fractura_amd import _abi
from fractura_amd.high_precision import GUARD_BITS, _fixed, compute_single_highprec_orbit
from oracle import oracle


def _restated(cx, cy, max_iter, ftype, jx=0.0, jy=0.0, bits=256, n=3, kr=0.25, ki=1.0):
    P = bits + GUARD_BITS
    if ftype in (2, 3):
        X, Y, CX, CY = _fixed(cx, P), _fixed(cy, P), _fixed(jx, P), _fixed(jy, P)
    else:
        X, Y, CX, CY = 0, 0, _fixed(cx, P), _fixed(cy, P)
    KR, KI = (_fixed(kr, P), _fixed(ki, P)) if ftype == 4 else (0, 0)
    return oracle.highprec_orbit(X, Y, CX, CY, P, max_iter, ftype, n, KR, KI)


def _same(a, b):
    ra, ia, na = a
    rb, ib, nb = b
    assert na == nb
    assert np.array_equal(ra.view(np.uint64), rb.view(np.uint64))
    assert np.array_equal(ia.view(np.uint64), ib.view(np.uint64))


def _points(seed, count, scale):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        x, y = rng.uniform(-2.0, 1.0), rng.uniform(-1.3, 1.3)
        # a double plus a perturbation at the given scale (exact in Decimal)
        dx, dy = rng.standard_normal(2) * scale
        out.append((Decimal(repr(float(x))) + Decimal(repr(float(dx))),
                    Decimal(repr(float(y))) + Decimal(repr(float(dy)))))
    return out


@pytest.mark.parametrize("ftype", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("bits", [256, 1024, 3449])
def test_native_equals_restatement(ftype, bits):
    pts = _points(100 * ftype + bits, 6, 1e-30)
    for cx, cy in pts:
        kw = dict(jx=-0.75, jy=0.11) if ftype in (2, 3) else {}
        for n in ((3, 2, 5) if ftype == 4 else (3,)):
            want = _restated(cx, cy, 400, ftype, bits=bits, n=n, **kw)
            got = compute_single_highprec_orbit(cx, cy, 400, ftype, kw.get("jx", 0.0),
                                                kw.get("jy", 0.0), bits, n, 0.25, 1.0)
            _same(got, want)


def test_native_equals_restatement_on_long_and_tiny_orbits():
    # A deep-zoom centre on the cardioid boundary (long, bounded orbit).
    cx = Decimal("-0.7436438870371587047521915061147745")
    cy = Decimal("0.1318259042053119704931320563851390")
    _same(compute_single_highprec_orbit(cx, cy, 3000, 0, prec_bits=3449),
          _restated(cx, cy, 3000, 0, bits=3449))
    # Values below 2^-1022: the stored doubles are subnormal or zero.
    for c in ("1e-310", "-3.3e-320", "4.9e-324", "2e-330"):
        d = Decimal(c)
        got = compute_single_highprec_orbit(d, d, 10, 0, prec_bits=1400)
        _same(got, _restated(d, d, 10, 0, bits=1400))
        assert got[0][1] == float(d)  # correctly rounded, as float(Decimal) is
    # Escape at the first steps, and a start outside the bailout (Julia).
    _same(compute_single_highprec_orbit(Decimal(3), Decimal(0), 50, 0, prec_bits=256),
          _restated(Decimal(3), Decimal(0), 50, 0))
    _same(compute_single_highprec_orbit(Decimal(300), Decimal(1), 50, 2, 0.1, 0.2, 256),
          _restated(Decimal(300), Decimal(1), 50, 2, jx=0.1, jy=0.2))


@pytest.mark.parametrize("c,period_from", [((-1, 0), 2), ((0, 1), 2), ((-2, 0), 1), ((0, 0), 1)])
def test_period_detection_on_exact_cycles(c, period_from):
    cx, cy = Decimal(c[0]), Decimal(c[1])
    re, im, n = compute_single_highprec_orbit(cx, cy, 100, 0, prec_bits=256)
    assert n == 100  # a detected period reports max_iter
    _same((re, im, n), _restated(cx, cy, 100, 0))
    # The fill repeats the cycle to the end of the arrays.
    z = [complex(0, 0)]
    for _ in range(100):
        z.append(z[-1] ** 2 + complex(c[0], c[1]))
    assert np.array_equal(re, np.array([v.real for v in z]))
    assert np.array_equal(im, np.array([v.imag for v in z]))


def _mp_orbit(cx, cy, max_iter, ftype, prec, jx=0.0, jy=0.0, n=3, kr=0.25, ki=1.0):
    mp.mp.prec = prec
    c = mp.mpc(mp.mpf(str(cx)), mp.mpf(str(cy)))
    if ftype in (2, 3):
        z, c = c, mp.mpc(mp.mpf(repr(jx)), mp.mpf(repr(jy)))
    else:
        z = mp.mpc(0)
    k_ = mp.mpc(mp.mpf(repr(kr)), mp.mpf(repr(ki)))
    out = []
    for k in range(max_iter):
        rx, ry = float(z.real), float(z.imag)
        out.append((rx, ry))
        if rx * rx + ry * ry > 65536.0:
            return out, k
        if ftype in (0, 2):
            z = z * z + c
        elif ftype in (1, 3):
            x, y = z.real, z.imag
            z = mp.mpc(x * x - y * y + c.real, 2 * abs(x) * abs(y) - c.imag)
        else:
            z = z ** n + k_ * z + c
    return out, max_iter


@pytest.mark.parametrize("ftype", [0, 1, 2, 4])
def test_restatement_against_mpmath(ftype):
    """Stored doubles within one ulp of mpmath's float(Z_k), same escape index
    (non-periodic seeded orbits of a few hundred steps at 320 bits)."""
    pts = _points(7 + ftype, 8, 1e-20)
    # Long escapes: the neck at -3/4 (about pi/0.02 steps) and a point near the
    # Julia set of -0.75 + 0.11i.
    pts += [(Decimal("-0.75"), Decimal("0.02")), (Decimal("0.3"), Decimal("0.5"))]
    compared = 0
    for cx, cy in pts:
        kw = dict(jx=-0.75, jy=0.11) if ftype in (2, 3) else {}
        re, im, n = _restated(cx, cy, 300, ftype, bits=256, **kw)
        if n == 300 and ftype in (0, 2):
            continue  # may have hit the period rule, which mpmath does not model
        ref, n_ref = _mp_orbit(cx, cy, 300, ftype, 320, **kw)
        assert n == n_ref
        for k, (rx, ry) in enumerate(ref[: n + 1]):
            assert abs(re[k] - rx) <= np.spacing(abs(rx)), (k, re[k], rx)
            assert abs(im[k] - ry) <= np.spacing(abs(ry)), (k, im[k], ry)
        compared += n + 1
    assert compared >= 100, compared


def test_entry_point_rejects_bad_arguments():
    lib = _abi.load()
    re = np.zeros(11)
    im = np.zeros(11)
    init = np.zeros((6, 1), dtype=np.uint64)
    signs = np.zeros(6, dtype=np.int32)
    rl = C.c_int32()
    dp, up, ip = C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_int32)
    args = lambda t, lim, n: (t, 10, 320, lim, init.ctypes.data_as(up), signs.ctypes.data_as(ip), n,
                              re.ctypes.data_as(dp), im.ctypes.data_as(dp), C.byref(rl))
    assert lib.ap_reference_orbit(*args(0, 1, 2)) == _abi.AP_OK
    assert lib.ap_reference_orbit(*args(5, 1, 2)) == _abi.AP_ERR_ARG       # type
    assert lib.ap_reference_orbit(*args(4, 1, 17)) == _abi.AP_ERR_ARG      # n
    assert lib.ap_reference_orbit(*args(0, 0, 2)) == _abi.AP_ERR_ARG       # limbs
    assert lib.ap_reference_orbit(*args(0, _abi.AP_MAX_ORBIT_LIMBS + 1, 2)) == _abi.AP_ERR_ARG
    assert lib.ap_reference_orbit(0, 10, 320, 1, None, None, 2, None, None, None) == _abi.AP_ERR_ARG


@pytest.mark.parametrize("max_iter", [0, 1, 2, 8, 9, 17])
def test_short_orbits_every_type(max_iter):
    """max_iter at and around the first period checkpoints (k = 8, 16)."""
    for ftype in range(5):
        kw = dict(jx=0.3, jy=-0.2) if ftype in (2, 3) else {}
        got = compute_single_highprec_orbit(Decimal("-0.1"), Decimal("0.65"), max_iter, ftype,
                                            kw.get("jx", 0.0), kw.get("jy", 0.0), 256)
        _same(got, _restated(Decimal("-0.1"), Decimal("0.65"), max_iter, ftype, **kw))


@pytest.mark.parametrize("bits", [2500, 10000, 33300])
def test_wide_precisions_split_products(bits):
    """Above 40 limbs the products split (Karatsuba, up to four levels at
    33300 bits): still the restatement's orbit, bit for bit."""
    cx = Decimal("-0.7436438870371587047521915061147745")
    cy = Decimal("0.1318259042053119704931320563851390")
    for ftype in (0, 1, 4):
        got = compute_single_highprec_orbit(cx, cy, 60, ftype, prec_bits=bits)
        _same(got, _restated(cx, cy, 60, ftype, bits=bits))
