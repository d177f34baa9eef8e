"""Host-side arbitrary-precision reference orbits (input to the GPU path).

The reference computes Z_k in MPFR (gmpy2) or Decimal and stores float(Z_k)
(math_core/high_precision.py:706-875).  Here the orbit runs in binary fixed
point (the library's host routine ap_reference_orbit, C++): Z = X / 2^P with
P = precision bits + 64 guard bits, products truncated by a right shift.
float(Z) is correctly rounded.  The algorithm (storage until |float Z|^2 > 65536,
period detection on exact equality at doubling checkpoints for types 0/2,
Z_ref_len stored after the loop) follows the reference; only the rounding of
the big-number arithmetic differs, far below the stored double's precision.
"""
import os
import threading
from concurrent.futures import ThreadPoolExecutor
from decimal import Decimal

import numpy as np

from .config import compute_required_prec_bits, decimal_context

GUARD_BITS = 64


def to_dd(dec_val):
    """Decimal -> double-double (hi, lo) (math_core/high_precision.py:17-21)."""
    d = Decimal(dec_val)
    hi = float(d)
    return hi, float(d - Decimal(hi))


def _fixed(dec, P):
    """Round a Decimal (or str/float) to a fixed-point integer with P fraction bits."""
    d = Decimal(str(dec)) if not isinstance(dec, Decimal) else dec
    # Exact: scale by 2^P as an integer, then round half-even.
    sign, digits, exp = d.as_tuple()
    n = int("".join(map(str, digits))) if digits else 0
    if exp >= 0:
        v = (n * 10 ** exp) << P
    else:
        num = n << P
        den = 10 ** (-exp)
        q, r = divmod(num, den)
        if 2 * r > den or (2 * r == den and q & 1):
            q += 1
        v = q
    return -v if sign else v


def _to_float(X, P):
    return X / (1 << P)


def _limbs_of(values):
    """Six fixed-point integers as (6, limbs) little-endian uint64 magnitudes
    and int32 signs."""
    nbits = max(max(abs(v).bit_length() for v in values), 1)
    limbs = nbits // 64 + 1
    mags = np.zeros((len(values), limbs), dtype=np.uint64)
    signs = np.zeros(len(values), dtype=np.int32)
    for i, v in enumerate(values):
        signs[i] = 1 if v < 0 else 0
        mags[i] = np.frombuffer(abs(v).to_bytes(limbs * 8, "little"), dtype=np.uint64)
    return mags, signs, limbs


def compute_single_highprec_orbit(cand_cx_dec, cand_cy_dec, max_iter, fractal_type=0,
                                  julia_cx=0.0, julia_cy=0.0, prec_bits=256,
                                  gen_n=3, gen_kr=0.25, gen_ki=1.0):
    """Reference orbit float64 arrays (re, im) of length max_iter+1 and its
    escape index ref_len (math_core/high_precision.py:706-875), iterated in
    binary fixed point with P = prec_bits + 64 guard bits by the library's
    host routine ap_reference_orbit."""
    import ctypes as C

    from . import _abi

    P = int(prec_bits) + GUARD_BITS
    max_iter = int(max_iter)
    fractal_type = int(fractal_type)
    if fractal_type in (2, 3):
        X, Y = _fixed(cand_cx_dec, P), _fixed(cand_cy_dec, P)
        CX, CY = _fixed(julia_cx, P), _fixed(julia_cy, P)
    else:
        X, Y = 0, 0
        CX, CY = _fixed(cand_cx_dec, P), _fixed(cand_cy_dec, P)
    KR, KI = (_fixed(gen_kr, P), _fixed(gen_ki, P)) if fractal_type == 4 else (0, 0)
    mags, signs, limbs = _limbs_of([X, Y, CX, CY, KR, KI])
    if limbs > _abi.AP_MAX_ORBIT_LIMBS:
        raise ValueError(f"orbit precision of {limbs * 64} bits exceeds the library's limit")
    ref_re = np.zeros(max_iter + 1, dtype=np.float64)
    ref_im = np.zeros(max_iter + 1, dtype=np.float64)
    rl = C.c_int32(0)
    lib = _abi.load()
    dp = C.POINTER(C.c_double)
    st = lib.ap_reference_orbit(fractal_type, max_iter, P, limbs,
                                mags.ctypes.data_as(C.POINTER(C.c_uint64)),
                                signs.ctypes.data_as(C.POINTER(C.c_int32)),
                                int(gen_n) if fractal_type == 4 else 2,
                                ref_re.ctypes.data_as(dp), ref_im.ctypes.data_as(dp), C.byref(rl))
    if st != _abi.AP_OK:
        raise ValueError(f"ap_reference_orbit rejected type={fractal_type} max_iter={max_iter} "
                         f"P={P} gen_n={gen_n} (status {st})")
    return ref_re, ref_im, int(rl.value)


def compute_highprec_escape_scalar(cx_dec, cy_dec, max_iter, fractal_type=0, julia_cx=0.0,
                                   julia_cy=0.0, prec_bits=256, gen_n=3, gen_kr=0.25, gen_ki=1.0):
    """Escape index of one point at high precision (max_iter if it stays
    bounded), the probe of math_core/high_precision.py:581-703."""
    _, _, n = compute_single_highprec_orbit(cx_dec, cy_dec, max_iter, fractal_type, julia_cx,
                                            julia_cy, prec_bits, gen_n, gen_kr, gen_ki)
    return float(n)


_CACHE = []
_CACHE_LOCK = threading.Lock()
_CACHE_SIZE = 16

_PROBE_UV = [
    (0.35, 0.35), (0.65, 0.35), (0.35, 0.65), (0.65, 0.65),
    (0.5, 0.25), (0.5, 0.75), (0.25, 0.5), (0.75, 0.5),
    (0.2, 0.2), (0.8, 0.2), (0.2, 0.8), (0.8, 0.8),
    (0.5, 0.35), (0.5, 0.65), (0.35, 0.5), (0.65, 0.5),
]


def compute_bbsa_coeffs(ref_re, ref_im, ref_len, max_r, fractal_type, bbsa_tol=1e-4, bbsa_order=4,
                        gen_n=3, gen_kr=0.25, gen_ki=1.0, E_scale=0):
    """(skip, A_re, A_im, ..., H_re, H_im): the series approximation that
    replaces the first `skip` steps of the FP64-delta perturbation loop, in
    u = dc / max_r (compute_bbsa_coeffs, math_core/high_precision.py:537-575).
    Host arithmetic in libapeiron_hip.so (ap_bbsa_coeffs); no device work."""
    from . import _abi

    ref_len = int(ref_len)
    out = np.zeros(17, dtype=np.float64)
    re = np.ascontiguousarray(ref_re, dtype=np.float64)
    im = np.ascontiguousarray(ref_im, dtype=np.float64)
    if ref_len > 0 and (len(re) < ref_len or len(im) < ref_len):
        raise ValueError("orbit shorter than ref_len")
    lib = _abi.load()
    st = lib.ap_bbsa_coeffs(re.ctypes.data, im.ctypes.data, ref_len, float(max_r),
                            int(fractal_type), float(bbsa_tol), int(bbsa_order), int(gen_n),
                            float(gen_kr), float(gen_ki), int(E_scale), out.ctypes.data)
    if st != _abi.AP_OK:
        raise _abi.ApError(st, "ap_bbsa_coeffs", "bad arguments")
    return (int(out[0]),) + tuple(float(x) for x in out[1:])


def clear_reference_orbit_cache():
    with _CACHE_LOCK:
        _CACHE.clear()


def cache_reference_orbit(cx_dec, cy_dec, pw_dec, max_iter, ref_re, ref_im, ref_len,
                          fractal_type=0, julia_cx=0.0, julia_cy=0.0, gen_n=3, gen_kr=0.25,
                          gen_ki=1.0, is_floatexp=False):
    """Insert an orbit at the front of the 16-entry LRU list
    (math_core/high_precision.py:893-916)."""
    with _CACHE_LOCK:
        _CACHE.insert(0, dict(is_floatexp=is_floatexp, fractal_type=fractal_type,
                              julia_cx=julia_cx, julia_cy=julia_cy, gen_n=gen_n, gen_kr=gen_kr,
                              gen_ki=gen_ki, pw_dec=pw_dec, cx_dec=cx_dec, cy_dec=cy_dec,
                              ref_len=ref_len, max_iter=max_iter, ref_re=ref_re, ref_im=ref_im))
        if len(_CACHE) > _CACHE_SIZE:
            _CACHE.pop()


def find_optimal_reference_orbit(cx_dec, cy_dec, pw_dec, ph_dec, max_iter, screen_w, screen_h,
                                 fractal_type=0, julia_cx=0.0, julia_cy=0.0, gen_n=3,
                                 gen_kr=0.25, gen_ki=1.0, is_floatexp=False, probe=True):
    """Reference point and orbit for a view: cache reuse, then the centre, then
    (if the centre escapes before 60% of max_iter) the best of 16 probes
    (math_core/high_precision.py:919-1097).  Returns
    (cx_dec, cy_dec, u, v, ref_re, ref_im, ref_len) with (u, v) the reference's
    position as a fraction of the view."""
    with decimal_context(pw_dec):
        x_min = cx_dec - pw_dec / Decimal(2)
        y_max = cy_dec + ph_dec / Decimal(2)
        best_entry, best_score = None, -1
        with _CACHE_LOCK:
            for e in _CACHE:
                if e["is_floatexp"] != is_floatexp:
                    continue
                if (e["fractal_type"], e["julia_cx"], e["julia_cy"], e["gen_n"], e["gen_kr"],
                        e["gen_ki"]) != (fractal_type, julia_cx, julia_cy, gen_n, gen_kr, gen_ki):
                    continue
                ratio = pw_dec / e["pw_dec"]
                if not (Decimal("0.25") <= ratio <= Decimal("4.0")):
                    continue
                cu = float((e["cx_dec"] - x_min) / pw_dec)
                cv = float((y_max - e["cy_dec"]) / ph_dec)
                if 0.10 <= cu <= 0.90 and 0.10 <= cv <= 0.90 and e["ref_len"] > best_score:
                    best_score, best_entry = e["ref_len"], e
                    if e["ref_len"] >= max_iter:
                        break
        can_reuse = False
        if best_entry is not None:
            natural = best_entry["ref_len"] < best_entry.get("max_iter", best_entry["ref_len"])
            ratio = pw_dec / best_entry["pw_dec"]
            panning = Decimal("0.75") <= ratio <= Decimal("1.33")
            full_or_natural = best_entry["ref_len"] >= max_iter or natural
            if best_score >= max_iter:
                can_reuse = True
            elif panning and best_score >= 100 and full_or_natural:
                can_reuse = True
            elif natural and best_score >= int(max_iter * 0.35):
                can_reuse = True
            elif (best_entry["cx_dec"], best_entry["cy_dec"], best_entry["pw_dec"]) == (cx_dec, cy_dec, pw_dec) \
                    and full_or_natural:
                can_reuse = True

        def reuse():
            e = best_entry
            u = float((e["cx_dec"] - x_min) / pw_dec)
            v = float((y_max - e["cy_dec"]) / ph_dec)
            return e["cx_dec"], e["cy_dec"], u, v, e["ref_re"], e["ref_im"], min(e["ref_len"], max_iter)

        if can_reuse:
            return reuse()

        bits = compute_required_prec_bits(pw_dec)
        ref_re, ref_im, ref_len = compute_single_highprec_orbit(
            cx_dec, cy_dec, max_iter, fractal_type, julia_cx, julia_cy, bits, gen_n, gen_kr, gen_ki)
        best = (cx_dec, cy_dec, 0.5, 0.5)
        if probe and ref_len < int(max_iter * 0.6) and pw_dec >= Decimal("1e-10000"):
            # The 16 probe orbits run in threads, as the reference's do: the
            # library's orbit routine runs without the interpreter lock.  The
            # points are formed here, under this thread's Decimal context.
            pts = [(x_min + Decimal(str(u)) * pw_dec, y_max - Decimal(str(v)) * ph_dec, u, v)
                   for u, v in _PROBE_UV]

            def escape_of(pt):
                return compute_highprec_escape_scalar(pt[0], pt[1], max_iter, fractal_type,
                                                      julia_cx, julia_cy, bits, gen_n, gen_kr,
                                                      gen_ki)

            workers = max(1, min(len(pts), os.cpu_count() or 1))
            with ThreadPoolExecutor(max_workers=workers) as ex:
                escs = list(ex.map(escape_of, pts))
            results = [(esc, px, py, u, v) for esc, (px, py, u, v) in zip(escs, pts)]
            best_esc = max(int(ref_len * 1.2), ref_len + 32)
            cand = None
            for esc, px, py, u, v in results:
                if esc > best_esc:
                    best_esc, cand = esc, (px, py, u, v)
            if cand is not None:
                c_re, c_im, c_len = compute_single_highprec_orbit(
                    cand[0], cand[1], max_iter, fractal_type, julia_cx, julia_cy, bits, gen_n,
                    gen_kr, gen_ki)
                if c_len > ref_len:
                    ref_re, ref_im, ref_len = c_re, c_im, c_len
                    best = cand
        if can_reuse and best_entry is not None and best_entry["ref_len"] > ref_len:
            return reuse()
        cache_reference_orbit(best[0], best[1], pw_dec, max_iter, ref_re, ref_im, ref_len,
                              fractal_type, julia_cx, julia_cy, gen_n, gen_kr, gen_ki, is_floatexp)
        return best[0], best[1], best[2], best[3], ref_re, ref_im, ref_len
