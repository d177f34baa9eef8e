"""Write tests/golden/configs.json: the seeded inputs of the five benchmark
configurations (SEED = 20261203, numpy PCG64).

  config 2: theta = 2 pi u, u the first 128-bit draw / 2^128;
            c = e^{i theta}/2 - e^{2 i theta}/4 at 60 digits (mpmath).
  config 3: the same construction from the second 128-bit draw, 1100 digits.
  config 4: centre uniform in [-1.8, -1.7] x [-0.1, 0.0] as 80-digit decimals.
  config 5: 64 frames of z^3 + k z + c (its own stream, SEED + 5): k uniform
            in the disc |k| < 0.3 (r = 0.3 sqrt(U), phi = 2 pi V); the target
            centre is a Misiurewicz point of that family found by a seeded
            search (0 strictly pre-periodic onto a repelling cycle, so its
            orbit never escapes), refined by Newton's method at 260 digits and
            verified there, then written at 220 digits; widths
            w_f = 4 (1e-200/4)^(f/63); centres by the reference's zoom-in focal
            tracking from (0, 0) (workers/video_workers.py:67-80); tile sizes
            drawn from {256x256, 640x360, 1280x720, 1920x1080};
            max_iter_f = round(10^(3 + 2 log10(4/w_f) / log10(4/1e-200))).
Run: python tests/golden/make_configs.py
"""
import json
import os

import mpmath
import numpy as np

SEED = 20261203
HERE = os.path.dirname(os.path.abspath(__file__))


def u128(rng):
    a, b = (int(x) for x in rng.integers(0, 2 ** 64, 2, dtype=np.uint64))
    return (a << 64) | b


def cardioid_point(u, digits):
    with mpmath.workdps(digits + 20):
        th = 2 * mpmath.pi * mpmath.mpf(u) / mpmath.mpf(2) ** 128
        c = mpmath.expj(th) / 2 - mpmath.expj(2 * th) / 4
        return (mpmath.nstr(c.real, digits, strip_zeros=False, min_fixed=-mpmath.inf, max_fixed=mpmath.inf),
                mpmath.nstr(c.imag, digits, strip_zeros=False, min_fixed=-mpmath.inf, max_fixed=mpmath.inf))


def decimal_uniform(rng, lo_int_str, digits):
    d = "".join(str(int(x)) for x in rng.integers(0, 10, digits))
    return d


def misiurewicz_point(k, rng, dps=60):
    """Seeded search for c with f_c^(m+p)(0) = f_c^m(0), f_c(z) = z^3 + k z + c,
    m the exact pre-period and p the exact period of a repelling cycle.  k is
    taken as the decimal value of its shortest float repr, as the reference's
    high-precision orbit takes it (mpfr(str(gen_kr)),
    math_core/high_precision.py:728-729)."""

    def orbit(c, n):
        z, zs = mpmath.mpc(0), [mpmath.mpc(0)]
        for _ in range(n):
            z = z ** 3 + kk * z + c
            zs.append(z)
        return zs

    with mpmath.workdps(dps):
        kk = mpmath.mpc(mpmath.mpf(repr(k.real)), mpmath.mpf(repr(k.imag)))
        while True:
            m, p = int(rng.integers(2, 5)), int(rng.integers(1, 4))
            c0 = mpmath.mpc(rng.uniform(-1.2, 1.2), rng.uniform(-1.2, 1.2))
            try:
                c = mpmath.findroot(lambda c: orbit(c, m + p)[m + p] - orbit(c, m + p)[m], c0,
                                    tol=mpmath.mpf(10) ** (5 - dps), maxsteps=200)
            except (ValueError, ZeroDivisionError):
                continue
            zs = orbit(c, m + p)
            tiny = mpmath.mpf(10) ** -20
            if abs(c) > 1.6 or abs(zs[m - 1 + p] - zs[m - 1]) < tiny:
                continue  # outside the view, or a shorter pre-period
            if min(abs(zs[j]) for j in range(m, m + p)) < tiny:
                continue  # 0 on the cycle: periodic, not pre-periodic
            if any(abs(zs[m + q] - zs[m]) < tiny for q in range(1, p)):
                continue  # period shorter than p
            mult = mpmath.mpf(1)
            for j in range(m, m + p):
                mult *= abs(3 * zs[j] ** 2 + kk)
            if mult <= 1:
                continue
            break
    # Newton at 260 digits, then verify the relation there.
    with mpmath.workdps(260):
        kk = mpmath.mpc(mpmath.mpf(repr(k.real)), mpmath.mpf(repr(k.imag)))
        c = mpmath.findroot(lambda c: orbit(c, m + p)[m + p] - orbit(c, m + p)[m], mpmath.mpc(c),
                            tol=mpmath.mpf(10) ** -250, maxsteps=100)
        zs = orbit(c, m + p)
        residual = abs(zs[m + p] - zs[m])
        assert residual < mpmath.mpf(10) ** -240, residual
        return c, m, p, float(mult), float(mpmath.log10(residual))


def config5():
    rng = np.random.Generator(np.random.PCG64(SEED + 5))
    U, V = rng.random(2)
    r, phi = 0.3 * np.sqrt(U), 2 * np.pi * V
    k = complex(r * np.cos(phi), r * np.sin(phi))
    c, m, p, mult, log_res = misiurewicz_point(k, rng)
    digits = 220
    fmt = dict(strip_zeros=False, min_fixed=-mpmath.inf, max_fixed=mpmath.inf)
    with mpmath.workdps(digits + 40):
        tx = mpmath.mpf(mpmath.nstr(c.real, digits, **fmt))
        ty = mpmath.mpf(mpmath.nstr(c.imag, digits, **fmt))
        w0, w1 = mpmath.mpf(4), mpmath.mpf("1e-200")
        sizes = [(256, 256), (640, 360), (1280, 720), (1920, 1080)]
        picks = rng.integers(0, 4, 64)
        frames = []
        for f in range(64):
            w = w0 * (w1 / w0) ** (mpmath.mpf(f) / 63)
            # Zoom-in focal tracking (workers/video_workers.py:67-72): the centre
            # moves from (0, 0) to the target in step with the width.
            ratio = (w - w1) / (w0 - w1)
            cx, cy = tx - ratio * tx, ty - ratio * ty
            mi = int(round(10 ** (3 + 2 * float(mpmath.log10(w0 / w)) / float(mpmath.log10(w0 / w1)))))
            W, H = sizes[int(picks[f])]
            wd = max(digits, int(-mpmath.log10(w)) + 30)
            frames.append({"width": mpmath.nstr(w, 25, **fmt) if f not in (0, 63) else ("4" if f == 0 else "1e-200"),
                           "center_x": mpmath.nstr(cx, wd, **fmt), "center_y": mpmath.nstr(cy, wd, **fmt),
                           "size": [W, H], "max_iter": mi})
    return {"gen_n": 3, "gen_kr": k.real, "gen_ki": k.imag,
            "target_x": mpmath.nstr(tx, digits, **fmt), "target_y": mpmath.nstr(ty, digits, **fmt),
            "preperiod": m, "period": p, "cycle_multiplier_abs": mult,
            "log10_residual_at_260_digits": log_res, "frames": frames}


def main():
    rng = np.random.Generator(np.random.PCG64(SEED))
    u2 = u128(rng)
    u3 = u128(rng)
    c2 = cardioid_point(u2, 60)
    c3 = cardioid_point(u3, 1100)
    dx = "".join(str(int(x)) for x in rng.integers(0, 10, 79))
    dy = "".join(str(int(x)) for x in rng.integers(0, 10, 79))
    # x in [-1.8, -1.7]: -1.8 + 0.1*U, y in [-0.1, 0]: -0.1 + 0.1*U, with U = 0.d1d2...
    with mpmath.workdps(100):
        x4 = mpmath.mpf("-1.8") + mpmath.mpf("0." + dx) / 10
        y4 = mpmath.mpf("-0.1") + mpmath.mpf("0." + dy) / 10
        c4 = (mpmath.nstr(x4, 80, strip_zeros=False, min_fixed=-mpmath.inf, max_fixed=mpmath.inf),
              mpmath.nstr(y4, 80, strip_zeros=False, min_fixed=-mpmath.inf, max_fixed=mpmath.inf))
    cfg = {
        "seed": SEED,
        "config2": {"u128": str(u2), "center_x": c2[0], "center_y": c2[1], "width": "1e-20",
                    "size": [1920, 1080], "max_iter": 20000},
        "config3": {"u128": str(u3), "center_x": c3[0], "center_y": c3[1], "width": "1e-1000",
                    "size": [3840, 2160], "max_iter": 200000},
        "config4": {"center_x": c4[0], "center_y": c4[1], "width": "1e-50", "size": [2560, 1440],
                    "max_iter": 50000, "ssaa": 4},
        "config5": config5(),
    }
    with open(os.path.join(HERE, "configs.json"), "w") as f:
        json.dump(cfg, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
