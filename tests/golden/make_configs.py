"""Write tests/golden/configs.json: the seeded inputs of the five benchmark
configurations (SEED = 20261203, numpy PCG64).

  config 2: theta = 2 pi u, u the first 128-bit draw / 2^128;
            c = e^{i theta}/2 - e^{2 i theta}/4 at 60 digits (mpmath).
  config 3: the same construction from the second 128-bit draw, 1100 digits.
  config 4: centre uniform in [-1.8, -1.7] x [-0.1, 0.0] as 80-digit decimals.
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
    }
    with open(os.path.join(HERE, "configs.json"), "w") as f:
        json.dump(cfg, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
