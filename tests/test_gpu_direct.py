"""GPU parity of the direct engines (FP64, DD) and their colouring against
the CPU oracle, through the C ABI (via fractura_amd.engine)."""
from decimal import Decimal

import numpy as np
import pytest

from fractura_amd import engine
from tests.helpers import config1_plan, lut, oracle_frame, to_np

pytestmark = pytest.mark.gpu


def _sync():
    import torch

    torch.cuda.synchronize()


def _check_maps(fr, m, rgb_exact_frac=0.999):
    counts = to_np(fr.counts)
    assert np.array_equal(counts, m.counts), f"{np.sum(counts != m.counts)} count mismatches"
    it = to_np(fr.iter_map)
    # float32 map: the same float64 log chain; libm and the device log may
    # differ in the last bit, which can move the float32 value by one ulp.
    ulps = np.abs(it.view(np.int32).astype(np.int64) - m.iter_map.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1
    assert np.mean(ulps == 0) >= 0.999
    sm = to_np(fr.smooth)
    esc = m.smooth >= 0
    assert np.array_equal(sm < 0, ~esc)
    rel = np.abs(sm[esc] - m.smooth[esc]) / np.maximum(1.0, np.abs(m.smooth[esc]))
    assert rel.max(initial=0.0) <= 1e-12
    rgb = to_np(fr.rgb).astype(np.int64)
    diff = np.abs(rgb - m.rgb.astype(np.int64))
    assert diff.max() <= 1
    assert np.mean(diff == 0) >= rgb_exact_frac


def test_fp64_config1_bit_exact():
    """Config 1 (1280x720, w=4, max_iter 1000): counts bit-exact."""
    plan = config1_plan()
    fr = engine.render_view(plan, "inferno", rebase=False)
    _sync()
    m = oracle_frame(plan, lut())
    _check_maps(fr, m)
    assert fr.host_stats()["iterations"] == m.iterations


@pytest.mark.parametrize("ftype", [1, 2, 3, 4])
def test_fp64_types(ftype):
    from fractura_amd.config import FRACTAL_DEFAULTS

    cx, cy, w = FRACTAL_DEFAULTS[ftype]
    plan = engine.plan_view(320, 180, 500, cx, cy, w, fractal_type=ftype, julia_cx=-0.67,
                            julia_cy=0.37, gen_n=3, gen_kr=0.25, gen_ki=1.0)
    fr = engine.render_view(plan, "magma", rebase=False)
    _sync()
    m = oracle_frame(plan, lut(name="magma"))
    _check_maps(fr, m)


@pytest.mark.parametrize("max_iter", [1, 2, 3, 5, 6, 7, 13, 101, 1002, 1003])
def test_fp64_max_iter_remainders(max_iter):
    """The FP64 loop tests escape once per four iterations and finishes the
    last max_iter mod 4 steps one at a time: every remainder, every type,
    bit-exact."""
    from fractura_amd.config import FRACTAL_DEFAULTS

    for ftype in range(5):
        cx, cy, w = ("-0.65", "0.0", "4.0") if ftype == 0 else FRACTAL_DEFAULTS[ftype]
        plan = engine.plan_view(96, 54, max_iter, cx, cy, w, fractal_type=ftype, julia_cx=-0.67,
                                julia_cy=0.37, gen_n=3 + ftype % 2, gen_kr=0.25, gen_ki=1.0)
        assert plan.engine == "fp64"
        fr = engine.render_view(plan, "magma", rebase=False)
        _sync()
        m = oracle_frame(plan, lut(name="magma"))
        _check_maps(fr, m, rgb_exact_frac=0.99)


@pytest.mark.parametrize("scheme", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_fp64_palettes(scheme):
    plan = config1_plan(256, 144, 300, scheme=scheme, palette_offset=0.3, color_density=1.7,
                        color_contrast=1.3)
    fr = engine.render_view(plan, "viridis", rebase=False)
    _sync()
    m = oracle_frame(plan, lut(name="viridis"))
    _check_maps(fr, m, rgb_exact_frac=0.995)


@pytest.mark.parametrize("factor,scheme,hdr", [(2, 0, False), (3, 11, True), (4, 11, False)])
def test_fp64_ssaa_and_histogram(factor, scheme, hdr):
    plan = config1_plan(320, 180, 400, factor=factor, scheme=scheme, is_hdr=hdr)
    fr = engine.render_view(plan, "inferno", rebase=False)
    _sync()
    m = oracle_frame(plan, lut(is_hdr=hdr))
    _check_maps(fr, m, rgb_exact_frac=0.995)
    assert fr.host_stats()["iterations"] == m.iterations


def test_dd_small():
    """DD engine at w = 1e-20 near the seeded cardioid-boundary point."""
    plan = engine.plan_view(192, 108, 3000, Decimal("-0.7436438870371587"),
                            Decimal("0.1318259042053119"), Decimal("1e-20"),
                            precision_mode="1e-24 (DD)")
    assert plan.engine == "dd"
    fr = engine.render_view(plan, "inferno", rebase=False)
    _sync()
    m = oracle_frame(plan, lut())
    _check_maps(fr, m)


@pytest.mark.parametrize("ftype", [1, 2, 4])
def test_dd_types(ftype):
    from fractura_amd.config import FRACTAL_DEFAULTS

    cx, cy, w = FRACTAL_DEFAULTS[ftype]
    plan = engine.plan_view(160, 90, 400, cx, cy, w * Decimal("1e-13"), fractal_type=ftype,
                            julia_cx=-0.67, julia_cy=0.37, precision_mode="1e-24 (DD)")
    assert plan.engine == "dd"
    fr = engine.render_view(plan, "inferno", rebase=False)
    _sync()
    m = oracle_frame(plan, lut())
    _check_maps(fr, m)


@pytest.mark.parametrize("max_iter", [1, 2, 3, 7, 101])
def test_dd_max_iter_remainders(max_iter):
    """The DD loop tests escape once per two iterations and finishes an odd
    last step alone: odd and even max_iter, every type, bit-exact."""
    from fractura_amd.config import FRACTAL_DEFAULTS

    for ftype in range(5):
        cx, cy, w = ("-0.65", "0.0", "4.0") if ftype == 0 else FRACTAL_DEFAULTS[ftype]
        plan = engine.plan_view(96, 54, max_iter, cx, cy, w, fractal_type=ftype, julia_cx=-0.67,
                                julia_cy=0.37, gen_n=3 + ftype % 2, gen_kr=0.25, gen_ki=1.0,
                                engine_override="dd")
        fr = engine.render_view(plan, "magma", rebase=False)
        _sync()
        m = oracle_frame(plan, lut(name="magma"))
        _check_maps(fr, m, rgb_exact_frac=0.99)
