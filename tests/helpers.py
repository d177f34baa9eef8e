"""Shared fixtures of the test-suite: reference views and device helpers."""
from decimal import Decimal

import numpy as np

from fractura_amd import engine
from fractura_amd.color import get_cpu_cmap_lut

SEED = 20261203


def config1_plan(width=1280, height=720, max_iter=1000, factor=1, scheme=0, fractal_type=0, **kw):
    """Config 1: 1280x720 at (-0.65, 0), w = 4, max_iter 1000 (FP64)."""
    return engine.plan_view(width, height, max_iter, Decimal("-0.65"), Decimal("0.0"),
                            Decimal("4.0"), ssaa_factor=factor, fractal_type=fractal_type,
                            color_scheme_id=scheme, **kw)


def lut(is_hdr=False, name="inferno"):
    return get_cpu_cmap_lut(name, 2048, is_hdr)


def to_np(t):
    return t.detach().cpu().numpy()


def oracle_frame(plan, lut_arr, edge_threshold=0.35, orbit=None, bla=None, levels=0):
    """The frame flow of engine.render_view (without rebasing), run on the CPU
    oracle: pass 1, histogram colouring for scheme 11, and the adaptive-SSAA
    edge pass."""
    from oracle import oracle
    from fractura_amd import _abi

    v = plan.view.copy()
    hist = v.scheme_id == 11
    eng = plan.engine
    v.mode = 0
    m = oracle.Maps(v, rgb=True)
    kw = dict(orbit=orbit, bla=bla, levels=levels) if eng in ("pert64", "fexp64") else {}
    oracle.render(eng, v, lut=None if hist else lut_arr, maps=m, **kw)
    cdf = None
    if hist:
        h = oracle.hist_build(m.iter_map, v.max_iter)
        cdf = oracle.cdf_from_hist(h)
        m.rgb[...] = oracle.colorize(v, m.iter_map, lut_arr, cdf)
    if v.factor > 1:
        edge = oracle.edge_mask(m.iter_map, edge_threshold)
        v2 = v.copy()
        v2.mode = _abi.AP_MODE_EDGE_SUBSAMPLES
        it_before = m.iterations
        oracle.render(eng, v2, lut=lut_arr, mask=edge, cdf=cdf, maps=m, **kw)
        m.iterations += it_before
        m.edge = edge
    m.cdf = cdf
    return m
