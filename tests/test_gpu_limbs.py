"""The GPU limb evaluator (ap_render_limbs) and the Strict-Metric fallback.

Bars:
  * against the CPU oracle on the same pixels: counts identical, float32 map
    values within one ulp and float64 smooth values within 1e-12 relative
    (the two math libraries' log may differ in the last bit), RGB identical
    where the map values are;
  * against the mpmath subsamples: counts on >= 99.9% of the pixels, smooth
    within 1e-9 where they agree;
  * Strict Metric end to end: every pixel still glitched after the rebase
    passes is re-evaluated, the glitch map ends all zero.
"""
import ctypes as C
import json
import os
from decimal import Decimal, localcontext

import numpy as np
import pytest

from fractura_amd import _abi, engine, exhaustive
from fractura_amd.color import get_cpu_cmap_lut, get_gpu_cmap_lut
from fractura_amd.config import INV_LN2_F32, LN_BAILOUT_LOG
from fractura_amd.device import get_device, ptr
from fractura_amd.high_precision import compute_single_highprec_orbit
from oracle import oracle
from tests.helpers import to_np

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "mpmath_subsamples.json")
GPU34, HIGHPREC = _abi.AP_LIMBS_GPU34, _abi.AP_LIMBS_HIGHPREC


def _fixture(name):
    with open(FIXTURE) as f:
        return {v["name"]: v for v in json.load(f)["views"]}[name]


def _height(w, W, H):
    with localcontext() as ctx:
        ctx.prec = 1200
        return Decimal(w) * Decimal(H) / Decimal(W)


def _buffers(W, H, hdr=False):
    import torch

    d = get_device(0).torch_device
    return dict(rgb=torch.zeros((H, W, 3), dtype=torch.uint16 if hdr else torch.uint8, device=d),
                glitch=torch.ones((H, W), dtype=torch.uint8, device=d),
                it=torch.full((H, W), -7.0, dtype=torch.float32, device=d),
                counts=torch.full((H, W), -7, dtype=torch.int32, device=d),
                smooth=torch.zeros((H, W), dtype=torch.float64, device=d))


def _compare(got, want, ys, xs, label):
    """GPU outputs (dict of numpy arrays) against an oracle Maps at (ys, xs)."""
    gc, wc = got["counts"][ys, xs], want.counts[ys, xs]
    assert np.array_equal(gc, wc), f"{label}: counts differ at {np.flatnonzero(gc != wc)[:8]}"
    gi, wi = got["it"][ys, xs], want.iter_map[ys, xs]
    ulps = np.abs(gi.view(np.int32).astype(np.int64) - wi.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1, f"{label}: float32 map differs by {ulps.max()} ulps"
    gs, ws = got["smooth"][ys, xs], want.smooth[ys, xs]
    assert np.all(np.abs(gs - ws) <= 1e-12 * np.maximum(1.0, np.abs(ws))), label
    same = ulps == 0
    if want.rgb is not None:
        assert np.array_equal(got["rgb"][ys, xs][same], want.rgb[ys, xs][same]), f"{label}: rgb"
    print(f"{label}: {len(xs)} px, float32 map bit-equal on {np.mean(same):.4f}")


def _gpu_fallback(view, xs, ys, semantics, scheme=0, hdr=False, cdf=None, auto_tier=False):
    W, H = view["W"], view["H"]
    w = Decimal(view["width"])
    n, kr, ki = view.get("gen", (3, 0.25, 1.0))
    b = _buffers(W, H, hdr)
    dev = get_device(0)
    lut = get_gpu_cmap_lut("inferno", 2048, hdr, dev.torch_device)
    tier = exhaustive.eval_glitched_pixels_fallback_cuda(
        b["rgb"], b["glitch"], lut, np.stack([ys, xs], 1), W, H, view["max_iter"],
        Decimal(view["cx"]), Decimal(view["cy"]), w, _height(w, W, H), scheme, view["type"], 0.0,
        0.0, gen_n=n, gen_kr=kr, gen_ki=ki, d_cdf_lut=cdf, num_bins=0 if cdf is None else len(cdf),
        is_hdr=hdr, d_iter_map=b["it"], d_counts=b["counts"], d_smooth=b["smooth"],
        tier=None if auto_tier else (3 if semantics == HIGHPREC else 2))
    import torch

    torch.cuda.synchronize()
    return tier, {k: to_np(t) for k, t in b.items()}


def _oracle_fallback(view, xs, ys, semantics, scheme=0, hdr=False, cdf=None):
    W, H = view["W"], view["H"]
    w = Decimal(view["width"])
    n, kr, ki = view.get("gen", (3, 0.25, 1.0))
    L = exhaustive.ML_LIMBS if semantics == GPU34 else exhaustive.highprec_limbs(w)
    coords, signs = exhaustive.pixel_coordinate_limbs(ys, xs, W, H, Decimal(view["cx"]),
                                                      Decimal(view["cy"]), w, _height(w, W, H), L)
    params, psigns = exhaustive._param_limbs(0.0, 0.0, kr, ki, L)
    v = _abi.ApView()
    v.width, v.height, v.factor, v.max_iter = W, H, 1, view["max_iter"]
    v.fractal_type, v.gen_n, v.gen_kr, v.gen_ki = view["type"], n, kr, ki
    v.inv_ln2, v.ln_bailout_log = INV_LN2_F32, LN_BAILOUT_LOG
    v.scheme_id, v.is_hdr = scheme, 1 if hdr else 0
    lut = get_cpu_cmap_lut("inferno", 2048, hdr)
    return oracle.render_limbs(v, semantics, L, ys * W + xs, coords, signs, params, psigns,
                               lut=lut, cdf=None if cdf is None else to_np(cdf), threads=16)


CASES = [("tier2_misiurewicz_i_1e-250", GPU34, None), ("tier2_bship_tip_1e-250", GPU34, None),
         ("config4", GPU34, None), ("tier2_misiurewicz_i_1e-250", HIGHPREC, None),
         ("deep_misiurewicz_i_1e-320", HIGHPREC, None),
         ("deep_misiurewicz_i_1e-1000", HIGHPREC, 192),
         ("config5_frame12", HIGHPREC, None), ("config5_frame31", HIGHPREC, None),
         ("config5_frame47", HIGHPREC, 256), ("config5_frame63", HIGHPREC, 256)]


@pytest.mark.parametrize("name,semantics,oracle_px", CASES)
def test_limbs_match_oracle_and_mpmath(name, semantics, oracle_px):
    view = _fixture(name)
    xs, ys = np.array(view["x"]), np.array(view["y"])
    tier, got = _gpu_fallback(view, xs, ys, semantics)
    assert tier == (3 if semantics == HIGHPREC else 2)
    # every listed pixel's glitch flag cleared
    assert np.all(got["glitch"][ys, xs] == 0)
    # against mpmath
    ref_c, ref_s = np.array(view["counts"]), np.array(view["smooth"])
    gc = got["counts"][ys, xs]
    agree = np.mean(gc == ref_c)
    esc = (gc == ref_c) & (ref_c < view["max_iter"])
    err = np.abs(got["smooth"][ys, xs][esc] - ref_s[esc])
    print(f"{name}: mpmath agreement {agree:.5f}, smooth max err {err.max() if err.size else 0:.2e}")
    assert agree >= 0.999
    assert err.size == 0 or err.max() <= 1e-9
    # against the oracle
    k = oracle_px or len(xs)
    want = _oracle_fallback(view, xs[:k], ys[:k], semantics)
    _compare(got, want, ys[:k], xs[:k], name)


def test_four_limbs_per_lane_match_oracle():
    # width 1e-2000: 1 + ceil(6771 / 32) = 213 limbs, four registers per lane.
    view = dict(name="i_1e-2000", cx="0", cy="1", width="1e-2000", W=640, H=360, max_iter=8000,
                type=0)
    assert exhaustive.highprec_limbs(Decimal("1e-2000")) == 213
    rng = np.random.Generator(np.random.PCG64(2000))
    xs, ys = rng.integers(0, 640, 16), rng.integers(0, 360, 16)
    _, got = _gpu_fallback(view, xs, ys, HIGHPREC)
    want = _oracle_fallback(view, xs, ys, HIGHPREC)
    _compare(got, want, ys, xs, "i_1e-2000")
    assert np.all(got["counts"][ys, xs] < 8000)


def _points(points, semantics, L, ftype, max_iter, n=3, kr=0.25, ki=1.0, julia=(0.0, 0.0)):
    """GPU and oracle evaluation of explicit points through ap_render_limbs."""
    import torch

    dev = get_device(0)
    W = len(points)
    coords = np.empty((W, 2, L), np.uint32)
    signs = np.empty((W, 2), np.int32)
    for i, (x, y) in enumerate(points):
        signs[i, 0], coords[i, 0] = exhaustive.decimal_to_limbs(Decimal(str(x)), L)
        signs[i, 1], coords[i, 1] = exhaustive.decimal_to_limbs(Decimal(str(y)), L)
    params, psigns = exhaustive._param_limbs(julia[0], julia[1], kr, ki, L)
    v = _abi.ApView()
    v.width, v.height, v.factor, v.max_iter = W, 1, 1, max_iter
    v.fractal_type, v.gen_n, v.gen_kr, v.gen_ki = ftype, n, kr, ki
    v.inv_ln2, v.ln_bailout_log = INV_LN2_F32, LN_BAILOUT_LOG
    q = np.arange(W, dtype=np.int32)
    td = dev.torch_device
    b = _buffers(W, 1)
    o = _abi.ApOutputs()
    o.counts, o.smooth, o.iter_map = b["counts"].data_ptr(), b["smooth"].data_ptr(), b["it"].data_ptr()
    dq = torch.from_numpy(q).to(td)
    dc = torch.from_numpy(coords.reshape(-1).view(np.int32)).to(td)
    ds = torch.from_numpy(signs.reshape(-1)).to(td)
    dp = torch.from_numpy(params.reshape(-1).view(np.int32)).to(td)
    dps = torch.from_numpy(psigns).to(td)
    dev.set_stream(None)
    dev.call("ap_render_limbs", C.byref(v), semantics, L, ptr(dq), W, ptr(dc), ptr(ds), ptr(dp),
             ptr(dps), C.byref(o))
    torch.cuda.synchronize()
    got = {k: to_np(t) for k, t in b.items()}
    want = oracle.render_limbs(v, semantics, L, q, coords, signs, params, psigns)
    return got, want


@pytest.mark.parametrize("ftype", [2, 3])
def test_julia_closed_form_on_gpu(ftype):
    pts = [(2, 0), (0, -2), (Decimal(2).sqrt(), Decimal(2).sqrt()), ("0.3", "0.4")]
    got, want = _points(pts, GPU34, 34, ftype, 100)
    assert got["counts"][0, 0] == 3 and got["counts"][0, 1] == 3
    _compare(got, want, np.zeros(4, int), np.arange(4), f"julia type {ftype} gpu34")
    got, want = _points(pts[:2], HIGHPREC, 8, ftype, 100)
    assert got["counts"][0, 0] == 4
    _compare(got, want, np.zeros(2, int), np.arange(2), f"julia type {ftype} highprec")


@pytest.mark.parametrize("n,kr,ki", [(2, 0.0, 0.0), (5, -0.2, 0.1), (8, 0.05, -0.3), (17, 0.0, 0.5)])
def test_type4_powers_on_gpu(n, kr, ki):
    rng = np.random.Generator(np.random.PCG64(n))
    pts = [(f"{x:.12f}", f"{y:.12f}") for x, y in rng.uniform(-1.3, 1.3, (64, 2))]
    got, want = _points(pts, HIGHPREC, 6, 4, 150, n, kr, ki)
    _compare(got, want, np.zeros(64, int), np.arange(64), f"type 4 n={n}")


@pytest.mark.parametrize("scheme,hdr", [(0, False), (7, True), (11, False)])
@pytest.mark.parametrize("semantics", [GPU34, HIGHPREC])
def test_fallback_colours(scheme, hdr, semantics):
    import torch

    view = _fixture("tier2_misiurewicz_i_1e-250")
    xs, ys = np.array(view["x"][:128]), np.array(view["y"][:128])
    cdf = None
    if scheme == 11:
        ramp = np.linspace(0.0, 1.0, 65536, dtype=np.float32) ** 0.5
        cdf = torch.from_numpy(ramp.astype(np.float32)).to(get_device(0).torch_device)
    _, got = _gpu_fallback(view, xs, ys, semantics, scheme=scheme, hdr=hdr, cdf=cdf)
    want = _oracle_fallback(view, xs, ys, semantics, scheme=scheme, hdr=hdr, cdf=cdf)
    _compare(got, want, ys, xs, f"scheme {scheme} hdr {hdr}")


def _short_reference_frame(cx, cy, width, ftype, engine_kw, scheme=0, size=20, ref_len=10):
    """A size x size Strict-Metric frame whose reference orbit is cut at
    ref_len iterations: every pixel outlives it and is flagged (the reference
    escaped before max_iter), and with at most 512 glitched pixels the Strict
    pass goes straight to the exhaustive fallback."""
    plan = engine.plan_view(size, size, 3000, Decimal(cx), Decimal(cy), Decimal(width),
                            fractal_type=ftype, glitch_mode="Strict Metric", color_scheme_id=scheme,
                            compute_orbit=False, **engine_kw)
    from fractura_amd.config import compute_required_prec_bits

    re, im, rl = compute_single_highprec_orbit(Decimal(cx), Decimal(cy), 3000, ftype, 0.0, 0.0,
                                               compute_required_prec_bits(Decimal(width)))
    assert rl > ref_len
    engine.set_reference(plan, Decimal(cx), Decimal(cy), 0.5, 0.5, re[: ref_len + 1],
                         im[: ref_len + 1], ref_len)
    return plan


@pytest.mark.parametrize("cx,cy,width,ftype,kw,scheme,tier", [
    ("0", "1", "1e-320", 0, dict(force_floatexp=True), 0, 3),
    ("0", "1", "1e-250", 0, dict(), 11, 2),
    ("-2", "0", "1e-250", 1, dict(), 0, 2),
    ("0", "1", "1e-20", 0, dict(), 0, 1),
])
def test_strict_metric_frame_ends_without_glitches(cx, cy, width, ftype, kw, scheme, tier):
    import torch

    plan = _short_reference_frame(cx, cy, width, ftype, kw, scheme)
    fr = engine.render_view(plan, "inferno", glitch_mode="Strict Metric")
    torch.cuda.synchronize()
    assert fr.glitch_counts and fr.glitch_counts[0] == 400
    assert fr.exhaustive_pixels == 400 and fr.exhaustive_tier == tier
    assert int(fr.glitch.sum().item()) == 0
    # Every pixel now holds the direct evaluation at the tier's precision.
    W = H = 20
    ys, xs = np.divmod(np.arange(400), W)
    w = Decimal(width)
    view = dict(W=W, H=H, cx=cx, cy=cy, width=width, max_iter=3000, type=ftype)
    got = dict(counts=to_np(fr.counts), it=to_np(fr.iter_map), smooth=to_np(fr.smooth),
               rgb=to_np(fr.rgb))
    if tier >= 2:
        sem = HIGHPREC if tier == 3 else GPU34
        cdf = None
        want = _oracle_fallback(view, xs, ys, sem, scheme=scheme, cdf=cdf) if scheme != 11 else None
        if want is not None:
            _compare(got, want, ys, xs, f"strict {width} type {ftype}")
        else:
            want = _oracle_fallback(view, xs, ys, sem)
            assert np.array_equal(got["counts"][ys, xs], want.counts[ys, xs])
    else:
        v = plan.view.copy()
        v.factor, v.mode, v.rebase_pass = 1, 0, 0
        with localcontext() as ctx:
            ctx.prec = 500
            h = w * Decimal(H) / Decimal(W)
            v.x_min, v.x_min_lo = (float(x) for x in _dd(Decimal(cx) - w / 2))
            v.y_max, v.y_max_lo = (float(x) for x in _dd(Decimal(cy) + h / 2))
            v.dx, v.dx_lo = (float(x) for x in _dd(w / (W - 1)))
            v.dy, v.dy_lo = (float(x) for x in _dd(h / (H - 1)))
        m = oracle.render("dd", v)
        assert np.array_equal(got["counts"], m.counts)


def _dd(d):
    hi = float(d)
    return hi, float(d - Decimal(hi))


def test_tier0_fallback_is_the_fp64_loop():
    import torch

    W, H, w = 64, 36, Decimal("1e-5")
    view = dict(W=W, H=H, cx="-0.743643887037151", cy="0.131825904205330", width="1e-5",
                max_iter=2000, type=0)
    ys, xs = np.divmod(np.arange(W * H), W)
    tier, got = _gpu_fallback(view, xs, ys, GPU34, auto_tier=True)
    assert tier == 0
    plan = engine.plan_view(W, H, 2000, Decimal(view["cx"]), Decimal(view["cy"]), w,
                            precision_mode="1e-12 (FP64)", compute_orbit=False)
    v = plan.view.copy()
    with localcontext() as ctx:
        ctx.prec = 500
        v.x_min = float(Decimal(view["cx"]) - w / 2)
        v.y_max = float(Decimal(view["cy"]) + _height(w, W, H) / 2)
        v.dx, v.dy = float(w / (W - 1)), float(_height(w, W, H) / (H - 1))
    m = oracle.render("fp64", v, lut=get_cpu_cmap_lut("inferno", 2048, False))
    _compare(got, m, ys, xs, "tier 0")
