"""Strict-Metric exhaustive fallback (math_core/engine.py:191-410).

Pixels still glitched after the rebase passes are evaluated again directly,
without perturbation, at a precision tier chosen from the width
(math_core/engine.py:282-293):

  tier 0   w >= 1e-12           FP64 direct loop   (ap_render_fp64 over a queue)
  tier 1   1e-30 <= w < 1e-12   double-double      (ap_render_dd over a queue)
  tier 2   1e-315 <= w < 1e-30  34-limb fixed point, compute_ml_sample
                                (ap_render_limbs, AP_LIMBS_GPU34)
  tier 3   w < 1e-315, and type 4 at tier 2: the reference evaluates these on
           the CPU in MPFR (compute_highprec_escape_scalar,
           math_core/high_precision.py:581-656); here they run on the GPU in
           1 + ceil(prec_bits / 32) limbs (ap_render_limbs, AP_LIMBS_HIGHPREC),
           prec_bits = compute_required_prec_bits(w) as the reference uses.

Every tier writes the listed pixels' float32 map value, RGB and glitch flag
(0), plus this library's int32 counts and float64 smooth values.
"""
import ctypes as C
from decimal import Decimal

import numpy as np

from . import _abi
from .color import get_gpu_cmap_lut
from .config import HISTOGRAM_BINS, HISTOGRAM_SCHEME, INV_LN2_F32, LN_BAILOUT_LOG, \
    compute_required_prec_bits, decimal_context
from .device import get_device, get_scratch_bytes, get_scratch_device_array, get_stream_slot, ptr
from .high_precision import to_dd

ML_LIMBS = 34

__all__ = ["ML_LIMBS", "decimal_to_limbs", "limbs_to_decimal", "precision_tier",
           "highprec_limbs", "pixel_coordinate_limbs", "eval_glitched_pixels_fallback_cuda",
           "eval_glitched_pixels_fallback"]


def decimal_to_limbs(d, num_limbs=ML_LIMBS):
    """(sign, uint32[num_limbs]) of a number: limb 0 its integer part, limbs
    1.. its fraction in base 2^32, truncated toward zero; sign 1 if negative
    (decimal_to_limbs, math_core/cuda_kernels.py:2241-2274, Decimal path).
    Computed exactly from the decimal digits; the reference's digit-by-digit
    Decimal loop gives the same limbs at its >= 500-digit precision."""
    d = d if isinstance(d, Decimal) else Decimal(str(d))
    if not d.is_finite():
        raise ValueError(f"cannot convert {d} to limbs")
    sign = 1 if d < 0 else 0
    t = d.as_tuple()
    n = int("".join(map(str, t.digits))) if t.digits else 0
    frac_bits = 32 * (num_limbs - 1)
    if t.exponent >= 0:
        x = (n * 10 ** t.exponent) << frac_bits
    else:
        x = (n << frac_bits) // (10 ** (-t.exponent))
    if x >> (32 * num_limbs):
        raise ValueError(f"|{d}| does not fit a 32-bit integer limb")
    limbs = np.empty(num_limbs, dtype=np.uint32)
    for i in range(num_limbs - 1, -1, -1):
        limbs[i] = x & 0xFFFFFFFF
        x >>= 32
    return sign, limbs


def limbs_to_decimal(sign, limbs):
    """Exact value of a limb vector (for checks and diagnostics)."""
    x = 0
    for v in limbs:
        x = (x << 32) | int(v)
    frac_bits = 32 * (len(limbs) - 1)
    with decimal_context(Decimal(2) ** -frac_bits):
        d = Decimal(x) / (Decimal(2) ** frac_bits)
    return -d if sign else d


def precision_tier(plot_width_dec):
    """math_core/engine.py:282-293."""
    w = Decimal(str(plot_width_dec))
    if w >= Decimal("1e-12"):
        return 0
    if w >= Decimal("1e-30"):
        return 1
    if w >= Decimal("1e-315"):
        return 2
    return 3


def highprec_limbs(plot_width_dec):
    """Limbs of the arbitrary-precision tier: one integer limb and enough
    fraction limbs for the reference's prec_bits."""
    bits = compute_required_prec_bits(plot_width_dec)
    L = 1 + -(-bits // 32)
    if L > _abi.AP_MAX_LIMBS:
        raise NotImplementedError(
            f"width {plot_width_dec} needs {bits} bits; the limb evaluator holds at most "
            f"{32 * (_abi.AP_MAX_LIMBS - 1)}")
    return L


def pixel_coordinate_limbs(rows, cols, target_w, target_h, center_x_dec, center_y_dec,
                           plot_width_dec, plot_height_dec, num_limbs):
    """Limbs of the pixel centres c = (x_min + col dx, y_max - row dy) with
    x_min = cx - w/2, y_max = cy + h/2, dx = w/(W-1), dy = h/(H-1) in Decimal
    at the frame's precision (math_core/engine.py:295-298, 351-356).
    Returns (coords uint32 (n, 2, L), signs int32 (n, 2))."""
    n = len(rows)
    coords = np.empty((n, 2, num_limbs), dtype=np.uint32)
    signs = np.empty((n, 2), dtype=np.int32)
    with decimal_context(plot_width_dec):
        w, h = Decimal(plot_width_dec), Decimal(plot_height_dec)
        x_min = Decimal(center_x_dec) - w / Decimal(2)
        y_max = Decimal(center_y_dec) + h / Decimal(2)
        dx = w / Decimal(max(1, int(target_w) - 1))
        dy = h / Decimal(max(1, int(target_h) - 1))
        for i, (r, c) in enumerate(zip(rows, cols)):
            signs[i, 0], coords[i, 0] = decimal_to_limbs(x_min + Decimal(int(c)) * dx, num_limbs)
            signs[i, 1], coords[i, 1] = decimal_to_limbs(y_max - Decimal(int(r)) * dy, num_limbs)
    return coords, signs


def _param_limbs(julia_cx, julia_cy, gen_kr, gen_ki, num_limbs):
    """julia_cx, julia_cy, gen_kr, gen_ki as the decimal values of their str()
    (math_core/engine.py:359-360; math_core/high_precision.py:593-601)."""
    params = np.empty((4, num_limbs), dtype=np.uint32)
    psigns = np.empty(4, dtype=np.int32)
    for i, x in enumerate((julia_cx, julia_cy, gen_kr, gen_ki)):
        psigns[i], params[i] = decimal_to_limbs(Decimal(str(x)), num_limbs)
    return params, psigns


def _stage(arr, tag, slot, device):
    """Copy a host array into a pooled device byte buffer (synchronous copy,
    so the data is on the device before any stream's launch reads it)."""
    import torch

    raw = np.ascontiguousarray(arr).reshape(-1).view(np.uint8)
    buf = get_scratch_bytes(raw.nbytes, f"fallback_{tag}_{slot}", device)
    buf[: raw.nbytes].copy_(torch.from_numpy(raw), non_blocking=False)
    return buf


def eval_glitched_pixels_fallback_cuda(
        d_out_rgb, d_glitch_out, d_lut, glitched_indices, target_w, target_h, max_iter,
        center_x_dec, center_y_dec, plot_width_dec, plot_height_dec, color_scheme_id,
        fractal_type, julia_cx, julia_cy, gen_n=3, gen_kr=0.25, gen_ki=1.0, palette_offset=0.0,
        color_density=1.0, color_contrast=1.0, d_cdf_lut=None, num_bins=2048,
        cmap_name="inferno", is_hdr=False, stream=0, d_iter_map=None, stream_slot=0,
        d_counts=None, d_smooth=None, device=0, tier=None):
    """Re-evaluate the (row, col) pixels of `glitched_indices` directly and
    write their RGB, float32 map value and glitch flag (0), plus counts and
    smooth values where those buffers are given
    (math_core/engine.py:270-410).  Device buffers are torch tensors;
    `d_cdf_lut` colours by the histogram CDF when the scheme is 11.  `tier`
    forces a precision tier (3: the arbitrary-precision semantics at any
    width); by default it follows the width.  Returns the tier used."""
    idx = np.asarray(glitched_indices, dtype=np.int64).reshape(-1, 2)
    n = len(idx)
    W, H = int(target_w), int(target_h)
    center_x_dec, center_y_dec = Decimal(str(center_x_dec)), Decimal(str(center_y_dec))
    plot_width_dec, plot_height_dec = Decimal(str(plot_width_dec)), Decimal(str(plot_height_dec))
    tier = precision_tier(plot_width_dec) if tier is None else int(tier)
    if n == 0:
        return tier
    if np.any(idx[:, 0] < 0) or np.any(idx[:, 0] >= H) or np.any(idx[:, 1] < 0) or np.any(idx[:, 1] >= W):
        raise ValueError("glitched pixel outside the frame")
    highprec = tier == 3 or (tier == 2 and int(fractal_type) == 4)
    if highprec and int(fractal_type) == 4 and not (abs(float(gen_kr)) <= 1024 and abs(float(gen_ki)) <= 1024):
        raise NotImplementedError("the limb evaluator needs |gen_kr|, |gen_ki| <= 1024")
    dev = get_device(device)
    with dev.lock:
        dev.set_stream(stream)
        slot = get_stream_slot(stream, stream_slot)
        v = _abi.ApView()
        v.width, v.height, v.factor, v.max_iter = W, H, 1, int(max_iter)
        v.fractal_type, v.gen_n = int(fractal_type), int(gen_n)
        v.gen_kr, v.gen_ki = float(gen_kr), float(gen_ki)
        v.inv_ln2, v.ln_bailout_log = INV_LN2_F32, LN_BAILOUT_LOG
        v.scheme_id, v.is_hdr = int(color_scheme_id), 1 if is_hdr else 0
        v.palette_offset, v.color_density = float(palette_offset), float(color_density)
        v.color_contrast = float(color_contrast)
        # One un-centred sample per listed pixel, map values written.
        v.mode = _abi.AP_MODE_ALL_SUBSAMPLES
        v.rebase_pass = 1
        it = d_iter_map if d_iter_map is not None else get_scratch_device_array(
            (H, W), np.float32, f"fallback_dummy_iter_{slot}", device)
        cdf = d_cdf_lut if (d_cdf_lut is not None and v.scheme_id == HISTOGRAM_SCHEME) else None
        o = _abi.ApOutputs()
        for name, t in (("counts", d_counts), ("smooth", d_smooth), ("iter_map", it),
                        ("glitch", d_glitch_out), ("rgb", d_out_rgb), ("lut", d_lut), ("cdf", cdf)):
            setattr(o, name, None if t is None else t.data_ptr())
        o.num_bins = int(num_bins) if cdf is not None else 0
        queue = _stage((idx[:, 0] * W + idx[:, 1]).astype(np.int32), "queue", slot, device)
        if tier <= 1:
            qn = _stage(np.array([n], dtype=np.int32), "count", slot, device)
            o.queue, o.queue_count = queue.data_ptr(), qn.data_ptr()
            with decimal_context(plot_width_dec):
                v.x_min, v.x_min_lo = to_dd(center_x_dec - plot_width_dec / Decimal(2))
                v.y_max, v.y_max_lo = to_dd(center_y_dec + plot_height_dec / Decimal(2))
                v.dx, v.dx_lo = to_dd(plot_width_dec / Decimal(max(1, W - 1)))
                v.dy, v.dy_lo = to_dd(plot_height_dec / Decimal(max(1, H - 1)))
                v.julia_cx, v.julia_cx_lo = to_dd(Decimal(str(julia_cx)))
                v.julia_cy, v.julia_cy_lo = to_dd(Decimal(str(julia_cy)))
            dev.call("ap_render_fp64" if tier == 0 else "ap_render_dd", C.byref(v), C.byref(o))
            return tier
        if highprec:
            semantics, L = _abi.AP_LIMBS_HIGHPREC, highprec_limbs(plot_width_dec)
        else:
            semantics, L = _abi.AP_LIMBS_GPU34, ML_LIMBS
        coords, signs = pixel_coordinate_limbs(idx[:, 0], idx[:, 1], W, H, center_x_dec,
                                               center_y_dec, plot_width_dec, plot_height_dec, L)
        params, psigns = _param_limbs(julia_cx, julia_cy, gen_kr, gen_ki, L)
        d_coords = _stage(coords, "coords", slot, device)
        d_signs = _stage(signs, "signs", slot, device)
        d_params = _stage(params, "params", slot, device)
        d_psigns = _stage(psigns, "psigns", slot, device)
        dev.call("ap_render_limbs", C.byref(v), semantics, L, ptr(queue), n, ptr(d_coords),
                 ptr(d_signs), ptr(d_params), ptr(d_psigns), C.byref(o))
        if highprec and d_glitch_out is not None:
            # The CPU tier clears the whole glitch map (math_core/engine.py:332-333).
            d_glitch_out.zero_()
    return 3 if highprec else tier


def eval_glitched_pixels_fallback(out_ary, glitched_indices, target_w, target_h, max_iter,
                                  center_x_dec, center_y_dec, plot_width_dec, plot_height_dec,
                                  cmap_name, color_scheme_id, fractal_type, julia_cx, julia_cy,
                                  gen_n=3, gen_kr=0.25, gen_ki=1.0, is_hdr=False,
                                  palette_offset=0.0, color_density=1.0, color_contrast=1.0,
                                  cdf_lut=None, num_bins=2048, iter_map=None, device=0):
    """The reference's arbitrary-precision evaluator over host arrays
    (math_core/engine.py:191-267): out_ary (H, W, 3) RGB and the optional
    float32 iter_map are updated in place for the listed (row, col) pixels.
    Here it runs the GPU limb evaluator (AP_LIMBS_HIGHPREC) at the
    reference's prec_bits for the width."""
    idx = np.asarray(glitched_indices, dtype=np.int64).reshape(-1, 2)
    if len(idx) == 0:
        return out_ary
    import torch

    dev = get_device(device)
    td = dev.torch_device
    W, H = int(target_w), int(target_h)
    rgb = torch.from_numpy(np.ascontiguousarray(out_ary)).to(td)
    it = torch.from_numpy(np.ascontiguousarray(iter_map, dtype=np.float32)).to(td) \
        if iter_map is not None else torch.zeros((H, W), dtype=torch.float32, device=td)
    lut = get_gpu_cmap_lut(cmap_name, 2048, bool(is_hdr), td)
    cdf = torch.from_numpy(np.ascontiguousarray(cdf_lut, dtype=np.float32)).to(td) \
        if cdf_lut is not None else None
    eval_glitched_pixels_fallback_cuda(
        rgb, None, lut, idx, W, H, max_iter, center_x_dec, center_y_dec, plot_width_dec,
        plot_height_dec, color_scheme_id, fractal_type, julia_cx, julia_cy, gen_n=gen_n,
        gen_kr=gen_kr, gen_ki=gen_ki, palette_offset=palette_offset, color_density=color_density,
        color_contrast=color_contrast, d_cdf_lut=cdf, num_bins=num_bins, is_hdr=is_hdr,
        d_iter_map=it, device=device, tier=3)
    torch.cuda.synchronize(td)
    out_ary[...] = rgb.cpu().numpy()
    if iter_map is not None:
        iter_map[...] = it.cpu().numpy()
    return out_ary
