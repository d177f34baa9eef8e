"""The reference-shaped render surface (render_frame / render_frame_ssaa /
recolor_frame) on the GPU, against the CPU oracle."""
from decimal import Decimal

import numpy as np
import pytest

from fractura_amd import RenderParameters, engine, recolor_frame, render_frame, render_frame_ssaa
from oracle import oracle
from tests.helpers import config1_plan, lut, oracle_frame

pytestmark = pytest.mark.gpu


def test_render_frame_parameters_roundtrip():
    import torch

    p = RenderParameters(target_w=320, target_h=180, max_iter=500, center_x_dec=Decimal("-0.65"),
                         center_y_dec=Decimal("0.0"), plot_width_dec=Decimal("4.0"),
                         cmap_name="inferno", color_scheme_id=0)
    d_iter = torch.empty((180, 320), dtype=torch.float32, device="cuda:0")
    h_iter = np.zeros((180, 320), np.float32)
    p.d_iter_buf, p.h_iter_buf = d_iter, h_iter
    rgb = render_frame(p)
    assert rgb.shape == (180, 320, 3) and rgb.dtype == np.uint8
    m = oracle_frame(config1_plan(320, 180, 500), lut())
    assert np.abs(rgb.astype(int) - m.rgb.astype(int)).max() <= 1
    assert np.array_equal(h_iter >= 0, m.iter_map >= 0)
    assert np.array_equal(d_iter.cpu().numpy(), h_iter)


def test_render_frame_ssaa_hdr_histogram():
    rgb = render_frame_ssaa(256, 144, 400, Decimal("-0.65"), Decimal("0.0"), Decimal("4.0"),
                            "magma", ssaa_factor=2, color_scheme_id=11, is_hdr=True)
    assert rgb.dtype == np.uint16 and rgb.shape == (144, 256, 3)
    m = oracle_frame(config1_plan(256, 144, 400, factor=2, scheme=11, is_hdr=True),
                     lut(is_hdr=True, name="magma"))
    d = np.abs(rgb.astype(int) - m.rgb.astype(int))
    assert d.max() <= 1 and np.mean(d == 0) >= 0.995


@pytest.mark.parametrize("scheme", [0, 5, 11])
def test_recolor_frame_from_cached_map(scheme):
    import torch

    W, H = 300, 170
    plan = config1_plan(W, H, 600)
    fr = engine.render_view(plan, "inferno", rebase=False)
    torch.cuda.synchronize()
    p = RenderParameters(target_w=W, target_h=H, max_iter=600, cmap_name="plasma",
                         color_scheme_id=scheme, palette_offset=0.2, color_density=1.5)
    rgb = recolor_frame(p, d_iter_buf=fr.iter_map)
    it = fr.iter_map.cpu().numpy()
    v = plan.view.copy()
    v.scheme_id, v.palette_offset, v.color_density, v.color_contrast = scheme, 0.2, 1.5, 1.0
    cdf = oracle.cdf_from_hist(oracle.hist_build(it, 600)) if scheme == 11 else None
    ref = oracle.colorize(v, it, lut(name="plasma"), cdf)
    d = np.abs(rgb.astype(int) - ref.astype(int))
    assert d.max() <= 1 and np.mean(d == 0) >= 0.995


def test_cpu_compute_device_is_refused():
    with pytest.raises(NotImplementedError):
        render_frame_ssaa(64, 36, 100, Decimal("0"), Decimal("0"), Decimal("4"), "inferno",
                          compute_device="CPU")
