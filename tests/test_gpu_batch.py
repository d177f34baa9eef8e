"""The batched persistent pass renders every frame exactly as the single-frame
FP32 pass does (same engine, same tables), for frames of uneven size and
depth and every tile shape; its HIP graph replays to the eager result."""
import ctypes as C
import json
import os
from decimal import Decimal

import numpy as np
import pytest

from fractura_amd import _abi, batch, engine
from tests.helpers import to_np

pytestmark = pytest.mark.gpu

CONFIGS = os.path.join(os.path.dirname(__file__), "golden", "configs.json")
SIZES = [(96, 54), (33, 17), (64, 64), (100, 7), (48, 80)]


def _specs():
    with open(CONFIGS) as f:
        c5 = json.load(f)["config5"]
    specs = batch.flythrough_specs(c5)
    picked = [specs[i] for i in (0, 12, 31, 47, 63)]
    for s, size in zip(picked, SIZES):
        s["size"] = list(size)
        s["max_iter"] = min(int(s["max_iter"]), 6000)
    return picked


def _single(plan):
    """The same frame through the single-frame FP32 pass, on a table built
    with the FP32 form."""
    import torch

    dev = engine.get_device(0)
    dev.set_stream(None)
    v = plan.view
    orbit = engine._Orbit(dev, plan, 0, tag="batchref", forms=_abi.AP_BLA_FORM_FP32)
    outs = dict(counts=torch.zeros((v.height, v.width), dtype=torch.int32, device=dev.torch_device),
                smooth=torch.zeros((v.height, v.width), dtype=torch.float64, device=dev.torch_device),
                iter_map=torch.zeros((v.height, v.width), dtype=torch.float32, device=dev.torch_device),
                glitch=torch.zeros((v.height, v.width), dtype=torch.uint8, device=dev.torch_device))
    engine._launch(dev, "fexp32", v.copy(), orbit,
                   engine._outs(outs["counts"], outs["smooth"], outs["iter_map"], outs["glitch"]))
    torch.cuda.synchronize()
    return {k: to_np(t) for k, t in outs.items()}


@pytest.fixture(scope="module")
def plans():
    return batch.plan_frames(_specs(), processes=1)


@pytest.mark.parametrize("tile", batch.TILE_SHAPES)
def test_batch_matches_single_frame_passes(plans, tile):
    import torch

    b = batch.FrameBatch(plans, tile=tile)
    lut = engine.get_gpu_cmap_lut("inferno", 2048, False, b.dev.torch_device)
    b.step(lut)
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        ref = _single(p)
        for k in ("counts", "smooth", "iter_map", "glitch"):
            got = to_np(b.outs[i][k])
            assert np.array_equal(got.view(np.uint8), ref[k].view(np.uint8)), (i, k, tile)
    assert any(np.mean(to_np(o["counts"]) < p.view.max_iter) > 0.5 for o, p in zip(b.outs, plans))


def test_batch_graph_replays_and_stats(plans):
    import torch

    b = batch.FrameBatch(plans, tile=(32, 8), blocks=3 * b_cus())
    lut = engine.get_gpu_cmap_lut("inferno", 2048, False, b.dev.torch_device)
    b.step(lut)
    b.render(stats=True)
    torch.cuda.synchronize()
    eager = [{k: to_np(t).copy() for k, t in o.items()} for o in b.outs]
    st = to_np(b.stats).view(np.uint64)
    assert int(st[_abi.AP_STAT_ITERATIONS]) == b.iterations()
    assert int(st[_abi.AP_STAT_SAMPLES]) == b.pixels()
    g = b.capture(lut)
    for rep in range(3):
        for o in b.outs:
            for t in o.values():
                t.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for o, e in zip(b.outs, eager):
            for k, t in o.items():
                assert np.array_equal(to_np(t).view(np.uint8), e[k].view(np.uint8)), (rep, k)
        b.render()  # eager work between replays
        torch.cuda.synchronize()


def b_cus():
    return engine.get_device(0).num_cus


def test_batch_autotune_picks_a_candidate(plans):
    b = batch.FrameBatch(plans)
    res = batch.autotune(b, reps=1, blocks_per_cu=(0, 4))
    assert tuple(res["tile"]) in batch.TILE_SHAPES
    assert len(res["candidates"]) == len(batch.TILE_SHAPES) * 2
    assert b.tile == tuple(res["tile"])


def test_batch_frames_against_mpmath():
    """Config-5 frames at their real size and depth through the batched pass
    (no rebasing) against the mpmath oracle's seeded subsamples."""
    import torch

    from tests.test_gpu_mpmath import _check, _views

    with open(CONFIGS) as f:
        c5 = json.load(f)["config5"]
    specs = batch.flythrough_specs(c5)
    names = ["config5_frame12", "config5_frame31", "config5_frame47", "config5_frame63"]
    idx = [12, 31, 47, 63]
    plans = batch.plan_frames([specs[i] for i in idx], processes=1)
    b = batch.FrameBatch(plans)
    lut = engine.get_gpu_cmap_lut("inferno", 2048, False, b.dev.torch_device)
    b.step(lut)
    torch.cuda.synchronize()
    views = _views()
    for name, o in zip(names, b.outs):
        _check(views[name], to_np(o["counts"]), to_np(o["smooth"]))
