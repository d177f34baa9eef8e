"""A frame's device work (BLA build, pixel pass, colouring) captured as a HIP
graph replays to the eager result, replay after replay, with eager work and
synchronisations between the replays."""
import ctypes as C
from decimal import Decimal

import numpy as np
import pytest

from fractura_amd import engine
from tests.helpers import to_np

pytestmark = pytest.mark.gpu


def _frame_fn(dev, name, plan, orbit, bufs, lut):
    v = plan.view.copy()

    def fn():
        if orbit is not None and orbit.table is not None:
            dev.call("ap_bla_build", C.byref(orbit.abi), v.fractal_type, v.gen_n, v.gen_kr,
                     v.gen_ki, plan.bla_tol, orbit.levels, engine.bla_forms(name),
                     engine.ptr(orbit.table))
        engine._launch(dev, name, v.copy(), orbit,
                       engine._outs(bufs["counts"], bufs["smooth"], bufs["iter_map"], bufs["glitch"]))
        dev.call("ap_colorize", C.byref(v), engine.ptr(bufs["iter_map"]), None, 0, engine.ptr(lut),
                 engine.ptr(bufs["rgb"]))

    return fn


@pytest.mark.parametrize("name,w,max_iter", [("fexp32", "1e-320", 4000), ("fexp64", "1e-320", 4000),
                                             ("fp64", "4.0", 500)])
def test_graph_replays_match_eager(name, w, max_iter):
    import torch

    W, H = 160, 90
    kw = {}
    if name in ("fexp32", "fexp64"):
        kw = dict(bbsa_order=32, bbsa_tol=1e-6, glitch_mode="Pauldelbrot",
                  fexp_mantissa=32 if name == "fexp32" else 64)
        cx, cy = Decimal("0"), Decimal("1")
    else:
        cx, cy = Decimal("-0.65"), Decimal("0")
    plan = engine.plan_view(W, H, max_iter, cx, cy, Decimal(w), **kw)
    assert plan.engine == name
    dev = engine.get_device(0)
    dev.set_stream(None)
    orbit = engine._Orbit(dev, plan, 0, tag="graphtest") if name != "fp64" else None
    lut = engine.get_gpu_cmap_lut("inferno", 2048, False, dev.torch_device)
    bufs = dict(counts=torch.empty((H, W), dtype=torch.int32, device=dev.torch_device),
                smooth=torch.empty((H, W), dtype=torch.float64, device=dev.torch_device),
                iter_map=torch.empty((H, W), dtype=torch.float32, device=dev.torch_device),
                glitch=torch.empty((H, W), dtype=torch.uint8, device=dev.torch_device),
                rgb=torch.empty((H, W, 3), dtype=torch.uint8, device=dev.torch_device))
    for t in bufs.values():
        t.zero_()
    fn = _frame_fn(dev, name, plan, orbit, bufs, lut)
    fn()
    torch.cuda.synchronize()
    ref = {k: to_np(t).copy() for k, t in bufs.items()}
    assert np.mean(ref["counts"] < max_iter) > 0.3  # escaping pixels, so the colours vary

    g = engine.capture_graph(dev, fn)
    for rep in range(3):
        for t in bufs.values():
            t.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, t in bufs.items():
            got = to_np(t)
            assert np.array_equal(got.view(np.uint8), ref[k].view(np.uint8)), f"replay {rep}: {k}"
        # Eager library work between replays.
        fn()
        torch.cuda.synchronize()


def test_two_frames_in_flight_on_two_streams():
    """Two slots (own table and outputs) replayed alternately on two streams,
    as the benchmark times them: every replay of either slot gives the eager
    frame, byte for byte."""
    import torch

    W, H = 160, 90
    plan = engine.plan_view(W, H, 4000, Decimal("0"), Decimal("1"), Decimal("1e-320"),
                            bbsa_order=32, bbsa_tol=1e-6, glitch_mode="Pauldelbrot",
                            fexp_mantissa=32)
    dev = engine.get_device(0)
    dev.set_stream(None)
    lut = engine.get_gpu_cmap_lut("inferno", 2048, False, dev.torch_device)
    slots = []
    for i in range(2):
        orbit = engine._Orbit(dev, plan, 0, tag=f"inflight{i}")
        bufs = dict(counts=torch.zeros((H, W), dtype=torch.int32, device=dev.torch_device),
                    smooth=torch.zeros((H, W), dtype=torch.float64, device=dev.torch_device),
                    iter_map=torch.zeros((H, W), dtype=torch.float32, device=dev.torch_device),
                    glitch=torch.zeros((H, W), dtype=torch.uint8, device=dev.torch_device),
                    rgb=torch.zeros((H, W, 3), dtype=torch.uint8, device=dev.torch_device))
        slots.append((_frame_fn(dev, "fexp32", plan, orbit, bufs, lut), bufs))
    slots[0][0]()
    torch.cuda.synchronize()
    ref = {k: to_np(t).copy() for k, t in slots[0][1].items()}
    assert np.mean(ref["counts"] < 4000) > 0.3
    graphs = [engine.capture_graph(dev, fn) for fn, _ in slots]
    streams = [torch.cuda.Stream(device=dev.torch_device) for _ in range(2)]
    for rep in range(3):
        for _, bufs in slots:
            for t in bufs.values():
                t.zero_()
        torch.cuda.synchronize()
        for st in streams:
            st.wait_stream(torch.cuda.current_stream())
        for i in range(6):
            with torch.cuda.stream(streams[i % 2]):
                graphs[i % 2].replay()
        torch.cuda.synchronize()
        for j, (_, bufs) in enumerate(slots):
            for k, t in bufs.items():
                got = to_np(t)
                assert np.array_equal(got.view(np.uint8), ref[k].view(np.uint8)), (rep, j, k)
