"""Colour-LUT contract of the path: a (2048, 3) float32 table, the colormap
sampled at linspace(0, 1, 2048), times 255 (or 65535 for 16-bit output)
(color_engine.py:118-132).  Colormap data comes from matplotlib; an unknown
name falls back to 'inferno' as resolve_colormap does (color_engine.py:
82-89)."""
import threading

import numpy as np

_CPU = {}
_GPU = {}
_LOCK = threading.Lock()


def _resolve(cmap_name):
    import matplotlib

    try:
        cmap = matplotlib.colormaps[cmap_name] if cmap_name else None
    except (KeyError, ValueError, TypeError):
        cmap = None
    return cmap if cmap is not None else matplotlib.colormaps["inferno"]


def get_cpu_cmap_lut(cmap_name, lut_size=2048, is_hdr=False):
    key = (cmap_name, lut_size, bool(is_hdr))
    with _LOCK:
        if key not in _CPU:
            scale = 65535.0 if is_hdr else 255.0
            rgb = (_resolve(cmap_name)(np.linspace(0, 1, lut_size))[:, :3] * scale).astype(np.float32)
            _CPU[key] = np.ascontiguousarray(rgb)
        return _CPU[key]


def get_gpu_cmap_lut(cmap_name, lut_size=2048, is_hdr=False, device="cuda:0"):
    """The LUT as a device tensor (cached per device)."""
    import torch

    key = (cmap_name, lut_size, bool(is_hdr), str(device))
    with _LOCK:
        hit = _GPU.get(key)
    if hit is not None:
        return hit
    t = torch.from_numpy(get_cpu_cmap_lut(cmap_name, lut_size, is_hdr)).to(device)
    with _LOCK:
        _GPU[key] = t
    return t
