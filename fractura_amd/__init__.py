"""fractura_amd: the per-pixel escape-time hot path of the Apeiron fractal
explorer (szabolcsmeszaro-arch/Fractura), rebuilt for MI355X (gfx950).

The render surface keeps math_core's names (render_frame, render_frame_ssaa,
recolor_frame, RenderParameters, ...); the kernels live in
libapeiron_hip.so (HIP, gfx950) behind the C ABI of include/apeiron_hip.h.
"""
from .config import RenderParameters, compute_floatexp_scale, compute_required_prec_bits
from .device import init_primary_cuda_device
from .engine import (capture_graph, clear_scratch_buffer_pool, compute_bbsa_coeffs_cuda,
                     compute_dynamic_max_iter, get_adaptive_max_iter_from_probe,
                     get_scratch_device_array, plan_view, probe_adaptive_max_iter, recolor_frame,
                     render_frame, render_frame_cpu, render_frame_ssaa, render_view,
                     select_engine)
from .exhaustive import eval_glitched_pixels_fallback, eval_glitched_pixels_fallback_cuda
from .high_precision import (cache_reference_orbit, compute_bbsa_coeffs,
                             compute_highprec_escape_scalar, compute_single_highprec_orbit,
                             find_optimal_reference_orbit, to_dd)

__all__ = [
    "HAS_CUDA", "init_primary_cuda_device",
    "RenderParameters", "render_frame", "render_frame_ssaa", "render_frame_cpu", "recolor_frame",
    "render_view", "plan_view", "select_engine", "compute_dynamic_max_iter",
    "probe_adaptive_max_iter", "get_adaptive_max_iter_from_probe", "capture_graph",
    "eval_glitched_pixels_fallback_cuda", "eval_glitched_pixels_fallback",
    "get_scratch_device_array", "clear_scratch_buffer_pool", "to_dd",
    "compute_bbsa_coeffs", "compute_bbsa_coeffs_cuda",
    "compute_single_highprec_orbit", "compute_highprec_escape_scalar",
    "find_optimal_reference_orbit", "cache_reference_orbit", "compute_floatexp_scale",
    "compute_required_prec_bits",
]


def __getattr__(name):
    # HAS_CUDA (math_core/cuda_kernels.py) is answered on first use, so that
    # importing the package does not start the GPU runtime (a caller may
    # still fork worker processes after the import).
    if name == "HAS_CUDA":
        from .device import has_gpu

        return has_gpu()
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
