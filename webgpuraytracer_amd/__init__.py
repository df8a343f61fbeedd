"""MI355X-native path-tracer hot path (shader.wgsl / denoiser.wgsl of
DragonFireGames/WebGPURaytracer), behind the pt_* C ABI of libpt.so.

    from webgpuraytracer_amd import SceneBuilder, Renderer, Denoiser
"""
from . import _abi as abi  # noqa: F401
from .renderer import Batch, Denoiser, LearnedDenoiser, Renderer  # noqa: F401
from .scene import SceneBuilder, build_sky_cdf, read_hdr, read_obj  # noqa: F401

__all__ = ["SceneBuilder", "Renderer", "Denoiser", "LearnedDenoiser", "Batch", "build_sky_cdf", "abi"]
