"""spectral_amd -- MI355X (gfx950) light-transport core of spectral-core.

Camera render, lighttrace and volume entry points as hand-written HIP kernels
behind the C ABI in include/spectral_hip.h (libspectral_hip.so, built in-tree).
"""
from . import _abi as abi
from .core import (Camera, Context, DeviceBuffer, Dielectric, Film, HipTracer, Lambertian, Scene,
                   axis_box, beam, display_quantile, display_resolve, glass_n, lighttrace_trace, opts,
                   photon_records, resolve_srgb8,
                   render_volume, render_volumetric, render_volumetric_scene,
                   render_volumetric_scene_point, rng_draws, screen, surface_rim, triangular_prism, wedge)

__all__ = ["abi", "Camera", "Context", "DeviceBuffer", "Dielectric", "Film", "HipTracer",
           "Lambertian", "Scene", "axis_box", "beam", "display_quantile", "display_resolve", "glass_n",
           "lighttrace_trace", "opts", "photon_records", "resolve_srgb8", "render_volume", "render_volumetric",
           "render_volumetric_scene", "render_volumetric_scene_point", "rng_draws", "screen", "surface_rim",
           "triangular_prism", "wedge"]
