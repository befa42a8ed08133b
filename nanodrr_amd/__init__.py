"""nanodrr_amd: MI355X-native fused differentiable DRR renderer.

A drop-in for nanodrr's fused ray-march backend: `render(...)` and `DRR`
keep nanodrr's signatures and reach hand-written HIP kernels for gfx950
through the C ABI of include/nanodrr_amd.h.
"""

from .camera import invert_rt_inv, make_k_inv, make_rt_inv
from .drr import DRR
from .geometry import convert, transform_point
from .renderer import render
from .subject import Subject, hu_to_mu

__all__ = [
    "DRR",
    "Subject",
    "convert",
    "hu_to_mu",
    "invert_rt_inv",
    "make_k_inv",
    "make_rt_inv",
    "render",
    "transform_point",
]
