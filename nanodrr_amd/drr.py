"""`DRR` module: a camera's intrinsics as module state, `forward(subject, rt_inv)` renders.

Same constructor, `forward` and `from_carm_intrinsics` as nanodrr's module
(src/nanodrr/drr/drr.py:9-106), with `k_inv` and `sdd` as persistent buffers
so state_dicts load in either package. The detector rays (`src`, `tgt` and
the per-ray step) are made on first use after any intrinsic changes
(including a device move of the buffers or an in-place edit) rather than at
every assignment: in one native launch for fp32 device intrinsics, as
`render` makes them, so `forward` renders bit for bit as `render` does. They
are not kept while `k_inv` or `sdd` require gradients: each render then
builds them inside the autograd graph.
"""

from __future__ import annotations

import torch

from ._memo import Derived
from ._ops import eager
from .camera import make_k_inv
from .renderer import _make_src, _make_tgt, _native_rays, check_render_args, render, render_hip

_DEFAULT_SAMPLES = 500  # render's default (renderer.py:28), for the `src` / `tgt` attributes


class DRR(torch.nn.Module):
    def __init__(self, k_inv: torch.Tensor, sdd: torch.Tensor, height: int, width: int, orthographic: bool = False):
        super().__init__()
        self._rays = Derived()
        self.register_buffer("k_inv", k_inv)
        self.register_buffer("sdd", sdd)
        self.height = int(height)
        self.width = int(width)
        self.orthographic = bool(orthographic)

    def _cacheable(self) -> bool:
        return not (self.k_inv.requires_grad or self.sdd.requires_grad or torch.compiler.is_compiling())

    def _ray_set(self, n_samples: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """(src, tgt, step) of the current intrinsics for n_samples samples per ray."""
        shape = (self.height, self.width, self.orthographic)

        def make():
            if _native_rays(self.k_inv, self.sdd, self.k_inv.dtype):
                tgt, src, step = eager.make_rays(self.k_inv, self.sdd, self.height, self.width, self.orthographic,
                                                 n_samples)
                return src, tgt, step
            tgt = _make_tgt(self.k_inv, self.sdd, self.height, self.width, self.k_inv.device, self.k_inv.dtype)
            src = _make_src(self.orthographic, tgt, self.sdd)
            return src, tgt, (tgt - src).norm(dim=-1) / float(n_samples - 1)

        if not self._cacheable():
            return make()
        # keyed on both buffers: the entry on k_inv holds a table keyed on sdd, one entry per sample count
        inner = self._rays.get("k_inv", self.k_inv, shape, Derived)
        return inner.get(f"sdd:{n_samples}", self.sdd, None, make)

    @property
    def src(self) -> torch.Tensor:
        """(B|1, 1, 3) perspective sources, or (B, H*W, 3) orthographic ones."""
        return self._ray_set(_DEFAULT_SAMPLES)[0]

    @property
    def tgt(self) -> torch.Tensor:
        """(B, H*W, 3) detector pixel centres in camera space."""
        return self._ray_set(_DEFAULT_SAMPLES)[1]

    def _apply(self, fn, *args, **kwargs):
        self._rays.drop()
        return super()._apply(fn, *args, **kwargs)

    def forward(self, subject, rt_inv: torch.Tensor, n_samples: int = _DEFAULT_SAMPLES, backend: str = "auto"):
        check_render_args(n_samples, backend)
        if not self._cacheable():
            return render(subject, self.k_inv, rt_inv, self.sdd, self.height, self.width, n_samples,
                          self.orthographic, None, None, backend)
        src, tgt, step = self._ray_set(n_samples)
        return render_hip(subject, rt_inv, src, tgt, step, n_samples, self.height, self.width)

    @classmethod
    def from_carm_intrinsics(cls, sdd, delx, dely, x0, y0, height, width, orthographic=False, dtype=None,
                             device=None) -> "DRR":
        """A DRR of a C-arm with source-to-detector distance `sdd`, pixel spacings (delx, dely)
        and principal-point offsets (x0, y0) (drr.py:90-106)."""
        k_inv = make_k_inv(sdd, delx, dely, x0, y0, height, width).to(dtype=dtype, device=device)
        return cls(k_inv, torch.full((1,), float(sdd), dtype=dtype, device=device), height, width, orthographic)
