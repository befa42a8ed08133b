"""CT subject: attenuation volume, optional label map and their affines.

Mirrors nanodrr's `Subject` (src/nanodrr/data/subject.py:43-242) and
`hu_to_mu` (src/nanodrr/data/preprocess.py:9-50): same constructor, buffers,
lazily cached HU -> mu image and `align_corners=False` voxel-to-grid affine.
File loading through torchio (subject.py:139-209) is replaced by
`Subject.from_volume`, which builds the affines of an axis-aligned volume from
its spacing. The kernels read the label map as int8; `label_i8` caches that
copy per device.
"""

from __future__ import annotations

import torch

MU_WATER = 0.0192
MU_BONE = 0.0573
HU_BONE = 1000.0

_MU_FIELDS = frozenset({"convert_to_mu", "mu_water", "mu_bone", "hu_bone"})


def hu_to_mu(data, mu_water=MU_WATER, mu_bone=MU_BONE, hu_bone=HU_BONE):
    """Piecewise-linear HU -> linear attenuation (1/mm): air-water below 0 HU,
    water-bone above, after clamping HU >= -1000; result clamped >= 0 (preprocess.py:44-50)."""
    hu = data.clamp(min=-1000.0)
    mu_low = mu_water * (1.0 + hu / 1000.0)
    mu_high = mu_water + (hu / hu_bone) * (mu_bone - mu_water)
    return torch.where(hu <= 0, mu_low, mu_high).clamp(min=0.0)


class _CachedParam:
    """Scalar that invalidates the cached mu image when it changes (subject.py:18-40)."""

    def __init__(self, default) -> None:
        self.default = default

    def __set_name__(self, owner, name) -> None:
        self.name = f"_{name}"

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self
        return getattr(obj, self.name, self.default)

    def __set__(self, obj, value) -> None:
        if getattr(obj, self.name, self.default) != value:
            setattr(obj, self.name, value)
            if hasattr(obj, "_image_mu_cache"):
                obj._image_mu_cache = None


class Subject(torch.nn.Module):
    convert_to_mu: bool = _CachedParam(True)  # type: ignore[assignment]
    mu_water: float = _CachedParam(MU_WATER)  # type: ignore[assignment]
    mu_bone: float = _CachedParam(MU_BONE)  # type: ignore[assignment]
    hu_bone: float = _CachedParam(HU_BONE)  # type: ignore[assignment]

    def __init__(self, imagedata, labeldata, voxel_to_world, world_to_voxel, voxel_to_grid, isocenter,
                 max_label=None, **mu) -> None:
        unknown = mu.keys() - _MU_FIELDS
        if unknown:
            raise TypeError(f"Unknown mu param(s): {', '.join(repr(k) for k in unknown)}")
        super().__init__()
        self.register_buffer("_image_hu", imagedata)
        self.register_buffer("_image_mu_cache", None, persistent=False)
        self.register_buffer("label", labeldata)
        self.register_buffer("world_to_grid", voxel_to_grid @ world_to_voxel)
        self.register_buffer("isocenter", isocenter)
        self.register_buffer("voxel_to_world", voxel_to_world)
        self.register_buffer("world_to_voxel", world_to_voxel)
        self.register_buffer("voxel_to_grid", voxel_to_grid)
        for k, v in mu.items():
            setattr(self, k, v)
        self.n_classes = int(max_label + 1) if max_label is not None else int(self.label.max().item()) + 1
        self._label_i8_key = None
        self._label_i8 = None
        self._packed_key = None
        self._packed = None
        self._packed_lab_key = None
        self._packed_lab = None

    # -- volume ---------------------------------------------------------------
    @property
    def image(self) -> torch.Tensor:
        """(1, 1, D, H, W) attenuation (or raw values when convert_to_mu is False), cached."""
        if not self.convert_to_mu:
            return self._image_hu
        if self._image_mu_cache is None:
            self._image_mu_cache = hu_to_mu(self._image_hu, self.mu_water, self.mu_bone, self.hu_bone)
        return self._image_mu_cache

    def set_mu(self, mu_water=None, mu_bone=None, hu_bone=None) -> None:
        """Recompute the mu cache from new conversion parameters (subject.py:118-137)."""
        self._image_mu_cache = hu_to_mu(
            self._image_hu,
            mu_water if mu_water is not None else self.mu_water,
            mu_bone if mu_bone is not None else self.mu_bone,
            hu_bone if hu_bone is not None else self.hu_bone,
        )

    def to(self, *args, **kwargs) -> "Subject":
        had_cache = self._image_mu_cache is not None
        result = super().to(*args, **kwargs)
        if had_cache:
            result._image_mu_cache = hu_to_mu(result._image_hu, result.mu_water, result.mu_bone, result.hu_bone)
        result._label_i8_key = None
        result._label_i8 = None
        result._packed_key = None
        result._packed = None
        result._packed_lab_key = None
        result._packed_lab = None
        return result

    def with_volume_dtype(self, dtype: torch.dtype) -> "Subject":
        """Store the attenuation volume in `dtype` (e.g. bfloat16) and keep the
        affines and labels untouched (torch's `.to(dtype)` would also round the
        affines, costing sampling accuracy)."""
        image = self.image.detach().to(dtype)
        out = Subject(image, self.label, self.voxel_to_world, self.world_to_voxel, self.voxel_to_grid,
                      self.isocenter, max_label=self.n_classes - 1, convert_to_mu=False)
        return out

    def world_to_voxel_f64(self) -> torch.Tensor:
        """`world_to_voxel` in float64 (the camera -> voxel composition runs in FP64),
        cached until the buffer changes."""
        w = self.world_to_voxel
        if w.dtype == torch.float64 or torch.compiler.is_compiling():
            return w.to(torch.float64).contiguous()
        key = (w.data_ptr(), w.device, w._version)
        if getattr(self, "_w2v64_key", None) != key:
            self._w2v64 = w.to(torch.float64).contiguous()
            self._w2v64_key = key
        return self._w2v64

    def packed_volume(self, layout: str) -> torch.Tensor:
        """`image` re-laid out for the kernels ("quad" or "stencil", see
        ndrr_pack_volume), cached until the image tensor or the layout changes."""
        from ._ops import pack_volume

        img = self.image
        if torch.compiler.is_compiling():  # traced graphs re-pack (the cache is Python state)
            return pack_volume(img[0, 0], layout)
        key = (img.data_ptr(), img.device, img.dtype, img._version, tuple(img.shape), layout)
        if self._packed_key != key:
            self._packed = None  # free the old copy first
            self._packed = pack_volume(img[0, 0], layout)
            self._packed_key = key
        return self._packed

    # -- labels ---------------------------------------------------------------
    def packed_labels(self, layout: str) -> torch.Tensor:
        """`label_i8` in the form `layout`'s march reads (label quads for "quad",
        see ndrr_pack_labels), cached until the labels or the layout change."""
        from ._ops import pack_labels

        if layout != "quad":
            return self.label_i8
        lab = self.label_i8
        if torch.compiler.is_compiling():
            return pack_labels(lab, layout)
        key = (lab.data_ptr(), self._label_i8_key, layout)
        if self._packed_lab_key != key:
            self._packed_lab = None
            self._packed_lab = pack_labels(lab, layout)
            self._packed_lab_key = key
        return self._packed_lab

    @property
    def label_i8(self) -> torch.Tensor:
        """(D, H, W) int8 class map used by the kernels: trunc(label) when it
        lies in [0, n_classes), else -1 (dropped). Cached until label changes."""
        from ._ops import labels_to_i8

        lab = self.label
        if torch.compiler.is_compiling():  # traced graphs convert (the cache is Python state)
            return labels_to_i8(lab[0, 0], self.n_classes)
        key = (lab.data_ptr(), lab.device, lab._version, tuple(lab.shape), self.n_classes)
        if self._label_i8_key != key:
            self._label_i8 = labels_to_i8(lab[0, 0], self.n_classes)
            self._label_i8_key = key
        return self._label_i8

    # -- construction -----------------------------------------------------------
    @staticmethod
    def _make_voxel_to_grid(shape) -> torch.Tensor:
        """Voxel index -> grid_sample(align_corners=False) coordinate: 2 i / S + 1 / S - 1 (subject.py:218-242)."""
        *_, D, H, W = shape
        size = torch.tensor([W, H, D], dtype=torch.float32)
        mat = torch.eye(4, dtype=torch.float32)
        mat[:3, :3] = torch.diag(2.0 / size)
        mat[:3, 3] = 1.0 / size - 1.0
        return mat

    @classmethod
    def from_volume(cls, image, spacing=(1.0, 1.0, 1.0), label=None, origin=None, max_label=None, **mu) -> "Subject":
        """Subject from a (D, H, W) or (1, 1, D, H, W) volume on an axis-aligned grid.

        `spacing` is (sx, sy, sz) in mm along (W, H, D). The world origin is the
        volume centre unless `origin` (world position of voxel (0, 0, 0)) is
        given; the isocentre is the volume centre.
        """
        image = image.reshape(1, 1, *image.shape[-3:])
        D, H, W = image.shape[-3:]
        sx, sy, sz = (float(s) for s in spacing)
        if origin is None:
            origin = (-(W - 1) * sx / 2.0, -(H - 1) * sy / 2.0, -(D - 1) * sz / 2.0)
        v2w = torch.eye(4, dtype=torch.float64)
        v2w[0, 0], v2w[1, 1], v2w[2, 2] = sx, sy, sz
        v2w[:3, 3] = torch.tensor(origin, dtype=torch.float64)
        w2v = torch.linalg.inv(v2w)
        centre = v2w @ torch.tensor([(W - 1) / 2.0, (H - 1) / 2.0, (D - 1) / 2.0, 1.0], dtype=torch.float64)
        if label is None:
            labeldata = torch.zeros(1, 1, D, H, W, dtype=torch.float32, device=image.device)
        else:
            labeldata = label.reshape(1, 1, D, H, W).to(torch.float32)
        dev = image.device
        return cls(image, labeldata, v2w.float().to(dev), w2v.float().to(dev),
                   cls._make_voxel_to_grid(image.shape).to(dev), centre[:3].float().to(dev),
                   max_label=max_label, **mu)
