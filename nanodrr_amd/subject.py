"""CT subject: attenuation volume, optional label map and the affines between
world, voxel-index and grid_sample coordinates.

The public surface is nanodrr's `Subject` (src/nanodrr/data/subject.py:43-242)
and `hu_to_mu` (src/nanodrr/data/preprocess.py:9-50): the constructor
arguments, the buffer names (so a state_dict loads in either package), the
HU -> mu parameters whose change recomputes `image`, `set_mu`, `n_classes` and
the `align_corners=False` voxel-to-grid affine. File loading through torchio
(subject.py:139-209) is replaced by `Subject.from_volume`, which builds the
affines of an axis-aligned volume from its spacing.

Everything the kernels read is derived from those buffers on first use and
kept in one `Derived` table (nanodrr_amd._memo): the mu image, the float64
world -> voxel affine, the int8 class map and the packed volume and label
layouts. An entry stays valid while its source is the same tensor object at
the same version, so replacing a buffer, an in-place edit, a device move or a
change of the HU -> mu parameters makes the next read rebuild it.
"""

from __future__ import annotations

import torch

from ._memo import Derived

MU_WATER = 0.0192
MU_BONE = 0.0573
HU_BONE = 1000.0

_MU_FIELDS = ("convert_to_mu", "mu_water", "mu_bone", "hu_bone")
_MU_DEFAULTS = {"convert_to_mu": True, "mu_water": MU_WATER, "mu_bone": MU_BONE, "hu_bone": HU_BONE}


def hu_to_mu(data, mu_water=MU_WATER, mu_bone=MU_BONE, hu_bone=HU_BONE):
    """Hounsfield units -> linear attenuation (1/mm) (preprocess.py:9-50).

    HU is clamped at -1000 (air). The map is linear on each side of water
    (0 HU -> mu_water): air to water below, mu_water / 1000 per HU; water to
    bone above, (mu_bone - mu_water) / hu_bone per HU. The result is clamped
    at 0.
    """
    hu = torch.clamp(data, min=-1000.0)
    below = mu_water * (1.0 + hu / 1000.0)
    above = mu_water + (mu_bone - mu_water) * (hu / hu_bone)
    return torch.clamp(torch.where(hu > 0, above, below), min=0.0)


class _Same:
    """A cache key part equal only to a key part holding the same object (which it keeps alive)."""

    __slots__ = ("obj",)

    def __init__(self, obj) -> None:
        self.obj = obj

    def __eq__(self, other) -> bool:
        return isinstance(other, _Same) and other.obj is self.obj

    __hash__ = object.__hash__


def _mu_property(name: str):
    """A HU -> mu parameter: changing its value drops the derived mu image (and
    what was made from it) and any `set_mu` override."""
    slot = "_" + name

    def fget(self):
        return self.__dict__.get(slot, _MU_DEFAULTS[name])

    def fset(self, value):
        if self.__dict__.get(slot, _MU_DEFAULTS[name]) != value:
            self.__dict__[slot] = value
            self._mu_changed()

    return property(fget, fset, doc=f"HU -> mu parameter `{name}`; setting a new value recomputes `image`.")


class Subject(torch.nn.Module):
    convert_to_mu = _mu_property("convert_to_mu")
    mu_water = _mu_property("mu_water")
    mu_bone = _mu_property("mu_bone")
    hu_bone = _mu_property("hu_bone")

    def __init__(self, imagedata, labeldata, voxel_to_world, world_to_voxel, voxel_to_grid, isocenter,
                 max_label=None, *, storage_dtype: torch.dtype | None = None, **mu) -> None:
        bad = sorted(set(mu) - set(_MU_FIELDS))
        if bad:
            raise TypeError(f"Unknown mu param(s): {', '.join(repr(k) for k in bad)}")
        super().__init__()
        self._derived = Derived()
        self.storage_dtype = storage_dtype  # None: `image` keeps the volume's own type
        self._mu_override = None  # (token, mu image) after set_mu, until a parameter changes
        for name, t in (("_image_hu", imagedata), ("label", labeldata),
                        ("world_to_grid", voxel_to_grid @ world_to_voxel), ("isocenter", isocenter),
                        ("voxel_to_world", voxel_to_world), ("world_to_voxel", world_to_voxel),
                        ("voxel_to_grid", voxel_to_grid)):
            self.register_buffer(name, t)
        for name in _MU_FIELDS:
            if name in mu:
                setattr(self, name, mu[name])
        top = max_label if max_label is not None else self.label.max().item()
        self.n_classes = int(top) + 1

    # -- HU -> mu ---------------------------------------------------------------
    def _mu_changed(self) -> None:
        self._mu_override = None
        self._derived.drop("mu", "stored")
        self._derived.drop_prefix("volume:")

    def precise_image(self) -> torch.Tensor:
        """(1, 1, D, H, W) attenuation before any rounding to `storage_dtype` (the values the
        coefficient layout's records are made from)."""
        if not self.convert_to_mu:
            return self._image_hu
        if self._mu_override is not None:
            return self._mu_override[1]
        params = (self.mu_water, self.mu_bone, self.hu_bone)
        if torch.compiler.is_compiling() or self._image_hu.requires_grad:  # no cached autograd graphs
            return hu_to_mu(self._image_hu, *params)
        return self._derived.get("mu", self._image_hu, params, lambda: hu_to_mu(self._image_hu, *params))

    @property
    def image(self) -> torch.Tensor:
        """(1, 1, D, H, W) attenuation (the stored values when convert_to_mu is False), in
        `storage_dtype` when one is set."""
        base = self.precise_image()
        dt = self.storage_dtype
        if dt is None or base.dtype == dt:
            return base
        if torch.compiler.is_compiling() or base.requires_grad:
            return base.to(dt)
        return self._derived.get("stored", base, dt, lambda: base.to(dt))

    def set_mu(self, mu_water=None, mu_bone=None, hu_bone=None) -> None:
        """Render with other HU -> mu parameters (floats or tensors, e.g. ones being
        optimised) without changing the stored ones (subject.py:118-137). The image
        made here stays in use until a stored parameter changes or the subject moves."""
        args = (self.mu_water if mu_water is None else mu_water,
                self.mu_bone if mu_bone is None else mu_bone,
                self.hu_bone if hu_bone is None else hu_bone)
        self._derived.drop("stored")
        self._derived.drop_prefix("volume:")
        self._mu_override = (object(), hu_to_mu(self._image_hu, *args))

    def _apply(self, fn, *args, **kwargs):
        # .to(), .cuda(), .half(), ...: buffers become new tensors, so every derived
        # tensor (which still references the old ones) is dropped; a set_mu override
        # does not survive the move (the stored parameters apply again).
        self._derived.drop()
        self._mu_override = None
        return super()._apply(fn, *args, **kwargs)

    def with_volume_dtype(self, dtype: torch.dtype) -> "Subject":
        """The same subject with `image` stored in `dtype` (e.g. bfloat16): it shares the
        volume, labels and affines (torch's `.to(dtype)` would also round the affines, costing
        sampling accuracy) and rounds the attenuation once, on first use. The unrounded values
        stay available (`precise_image`) for the coefficient layout."""
        mu = {name: getattr(self, name) for name in _MU_FIELDS}
        out = Subject(self._image_hu, self.label, self.voxel_to_world, self.world_to_voxel, self.voxel_to_grid,
                      self.isocenter, max_label=self.n_classes - 1, storage_dtype=dtype, **mu)
        if self._mu_override is not None:
            out._mu_override = self._mu_override
        return out

    # -- what the kernels read ----------------------------------------------------
    def world_to_voxel_f64(self) -> torch.Tensor:
        """`world_to_voxel` in float64 (the camera -> voxel composition runs in FP64)."""
        w = self.world_to_voxel
        if w.dtype == torch.float64 or torch.compiler.is_compiling():
            return w.to(torch.float64).contiguous()
        return self._derived.get("w2v64", w, None, lambda: w.to(torch.float64).contiguous())

    def packed_volume(self, layout: str) -> torch.Tensor:
        """`image` re-laid out for the kernels ("quad", "stencil", "coef", "lcoef" or "qcoef", see
        ndrr_pack_volume). Coefficient records of a 16-bit `image` are made from `precise_image` and
        rounded once (ndrr_pack_coef), so they carry the unrounded volume's edge differences; the
        labelled ones ("lcoef", ndrr_pack_coef_labels) also carry `label_i8` (ValueError when the
        volume or classes do not fit them: a negative value, more than 4 classes, fp32 storage)."""
        from ._ops import pack_coef, pack_volume

        if layout == "lcoef":
            value = self._labelled_coef()
            if isinstance(value, ValueError):
                raise ValueError(str(value))
            return value
        img = self.image
        if layout == "qcoef":  # fp16 records from the unrounded values
            src = self.precise_image()

            def make_q():
                return pack_volume(src[0, 0].detach().to(torch.float32), "qcoef")

            if torch.compiler.is_compiling():
                return make_q()
            return self._derived.get("volume:qcoef", src, None, make_q)
        if layout == "coef":
            src = self.precise_image()
            if src is not img:
                def make():
                    return pack_coef(src[0, 0].detach().to(torch.float32), img.dtype)

                if torch.compiler.is_compiling():
                    return make()
                return self._derived.get("volume:coef", src, img.dtype, make)
        if torch.compiler.is_compiling():  # traced graphs re-pack
            return pack_volume(img[0, 0], layout)
        return self._derived.get("volume:" + layout, img, None, lambda: pack_volume(img[0, 0], layout))

    def _labelled_coef(self):
        """Labelled coefficient records of (precise_image, label_i8), or the ValueError saying why
        they cannot hold them; either is kept until the volume or labels change."""
        from ._ops import pack_coef_labels

        img, src, lab, C = self.image, self.precise_image(), self.label_i8, self.n_classes

        def make():
            try:
                if img.dtype not in (torch.bfloat16, torch.float16):
                    raise ValueError(f"labelled coefficient records are 16-bit; the volume is stored in {img.dtype}")
                if C > 4:
                    raise ValueError(f"labelled coefficient records hold at most 4 classes, not {C}")
                return pack_coef_labels(src[0, 0].detach().to(torch.float32), lab, C, img.dtype)
            except ValueError as e:
                return e

        if torch.compiler.is_compiling():
            return make()
        return self._derived.get("volume:lcoef", src, (img.dtype, C, _Same(lab), lab._version), make)

    def labelled_coef_fits(self) -> bool:
        """Whether labelled coefficient records ("lcoef") can hold this subject (packs them once)."""
        return not isinstance(self._labelled_coef(), ValueError)

    @property
    def label_i8(self) -> torch.Tensor:
        """(D, H, W) int8 class map of the kernels: trunc(label) when it lies in
        [0, n_classes), else -1 (dropped)."""
        from ._ops import labels_to_i8

        lab = self.label
        if torch.compiler.is_compiling():
            return labels_to_i8(lab[0, 0], self.n_classes)
        return self._derived.get("label_i8", lab, self.n_classes, lambda: labels_to_i8(lab[0, 0], self.n_classes))

    def packed_labels(self, layout: str) -> torch.Tensor:
        """`label_i8` in the form `layout`'s march reads (label quads for "quad", the padded map
        for "stencil" and "coef", see ndrr_pack_labels; the map itself for "linear")."""
        from ._ops import pack_labels

        if layout == "linear":
            return self.label_i8
        lab = self.label_i8
        if torch.compiler.is_compiling():
            return pack_labels(lab, layout)
        return self._derived.get("labels:" + layout, lab, None, lambda: pack_labels(lab, layout))

    # -- construction ---------------------------------------------------------------
    @staticmethod
    def _make_voxel_to_grid(shape) -> torch.Tensor:
        """Voxel index -> grid_sample(align_corners=False) coordinate per axis (x: W, y: H,
        z: D): g = (2 i + 1) / n - 1, so the voxel extent [-1/2, n - 1/2] spans [-1, 1]
        (subject.py:218-242)."""
        *_, D, H, W = shape
        n = torch.tensor([W, H, D], dtype=torch.float32)
        out = torch.eye(4, dtype=torch.float32)
        out[0, 0], out[1, 1], out[2, 2] = 2.0 / n
        out[:3, 3] = 1.0 / n - 1.0
        return out

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
