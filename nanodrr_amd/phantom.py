"""Seeded synthetic CT phantoms and pose batches for the benchmark workloads.

The phantom (no counterpart in nanodrr, whose benchmark loads a CT file):
an ellipsoidal soft-tissue body (40 HU) holding two lung ellipsoids
(-800 HU) and a vertical column of 12 stacked vertebral ellipsoids (1000 HU
cortical shell around a 300 HU core), in air (-1000 HU). Gaussian noise
(sigma 20 HU) is added inside the body and the result is clipped to
[-1000, 2000] HU; the background stays noise-free air. Attenuation is
mu = 0.02 (HU + 1000) / 1000 per mm. Labels: 0 background, 1 body, 2 lungs,
3 bones. Shapes are fractions of the field of view, so any grid holds the
same anatomy; each seed jitters centres and semi-axes by up to 5 %.

Poses follow nanodrr's benchmark set-up (tests/benchmark/benchmark.py:27-38):
`make_rt_inv(rotation, (dx, distance + dy, dz), "AP", isocenter)` with ZXY
Euler rotations drawn uniformly in +-rot_deg and offsets in +-trans_mm.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from .camera import make_k_inv, make_rt_inv
from .subject import Subject

MU_PER_HU = 0.02 / 1000.0  # mu = 0.02 (HU + 1000) / 1000 1/mm


def hu_to_mu_linear(hu: torch.Tensor) -> torch.Tensor:
    return (hu + 1000.0) * MU_PER_HU


@dataclass
class Phantom:
    hu: torch.Tensor      # (D, H, W) float32
    labels: torch.Tensor  # (D, H, W) int8
    spacing: tuple[float, float, float]

    @property
    def mu(self) -> torch.Tensor:
        return hu_to_mu_linear(self.hu)


def make_phantom(shape=(256, 256, 256), spacing=(1.0, 1.0, 1.0), seed: int = 0, noise_hu: float = 20.0,
                 jitter: float = 0.05) -> Phantom:
    """Phantom on a (D, H, W) grid with `spacing` = (sx, sy, sz) mm; generated on the CPU
    with a seeded torch.Generator so every device sees the same volume."""
    D, H, W = shape
    sx, sy, sz = spacing
    g = torch.Generator().manual_seed(int(seed))

    def u(n=1):
        return (torch.rand(n, generator=g, dtype=torch.float64) * 2 - 1) * jitter

    fov = torch.tensor([W * sx, H * sy, D * sz], dtype=torch.float64)
    z = (torch.arange(D, dtype=torch.float32) - (D - 1) / 2) * sz
    y = (torch.arange(H, dtype=torch.float32) - (H - 1) / 2) * sy
    x = (torch.arange(W, dtype=torch.float32) - (W - 1) / 2) * sx
    Z, Y, X = z[:, None, None], y[None, :, None], x[None, None, :]

    def ellipsoid(centre_frac, axes_frac):
        c = (torch.tensor(centre_frac, dtype=torch.float64) + u(3) * 0.5) * fov
        a = torch.tensor(axes_frac, dtype=torch.float64) * (1 + u(3)) * fov
        return ((X - c[0]) / a[0]) ** 2 + ((Y - c[1]) / a[1]) ** 2 + ((Z - c[2]) / a[2]) ** 2

    hu = torch.full((D, H, W), -1000.0)
    labels = torch.zeros((D, H, W), dtype=torch.int8)

    body = ellipsoid((0.0, 0.0, 0.0), (0.40, 0.30, 0.45)) <= 1.0
    hu[body] = 40.0
    labels[body] = 1
    for side in (-1.0, 1.0):
        lung = ellipsoid((side * 0.17, -0.03, 0.10), (0.12, 0.16, 0.22)) <= 1.0
        lung &= body
        hu[lung] = -800.0
        labels[lung] = 2
    n_vert = 12
    for k in range(n_vert):
        zc = -0.38 + 0.76 * k / (n_vert - 1)
        r2 = ellipsoid((0.0, 0.19, zc), (0.065, 0.055, 0.028))
        shell = r2 <= 1.0
        core = r2 <= 0.49  # inner ellipsoid at 70 % of the semi-axes
        hu[shell] = 1000.0
        hu[core] = 300.0
        labels[shell] = 3
    noise = torch.randn((D, H, W), generator=g) * noise_hu
    hu = torch.where(labels > 0, hu + noise, hu).clamp(-1000.0, 2000.0)
    return Phantom(hu=hu, labels=labels, spacing=(float(sx), float(sy), float(sz)))


def phantom_subject(ph: Phantom, n_classes: int = 1, dtype=torch.float32, device="cpu") -> Subject:
    """Subject holding mu (convert_to_mu=False), stored in `dtype` (rounded from the fp32
    values, which the subject keeps), and, for n_classes > 1, the label map (labels >=
    n_classes fold into the last class)."""
    label = None
    if n_classes > 1:
        label = ph.labels.to(torch.float32).clamp(max=n_classes - 1)
    subject = Subject.from_volume(ph.mu, spacing=ph.spacing, label=label, max_label=n_classes - 1,
                                  convert_to_mu=False)
    if dtype != torch.float32:
        subject = subject.with_volume_dtype(dtype)
    return subject.to(device)


def sample_poses(batch: int, seed: int, rot_deg: float = 15.0, trans_mm: float = 20.0, distance: float = 600.0,
                 isocenter: torch.Tensor | None = None, dtype=torch.float32):
    """(rt_inv (B, 4, 4), rotation (B, 3) degrees, translation (B, 3) mm), AP about the isocentre."""
    g = torch.Generator().manual_seed(int(seed))
    rot = (torch.rand(batch, 3, generator=g, dtype=torch.float64) * 2 - 1) * rot_deg
    off = (torch.rand(batch, 3, generator=g, dtype=torch.float64) * 2 - 1) * trans_mm
    trans = off + torch.tensor([0.0, distance, 0.0], dtype=torch.float64)
    iso = None if isocenter is None else isocenter.detach().cpu().to(torch.float64)
    rt_inv = make_rt_inv(rot, trans, "AP", isocenter=iso)
    return rt_inv.to(dtype), rot.to(dtype), trans.to(dtype)


def detector(sdd: float = 1020.0, spacing: float = 1.0, height: int = 200, width: int = 200, dtype=torch.float32):
    """(k_inv (1, 3, 3), sdd (1,)) for a centred pinhole detector (make_k_inv with x0 = y0 = 0)."""
    return make_k_inv(sdd, spacing, spacing, 0.0, 0.0, height, width, dtype=dtype), torch.tensor([sdd], dtype=dtype)


@dataclass
class RegistrationProblems:
    rt_true: torch.Tensor     # (B, 4, 4) float64 camera-to-world target poses
    rt_init: torch.Tensor     # (B, 4, 4) float64 perturbed starting poses
    fiducials: torch.Tensor   # (K, 3) float64 world points for mTRE
    k_inv: torch.Tensor       # (1, 3, 3)
    sdd: torch.Tensor         # (1,)
    height: int
    width: int


def registration_problems(subject: Subject, batch: int, seed: int, target_rot_deg: float = 30.0,
                          target_trans_mm: float = 40.0, start_rot_deg: float = 10.0, start_trans_mm: float = 15.0,
                          n_fiducials: int = 32, distance: float = 600.0, sdd: float = 1020.0, height: int = 200,
                          width: int = 200) -> RegistrationProblems:
    """Batched 2D/3D registration problems on `subject`.

    Targets: ZXY Euler rotations in +-target_rot_deg and offsets in
    +-target_trans_mm about the isocentre at `distance` (as sample_poses).
    Starts: each target composed with a rigid perturbation about the
    isocentre, rotation vector in +-start_rot_deg per axis and translation
    in +-start_trans_mm per axis. Fiducials: uniform in the central 80 % of
    the volume's world bounding box.
    """
    from .geometry import so3_exp

    iso = subject.isocenter.detach().cpu().to(torch.float64)
    rt_true, _, _ = sample_poses(batch, seed, target_rot_deg, target_trans_mm, distance, iso, torch.float64)
    g = torch.Generator().manual_seed(int(seed) + 7919)
    w = (torch.rand(batch, 3, generator=g, dtype=torch.float64) * 2 - 1) * (start_rot_deg * torch.pi / 180.0)
    t = (torch.rand(batch, 3, generator=g, dtype=torch.float64) * 2 - 1) * start_trans_mm
    P = torch.eye(4, dtype=torch.float64).repeat(batch, 1, 1)
    R = so3_exp(w)
    P[:, :3, :3] = R
    P[:, :3, 3] = t + iso - (R @ iso[:, None])[..., 0]  # rotate about the isocentre, then shift
    rt_init = P @ rt_true
    fid = fiducials(subject, n_fiducials, g)
    k_inv, sdd_t = detector(sdd, 1.0, height, width, dtype=torch.float64)
    return RegistrationProblems(rt_true, rt_init, fid, k_inv, sdd_t, height, width)


def fiducials(subject: Subject, n: int, generator: torch.Generator) -> torch.Tensor:
    """(n, 3) float64 world points uniform in the central 80 % of the volume's box."""
    v2w = subject.voxel_to_world.detach().cpu().to(torch.float64)
    D, H, W = subject.image.shape[-3:]
    frac = torch.rand(n, 3, generator=generator, dtype=torch.float64) * 0.8 + 0.1
    idx = frac * torch.tensor([W - 1, H - 1, D - 1], dtype=torch.float64)
    return (v2w[:3, :3] @ idx.T).T + v2w[:3, 3]


def nominal_pose(subject: Subject, distance: float = 600.0, dtype=torch.float64) -> torch.Tensor:
    """(1, 4, 4) AP camera-to-world pose with no rotation or offset, isocentre `distance` from the source."""
    iso = subject.isocenter.detach().cpu().to(torch.float64)
    rt = make_rt_inv(torch.zeros(1, 3, dtype=torch.float64), torch.tensor([[0.0, distance, 0.0]], dtype=torch.float64),
                     "AP", isocenter=iso)
    return rt.to(dtype)
