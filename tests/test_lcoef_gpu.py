"""Labelled coefficient records ("lcoef", NDRR_LAYOUT_LCOEF) on the GPU.

* ndrr_pack_coef_labels records bit for bit against tests.helpers.lcoef_expected (the fp32
  coefficients of oracle.raymarch.coef_records at the labelled scale, rounded once, with the
  corners' class codes), header included; volumes or labels the records cannot hold are refused.
* The march (forward, single-pass and two-pass adjoints) against the float64 oracle evaluating the
  records it reads (decoded) with the volume's own labels: the kernels' arithmetic, and routing by
  the class codes in the records equal to nearest-label routing.
* The renderer picks the layout for 16-bit subjects with 2 to 4 classes and a volume that is not
  negative, and keeps "coef" otherwise.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from nanodrr_amd import _lib
from nanodrr_amd._ops import (fused_raymarch, fused_raymarch_bwd, fused_raymarch_bwd_jac, fused_raymarch_jac,
                              pack_coef_labels, pack_volume)
from oracle import raymarch as orm
from tests.helpers import (lcoef_decode, lcoef_expected, oracle_inputs, random_subject, random_volume, ref_camera_np,
                           rel_to_max)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
SRC_DTYPES = [torch.float32, torch.bfloat16, torch.float16]
REC_DTYPES = [torch.bfloat16, torch.float16]


def _labels(shape, C, seed, dropped=True):
    g = torch.Generator().manual_seed(seed)
    lo = -1 if dropped and C < 4 else 0
    return torch.randint(lo, C, shape, generator=g).to(torch.int8)


@pytest.mark.parametrize("src_dtype", SRC_DTYPES)
@pytest.mark.parametrize("rec_dtype", REC_DTYPES)
@pytest.mark.parametrize("C", [1, 2, 3, 4])
@pytest.mark.parametrize("magnitude", [1e-4, 3.0, 2e4])
def test_lcoef_records_bit_for_bit(src_dtype, rec_dtype, C, magnitude):
    torch.manual_seed(C)
    vol = (torch.randn(5, 6, 7).abs() * magnitude).to(src_dtype)
    vol[1, 2] = 0.0
    lab = _labels(vol.shape, C, 7)
    packed = pack_coef_labels(vol.to(DEV), lab.to(DEV), C, rec_dtype)
    expect, s, m = lcoef_expected(vol.float().numpy(), lab.numpy(), rec_dtype)
    hdr = packed[:24].cpu()
    f, u = hdr.view(torch.float32), hdr.view(torch.int32)
    assert float(f[0]) == m and float(f[1]) == s and float(f[2]) == 1.0 / s
    assert (int(u[3]), int(u[4]), int(u[5])) == ({torch.bfloat16: 1, torch.float16: 2}[rec_dtype], 0, C)
    got = packed[256:].cpu().view(torch.int16).reshape(expect.shape)
    assert torch.equal(got, expect)


def test_lcoef_refuses_what_the_records_cannot_hold():
    vol = torch.rand(4, 5, 6, device=DEV)
    lab = _labels(vol.shape, 3, 1).to(DEV)
    with pytest.raises(ValueError, match="negative"):
        pack_coef_labels(vol - 0.5, lab, 3, torch.bfloat16)
    with pytest.raises(ValueError, match="class"):
        pack_coef_labels(vol, lab, 2, torch.bfloat16)  # class 2 with two classes
    lab4 = _labels(vol.shape, 4, 2).to(DEV)
    pack_coef_labels(vol, lab4, 4, torch.float16)
    lab4[1, 2, 3] = -1
    with pytest.raises(ValueError, match="class"):
        pack_coef_labels(vol, lab4, 4, torch.float16)  # a dropped voxel with four classes
    with pytest.raises(ValueError, match="not finite"):
        bad = vol.clone()
        bad[0, 0, 0] = float("inf")
        pack_coef_labels(bad, lab, 3, torch.bfloat16)
    with pytest.raises(ValueError, match="bfloat16 or float16"):
        pack_coef_labels(vol, lab, 3, torch.float32)
    with pytest.raises(ValueError, match="at most 4"):
        pack_coef_labels(vol, lab, 5, torch.bfloat16)
    with pytest.raises(ValueError, match="pack_coef_labels"):
        pack_volume(vol, "lcoef")
    lib = _lib.lib()
    out = torch.empty(lib.ndrr_packed_bytes(4, 5, 6, _lib.NDRR_BF16, _lib.LAYOUT_LCOEF), dtype=torch.uint8, device=DEV)
    assert lib.ndrr_packed_bytes(4, 5, 6, _lib.NDRR_F32, _lib.LAYOUT_LCOEF) == 0
    assert lib.ndrr_pack_coef_labels(vol.data_ptr(), 0, lab.data_ptr(), 5, 1, 4, 5, 6, out.data_ptr(), None) == -3
    assert lib.ndrr_pack_coef_labels(vol.data_ptr(), 0, lab.data_ptr(), 3, 0, 4, 5, 6, out.data_ptr(), None) == -2
    assert lib.ndrr_pack_labels(lab.data_ptr(), _lib.LAYOUT_LCOEF, 4, 5, 6, out.data_ptr(), None) == -1


def _golden(C, S=200):
    image, label = random_volume(32, C, 0)
    label = label.clone()
    if C < 4:  # dropped voxels (class -1) travel as code 3
        label[..., ::5, :, :] = -1.0
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    o = oracle_inputs(image, label, C, k_inv, rt_inv, sdd, h, w, S)
    t = {k: torch.as_tensor(o[k], dtype=torch.float32, device=DEV).contiguous() for k in ("M", "src", "tgt", "step")}
    return image, label, o, t


@pytest.mark.parametrize("dtype", REC_DTYPES)
@pytest.mark.parametrize("C", [1, 2, 3, 4])
def test_lcoef_march_matches_oracle_of_its_records(dtype, C):
    """Forward, single-pass and two-pass pose and ray gradients against the float64 oracle of the
    decoded records, routed by the volume's nearest labels: 1e-5 / 1e-4 as for the other layouts."""
    image, label, o, t = _golden(C)
    vol = image[0, 0].to(DEV)
    lab = label[0, 0].to(torch.int8).to(DEV)
    packed = pack_coef_labels(vol, lab, C, dtype)
    vol_t = vol.to(dtype)
    none = lab.new_empty(0)
    rec, codes = lcoef_decode(packed, dtype, vol.shape)
    out = fused_raymarch(vol_t, none, t["M"], t["src"], t["tgt"], t["step"], 200, C, 16, packed, "lcoef")
    ref = orm.forward(**o, coef=rec)
    out = out.double().cpu().numpy()
    assert rel_to_max(ref.sum(1), out.sum(1)) < 1e-5
    # per class: routing by the records' codes is the oracle's nearest-label routing (no pixel may
    # differ beyond a single sample flipped at a tie by a ~1e-6 coordinate difference)
    assert ((np.abs(ref - out).max(axis=1) / np.abs(ref).max()) > 1e-5).mean() < 1e-3
    go = torch.randn(1, C, 256, generator=torch.Generator().manual_seed(3)).to(DEV)
    g = orm.backward(go.double().cpu().numpy(), **o, need_gvol=False, coef=rec)
    _, jac = fused_raymarch_jac(vol_t, none, t["M"], t["src"], t["tgt"], t["step"], 200, C, 16, 7, packed, "lcoef")
    gM1, gsrc1, gtgt1, gstep1 = fused_raymarch_bwd_jac(go, jac, t["M"], t["src"], t["tgt"], t["step"], 16, True)
    _, gM2, _, gtgt2, gstep2 = fused_raymarch_bwd(go, vol_t, none, t["M"], t["src"], t["tgt"], t["step"], 200, C, 16,
                                                  False, True, packed, "lcoef")
    for got in (gM1, gM2):
        assert rel_to_max(g["gM"], got.double().cpu().numpy()) < 1e-4
    for got in (gtgt1, gtgt2):
        assert rel_to_max(g["gtgt"], got.double().cpu().numpy()) < 1e-4
    for got in (gstep1, gstep2):
        assert rel_to_max(g["gstep"], got.double().cpu().numpy()) < 1e-4


def test_renderer_picks_labelled_records():
    """16-bit storage with 2-4 classes and a volume that is not negative renders from "lcoef" (no
    label map); a negative volume, one class, more than 4 classes or fp32 storage keep the others."""
    from nanodrr_amd import renderer

    def picked(subject):
        M = torch.eye(4, device=DEV)[None, :3, :].contiguous()
        tgt = torch.zeros(1, 16, 3, device=DEV)
        src = torch.zeros(1, 1, 3, device=DEV)
        step = torch.ones(1, 16, device=DEV)
        return renderer.march_inputs(subject, M, src, tgt, step, 8, 4, 4)[2]

    for C, dt, shift, want in ((4, torch.bfloat16, 0.0, "lcoef"), (3, torch.float16, 0.0, "lcoef"),
                               (2, torch.bfloat16, -0.5, "coef"), (1, torch.bfloat16, 0.0, "coef"),
                               (6, torch.bfloat16, 0.0, "coef"), (4, torch.float32, 0.0, "quad")):
        sub = random_subject(16, C, 0, device=DEV)
        sub._image_hu.add_(shift)
        sub = sub.with_volume_dtype(dt) if dt != torch.float32 else sub
        assert picked(sub) == want, (C, dt, shift)
