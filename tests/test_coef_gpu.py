"""The coefficient layout ("coef", NDRR_LAYOUT_COEF) on the GPU.

* ndrr_pack_coef records bit for bit against oracle.raymarch.coef_records (fp32
  arithmetic in the documented order, rounded once), for every source and
  record type.
* The march (forward, single-pass and two-pass adjoints) against the float64
  oracle evaluating the same records: the kernels' own arithmetic.
* 16-bit records made from the fp32 volume against the float64 oracle of that
  fp32 volume at the benchmark workloads: what 2-byte storage costs, per dtype,
  for the image and for the pose gradient under three upstream gradients (a
  random one per pixel and class, the image sum, and the ZNCC gradient against
  a target rendered at another pose), on the rays free of fp32/float64
  discontinuity flips (see _accuracy) and, printed, on all rays. The printed
  figures are the ones quoted in the design notes; each bound is about twice
  the figure measured.
"""

from __future__ import annotations

import functools

import numpy as np
import pytest
import torch

from nanodrr_amd._ops import (fused_raymarch, fused_raymarch_bwd, fused_raymarch_bwd_jac, fused_raymarch_jac, pack_labels,
                              pack_coef, pack_volume)
from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses
from nanodrr_amd.renderer import _make_src, _make_tgt
from oracle import raymarch as orm
from oracle import registration as oreg
from tests.helpers import (coef_decode, coef_expected, coef_header, lcoef_decode, oracle_inputs, qcoef_decode,
                           qcoef_to_coef, random_volume, ref_camera_np, rel_to_max)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
DTYPES = [torch.float32, torch.bfloat16, torch.float16]


def _records(packed: torch.Tensor, dtype, shape) -> np.ndarray:
    return coef_decode(packed, dtype, shape)


@pytest.mark.parametrize("src_dtype", DTYPES)
@pytest.mark.parametrize("rec_dtype", DTYPES)
@pytest.mark.parametrize("magnitude", [1e-4, 3.0, 2e4])
def test_coef_records_bit_for_bit(src_dtype, rec_dtype, magnitude):
    """Header and every record slot, bit for bit; the scale keeps all slots finite and normal at
    any volume magnitude (tiny attenuation values, HU-scale values, large raw intensities)."""
    torch.manual_seed(0)
    vol = (torch.randn(5, 6, 7) * magnitude).to(src_dtype)
    vol[1, 2] = 0.0
    packed = pack_coef(vol.to(DEV), rec_dtype)
    expect, scale = coef_expected(vol.float().numpy(), rec_dtype)
    s, inv, code = coef_header(packed)
    assert (s, inv, code) == (scale, 1.0 / scale, {torch.float32: 0, torch.bfloat16: 1, torch.float16: 2}[rec_dtype])
    slot = torch.int32 if rec_dtype == torch.float32 else torch.int16
    got = packed[256:].cpu().view(slot).reshape(expect.shape)
    assert torch.equal(got, expect)
    assert np.isfinite(coef_decode(packed, rec_dtype, vol.shape)).all()
    if rec_dtype == src_dtype:  # ndrr_pack_volume's coefficient layout is the same pack
        assert torch.equal(pack_volume(vol.to(DEV), "coef").cpu(), packed.cpu())


def _golden_geometry(n_classes, S=200):
    image, label = random_volume(32, n_classes, 0)
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    o = oracle_inputs(image, label, n_classes, k_inv, rt_inv, sdd, h, w, S)
    t = {k: torch.as_tensor(o[k], dtype=torch.float32, device=DEV).contiguous() for k in ("M", "src", "tgt", "step")}
    return image, label, o, t


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n_classes", [1, 3])
def test_coef_march_matches_oracle_of_its_records(dtype, n_classes):
    """Forward, single-pass and two-pass pose gradients against the float64 oracle evaluating the
    records the kernels read (decoded): the kernels' own arithmetic, to the usual 1e-5 / 1e-4."""
    image, label, o, t = _golden_geometry(n_classes)
    vol = image[0, 0].to(DEV)
    packed = pack_coef(vol, dtype)
    vol_t = vol.to(dtype)
    lab = label[0, 0].to(torch.int8).to(DEV)
    lab_t = pack_labels(lab, "coef") if n_classes > 1 else lab.new_empty(0)
    rec = _records(packed, dtype, vol.shape)
    out = fused_raymarch(vol_t, lab_t, t["M"], t["src"], t["tgt"], t["step"], 200, n_classes, 16, packed, "coef")
    ref = orm.forward(**o, coef=rec)
    assert rel_to_max(ref.sum(1), out.double().cpu().numpy().sum(1)) < 1e-5
    go = torch.randn(1, n_classes, 256, generator=torch.Generator().manual_seed(3)).to(DEV)
    g = orm.backward(go.double().cpu().numpy(), **o, need_gvol=False, coef=rec)
    _, jac = fused_raymarch_jac(vol_t, lab_t, t["M"], t["src"], t["tgt"], t["step"], 200, n_classes, 16, 7, packed,
                                "coef")
    gM1, gsrc1, gtgt1, gstep1 = fused_raymarch_bwd_jac(go, jac, t["M"], t["src"], t["tgt"], t["step"], 16, True)
    _, gM2, _, gtgt2, _ = fused_raymarch_bwd(go, vol_t, lab_t, t["M"], t["src"], t["tgt"], t["step"], 200, n_classes,
                                             16, False, True, packed, "coef")
    for got in (gM1, gM2):
        assert rel_to_max(g["gM"], got.double().cpu().numpy()) < 1e-4
    assert rel_to_max(g["gtgt"], gtgt1.double().cpu().numpy()) < 1e-4
    assert rel_to_max(g["gtgt"], gtgt2.double().cpu().numpy()) < 1e-4
    assert rel_to_max(g["gstep"], gstep1.double().cpu().numpy()) < 1e-4


def test_fp32_coef_matches_the_corner_oracle():
    """fp32 records hold the interpolant to fp32 rounding: against the corner-form oracle of the
    volume itself the image is within 1e-5 and the pose gradient within 1e-4."""
    image, label, o, t = _golden_geometry(3)
    vol = image[0, 0].to(DEV)
    lab_t = pack_labels(label[0, 0].to(torch.int8).to(DEV), "coef")
    packed = pack_coef(vol, torch.float32)
    out = fused_raymarch(vol, lab_t, t["M"], t["src"], t["tgt"], t["step"], 200, 3, 16, packed, "coef")
    assert rel_to_max(orm.forward(**o).sum(1), out.double().cpu().numpy().sum(1)) < 1e-5
    go = torch.randn(1, 3, 256, generator=torch.Generator().manual_seed(4)).to(DEV)
    _, jac = fused_raymarch_jac(vol, lab_t, t["M"], t["src"], t["tgt"], t["step"], 200, 3, 16, 6, packed, "coef")
    gM = fused_raymarch_bwd_jac(go, jac, t["M"], t["src"], t["tgt"], t["step"], 16, False)[0]
    g = orm.backward(go.double().cpu().numpy(), **o, need_gvol=False)
    assert rel_to_max(g["gM"], gM.double().cpu().numpy()) < 1e-4


# --- accuracy of 16-bit records against the float64 oracle of the fp32 volume ----------------------

@functools.cache
def _phantom(shape, spacing):
    return make_phantom(shape, spacing, seed=0)


def _op_inputs(subject, k_inv, rt_inv, sdd, det, S):
    tgt = _make_tgt(k_inv, sdd, det, det, DEV, torch.float32)
    src = _make_src(False, tgt, sdd)
    step = (tgt - src).norm(dim=-1) / float(S - 1)
    M = (subject.world_to_voxel.double() @ rt_inv.double())[:, :3, :].float().contiguous()
    return M, src.contiguous(), tgt.contiguous(), step.contiguous()


def _zncc_go(img, tgt_img, C):
    """(B, C, R) upstream gradient of sum_b ZNCC(sum_c img, target) (the registration loss)."""
    g = oreg.zncc(img.sum(1), tgt_img.sum(1), 1e-7)[1]
    return np.repeat(g[:, None, :], C, axis=1)


def _accuracy(ph, C, dtype, det, det_mm, S, n_rays, ray_seed, pose_seed, bounds, layout="coef"):
    """Errors on `n_rays` sampled rays of 2 poses. "Clean" rays have no sample within 1e-4 voxel of a
    cell face or nearest-label tie, where the march is discontinuous: there an fp32 sample position
    (within ~1 ulp of the float64 one) may fall on the other side and change a derivative or a
    class, whatever the storage. On clean rays the comparison measures arithmetic and storage only;
    the bounds apply there. All-ray figures are printed beside them."""
    subject = phantom_subject(ph, n_classes=C, dtype=dtype, device=DEV)
    rt_inv, _, _ = sample_poses(4, seed=pose_seed, isocenter=subject.isocenter)
    k_inv, sdd = detector(1020.0, det_mm, det, det)
    k_inv, sdd, rt_inv = k_inv.to(DEV), sdd.to(DEV), rt_inv.to(DEV)
    packed = subject.packed_volume(layout)  # from the fp32 values, rounded once
    lab_t = subject.packed_labels(layout)
    vol_t = subject.image[0, 0]
    M, src, tgt, step = _op_inputs(subject, k_inv, rt_inv, sdd, det, S)
    Ms = M[:2].contiguous()
    vol = ph.mu.double().numpy()
    lab = ph.labels.long().clamp(max=C - 1).numpy()
    rec = {"coef": lambda: _records(packed, dtype, vol.shape), "lcoef": lambda: lcoef_decode(packed, dtype, vol.shape)[0],
           "qcoef": lambda: qcoef_to_coef(qcoef_decode(packed, vol.shape))}[layout]()
    Mn, srcn = Ms.double().cpu().numpy(), src.double().cpu().numpy()
    rays = np.sort(np.random.default_rng(ray_seed).choice(det * det, size=n_rays, replace=False))
    dist = orm.discontinuity_distance(Mn, srcn, tgt.double().cpu().numpy(), S, vol.shape, rays, labels=C > 1)
    clean = rays[(dist > 1e-4).all(axis=0)]
    assert len(clean) > n_rays // 2
    report = {}
    for name, rs in (("clean", clean), ("all", rays)):
        idx = torch.as_tensor(rs, device=DEV)
        tsub, ssub = tgt[:, idx].contiguous(), step[:, idx].contiguous()
        tn, sn = tsub.double().cpu().numpy(), ssub.double().cpu().numpy()
        out, jac = fused_raymarch_jac(vol_t, lab_t, Ms, src, tsub, ssub, S, C, 0, 6, packed, layout)
        img = out.double().cpu().numpy().sum(1)
        ref = orm.forward(vol, lab, C, Mn, srcn, tn, sn, S)
        pix = np.abs(ref.sum(1) - img) / np.abs(ref.sum(1)).max()
        e = {"image max": pix.max(), "image mean": pix.mean(),
             "own image": rel_to_max(orm.forward(vol, lab, C, Mn, srcn, tn, sn, S, coef=rec).sum(1), img)}
        target = orm.forward(vol, lab, C, M[2:4].double().cpu().numpy(), srcn, tn, sn, S)
        gos = {"random": np.random.default_rng(7).standard_normal((2, C, len(rs))), "ones": np.ones((2, C, len(rs))),
               "zncc": _zncc_go(ref, target, C)}
        for gname, go in gos.items():
            gM = fused_raymarch_bwd_jac(torch.as_tensor(go, dtype=torch.float32, device=DEV), jac, Ms, src, tsub,
                                        ssub, 0, False)[0].double().cpu().numpy()
            e[f"pose gradient ({gname} go)"] = rel_to_max(
                orm.backward(go, vol, lab, C, Mn, srcn, tn, sn, S, need_gvol=False)["gM"], gM)
            e[f"own gradient ({gname} go)"] = rel_to_max(
                orm.backward(go, vol, lab, C, Mn, srcn, tn, sn, S, need_gvol=False, coef=rec)["gM"], gM)
        report[name] = e
        print(f"\n{dtype} {layout} records, {name} rays ({len(rs)}/{n_rays}): vs float64 oracle of the fp32 volume: "
              + ", ".join(f"{k} {v:.2e}" for k, v in e.items() if not k.startswith("own"))
              + "; kernel vs float64 oracle of its own records: "
              + ", ".join(f"{k[4:]} {v:.2e}" for k, v in e.items() if k.startswith("own")))
    c = report["clean"]
    assert c["own image"] < 1e-5
    for gname in ("random", "ones", "zncc"):
        assert c[f"own gradient ({gname} go)"] < 1e-4
    for key, bound in bounds.items():
        assert c[key] < bound, (key, c[key], bound)


@pytest.mark.parametrize("dtype, bounds", [
    (torch.bfloat16, {"image max": 1e-3, "pose gradient (random go)": 8e-4, "pose gradient (ones go)": 6e-4,
                      "pose gradient (zncc go)": 3e-4}),
    (torch.float16, {"image max": 2e-4, "pose gradient (random go)": 8e-4, "pose gradient (ones go)": 6e-4,
                     "pose gradient (zncc go)": 3e-4}),
])
def test_coef_accuracy_benchmark_workload(dtype, bounds):
    """256^3 at 1 mm, 4 labels, 200 x 200 at 1 mm, S = 256: 2 poses x 400 sampled rays."""
    _accuracy(_phantom((256,) * 3, (1.0,) * 3), 4, dtype, 200, 1.0, 256, 400, 6, 1, bounds)


@pytest.mark.parametrize("dtype, bounds", [
    (torch.bfloat16, {"image max": 1e-3, "pose gradient (random go)": 8e-4, "pose gradient (ones go)": 8e-4,
                      "pose gradient (zncc go)": 3e-4}),
    (torch.float16, {"image max": 2e-4, "pose gradient (random go)": 8e-4, "pose gradient (ones go)": 8e-4,
                     "pose gradient (zncc go)": 3e-4}),
])
def test_labelled_coef_accuracy_benchmark_workload(dtype, bounds):
    """The same workload from labelled coefficient records (the renderer's choice for it): the
    smaller scale makes differences below 1/128 of the largest coefficient fp16 subnormals."""
    _accuracy(_phantom((256,) * 3, (1.0,) * 3), 4, dtype, 200, 1.0, 256, 400, 6, 1, bounds, "lcoef")


@pytest.mark.parametrize("dtype, bounds", [
    (torch.bfloat16, {"image max": 2e-4, "pose gradient (random go)": 5e-3, "pose gradient (ones go)": 2.5e-3,
                      "pose gradient (zncc go)": 1.2e-3}),
])
def test_quad_coef_accuracy_benchmark_workload(dtype, bounds):
    """The same workload from quad coefficient records (fp16, the registration's layout): the y and z
    differences are stored, the x difference is taken from two fp16 values."""
    _accuracy(_phantom((256,) * 3, (1.0,) * 3), 4, dtype, 200, 1.0, 256, 400, 6, 1, bounds, "qcoef")


@pytest.mark.parametrize("dtype, bounds", [
    (torch.bfloat16, {"image max": 1e-3, "pose gradient (random go)": 9e-4, "pose gradient (ones go)": 8e-4,
                      "pose gradient (zncc go)": 3e-4}),
    (torch.float16, {"image max": 1e-4, "pose gradient (random go)": 9e-4, "pose gradient (ones go)": 8e-4,
                     "pose gradient (zncc go)": 3e-4}),
])
def test_coef_accuracy_mixed_precision_workload(dtype, bounds):
    """512 x 512 x 400 at 0.8 x 0.8 x 1.0 mm, 2 labels, 512 x 512 at 0.8 mm, S = 512 (compensated
    sums): 2 poses x 300 sampled rays."""
    _accuracy(_phantom((400, 512, 512), (0.8, 0.8, 1.0)), 2, dtype, 512, 0.8, 512, 300, 21, 11, bounds)
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype, bounds", [
    (torch.bfloat16, {"image max": 1e-3, "pose gradient (random go)": 9e-4, "pose gradient (ones go)": 8e-4,
                      "pose gradient (zncc go)": 3e-4}),
    (torch.float16, {"image max": 1e-4, "pose gradient (random go)": 9e-4, "pose gradient (ones go)": 8e-4,
                     "pose gradient (zncc go)": 3e-4}),
])
def test_labelled_coef_accuracy_mixed_precision_workload(dtype, bounds):
    """The mixed-precision workload from labelled coefficient records (the renderer's choice for it)."""
    _accuracy(_phantom((400, 512, 512), (0.8, 0.8, 1.0)), 2, dtype, 512, 0.8, 512, 300, 21, 11, bounds, "lcoef")
    torch.cuda.empty_cache()
