"""Host-side checks that need no GPU: the C ABI library loads and exports what
include/*.h declares, argument errors come back before any launch, and the
render() front end validates like the reference (tests/drr/test_fused.py:229-251)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest
import torch

from nanodrr_amd import _lib, render
from nanodrr_amd.camera import make_k_inv, make_rt_inv
from tests.helpers import random_subject

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"\b(ndrr_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared_symbols()
    for n in ("ndrr_raymarch_fwd", "ndrr_raymarch_bwd", "ndrr_labels_to_i8", "ndrr_strerror"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    for name in _declared_symbols():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_error_strings():
    lib = _lib.lib()
    assert lib.ndrr_version() >= 100
    for code in (0, -1, -2, -3, -4, -5):
        assert lib.ndrr_strerror(code)
    assert b"invalid" in lib.ndrr_strerror(-1)


def test_workspace_size_is_12_floats_per_tile_part():
    lib = _lib.lib()
    tw, th, tr = _lib.launch_tile()
    assert tw % 8 == 0 and th % 8 == 0 and tr == tw * th
    tiles = -(-200 // tw) * -(-200 // th)
    # 12 floats per pose per partial; one partial per tile or per wave (tiles of tr rays, runs without a width)
    per_tile = lib.ndrr_bwd_workspace_floats(16, 40000, 200) // (16 * tiles * 12)
    assert lib.ndrr_bwd_workspace_floats(16, 40000, 200) == 16 * tiles * 12 * per_tile
    assert lib.ndrr_bwd_workspace_floats(2, 1000, 0) == 2 * -(-1000 // tr) * 12 * per_tile
    assert per_tile >= 1
    assert lib.ndrr_bwd_workspace_floats(0, 10, 0) == 0


def test_packed_sizes():
    lib = _lib.lib()
    D, H, W = 5, 6, 7
    assert lib.ndrr_packed_bytes(D, H, W, 0, 0) == D * H * W * 4
    assert lib.ndrr_packed_bytes(D, H, W, 1, 1) == (D + 1) * (H + 1) * (W + 1) * 16
    assert lib.ndrr_packed_bytes(D, H, W, 0, 2) == (D + 1) * (H + 1) * (W + 2) * 16
    assert lib.ndrr_packed_bytes(D, H, W, 2, 2) == (D + 1) * (H + 1) * (W + 2) * 8
    assert lib.ndrr_packed_bytes(D, H, W, 1, 3) == 256 + (D + 1) * (H + 1) * (W + 1) * 16  # header + records
    assert lib.ndrr_packed_bytes(D, H, W, 0, 3) == 256 + (D + 1) * (H + 1) * (W + 1) * 32
    assert lib.ndrr_packed_label_bytes(D, H, W, 3) == (D + 4) * (H + 4) * (W + 4)  # the padded map
    assert lib.ndrr_packed_label_bytes(D, H, W, 1) == (D + 4) * (H + 4) * (W + 4)
    assert lib.ndrr_packed_bytes(D, H, W, 1, 7) == 0
    assert lib.ndrr_packed_label_bytes(D, H, W, 2) == (D + 1) * (H + 1) * (W + 2) * 4
    assert lib.ndrr_packed_label_bytes(D, H, W, 0) == D * H * W


def test_packed_layouts_limited_to_2gib():
    """Packed layouts are read through 32-bit buffer offsets: larger copies are refused."""
    lib = _lib.lib()
    assert _fwd(lib, layout=1, dt=0, D=512, H=512, W=400) == -4  # fp32 stencil: 3.4 GB


def test_packed_layouts_need_24_bit_record_index():
    """The packed record index is formed with 24-bit multiplies: (D + 1)(H + 1) < 2^24."""
    from nanodrr_amd.renderer import packed_index_fits

    lib = _lib.lib()
    assert _fwd(lib, layout=2, dt=1, D=4096, H=4096, W=2, C=1) == -4  # 402 MB of quads, plane too large
    assert _fwd(lib, layout=1, dt=1, D=4096, H=4096, W=2, C=1) == -4
    assert _fwd(lib, layout=0, dt=1, D=4096, H=4096, W=2, C=1, vol=None) == -1  # linear: only the null volume
    assert packed_index_fits(512, 512, 512) and packed_index_fits(4094, 4095, 100)
    assert not packed_index_fits(4095, 4096, 2) and not packed_index_fits(1, 1, 2**24 - 2)
    # the padded label map's offsets use signed 24-bit multiplies: (H + 4)(W + 4) < 2^23
    assert packed_index_fits(10, 2800, 2800) and not packed_index_fits(10, 2900, 2900)


def _fwd(lib, **kw):
    args = dict(vol=1, dt=0, layout=0, lab=None, M=1, src=1, tgt=1, step=1, out=1, B=1, N=16, D=4, H=4, W=4, S=8, C=1,
                width=4, sbs=0, srs=0, tbs=0, pbs=0, stream=None)
    args.update(kw)
    return lib.ndrr_raymarch_fwd(*args.values())


def test_argument_errors_return_before_launch():
    """Invalid arguments are rejected on the host (no pointer is dereferenced)."""
    lib = _lib.lib()
    assert _fwd(lib, vol=None) == -1
    assert _fwd(lib, S=1) == -1
    assert _fwd(lib, width=5) == -1  # 16 rays are not a multiple of width 5
    assert _fwd(lib, dt=7) == -2
    assert _fwd(lib, C=9) == -3
    assert _fwd(lib, C=3, lab=None) == -1
    assert _fwd(lib, B=70000) == -4
    assert _fwd(lib, D=2048, H=2048, W=1024) == -4
    assert lib.ndrr_labels_to_i8(None, None, 10, 2, None) == -1
    assert lib.ndrr_labels_to_i8(1, 1, 10, 0, None) == -3


def test_check_raises_native_error():
    with pytest.raises(_lib.NativeError, match="n_classes"):
        _lib.check(-3, "probe")


def _camera():
    k_inv = make_k_inv(100.0, 1.0, 1.0, 0.0, 0.0, 16, 16)
    rt_inv = make_rt_inv(torch.tensor([[5.0, -3.0, 8.0]]), torch.tensor([[1.0, 80.0, -2.0]]), "AP", torch.zeros(3))
    return k_inv, rt_inv, torch.tensor([100.0]), 16, 16


def test_invalid_n_samples_raises():
    subject = random_subject()
    k_inv, rt_inv, sdd, h, w = _camera()
    with pytest.raises(ValueError, match="n_samples"):
        render(subject, k_inv, rt_inv, sdd, h, w, n_samples=1)


def test_unknown_backend_raises():
    subject = random_subject()
    k_inv, rt_inv, sdd, h, w = _camera()
    with pytest.raises(ValueError, match="backend"):
        render(subject, k_inv, rt_inv, sdd, h, w, backend="cuda")


def test_torch_backend_is_not_shipped():
    subject = random_subject()
    k_inv, rt_inv, sdd, h, w = _camera()
    with pytest.raises(ValueError, match="not part of nanodrr_amd"):
        render(subject, k_inv, rt_inv, sdd, h, w, backend="torch")


def test_cpu_tensors_fail_loudly():
    """No silent CPU fallback: host tensors are an error, not a slow path."""
    subject = random_subject()
    k_inv, rt_inv, sdd, h, w = _camera()
    for backend in ("auto", "hip", "triton"):
        with pytest.raises(ValueError, match="GPU"):
            render(subject, k_inv, rt_inv, sdd, h, w, n_samples=64, backend=backend)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libnanodrr_amd.so")
    _lib.lib.cache_clear()
    try:
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            _lib.lib()
    finally:
        monkeypatch.undo()
        _lib.lib.cache_clear()
    assert isinstance(_lib.lib(), ctypes.CDLL)


def test_zncc_workspace_is_one_record_per_chunk():
    lib = _lib.lib()
    # eight floats of chunk moments per 4096-pixel chunk per image
    assert lib.ndrr_zncc_workspace_floats(3, 40000) == 3 * 10 * 8
    assert lib.ndrr_zncc_workspace_floats(1, 4096) == 8
    assert lib.ndrr_zncc_workspace_floats(2, 4097) == 2 * 2 * 8
    assert lib.ndrr_zncc_workspace_floats(0, 100) == 0


def test_conv_weight_workspace_holds_whole_split_records():
    lib = _lib.lib()
    for N, H, W, Cin, Cout, s in ((2, 16, 16, 32, 64, 2), (256, 32, 32, 128, 256, 2), (4, 9, 7, 1, 32, 1)):
        rec = 9 * Cout * Cin + Cout  # weight partials, then the bias partials
        n = lib.ndrr_conv3x3_bwd_weight_workspace_floats(N, H, W, Cin, Cout, s)
        assert n >= rec and n % rec == 0, (N, H, W, Cin, Cout, s, n)
    assert lib.ndrr_conv3x3_bwd_weight_workspace_floats(2, 16, 16, 48, 64, 1) == 0  # Cin not 1 or 32k


def test_fused_pose_step_adjoint_rejects_bad_arguments_before_launch():
    """ndrr_raymarch_bwd_jac_se3: null pose-step inputs, a scale other than +-1, a short stride."""
    lib = _lib.lib()
    args = dict(go=1, gbs=16, jac=1, planes=6, M=1, src=1, tgt=1, step=1, ws=16, pre=1, suf=1, rot=1, xyz=1, ps=3,
                scale=-1.0, grot=1, gxyz=1, B=1, N=16, C=1, width=4, sbs=0, srs=0, tbs=0, pbs=0, stream=None)
    for key, bad in (("pre", None), ("rot", None), ("grot", None), ("scale", 0.5), ("ps", 2), ("M", None)):
        kw = dict(args, **{key: bad})
        assert lib.ndrr_raymarch_bwd_jac_se3(*kw.values()) == -1, key


def test_new_entry_points_reject_bad_arguments_before_launch():
    lib = _lib.lib()
    EINVAL = -1
    fake = 4096  # never dereferenced: each call returns before any launch
    # betas outside [0, 1), bias corrections <= 0
    assert lib.ndrr_adam_step(fake, fake, fake, fake, 8, 1e-3, 1.0, 0.999, 1e-8, 0.1, 0.1, None) == EINVAL
    assert lib.ndrr_adam_step(fake, fake, fake, fake, 8, 1e-3, 0.9, 0.999, 1e-8, 0.0, 0.1, None) == EINVAL
    assert lib.ndrr_adam_step(None, fake, fake, fake, 8, 1e-3, 0.9, 0.999, 1e-8, 0.1, 0.1, None) == EINVAL
    # negative eps, missing workspace
    assert lib.ndrr_zncc(fake, fake, 2, 100, -1.0, fake, fake, None, None) == EINVAL
    assert lib.ndrr_zncc(fake, fake, 2, 100, 1e-7, None, fake, None, None) == EINVAL
    # a workspace that is not 16-byte aligned
    assert lib.ndrr_conv3x3_bwd_weight(fake, fake, fake + 4, fake, None, 2, 16, 16, 32, 64, 2, None) == EINVAL
    # Cout not a multiple of 64 for the MFMA weight gradient
    assert lib.ndrr_conv3x3_bwd_weight(fake, fake, fake, fake, None, 2, 16, 16, 32, 96, 2, None) == EINVAL


def test_noalloc_contraction_rejects_oversized_work_on_the_host():
    """ndrr_raymarch_bwd_jac_pose_noalloc checks the fixed scratch array's size before any HIP call."""
    lib = _lib.lib()
    assert lib.ndrr_raymarch_bwd_jac_pose_noalloc(1, 0, 1, 7, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2000, 40000, 1, 200,
                                                  0, 0, 0, 0, None) == -5
    assert lib.ndrr_raymarch_bwd_jac_pose_noalloc(1, 0, 1, 7, None, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 400, 1, 20,
                                                  0, 0, 0, 0, None) == -1
    assert b"scratch" in lib.ndrr_strerror(-5)


def test_patch_zncc_rejects_bad_arguments_before_launch():
    lib = _lib.lib()
    assert lib.ndrr_patch_zncc_workspace_floats(2, 20, 20, 5) == 2 * (16 * 16 * 4 + 4)
    assert lib.ndrr_patch_zncc_workspace_floats(2, 20, 20, 33) == 0
    assert lib.ndrr_patch_zncc(1, 1, 2, 20, 20, 33, 1e-4, 16, 1, 1, None) == -1  # window beyond 32
    assert lib.ndrr_patch_zncc(1, 1, 2, 4, 20, 5, 1e-4, 16, 1, 1, None) == -1   # window taller than the image
    assert lib.ndrr_patch_zncc(1, 1, 2, 20, 20, 5, 1e-4, 8, 1, 1, None) == -1   # misaligned workspace
    assert lib.ndrr_patch_zncc(1, 1, 70000, 20, 20, 5, 1e-4, 16, 1, 1, None) == -4
