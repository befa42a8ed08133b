"""Host-side pieces of bench.py: case parsing, the byte model's volume-gradient term, the
guard on committed counter traffic and the host CPU record (no GPU needed)."""

import pytest
import torch

import bench
from nanodrr_amd.roofline import march_bytes


def test_parse_case():
    assert bench.parse_case("bf16:coef:single") == ("bf16", "coef", "single", False)
    assert bench.parse_case("bf16:quad") == ("bf16", "quad", "single", False)
    assert bench.parse_case("fp32:linear:two+gvol") == ("fp32", "linear", "two", True)
    for bad in ("bf16", "int8:quad:single", "bf16:quad:three", "bf16:quad:two+gvol", "fp32:linear:two+x"):
        with pytest.raises(ValueError):
            bench.parse_case(bad)
    for case in bench.DEFAULT_CASES.split(","):
        bench.parse_case(case)
    assert bench.DEFAULT_CASES.split(",")[0] == "bf16:lcoef:single"


def test_volume_gradient_bytes():
    base = march_bytes(1000, 10, 2, 4, torch.float32, backward=True)
    assert march_bytes(1000, 10, 2, 4, torch.float32, backward=True, need_gvol=True) == base + 1000 * 64
    # forward launches carry no gradient traffic
    assert march_bytes(1000, 10, 2, 4, torch.float32, need_gvol=True) == march_bytes(1000, 10, 2, 4, torch.float32)


def _summary(tmp_path, digest, fwd_us):
    text = "\n".join([
        "# counters",
        f"Kernel source digest: `{digest}`",
        "",
        "| role | kernel | avg us | HBM read MB | HBM write MB |",
        "|---|---|---|---|---|",
        f"| fwd | k_fwd<...> | {fwd_us} | 600.0 | 80.0 |",
        "| bwd | k_bwd_jac<...> | 15.0 | 70.0 | 0.5 |",
    ])
    (tmp_path / "raymarch_bf16_coef_single_counters_bench256_c4_b16.md").write_text(text)


def test_measured_traffic_guard(tmp_path):
    live = {"fwd": 0.145, "bwd": 0.015}
    traffic, note = bench.measured_traffic("bf16:coef:single", live, profiles=tmp_path)
    assert traffic is None and "no counter summary" in note
    _summary(tmp_path, "0" * 16, 145.0)
    traffic, note = bench.measured_traffic("bf16:coef:single", live, profiles=tmp_path)
    assert traffic is None and "another kernel source" in note
    _summary(tmp_path, bench.kernel_source_digest(), 145.0)
    traffic, _ = bench.measured_traffic("bf16:coef:single", live, profiles=tmp_path)
    assert traffic == {"k_fwd": pytest.approx(680e6), "k_bwd": pytest.approx(70.5e6)}
    _summary(tmp_path, bench.kernel_source_digest(), 190.0)  # a stale kernel time drops that kernel only
    traffic, note = bench.measured_traffic("bf16:coef:single", live, profiles=tmp_path)
    assert traffic == {"k_bwd": pytest.approx(70.5e6)} and "fwd: profiled at 190.0" in note


def test_host_cpu_record():
    host = bench.host_cpu()
    assert host["os_cpu_count"] >= 1 and 1 <= host["nproc"] <= host["os_cpu_count"]
    assert set(host) == {"os_cpu_count", "nproc", "cgroup_cpus", "model"}
