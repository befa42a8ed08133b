"""Benchmark: DRR/s for forward + analytic backward (pose gradients) on one MI355X.

Workload (fwd+bwd on a 256^3 volume): seeded 256^3 phantom at 1 mm, 4 label
channels (int8 labels), 16 poses (ZXY +-15 deg, +-20 mm about the isocentre
600 mm from the source), 200 x 200 detector at 1 mm, sdd 1020 mm, 256
samples per ray. One step = render(...) of the 16 poses plus the backward of
a weighted sum to the poses, captured once in a HIP graph and replayed (the
eager, uncaptured step is reported too); DRR/s = 16 * steps / elapsed.
Cases (volume dtype : layout the kernels read : adjoint[+gvol]):
  bf16:lcoef:single    the headline `value`: bf16 labelled coefficient records
                       (each cell's interpolant coefficients, made from the
                       fp32 volume, with the classes of its eight corners in
                       bits the record's scale leaves free: one 16-byte load
                       per sample, no label map; the renderer's own choice
                       for this subject), single-pass adjoint (the forward
                       march stores per-ray sums, the backward contracts them);
  bf16:coef:single     the same without the classes in the records (the label
                       map read beside them: the choice for more than 4
                       classes or a volume with negative values);
  fp32:linear:two      the plain fp32 kernel: the (D, H, W) volume, two-pass
                       adjoint (the backward marches the volume again, as
                       nanodrr's kernel does) - the baseline of the speed-ups;
  fp32:linear:two+gvol the same with volume gradients (fp32 atomics into a
                       (D, H, W) gradient volume, zeroed every step);
  bf16:quad:single, fp32:quad:single, bf16:stencil:single  other corners.
Every case accumulates in float32.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cases bf16:lcoef:single,...]
                    [--no-cpu-baseline] [--no-registration] [--no-mixed]

Prints one JSON line. `roofline` covers the dominant kernel of the headline
case, timed with HIP events on the stream it runs on, with the HBM traffic of
the committed rocprofv3 counter summary when that summary was measured on the
current kernel source at the same kernel time (else null); `cpu_baseline`
times the pure-PyTorch grid_sample restatement (fp32) of the workload and of
the small CPU case on the host cores for bounded samples; `mixed_precision`
is the 512 x 512 x 400 bf16 / fp16 workload; `registration` is the batched
registration leg (registrations/s reaching mTRE < 1 mm).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import torch  # noqa: E402

HBM_PEAK_GBS = 8000.0
SPIN_CYCLES = 20_000_000  # ~10 ms of GPU clock cycles: longer than the host needs to enqueue the timed launches
SIZE, SPACING, N_CLASSES, BATCH, DET, S = 256, 1.0, 4, 16, 200, 256
GRAPH_ITERS = 10  # registration iterations per captured HIP graph
# Layout of the registration renders (single channel): quad coefficient records (fp16, 136 MB at 256^3)
# stay in the Infinity Cache like the quad records and need fewer operations per sample (473 against 460
# registrations/s, the same 57 / 64 converged); the coefficient records (271 MB) do not fit, and ran 25 %
# slower with the same convergence.
REGISTRATION_LAYOUT = "qcoef"


DEFAULT_CASES = "bf16:lcoef:single,fp32:linear:two,fp32:linear:two+gvol,bf16:coef:single,bf16:quad:single," \
                "fp32:quad:single,bf16:stencil:single"


def make_workload(dtype, device, gvol: bool = False):
    """(subject, k_inv, rt_inv, sdd, phantom) of the 256^3 / C=4 / B=16 workload; with gvol the
    subject's fp32 volume requires grad."""
    from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses

    ph = make_phantom((SIZE,) * 3, (SPACING,) * 3, seed=0)
    subject = phantom_subject(ph, n_classes=N_CLASSES, dtype=dtype, device=device)
    if gvol:
        subject.image.requires_grad_(True)
    rt_inv, _, _ = sample_poses(BATCH, seed=1, isocenter=subject.isocenter)
    k_inv, sdd = detector(1020.0, 1.0, DET, DET)
    return subject, k_inv.to(device), rt_inv.to(device), sdd.to(device), ph


DTYPES = {"bf16": torch.bfloat16, "fp32": torch.float32, "fp16": torch.float16}
ADJOINTS = {"single": "single_pass", "two": "two_pass"}


def parse_case(case: str):
    """"dtype:layout[:adjoint[+gvol]]" -> (dtype name, layout, adjoint name, volume gradients);
    the adjoint defaults to the single pass. Volume gradients need fp32:linear:two."""
    parts = case.split(":")
    if len(parts) not in (2, 3):
        raise ValueError(f"bad case {case!r}: expected dtype:layout[:single|two[+gvol]]")
    name, layout = parts[0], parts[1]
    adjoint, _, extra = (parts[2] if len(parts) > 2 else "single").partition("+")
    gvol = extra == "gvol"
    if name not in DTYPES or adjoint not in ADJOINTS or extra not in ("", "gvol"):
        raise ValueError(f"bad case {case!r}: expected dtype:layout[:single|two[+gvol]]")
    if gvol and (name, layout, adjoint) != ("fp32", "linear", "two"):
        raise ValueError(f"bad case {case!r}: volume gradients are benchmarked as fp32:linear:two+gvol")
    return name, layout, adjoint, gvol


def _clear_grads(subject, rt):
    rt.grad = None
    if subject.image.requires_grad:
        subject.image.grad = None


def step_fn(subject, k_inv, rt_inv, sdd, wts):
    from nanodrr_amd import render

    rt = rt_inv.detach().requires_grad_(True)
    _clear_grads(subject, rt)
    out = render(subject, k_inv, rt, sdd, DET, DET, n_samples=S)
    out.backward(wts.expand_as(out))  # d(sum(out * wts))/d(out) = wts, without a reduction kernel
    return rt.grad


def graph_step(subject, k_inv, rt_inv, sdd, wts):
    """The same step captured once in a HIP graph (torch.cuda.CUDAGraph) and
    replayed: the counterpart of the reference benchmark's CUDA-graph rows
    (torch.compile mode="reduce-overhead", tests/benchmark/benchmark.py:226)."""
    from nanodrr_amd import render

    rt = rt_inv.detach().clone().requires_grad_(True)

    def body():
        _clear_grads(subject, rt)
        out = render(subject, k_inv, rt, sdd, DET, DET, n_samples=S)
        out.backward(wts.expand_as(out))

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            body()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    _clear_grads(subject, rt)
    with torch.cuda.graph(graph):
        out = render(subject, k_inv, rt, sdd, DET, DET, n_samples=S)
        out.backward(wts.expand_as(out))
    return graph.replay


def time_steps(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        fn()
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps  # ms per step


def _gvol_backward(go, vol, lab, M, src, tgt, step):
    """A closure launching ndrr_raymarch_bwd_bricked (the op's volume-gradient path: the march's
    scatter into a 4 x 4 x 4-voxel brick buffer, then its addition into gvol, which clears the
    buffer) with buffers allocated once (the op allocates and zero-fills them on every call; this
    times the kernels)."""
    from nanodrr_amd import _lib
    from nanodrr_amd._ops import _DTYPE_CODE, _geometry

    B, N, D, H, W, strides = _geometry(vol, lab, M, src, tgt, step, S, N_CLASSES, DET, "linear")
    lib = _lib.lib()
    dev = vol.device
    ws = torch.empty(lib.ndrr_bwd_workspace_floats(B, N, DET), device=dev, dtype=torch.float32)
    gM = torch.empty(B, 3, 4, device=dev, dtype=torch.float32)
    gvol = torch.zeros(vol.shape, device=dev, dtype=torch.float32)  # accumulates over the timed calls
    bricks = torch.zeros(lib.ndrr_gvol_brick_floats(D, H, W), device=dev, dtype=torch.float32)  # left zero
    stream = torch.cuda.current_stream(dev).cuda_stream

    def launch():
        rc = lib.ndrr_raymarch_bwd_bricked(go.data_ptr(), vol.data_ptr(), _DTYPE_CODE[vol.dtype], _lib.LAYOUT_LINEAR,
                                           lab.data_ptr(), M.data_ptr(), src.data_ptr(), tgt.data_ptr(),
                                           step.data_ptr(), ws.data_ptr(), gM.data_ptr(), None, None, None,
                                           gvol.data_ptr(), bricks.data_ptr(), 0, B, N, D, H, W, S, N_CLASSES, DET,
                                           *strides, stream)
        _lib.check(rc, "ndrr_raymarch_bwd_bricked")

    return launch


def kernel_times(subject, k_inv, rt_inv, sdd, wts, reps, layout, adjoint="two", gvol=False):
    """Average duration (ms) of one forward and one backward op launch, HIP events
    recorded on the stream the ops launch on (torch's current stream). The
    single-pass forward also writes the adjoint's per-ray sums and its backward
    is the contraction; with gvol the two-pass backward also forms the volume gradient."""
    from nanodrr_amd._ops import fused_raymarch, fused_raymarch_bwd, fused_raymarch_bwd_jac, fused_raymarch_jac
    from nanodrr_amd.renderer import _make_src, _make_tgt

    vol = subject.image[0, 0].detach()
    packed = None if layout == "linear" else subject.packed_volume(layout)
    lab = subject.packed_labels(layout)
    M = (subject.world_to_voxel.double() @ rt_inv.double())[:, :3, :].float().contiguous()
    tgt = _make_tgt(k_inv, sdd, DET, DET, k_inv.device, torch.float32).contiguous()
    src = _make_src(False, tgt, sdd).contiguous()
    step = ((tgt - src).norm(dim=-1) / (S - 1)).contiguous()
    go = wts.reshape(1, N_CLASSES, -1).expand(M.shape[0], -1, -1).contiguous()
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        # A bounded spin kernel ahead of the start event keeps the GPU busy while the host enqueues
        # every timed launch, so the events bracket back-to-back kernels and not the host's launch
        # latency (which exceeds the 20 us contraction and delays the first forward launches).
        torch.cuda._sleep(SPIN_CYCLES)
        t0.record(stream)
        for _ in range(reps):
            fn()
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / reps

    if adjoint == "single":
        jac = fused_raymarch_jac(vol, lab, M, src, tgt, step, S, N_CLASSES, DET, 6, packed, layout)[1]
        ms_fwd = timed(lambda: fused_raymarch_jac(vol, lab, M, src, tgt, step, S, N_CLASSES, DET, 6, packed, layout))
        ms_bwd = timed(lambda: fused_raymarch_bwd_jac(go, jac, M, src, tgt, step, DET, False))
    else:
        ms_fwd = timed(lambda: fused_raymarch(vol, lab, M, src, tgt, step, S, N_CLASSES, DET, packed, layout))
        if gvol:
            ms_bwd = timed(_gvol_backward(go, vol, lab, M, src, tgt, step))
        else:
            ms_bwd = timed(lambda: fused_raymarch_bwd(go, vol, lab, M, src, tgt, step, S, N_CLASSES, DET, False,
                                                      False, packed, layout))
    return ms_fwd, ms_bwd, (M, src, tgt)


def algorithmic_bytes(subject, geom, dtype, adjoint="two", gvol=False):
    from nanodrr_amd.roofline import contraction_bytes, inside_samples, march_bytes

    M, src, tgt = geom
    n_in = int(inside_samples(M, src, tgt, S, tuple(subject.image.shape[-3:])).sum())
    n_rays = M.shape[0] * tgt.shape[1]
    if adjoint == "single":
        fwd = march_bytes(n_in, n_rays, M.shape[0], N_CLASSES, dtype, jac_planes=6)
        bwd = contraction_bytes(n_rays, M.shape[0], N_CLASSES, 6)
    else:
        fwd = march_bytes(n_in, n_rays, M.shape[0], N_CLASSES, dtype)
        bwd = march_bytes(n_in, n_rays, M.shape[0], N_CLASSES, dtype, backward=True, need_gvol=gvol)
    return n_in, fwd, bwd


def _cpu_drr_rate(ph, n_classes, threads, budget_s, max_runs, n_samples=S, backward=True):
    """(DRR/s, runs): median rate of one DRR (1 pose; forward + pose backward, or forward only)
    through the pure-PyTorch grid_sample restatement (oracle.torch_ref, fp32) on `threads`
    threads, after one warm-up, repeated for about `budget_s` seconds (at least 2 runs, at most
    `max_runs`)."""
    from nanodrr_amd.phantom import detector, phantom_subject, sample_poses
    from oracle import torch_ref

    torch.set_num_threads(threads)
    subject = phantom_subject(ph, n_classes=n_classes)
    rt_all, _, _ = sample_poses(BATCH, seed=1, isocenter=subject.isocenter)
    k_inv, sdd = detector(1020.0, 1.0, DET, DET)
    image, label, w2g = subject.image, subject.label, subject.world_to_grid
    wts = torch.randn(1, n_classes, DET, DET, generator=torch.Generator().manual_seed(0))

    def one(i):
        rt = rt_all[i % BATCH: i % BATCH + 1].clone()
        if not backward:
            with torch.no_grad():
                torch_ref.render_torch(image, label, n_classes, w2g, k_inv, rt, sdd, DET, DET, n_samples)
            return
        rt.requires_grad_(True)
        out = torch_ref.render_torch(image, label, n_classes, w2g, k_inv, rt, sdd, DET, DET, n_samples)
        (out * wts).sum().backward()

    one(0)  # warm-up
    times = []
    t_start = time.perf_counter()
    i = 1
    while (time.perf_counter() - t_start < budget_s and i <= max_runs) or len(times) < 2:
        t0 = time.perf_counter()
        one(i)
        times.append(time.perf_counter() - t0)
        i += 1
    return 1.0 / statistics.median(times), len(times)


def host_cpu() -> dict:
    """This host's CPU: os.cpu_count(), the cores this process may run on (what `nproc`
    prints), the cgroup CPU quota if one is set, and the model name (lscpu's "Model name",
    read from /proc/cpuinfo)."""
    info = {"os_cpu_count": os.cpu_count() or 1}
    try:
        info["nproc"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["nproc"] = info["os_cpu_count"]
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        info["cgroup_cpus"] = None if quota == "max" else int(quota) / int(period)
    except (OSError, ValueError):
        info["cgroup_cpus"] = None
    info["model"] = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return info


def cpu_baseline(ph):
    """The CPU restatement on the host's cores: the workload's DRRs (fwd + pose backward) with
    torch.set_num_threads(os.cpu_count()), and - when fewer cores are usable (affinity or cgroup
    quota) - on that many threads too, `value` being the faster (its thread count in `cores`);
    a one-thread rate; and the small CPU case (128^3 at 1.5 mm, C=1, 200 x 200, S=256) forward
    only and fwd + backward on the same thread counts. About 30 s in all."""
    from nanodrr_amd.phantom import make_phantom

    host = host_cpu()
    counts = [host["os_cpu_count"]]
    usable = host["nproc"]
    if host["cgroup_cpus"]:
        usable = min(usable, max(1, int(host["cgroup_cpus"])))
    if usable != counts[0]:
        counts.append(usable)
    per_count = {}
    for n in counts:
        rate, runs = _cpu_drr_rate(ph, N_CLASSES, n, 8.0, 40)
        per_count[n] = (rate, runs)
    best = max(per_count, key=lambda n: per_count[n][0])
    rate, runs = per_count[best]
    rate1, runs1 = _cpu_drr_rate(ph, N_CLASSES, 1, 4.0, 3)
    ph0 = make_phantom((128,) * 3, (1.5,) * 3, seed=0)
    f0, f0_runs = _cpu_drr_rate(ph0, 1, best, 3.0, 40, backward=False)
    f01, f01_runs = _cpu_drr_rate(ph0, 1, 1, 2.0, 10, backward=False)
    b0, b0_runs = _cpu_drr_rate(ph0, 1, best, 3.0, 20)
    b01, b01_runs = _cpu_drr_rate(ph0, 1, 1, 2.0, 5)
    torch.set_num_threads(best)
    return {"value": rate, "unit": "DRR/s", "cores": best, "kind": "port",
            "sample": f"{runs} DRRs (1 pose each) of the 256^3/C=4/200x200/S=256 workload, fwd + pose "
                      f"backward, oracle.torch_ref fp32 grid_sample on {best} threads, median of runs",
            "host": host,
            "per_thread_count": {str(n): {"value": r, "runs": k} for n, (r, k) in per_count.items()},
            "one_thread": {"value": rate1, "sample": f"{runs1} DRRs of the same workload on 1 thread"},
            "cpu_case_128": {
                "unit": "DRR/s", "threads": best,
                "forward": f0, "forward_one_thread": f01, "fwd_bwd": b0, "fwd_bwd_one_thread": b01,
                "sample": f"128^3 at 1.5 mm, C=1, 200x200 at 1.0 mm, S=256, 1 pose per DRR: {f0_runs} / {f01_runs} "
                          f"forward-only and {b0_runs} / {b01_runs} fwd + pose backward DRRs on {best} / 1 threads"}}


def mixed_precision_leg(steps: int = 20, warmup: int = 5):
    """The mixed-precision workload at full size: a 512 x 512 x 400 phantom at 0.8 x 0.8 x 1.0 mm
    stored in bf16 and in fp16, 2 labels, 8 poses, 512 x 512 detector at 0.8 mm, S = 512
    (compensated sums), fwd + pose backward (single-pass adjoint; labelled coefficient records,
    the 16-bit default for this subject, plain coefficient records and quad records),
    graph-replayed steps timed with HIP events, plus the plain fp32 kernel on the fp32 volume."""
    import nanodrr_amd.renderer as renderer
    from nanodrr_amd import render
    from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses

    dev = torch.device("cuda:0")
    ph = make_phantom((400, 512, 512), (0.8, 0.8, 1.0), seed=0)
    B, det, n_s = 8, 512, 512
    res = {"workload": "512x512x400 @0.8x0.8x1.0 mm, C=2, B=8, 512x512 @0.8 mm, S=512 (compensated), fwd + pose "
                       "backward, graph-replayed step"}
    for name, dt, layout, adjoint in (("bf16_lcoef", torch.bfloat16, "lcoef", "single_pass"),
                                      ("fp16_lcoef", torch.float16, "lcoef", "single_pass"),
                                      ("bf16_coef", torch.bfloat16, "coef", "single_pass"),
                                      ("fp16_coef", torch.float16, "coef", "single_pass"),
                                      ("bf16_quad", torch.bfloat16, "quad", "single_pass"),
                                      ("fp16_quad", torch.float16, "quad", "single_pass"),
                                      ("fp32_plain", torch.float32, "linear", "two_pass")):
        renderer.VOLUME_LAYOUT, renderer.ADJOINT = layout, adjoint
        subject = phantom_subject(ph, n_classes=2, dtype=dt, device=dev)
        rt_inv, _, _ = sample_poses(B, seed=11, isocenter=subject.isocenter)
        k_inv, sdd = detector(1020.0, 0.8, det, det)
        k_inv, sdd, rt_inv = k_inv.to(dev), sdd.to(dev), rt_inv.to(dev)
        wts = torch.randn(1, 2, det, det, generator=torch.Generator().manual_seed(0)).to(dev)
        rt = rt_inv.detach().clone().requires_grad_(True)

        def body():
            rt.grad = None
            out = render(subject, k_inv, rt, sdd, det, det, n_samples=n_s)
            out.backward(wts.expand_as(out))

        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(3):
                body()
        torch.cuda.current_stream().wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        rt.grad = None
        with torch.cuda.graph(graph):
            out = render(subject, k_inv, rt, sdd, det, det, n_samples=n_s)
            out.backward(wts.expand_as(out))
        ms = time_steps(graph.replay, steps, warmup)
        res[name] = {"ms_per_step": ms, "drr_per_s": B * 1000.0 / ms, "layout": layout, "adjoint": adjoint}
        del graph, subject, out
        gc.collect()
        torch.cuda.empty_cache()
    renderer.VOLUME_LAYOUT, renderer.ADJOINT = "auto", "single_pass"
    res["speedup_bf16_lcoef_over_fp32_plain"] = res["bf16_lcoef"]["drr_per_s"] / res["fp32_plain"]["drr_per_s"]
    return res


def reference_shape_leg(reps: int = 200):
    """Forward-only DRRs at the reference benchmark's own settings (tests/benchmark/benchmark.py:
    27-38): one 200 x 200 DRR at 2.0 mm pixels, sdd 1020, AP pose 850 mm from the isocentre,
    n_samples 500, one channel, on the 256^3 phantom (the reference's benchmark volume is not
    shipped). FPS from HIP events: the GPU time of back-to-back render() calls replayed from a
    HIP graph (queued behind a spin so the host's launches stay off the clock), and the eager
    calls' rate with the host in the loop."""
    from nanodrr_amd import render
    from nanodrr_amd.camera import make_k_inv, make_rt_inv
    from nanodrr_amd.phantom import make_phantom, phantom_subject

    dev = torch.device("cuda:0")
    ph = make_phantom((SIZE,) * 3, (SPACING,) * 3, seed=0)
    res = {"settings": "B=1, C=1, 200x200 at 2.0 mm, sdd 1020, AP at 850 mm, S=500, forward only, 256^3 phantom"}
    for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16)):
        subject = phantom_subject(ph, n_classes=1, dtype=dt, device=dev)
        k_inv = make_k_inv(1020.0, 2.0, 2.0, 0.0, 0.0, 200, 200).to(dev)
        rt = make_rt_inv(torch.zeros(1, 3), torch.tensor([[0.0, 850.0, 0.0]]), "AP",
                         subject.isocenter.cpu()).to(dev)
        sdd = torch.tensor([1020.0], device=dev)
        with torch.no_grad():
            def call():
                return render(subject, k_inv, rt, sdd, 200, 200, n_samples=500)

            ms_eager = time_steps(call, reps, 25)
            for _ in range(3):
                call()
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                call()
            for _ in range(10):
                graph.replay()
            torch.cuda.synchronize()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda._sleep(SPIN_CYCLES)
            t0.record()
            for _ in range(reps):
                graph.replay()
            t1.record()
            torch.cuda.synchronize()
            ms_graph = t0.elapsed_time(t1) / reps
        res[name] = {"fps_graph_gpu_time": 1000.0 / ms_graph, "us_per_drr_graph": ms_graph * 1000.0,
                     "fps_eager": 1000.0 / ms_eager, "us_per_drr_eager": ms_eager * 1000.0}
        del graph, subject
        torch.cuda.empty_cache()
    res["reference_published"] = ("Triton fwd on RTX 6000 Ada, its own volume: 13,332 FPS fp32 eager, "
                                  "17,290 FPS bf16 + torch.compile (GPU time; benchmark.csv:76,81)")
    return res


def cnn_layer_leg(batch: int = 256, reps: int = 20):
    """Per-layer roofline of the pose-proposal CNN at its training shape (128 x 128 inputs, batch
    256): each convolution's forward, input gradient (with the ReLU mask of the layer below) and
    weight + bias gradient, and the pooled MLP head's forward and backward, timed with HIP events
    over back-to-back launches on the launch stream (queued behind a spin kernel). Algorithmic
    bytes: every operand read once and every result written once (bf16 activations and weights,
    fp32 bias and weight gradients); flops: 2 per multiply-add. The bound is the larger of
    bytes / 8 TB/s and flops / 2.5 PFLOP/s (bf16 dense); frac = bound time / measured time."""
    from nanodrr_amd import posenet as pn

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev).manual_seed(0)
    stream = torch.cuda.current_stream()
    MFMA_PEAK = 2.5e15

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(SPIN_CYCLES)
        t0.record(stream)
        for _ in range(reps):
            fn()
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / reps * 1e-3  # seconds

    def row(name, sec, nbytes, flops):
        t_mem, t_mfma = nbytes / (HBM_PEAK_GBS * 1e9), flops / MFMA_PEAK
        bound = "hbm" if t_mem >= t_mfma else "mfma"
        return {"op": name, "us": sec * 1e6, "bytes": nbytes, "flops": flops, "bound": bound,
                "tb_per_s": nbytes / sec / 1e12, "tflop_per_s": flops / sec / 1e12,
                "frac": max(t_mem, t_mfma) / sec}

    rows = []
    H = W = 128
    x = torch.randn(batch, H, W, 1, device=dev, generator=gen).to(torch.bfloat16)
    for i, (cin, cout, st) in enumerate(pn.LAYERS):
        Ho, Wo = (H - 1) // st + 1, (W - 1) // st + 1
        w = (torch.randn(cout, 3, 3, cin, device=dev, generator=gen) * (2.0 / (9 * cin)) ** 0.5).to(torch.bfloat16)
        b = torch.zeros(cout, device=dev)
        x_bytes, y_bytes, w_bytes = x.numel() * 2, batch * Ho * Wo * cout * 2, w.numel() * 2
        flops = 2.0 * batch * Ho * Wo * cout * 9 * cin
        y = pn.conv3x3(x, w, b, st, True)
        dy = torch.randn(y.shape, device=dev, generator=gen).to(torch.bfloat16)
        tag = f"conv{i + 1} {cin}->{cout}{' /2' if st == 2 else ''} {H}x{W}"
        rows.append(row(f"{tag} forward", timed(lambda: pn.conv3x3(x, w, b, st, True)),
                        x_bytes + w_bytes + 4 * cout + y_bytes, flops))
        if i > 0:  # the input of the first layer needs no gradient
            rows.append(row(f"{tag} input gradient", timed(lambda: pn.conv3x3_bwd_data(dy, w, x, st, (H, W))),
                            y_bytes + w_bytes + 2 * x_bytes, flops))
        rows.append(row(f"{tag} weight gradient", timed(lambda: pn.conv3x3_bwd_weight(dy, x, st, bias=True)),
                        y_bytes + x_bytes + 4 * (w.numel() + cout), flops))
        x, H, W = y, Ho, Wo
    feat = x
    K = (H // 2) * (W // 2) * feat.shape[-1]
    w1 = torch.randn(pn.HIDDEN, K, device=dev, generator=gen) / K ** 0.5
    b1 = torch.zeros(pn.HIDDEN, device=dev)
    w2 = torch.randn(pn.OUTPUTS, pn.HIDDEN, device=dev, generator=gen) * 0.01
    b2 = torch.zeros(pn.OUTPUTS, device=dev)
    out, saved = pn.head_forward(feat, w1, b1, w2, b2)
    dout = torch.randn(out.shape, device=dev, generator=gen)
    head_flops = 2.0 * batch * K * pn.HIDDEN
    rows.append(row(f"head pool {H}x{W}x{feat.shape[-1]} + {K}-{pn.HIDDEN}-{pn.OUTPUTS} forward",
                    timed(lambda: pn.head_forward(feat, w1, b1, w2, b2)),
                    feat.numel() * 2 + w1.numel() * 4 + batch * pn.OUTPUTS * 4, head_flops))
    rows.append(row(f"head pool {H}x{W}x{feat.shape[-1]} + {K}-{pn.HIDDEN}-{pn.OUTPUTS} backward",
                    timed(lambda: pn.head_backward(dout, feat, w2, saved)),
                    2 * feat.numel() * 2 + saved[1].numel() * 2 + saved[2].numel() * 2 + w1.numel() * 4,
                    2 * head_flops))
    total = sum(r["us"] for r in rows)
    return {"workload": f"pose-proposal CNN, batch {batch}, 128x128 bf16 inputs, one training step's kernels "
                        f"(weights transposed per input-gradient call; Adam and the DRR renders excluded)",
            "layers": rows, "total_us": total}


def counter_probe(case: str, reps: int = 4):
    """Launch the fwd and bwd ops of one case `reps` times, then copy a 256 MiB
    device buffer three times: the run that rocprofv3 --pmc passes profile (the
    copy calibrates FETCH_SIZE / WRITE_SIZE against known bytes)."""
    import nanodrr_amd.renderer as renderer

    name, layout, adjoint, gvol = parse_case(case)
    renderer.VOLUME_LAYOUT = layout
    subject, k_inv, rt_inv, sdd, _ = make_workload(DTYPES[name], torch.device("cuda:0"), gvol)
    wts = torch.randn(1, N_CLASSES, DET, DET, generator=torch.Generator().manual_seed(0)).cuda()
    f, b, _ = kernel_times(subject, k_inv, rt_inv, sdd, wts, reps, layout, adjoint, gvol)
    print(f"probe {case}: fwd {f * 1000:.1f} us, bwd {b * 1000:.1f} us", flush=True)
    x = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("probe: 3 copies of 256 MiB done", flush=True)


KERNEL_SOURCE = ROOT / "nanodrr_amd" / "csrc" / "raymarch.hip"


def kernel_source_digest() -> str:
    """First 16 hex digits of the sha256 of the ray-march kernel source."""
    import hashlib

    return hashlib.sha256(KERNEL_SOURCE.read_bytes()).hexdigest()[:16]


def measured_traffic(case: str, live_ms: dict, tolerance: float = 0.10, profiles: Path = ROOT / "profiles"):
    """(traffic, note): HBM bytes per launch (read + write) of the case's fwd and bwd kernels from
    the committed rocprofv3 counter summary of this workload, taken only when that summary was
    measured on the current kernel source (its recorded digest of raymarch.hip matches) and its
    kernel-trace average duration for the kernel agrees with this run's HIP-event time within
    `tolerance`; else (None, why)."""
    name, layout, adjoint, gvol = parse_case(case)
    tag = f"{name}_{layout}_{adjoint}" + ("_gvol" if gvol else "")
    path = profiles / f"raymarch_{tag}_counters_bench256_c4_b16.md"
    if not path.exists():
        return None, f"no counter summary {path.name}"
    digest, header, rows = None, None, {}
    for line in path.read_text().splitlines():
        if line.startswith("Kernel source digest:"):
            digest = line.split(":", 1)[1].strip().strip("`")
        elif line.startswith("| role |"):
            header = [c.strip() for c in line.strip().strip("|").split("|")]
        elif header and line.startswith(("| fwd |", "| bwd |")):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            rows[cells[0]] = dict(zip(header, cells))
    if digest != kernel_source_digest():
        return None, f"{path.name} was measured on another kernel source (digest {digest})"
    out, notes = {}, []
    for role, row in rows.items():
        try:
            prof_us = float(row["avg us"])
            traffic = (float(row["HBM read MB"]) + float(row["HBM write MB"])) * 1e6
        except (KeyError, ValueError):
            notes.append(f"{role}: unreadable row")
            continue
        live_us = live_ms[role] * 1000.0
        if abs(prof_us - live_us) > tolerance * live_us:
            notes.append(f"{role}: profiled at {prof_us:.1f} us, {live_us:.1f} us here")
            continue
        out[f"k_{role}"] = traffic
    return (out or None), ("; ".join(notes) or f"{path.name}, kernel source digest {digest}")


def registration_leg(n_iters: int = 300, batch: int = 64, dtype: torch.dtype = torch.bfloat16,
                     layout: str = REGISTRATION_LAYOUT):
    """Batched registration on the 256^3 phantom stored in `dtype` (bf16 by default): 64 problems with targets at +-30 deg /
    +-40 mm, starts perturbed +-10 deg / +-15 mm about the isocentre, 200 x 200 DRRs, -ZNCC
    loss, Adam (1e-2 rad, 1 mm), 300 iterations. Targets are rendered with S=1024 (a smooth
    stand-in for a radiograph); the optimised DRRs use S=256 with the sample lattice jittered
    by a random sub-step offset each iteration. Registrations/s = problems ending with
    mTRE < 1 mm (32 fiducials) / wall time of the optimisation loop."""
    from nanodrr_amd import render
    from nanodrr_amd.metrics import mtre
    from nanodrr_amd.phantom import make_phantom, phantom_subject, registration_problems
    from nanodrr_amd.registration import Registration, register

    dev = torch.device("cuda:0")
    ph = make_phantom((SIZE,) * 3, (SPACING,) * 3, seed=2)
    subject = phantom_subject(ph, n_classes=1, dtype=dtype, device=dev)
    prob = registration_problems(subject, batch, seed=3, height=DET, width=DET)
    k_inv, sdd = prob.k_inv.float().to(dev), prob.sdd.float().to(dev)
    with torch.no_grad():
        target = render(subject, k_inv, prob.rt_true.float().to(dev), sdd, DET, DET, n_samples=4 * S)
    torch.manual_seed(0)
    reg = Registration(subject, prob.rt_init.float().to(dev), k_inv, sdd, DET, DET, per_pose=True, layout=layout)
    fid = prob.fiducials.to(dev)
    before = mtre(reg.pose.detach(), prob.rt_true.to(dev), fid)
    reg(S)  # builds the packed volume once, outside the timed loop
    # warm-up: a few iterations of the same loop on a copy of the problems (first launches, allocations)
    warm = Registration(subject, prob.rt_init.float().to(dev), k_inv, sdd, DET, DET, per_pose=True, layout=layout)
    register(warm, target, n_iters=12, n_samples=S, jitter=True, seed=1, graph=True, iters_per_graph=GRAPH_ITERS)
    del warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    score = register(reg, target, n_iters=n_iters, n_samples=S, jitter=True, seed=0, graph=True,
                     iters_per_graph=GRAPH_ITERS)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    after = mtre(reg.pose.detach(), prob.rt_true.to(dev), fid)
    ok = int((after < 1.0).sum())
    return {
        "workload": f"{batch} problems, 256^3 {str(dtype).replace('torch.', '')}, 200x200, S={S} jittered (targets S={4 * S}), "
                    f"targets +-30deg/+-40mm, starts +-10deg/+-15mm, Adam 1e-2 rad / 1 mm, {n_iters} iterations, "
                    f"fwd+bwd+Adam step captured in a HIP graph of {GRAPH_ITERS} iterations and replayed "
                    f"(capture inside the timed call), {layout} volume layout",
        "registrations_per_s": ok / wall,
        "converged": ok, "problems": batch, "wall_s": wall, "iterations_per_s": n_iters / wall,
        "mtre_mm_median_before": float(before.median()), "mtre_mm_median_after": float(after.median()),
        "zncc_median_after": float(score.median()),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cases", default=DEFAULT_CASES,
                    help="dtype:layout:adjoint[+gvol] cases; the first is the headline, fp32:linear:two the plain "
                         "kernel")
    ap.add_argument("--probe", default=None,
                    help="run only the counter probe of one case, e.g. bf16:coef:single")
    ap.add_argument("--probe-reps", type=int, default=4, help="launches per op in the counter probe")
    ap.add_argument("--no-registration", action="store_true", help="skip the batched registration leg")
    ap.add_argument("--cnn-layers", action="store_true",
                    help="run only the per-layer roofline of the pose-proposal CNN at batch 256")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--registration-only", action="store_true", help="run only the registration leg")
    ap.add_argument("--reference-shape", action="store_true",
                    help="run only the forward at the reference benchmark's shape (B=1, C=1, 200x200 @2 mm, S=500)")
    ap.add_argument("--mixed", action="store_true",
                    help="run only the mixed-precision workload (512x512x400, C=2, B=8, 512x512, S=512)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every figure is one GPU's; --gpus must be 1"}))
        sys.exit(2)
    if args.probe:
        counter_probe(args.probe, args.probe_reps)
        return
    if args.cnn_layers:
        print(json.dumps(cnn_layer_leg()), flush=True)
        return
    if args.registration_only:
        print(json.dumps(registration_leg()), flush=True)
        return
    if args.mixed:
        print(json.dumps(mixed_precision_leg()), flush=True)
        return
    if args.reference_shape:
        print(json.dumps(reference_shape_leg()), flush=True)
        return
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(2)
    import nanodrr_amd.renderer as renderer

    dev = torch.device("cuda:0")
    results = {}
    ph = None
    for case in args.cases.split(","):
        name, layout, adjoint, gvol = parse_case(case)
        dtype = DTYPES[name]
        renderer.VOLUME_LAYOUT = layout
        renderer.ADJOINT = ADJOINTS[adjoint]
        subject, k_inv, rt_inv, sdd, ph = make_workload(dtype, dev, gvol)
        wts = torch.randn(1, N_CLASSES, DET, DET, generator=torch.Generator().manual_seed(0)).to(dev)
        ms_eager = time_steps(lambda: step_fn(subject, k_inv, rt_inv, sdd, wts), args.steps, args.warmup)
        ms = time_steps(graph_step(subject, k_inv, rt_inv, sdd, wts), args.steps, args.warmup)
        ms_fwd, ms_bwd, geom = kernel_times(subject, k_inv, rt_inv, sdd, wts, max(10, args.steps), layout, adjoint,
                                            gvol)
        n_in, b_fwd, b_bwd = algorithmic_bytes(subject, geom, dtype, adjoint, gvol)
        results[case] = dict(ms_per_step=ms, drr_per_s=BATCH * 1000.0 / ms, ms_eager=ms_eager,
                             drr_per_s_eager=BATCH * 1000.0 / ms_eager, ms_fwd=ms_fwd, ms_bwd=ms_bwd,
                             inside_samples=n_in, bytes_fwd=b_fwd, bytes_bwd=b_bwd,
                             gbs_fwd=b_fwd / ms_fwd / 1e6, gbs_bwd=b_bwd / ms_bwd / 1e6)
        print(f"# {case}: {results[case]}", file=sys.stderr, flush=True)
        del subject
        gc.collect()
        torch.cuda.empty_cache()

    renderer.ADJOINT = "single_pass"
    head = next(iter(results))
    head_dtype, head_layout, head_adjoint, _ = parse_case(head)
    r = results[head]
    dominant = "bwd" if r["ms_bwd"] >= r["ms_fwd"] else "fwd"
    achieved = r[f"gbs_{dominant}"]
    traffic, traffic_note = measured_traffic(head, {"fwd": r["ms_fwd"], "bwd": r["ms_bwd"]})
    line = {
        "metric": "DRR/s fwd+bwd 200x200 @256^3 (fp32 vs bf16)",
        "value": r["drr_per_s"],
        "unit": "DRR/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head_dtype,
        "data": "synthetic (seeded 256^3 phantom, 4 labels, 16 random poses)",
        "config": {
            "workload": "fwd+bwd (pose grads) 256^3 @1mm, C=4, B=16, 200x200, S=256",
            "volume_storage": head_dtype, "volume_layout": head_layout, "adjoint": ADJOINTS[head_adjoint],
            "accumulate": "fp32", "global_batch": BATCH, "n_samples": S,
            "step": "render + weighted-sum backward captured in one HIP graph, replayed (eager also given)",
            "per_case": {k: {"drr_per_s": v["drr_per_s"], "ms_per_step": v["ms_per_step"],
                             "drr_per_s_eager": v["drr_per_s_eager"], "ms_per_step_eager": v["ms_eager"],
                             "ms_fwd_kernel": v["ms_fwd"], "ms_bwd_kernel": v["ms_bwd"],
                             "bytes_fwd": v["bytes_fwd"], "bytes_bwd": v["bytes_bwd"],
                             "gbs_fwd": v["gbs_fwd"], "gbs_bwd": v["gbs_bwd"],
                             "frac_fwd": v["gbs_fwd"] / HBM_PEAK_GBS, "frac_bwd": v["gbs_bwd"] / HBM_PEAK_GBS}
                         for k, v in results.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{dominant} ({head_dtype} volume, {head_layout} layout, {ADJOINTS[head_adjoint]} adjoint)",
            "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if traffic is None else traffic.get(f"k_{dominant}"),
            "traffic_source": traffic_note,
            "algorithmic_bytes_per_launch": r[f"bytes_{dominant}"], "inside_samples": r["inside_samples"],
        },
    }
    if "fp32:linear:two" in results and head != "fp32:linear:two":
        plain = results["fp32:linear:two"]
        line["config"]["speedup_over_fp32_plain"] = {
            "step": r["drr_per_s"] / plain["drr_per_s"],
            "kernels": (plain["ms_fwd"] + plain["ms_bwd"]) / (r["ms_fwd"] + r["ms_bwd"]),
        }
    # the cases set the renderer's layout and adjoint per case: back to the defaults
    renderer.VOLUME_LAYOUT, renderer.ADJOINT = "auto", "single_pass"
    gc.collect()
    torch.cuda.empty_cache()  # the march cases' volumes and graph pools
    line["config"]["reference_benchmark_shape"] = reference_shape_leg()
    if not args.no_mixed:
        gc.collect()
        torch.cuda.empty_cache()
        line["mixed_precision"] = mixed_precision_leg()
    if not args.no_registration:
        gc.collect()
        torch.cuda.empty_cache()
        line["registration"] = registration_leg()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(ph)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
