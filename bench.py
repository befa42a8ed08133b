"""Benchmark: DRR/s for forward + analytic backward (pose gradients) on one MI355X.

Workload (fwd+bwd on a 256^3 volume): seeded 256^3 phantom at 1 mm, 4 label
channels (int8 labels), 16 poses (ZXY +-15 deg, +-20 mm about the isocentre
600 mm from the source), 200 x 200 detector at 1 mm, sdd 1020 mm, 256
samples per ray. One step = render(...) of the 16 poses plus the backward of
a weighted sum to the poses, captured once in a HIP graph and replayed (the
eager, uncaptured step is reported too); DRR/s = 16 * steps / elapsed.
Cases (volume dtype : layout the kernels read : adjoint):
  bf16:quad:single     the headline `value`: bf16 quad records, single-pass
                       adjoint (the forward march stores per-ray sums, the
                       backward contracts them);
  fp32:linear:two      the plain fp32 kernel: the (D, H, W) volume, two-pass
                       adjoint (the backward marches the volume again, as
                       nanodrr's kernel does) - the baseline of the speed-ups;
  fp32:quad:single, bf16:stencil:single, bf16:stencil:two  other corners.
Every case accumulates in float32.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cases bf16:quad:single,...] [--no-cpu-baseline]

Prints one JSON line. `roofline` covers the dominant kernel of the headline
case, timed with HIP events on the stream it runs on; `cpu_baseline` times
the pure-PyTorch grid_sample restatement (fp32) of the same workload on the
host cores for a bounded sample of DRRs; `registration` is the batched
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
# Layout of the registration renders (single channel): the quad records (136 MB at 256^3 bf16) stay in
# the Infinity Cache; the coefficient records (271 MB) do not, and ran 25 % slower with the same convergence.
REGISTRATION_LAYOUT = "quad"


def make_workload(dtype, device):
    from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses

    ph = make_phantom((SIZE,) * 3, (SPACING,) * 3, seed=0)
    subject = phantom_subject(ph, n_classes=N_CLASSES, dtype=dtype, device=device)
    rt_inv, _, _ = sample_poses(BATCH, seed=1, isocenter=subject.isocenter)
    k_inv, sdd = detector(1020.0, 1.0, DET, DET)
    return subject, k_inv.to(device), rt_inv.to(device), sdd.to(device), ph


DTYPES = {"bf16": torch.bfloat16, "fp32": torch.float32, "fp16": torch.float16}
ADJOINTS = {"single": "single_pass", "two": "two_pass"}


def parse_case(case: str):
    """"dtype:layout[:adjoint]" -> (dtype name, layout, adjoint name); the adjoint
    defaults to the single pass."""
    parts = case.split(":")
    name, layout = parts[0], parts[1]
    adjoint = parts[2] if len(parts) > 2 else "single"
    if name not in DTYPES or adjoint not in ADJOINTS:
        raise ValueError(f"bad case {case!r}: expected dtype:layout[:single|two]")
    return name, layout, adjoint


def step_fn(subject, k_inv, rt_inv, sdd, wts):
    from nanodrr_amd import render

    rt = rt_inv.detach().requires_grad_(True)
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
        rt.grad = None
        out = render(subject, k_inv, rt, sdd, DET, DET, n_samples=S)
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


def kernel_times(subject, k_inv, rt_inv, sdd, wts, reps, layout, adjoint="two"):
    """Average duration (ms) of one forward and one backward op launch, HIP events
    recorded on the stream the ops launch on (torch's current stream). The
    single-pass forward also writes the adjoint's per-ray sums and its backward
    is the contraction."""
    from nanodrr_amd._ops import fused_raymarch, fused_raymarch_bwd, fused_raymarch_bwd_jac, fused_raymarch_jac
    from nanodrr_amd.renderer import _make_src, _make_tgt

    vol = subject.image[0, 0]
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
        ms_bwd = timed(lambda: fused_raymarch_bwd(go, vol, lab, M, src, tgt, step, S, N_CLASSES, DET, False, False,
                                                  packed, layout))
    return ms_fwd, ms_bwd, (M, src, tgt)


def algorithmic_bytes(subject, geom, dtype, adjoint="two"):
    from nanodrr_amd.roofline import contraction_bytes, inside_samples, march_bytes

    M, src, tgt = geom
    n_in = int(inside_samples(M, src, tgt, S, tuple(subject.image.shape[-3:])).sum())
    n_rays = M.shape[0] * tgt.shape[1]
    if adjoint == "single":
        fwd = march_bytes(n_in, n_rays, M.shape[0], N_CLASSES, dtype, jac_planes=6)
        bwd = contraction_bytes(n_rays, M.shape[0], N_CLASSES, 6)
    else:
        fwd = march_bytes(n_in, n_rays, M.shape[0], N_CLASSES, dtype)
        bwd = march_bytes(n_in, n_rays, M.shape[0], N_CLASSES, dtype, backward=True)
    return n_in, fwd, bwd


def _cpu_drr_rate(ph, n_classes, threads, budget_s, max_runs, n_samples=S):
    """(DRR/s, runs): median rate of one DRR (fwd + pose backward, 1 pose) through the
    pure-PyTorch grid_sample restatement (oracle.torch_ref, fp32) on `threads` threads,
    repeated for about `budget_s` seconds (at least 2 runs, at most `max_runs`)."""
    from nanodrr_amd.phantom import detector, phantom_subject, sample_poses
    from oracle import torch_ref

    torch.set_num_threads(threads)
    subject = phantom_subject(ph, n_classes=n_classes)
    rt_all, _, _ = sample_poses(BATCH, seed=1, isocenter=subject.isocenter)
    k_inv, sdd = detector(1020.0, 1.0, DET, DET)
    image, label, w2g = subject.image, subject.label, subject.world_to_grid
    wts = torch.randn(1, n_classes, DET, DET, generator=torch.Generator().manual_seed(0))

    def one(i):
        rt = rt_all[i % BATCH: i % BATCH + 1].clone().requires_grad_(True)
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


def cpu_baseline(ph, budget_s=12.0):
    """The workload's DRRs through the pure-PyTorch CPU restatement on the host's cores (up to
    16), plus a one-thread rate and the small single-channel CPU case (128^3 at 1.5 mm, C=1,
    200 x 200, S=256) on both thread counts."""
    from nanodrr_amd.phantom import make_phantom

    threads = min(16, os.cpu_count() or 1)
    rate, runs = _cpu_drr_rate(ph, N_CLASSES, threads, budget_s, 40)
    rate1, runs1 = _cpu_drr_rate(ph, N_CLASSES, 1, 4.0, 3)
    ph0 = make_phantom((128,) * 3, (1.5,) * 3, seed=0)
    rate0, runs0 = _cpu_drr_rate(ph0, 1, threads, 3.0, 20)
    rate01, runs01 = _cpu_drr_rate(ph0, 1, 1, 3.0, 5)
    torch.set_num_threads(threads)
    return {"value": rate, "unit": "DRR/s", "cores": threads, "kind": "port",
            "sample": f"{runs} DRRs (1 pose each) of the 256^3/C=4/200x200/S=256 workload, fwd + pose "
                      f"backward, oracle.torch_ref fp32 grid_sample on {threads} threads, median of runs",
            "one_thread": {"value": rate1, "sample": f"{runs1} DRRs of the same workload on 1 thread"},
            "cpu_case_128": {"value": rate0, "value_one_thread": rate01, "unit": "DRR/s",
                             "sample": f"{runs0} / {runs01} DRRs, 128^3 at 1.5 mm, C=1, 200x200 at 1.0 mm, "
                                       f"S=256, fwd + pose backward, on {threads} / 1 threads"}}


def mixed_precision_leg(steps: int = 20, warmup: int = 5):
    """The mixed-precision workload at full size: a 512 x 512 x 400 phantom at 0.8 x 0.8 x 1.0 mm
    stored in bf16 and in fp16, 2 labels, 8 poses, 512 x 512 detector at 0.8 mm, S = 512
    (compensated sums), fwd + pose backward (single-pass adjoint, quad layout), graph-replayed
    steps timed with HIP events, plus the plain fp32 kernel on the fp32 volume."""
    import nanodrr_amd.renderer as renderer
    from nanodrr_amd import render
    from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses

    dev = torch.device("cuda:0")
    ph = make_phantom((400, 512, 512), (0.8, 0.8, 1.0), seed=0)
    B, det, n_s = 8, 512, 512
    res = {"workload": "512x512x400 @0.8x0.8x1.0 mm, C=2, B=8, 512x512 @0.8 mm, S=512 (compensated), fwd + pose "
                       "backward, graph-replayed step"}
    for name, dt, layout, adjoint in (("bf16", torch.bfloat16, "quad", "single_pass"),
                                      ("fp16", torch.float16, "quad", "single_pass"),
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
        res[name] = {"ms_per_step": ms, "drr_per_s": B * 1000.0 / ms}
        del graph, subject, out
        gc.collect()
        torch.cuda.empty_cache()
    renderer.VOLUME_LAYOUT, renderer.ADJOINT = "auto", "single_pass"
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


def counter_probe(case: str, reps: int = 4):
    """Launch the fwd and bwd ops of one case `reps` times, then copy a 256 MiB
    device buffer three times: the run that rocprofv3 --pmc passes profile (the
    copy calibrates FETCH_SIZE / WRITE_SIZE against known bytes)."""
    import nanodrr_amd.renderer as renderer

    name, layout, adjoint = parse_case(case)
    renderer.VOLUME_LAYOUT = layout
    subject, k_inv, rt_inv, sdd, _ = make_workload(DTYPES[name], torch.device("cuda:0"))
    wts = torch.randn(1, N_CLASSES, DET, DET, generator=torch.Generator().manual_seed(0)).cuda()
    f, b, _ = kernel_times(subject, k_inv, rt_inv, sdd, wts, reps, layout, adjoint)
    print(f"probe {case}: fwd {f * 1000:.1f} us, bwd {b * 1000:.1f} us", flush=True)
    x = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("probe: 3 copies of 256 MiB done", flush=True)


def measured_traffic(case: str):
    """HBM bytes per launch (read + write, MB) of the case's fwd and bwd kernels, from
    the committed rocprofv3 counter summary of this workload, or None."""
    name, layout, adjoint = parse_case(case)
    path = ROOT / "profiles" / f"raymarch_{name}_{layout}_{adjoint}_counters_bench256_c4_b16.md"
    if not path.exists():
        return None
    out = {}
    for line in path.read_text().splitlines():
        if line.startswith("| k_fwd") or line.startswith("| k_bwd"):
            cells = [c.strip() for c in line.strip("|").split("|")]
            out[cells[0][:5]] = (float(cells[2]) + float(cells[4])) * 1e6
    return out or None


def registration_leg(n_iters: int = 300, batch: int = 64):
    """Batched registration on the 256^3 bf16 phantom: 64 problems with targets at +-30 deg /
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
    subject = phantom_subject(ph, n_classes=1, dtype=torch.bfloat16, device=dev)
    prob = registration_problems(subject, batch, seed=3, height=DET, width=DET)
    k_inv, sdd = prob.k_inv.float().to(dev), prob.sdd.float().to(dev)
    with torch.no_grad():
        target = render(subject, k_inv, prob.rt_true.float().to(dev), sdd, DET, DET, n_samples=4 * S)
    torch.manual_seed(0)
    reg = Registration(subject, prob.rt_init.float().to(dev), k_inv, sdd, DET, DET, per_pose=True,
                       layout=REGISTRATION_LAYOUT)
    fid = prob.fiducials.to(dev)
    before = mtre(reg.pose.detach(), prob.rt_true.to(dev), fid)
    reg(S)  # builds the packed volume once, outside the timed loop
    # warm-up: a few iterations of the same loop on a copy of the problems (first launches, allocations)
    warm = Registration(subject, prob.rt_init.float().to(dev), k_inv, sdd, DET, DET, per_pose=True,
                        layout=REGISTRATION_LAYOUT)
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
        "workload": f"{batch} problems, 256^3 bf16, 200x200, S={S} jittered (targets S={4 * S}), "
                    f"targets +-30deg/+-40mm, starts +-10deg/+-15mm, Adam 1e-2 rad / 1 mm, {n_iters} iterations, "
                    f"fwd+bwd+Adam step captured in a HIP graph of {GRAPH_ITERS} iterations and replayed "
                    f"(capture inside the timed call), {REGISTRATION_LAYOUT} volume layout",
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
    ap.add_argument("--cases",
                    default="bf16:quad:single,fp32:linear:two,fp32:quad:single,bf16:stencil:single,bf16:stencil:two",
                    help="dtype:layout:adjoint cases; the first is the headline, fp32:linear:two the plain kernel")
    ap.add_argument("--probe", default=None,
                    help="run only the counter probe of one case, e.g. bf16:stencil:single")
    ap.add_argument("--no-registration", action="store_true", help="skip the batched registration leg")
    ap.add_argument("--registration-only", action="store_true", help="run only the registration leg")
    ap.add_argument("--reference-shape", action="store_true",
                    help="run only the forward at the reference benchmark's shape (B=1, C=1, 200x200 @2 mm, S=500)")
    ap.add_argument("--mixed", action="store_true",
                    help="run only the mixed-precision workload (512x512x400, C=2, B=8, 512x512, S=512)")
    args = ap.parse_args()
    if args.probe:
        counter_probe(args.probe)
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
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(2)
    import nanodrr_amd.renderer as renderer

    dev = torch.device("cuda:0")
    results = {}
    ph = None
    for case in args.cases.split(","):
        name, layout, adjoint = parse_case(case)
        dtype = DTYPES[name]
        renderer.VOLUME_LAYOUT = layout
        renderer.ADJOINT = ADJOINTS[adjoint]
        subject, k_inv, rt_inv, sdd, ph = make_workload(dtype, dev)
        wts = torch.randn(1, N_CLASSES, DET, DET, generator=torch.Generator().manual_seed(0)).to(dev)
        ms_eager = time_steps(lambda: step_fn(subject, k_inv, rt_inv, sdd, wts), args.steps, args.warmup)
        ms = time_steps(graph_step(subject, k_inv, rt_inv, sdd, wts), args.steps, args.warmup)
        ms_fwd, ms_bwd, geom = kernel_times(subject, k_inv, rt_inv, sdd, wts, max(10, args.steps), layout, adjoint)
        n_in, b_fwd, b_bwd = algorithmic_bytes(subject, geom, dtype, adjoint)
        results[case] = dict(ms_per_step=ms, drr_per_s=BATCH * 1000.0 / ms, ms_eager=ms_eager,
                             drr_per_s_eager=BATCH * 1000.0 / ms_eager, ms_fwd=ms_fwd, ms_bwd=ms_bwd,
                             inside_samples=n_in, bytes_fwd=b_fwd, bytes_bwd=b_bwd,
                             gbs_fwd=b_fwd / ms_fwd / 1e6, gbs_bwd=b_bwd / ms_bwd / 1e6)
        print(f"# {case}: {results[case]}", file=sys.stderr, flush=True)
        del subject
        torch.cuda.empty_cache()

    renderer.ADJOINT = "single_pass"
    head = next(iter(results))
    head_dtype, head_layout, head_adjoint = parse_case(head)
    r = results[head]
    dominant = "bwd" if r["ms_bwd"] >= r["ms_fwd"] else "fwd"
    achieved = r[f"gbs_{dominant}"]
    traffic = measured_traffic(head)
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
                              "gbs_fwd": v["gbs_fwd"], "gbs_bwd": v["gbs_bwd"]} for k, v in results.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{dominant} ({head_dtype} volume, {head_layout} layout, {ADJOINTS[head_adjoint]} adjoint)",
            "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if traffic is None else traffic.get(f"k_{dominant}"),
            "traffic_source": "profiles/ rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated" if traffic else None,
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
    if not args.no_registration:
        gc.collect()
        torch.cuda.empty_cache()
        line["registration"] = registration_leg()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(ph)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
