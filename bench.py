"""Benchmark of the vd.fft FFT-convolution path on one MI355X.

Default workload (the metric's configuration): FP32 2-D FFT-convolution,
64 real float32 images of 512x512 drawn U(-1, 1) from seed 0, one 512x512
kernel = isotropic Gaussian (sigma 4 px, circular, peak 1 at (0,0)) + 1%
N(0,1) noise, pre-transformed K = rfft2(k) (not timed), broadcast over the 64
images.  One step = ``vd.fft.convolve2DR(images, K, kernel_inner_only=True)``:
irfft2(rfft2(x) * conj(K)) in place.  Every step runs on its own resident copy
of the input (all copies in HBM before timing), so repeated in-place steps do
not feed results back into inputs.

Output: one JSON line with the metric (GB/s of algorithmic bytes: 64 MiB in,
64 MiB out, 1 MiB kernel spectrum per step), the roofline of the dominant
kernel (HIP events on the launch stream), the FP64-oracle parity of the
result, and the CPU baseline (the FP64 oracle, single core, on a bounded
sample; plus multithreaded pocketfft via scipy).

Other workloads (``--workload c2c4096``) print their own line for reference.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_FLOPS = 2.5e15  # MI355X dense bf16 / fp16 matrix peak


def conv_inputs(batch=64, n=512, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (batch, n, n)).astype(np.float32)
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    dy, dx = np.minimum(yy, n - yy), np.minimum(xx, n - xx)
    g = np.exp(-(dy.astype(np.float64) ** 2 + dx ** 2) / (2 * 4.0 ** 2))
    k = (g + 0.01 * rng.standard_normal((n, n))).astype(np.float32)
    return x, k


def _cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline_conv(x, K, budget_s=12.0):
    """FP64 oracle (single core) on a bounded sample of images, scaled to GB/s."""
    import oracle as O
    n = x.shape[-1]
    per_img_bytes = 2 * n * n * 4
    imgs, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s and imgs < x.shape[0]:
        O.convolve2DR(x[imgs:imgs + 1], K, kernel_inner_only=True)
        imgs += 1
    dt = time.perf_counter() - t0
    oracle_gbs = imgs * per_img_bytes / dt / 1e9
    extra = None
    try:
        import scipy.fft as sf
        w = os.cpu_count() or 1
        Kc = K.astype(np.complex64)
        sf.irfft2(sf.rfft2(x[:4], workers=w) * np.conj(Kc), s=(n, n), workers=w)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 5.0 and reps < 20:
            sf.irfft2(sf.rfft2(x, workers=w) * np.conj(Kc), s=(n, n), workers=w)
            reps += 1
        dt2 = (time.perf_counter() - t0) / reps
        extra = {"value": (x.nbytes * 2 + Kc.nbytes) / dt2 / 1e9, "unit": "GB/s", "cores": w,
                 "kind": "pocketfft (scipy.fft, float32)", "sample": f"{x.shape[0]} images x {reps} reps"}
    except Exception as e:  # scipy missing: report only the oracle
        extra = {"error": str(e)}
    return {"value": oracle_gbs, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{imgs} of {x.shape[0]} images through the FP64 oracle in {dt:.1f} s"}, extra


def committed_traffic(kernel_name):
    """HBM bytes per dispatch of ``kernel_name`` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries committed under profiles/ (lines 'traffic_bytes <kernel> <bytes>',
    already scaled by the copy-kernel calibration), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"traffic_bytes (\S+) (\d+)", line.strip())
                if m and m.group(1) == kernel_name:
                    return int(m.group(2))
    return None


def bench_conv(steps, warmup, args):
    import vkdispatch_amd as vd
    from vkdispatch_amd.fft import fused
    import oracle as O

    batch, n = 64, 512
    x, k = conv_inputs(batch, n, 0)
    K = np.fft.rfft2(k.astype(np.float64)).astype(np.complex64)
    kbuf = vd.RFFTBuffer((n, n))
    kbuf.write_fourier(K)
    src = vd.asrfftbuffer(x)
    bufs = []
    for _ in range(warmup + steps):
        b = vd.RFFTBuffer((batch, n, n))
        vd.runtime.dtod(b.ptr, src.ptr, src.mem_size)
        bufs.append(b)
    vd.synchronize()

    # each step convolves a fresh copy of the input (no step re-reads another's data);
    # launched directly (back to back on the stream) and as one replayed hipGraph per
    # step, the faster of the two is the bench value (both reported)
    def run_direct(i):
        vd.fft.convolve2DR(bufs[i], kbuf, kernel_inner_only=True)

    graphs = []
    for b in bufs:
        g = vd.CommandGraph()
        vd.fft.convolve2DR(b, kbuf, kernel_inner_only=True, graph=g)
        graphs.append(g)
    for g in graphs:
        g.capture()

    def timed_wall(run):
        for i in range(warmup):
            run(i)
        vd.synchronize()
        t0 = time.perf_counter()
        for i in range(warmup, warmup + steps):
            run(i)
        vd.synchronize()
        return time.perf_counter() - t0

    wall_graph = timed_wall(lambda i: graphs[i].submit())
    for b in bufs:  # restore the inputs
        vd.runtime.dtod(b.ptr, src.ptr, src.mem_size)
    vd.synchronize()
    wall_direct = timed_wall(run_direct)
    wall = min(wall_graph, wall_direct)
    ms = wall / steps * 1e3

    alg_bytes = batch * n * n * 4 * 2 + n * (n // 2 + 1) * 8
    value = alg_bytes / (wall / steps) / 1e9

    # parity on two images of the last step (same input as every step)
    res = bufs[-1].read_real(0)
    ref = O.convolve2DR(x[:2], K, kernel_inner_only=True)
    rel = O.rel_l2(res[:2], ref)

    # dominant-kernel timing with HIP events on the launch stream
    prof = fused.profile_convolve2DR(bufs[0], kbuf, reps=20)
    dom = max(prof, key=lambda p: p["ms"])
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": committed_traffic(dom["name"]),
                "kernel": dom["name"],
                "kernel_ms": round(dom["ms"], 5), "kernel_bytes": dom["bytes"]}
    out = {
        "metric": "FFT-conv GB/s & %HBM roofline; BF16-MFMA vs FP32 speedup at rel-L2<=1e-3",
        "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "fp32 2-D FFT-convolution, 64 x 512x512 real images, one 512x512 kernel "
                               "(Gaussian sigma=4 + 1% noise) broadcast, convolve2DR in place",
                   "batch": batch, "shape": [n, n], "algorithmic_bytes_per_step": alg_bytes},
        "roofline": roofline,
        "hbm_fraction_of_step": round(value / HBM_PEAK_GBS, 4),
        "ms_per_step_direct": round(wall_direct / steps * 1e3, 5),
        "ms_per_step_graph": round(wall_graph / steps * 1e3, 5),
        "parity_rel_l2_vs_fp64": rel,
        "kernels": prof,
    }
    if not args.no_cpu:
        base, extra = cpu_baseline_conv(x, K)
        out["cpu_baseline"] = base
        out["cpu_baseline_pocketfft"] = extra
        model, ncpu = _cpu_info()
        out["cpu_host"] = {"model": model, "os_cpu_count": ncpu}
    return out


def bench_c2c(steps, warmup, args):
    import vkdispatch_amd as vd
    import oracle as O
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((1024, 4096)) + 1j * rng.standard_normal((1024, 4096))).astype(np.complex64)
    buf = vd.asbuffer(x)
    g = vd.CommandGraph()
    vd.fft.fft(buf, graph=g)
    vd.fft.ifft(buf, graph=g)
    g.capture()
    for _ in range(warmup):
        g.submit()
    vd.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.submit()
    vd.synchronize()
    wall = (time.perf_counter() - t0) / steps
    alg = 2 * 2 * x.nbytes
    res = buf.read(0)
    k_fwd = vd.fft.make_fft_shader((1024, 4096), None)
    ms_fwd = _time_kernel(k_fwd, [buf], 20)
    return {"metric": "C2C N=4096 x 1024 fwd+inv GB/s", "value": round(alg / wall / 1e9, 2), "unit": "GB/s",
            "ms_per_step": round(wall * 1e3, 5), "steps": steps, "warmup": warmup,
            "fwd_kernel_ms": ms_fwd, "fwd_kernel_gbs": round(2 * x.nbytes / (ms_fwd * 1e-3) / 1e9, 1),
            "roundtrip_rel_l2": O.rel_l2(res, x)}


def bench_ragged(steps, warmup, args):
    """4096 ragged C2C transforms (60% pow2, 30% 7-smooth, 10% primes): the persistent
    ragged kernel, one launch per tile level, plan picked by the autotuner, replayed as
    a captured graph."""
    import vkdispatch_amd as vd
    from vkdispatch_amd.fft import ragged as R
    import oracle as O
    sizes = R.mixed_sizes(4096, seed=0)
    parts, packed = R.packed_normal(sizes, 0)
    src = vd.asbuffer(packed)
    buf = vd.Buffer((packed.size,), vd.complex64)
    tuned = R.tune(sizes, buf, reps=3, cache=bool(getattr(args, "cached_tune", False)))
    best = tuned["best"]
    plan = R.RaggedFFT(sizes, **{k: best[k] for k in R.TUNE_KEYS})
    g = vd.CommandGraph()
    g.record_callable(plan.bind(buf, src))
    g.capture()
    for _ in range(warmup):
        g.submit()
    vd.synchronize()
    e0, e1 = vd.Event(), vd.Event()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        g.submit()
    e1.record()
    vd.synchronize()
    wall = (time.perf_counter() - t0) / steps
    gpu_ms = e0.elapsed_ms(e1) / steps
    out = buf.read(0)
    worst, off = 0.0, 0
    for n, x in zip(sizes, parts):
        worst = max(worst, O.rel_l2(out[off:off + n], np.fft.fft(x.astype(np.complex64).astype(np.complex128))))
        off += n
    alg = 2 * packed.nbytes
    return {"metric": "ragged 4096 x C2C (64..65536) GFLOP/s", "value": round(plan.flops() / wall / 1e9, 1),
            "unit": "GFLOP/s", "ms_per_step": round(wall * 1e3, 4), "gpu_ms_per_step": round(gpu_ms, 4),
            "GB_s": round(alg / wall / 1e9, 1), "points": int(sum(sizes)), "tuned": tuned,
            "breakdown": plan.breakdown(), "worst_rel_l2_vs_fp64": worst}


def bench_mixed3d(steps, warmup, args):
    """3-D C2C 256^3: bf16/fp16 storage with MFMA radix-16 stages vs the FP32 VALU path."""
    import vkdispatch_amd as vd
    from vkdispatch_amd.fft.mixed import Mixed3D
    import oracle as O
    shape = (256, 256, 256)
    rng = np.random.default_rng(0)
    z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)

    def timed(run):
        for _ in range(warmup):
            run()
        vd.synchronize()
        e0, e1 = vd.Event(), vd.Event()
        e0.record()
        for _ in range(steps):
            run()
        e1.record()
        e1.synchronize()
        return e0.elapsed_ms(e1) / steps

    res = {}
    # FP32 baseline: vd.fft.fft3 in place on complex64 (three LDS-butterfly passes)
    b32 = vd.asbuffer(z)
    ref_unq = np.fft.fftn(z.astype(np.complex128))
    legs = {"fp32": lambda graph=None: vd.fft.fft3(b32, graph=graph)}
    bufs = {}
    for store, vt in (("bf16", vd.complex_bf16), ("fp16", vd.complex32)):
        inp = vd.Buffer(shape, vt)
        inp.write(z)
        out = vd.Buffer(shape, vd.complex64)
        plan = Mixed3D(shape, vt)
        bufs[store] = (inp, out, plan)
        legs[store] = (lambda p, o, i: (lambda graph=None: p(o, i, graph=graph)))(plan, out, inp)
    # the legs are timed in interleaved rounds (same clocks and cache history for all);
    # each leg's time is the median over the rounds of back-to-back direct launches
    for run in legs.values():
        for _ in range(warmup):
            run()
    vd.synchronize()
    rounds = 7
    samples = {k: [] for k in legs}
    for _ in range(rounds):
        for k, run in legs.items():
            e0, e1 = vd.Event(), vd.Event()
            e0.record()
            for _ in range(steps):
                run()
            e1.record()
            e1.synchronize()
            samples[k].append(e0.elapsed_ms(e1) / steps)
    med = {k: float(np.median(v)) for k, v in samples.items()}
    graph_ms = {}
    for k, run in legs.items():  # one replayed one-transform hipGraph per leg, for reference
        g = vd.CommandGraph()
        run(graph=g)
        g.capture()
        graph_ms[k] = timed(g.submit)
    ms32 = med["fp32"]
    b32.write(z)
    vd.fft.fft3(b32)
    res["fp32"] = {"ms": ms32, "ms_rounds": samples["fp32"], "ms_graph": graph_ms["fp32"],
                   "rel_l2_vs_fp64": O.rel_l2(b32.read(0), ref_unq), "bytes": 3 * 2 * z.nbytes}
    del b32
    for store in ("bf16", "fp16"):
        inp, out, plan = bufs[store]
        xq = inp.read(0)
        plan(out, inp)
        r = out.read(0)
        ms = med[store]
        res[store] = {"ms": ms, "ms_rounds": samples[store], "ms_graph": graph_ms[store],
                      "rel_l2_vs_fp64_quantized": O.rel_l2(r, np.fft.fftn(xq.astype(np.complex128))),
                      "rel_l2_vs_fp64_unquantized": O.rel_l2(r, ref_unq),
                      "passes": plan.npasses, "bytes": z.size * (4 + 8 * (plan.npasses - 1) + 8),
                      "mfma_tflops": plan.mfma_count() * 16384 / (ms * 1e-3) / 1e12,
                      "mfma_utilisation": plan.mfma_count() * 16384 / (ms * 1e-3) / MFMA_PEAK_FLOPS,
                      "speedup_vs_fp32": ms32 / ms}
    del bufs
    return {"metric": "3-D C2C 256^3: BF16-MFMA split-precision vs FP32 speedup at rel-L2<=1e-3",
            "value": round(res["bf16"]["speedup_vs_fp32"], 3), "unit": "x", "steps": steps, "warmup": warmup,
            "detail": res}


def bench_autograd(steps, warmup, args):
    """Differentiable convolution, 32 x 1024^2 float32 with a learnable 1024^2 kernel:
    forward + loss 0.5||y - t||^2 + backward (dx, dk) per step, against the same step
    written with torch.fft (rocFFT) autograd."""
    import torch
    from vkdispatch_amd.fft.autograd import fft_convolve2d
    B, H, W = 32, 1024, 1024
    rng = np.random.default_rng(0)
    dev = "cuda"
    x = torch.tensor(rng.standard_normal((B, H, W), dtype=np.float32), device=dev, requires_grad=True)
    k = torch.tensor((0.01 * rng.standard_normal((H, W))).astype(np.float32), device=dev, requires_grad=True)
    t = torch.tensor(rng.standard_normal((B, H, W), dtype=np.float32), device=dev)

    def ours():
        y = fft_convolve2d(x, k)
        return 0.5 * ((y - t) ** 2).sum()

    def vendor():
        y = torch.fft.irfft2(torch.fft.rfft2(x) * torch.conj(torch.fft.rfft2(k)), s=(H, W))
        return 0.5 * ((y - t) ** 2).sum()

    def timed(loss_fn):
        for _ in range(warmup):
            x.grad = k.grad = None
            loss_fn().backward()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            x.grad = k.grad = None
            loss_fn().backward()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    ms = timed(ours)
    dx, dk = x.grad.detach().clone(), k.grad.detach().clone()
    ms_vendor = timed(vendor)
    dx_v, dk_v = x.grad.detach(), k.grad.detach()
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))  # noqa: E731
    img = B * H * W * 4
    # FFT path bytes per step: forward reads x and k, writes y; backward reads g and x,
    # writes dx and dk (the loss's own elementwise kernels are not counted)
    alg = (img + H * W * 4 + img) + (2 * img + img + H * W * 4)
    return {"metric": "differentiable 2-D FFT convolution, forward + backward per step", "value": round(ms, 4),
            "unit": "ms", "higher_is_better": False, "steps": steps, "warmup": warmup,
            "config": {"workload": "32 x 1024^2 float32 images, learnable 1024^2 kernel, loss 0.5||y - t||^2"},
            "fft_path_bytes_per_step": alg, "fft_path_gbs_lower_bound": round(alg / (ms * 1e-3) / 1e9, 1),
            "torch_fft_autograd_ms": round(ms_vendor, 4), "speedup_vs_torch_fft": round(ms_vendor / ms, 3),
            "rel_l2_dx_vs_torch_fft": rel(dx, dx_v), "rel_l2_dk_vs_torch_fft": rel(dk, dk_v)}


def bench_copy(steps, warmup, args):
    """Plain device-to-device copy of 128 MiB (calibration of the HBM byte counters)."""
    import vkdispatch_amd as vd
    nbytes = 128 << 20
    a, b = vd.Buffer((nbytes // 8,), vd.complex64), vd.Buffer((nbytes // 8,), vd.complex64)
    vd.runtime.memset(a.ptr, 1, nbytes)
    for _ in range(warmup):
        vd.runtime.dtod(b.ptr, a.ptr, nbytes)
    vd.synchronize()
    e0, e1 = vd.Event(), vd.Event()
    e0.record()
    for _ in range(steps):
        vd.runtime.dtod(b.ptr, a.ptr, nbytes)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    return {"metric": "device copy 128 MiB", "value": round(2 * nbytes / (ms * 1e-3) / 1e9, 1), "unit": "GB/s",
            "ms": ms, "bytes_read": nbytes, "bytes_written": nbytes}


def _time_kernel(kernel, values, reps):
    import vkdispatch_amd as vd
    e0, e1 = vd.Event(), vd.Event()
    kernel.launch(values)
    e0.record()
    for _ in range(reps):
        kernel.launch(values)
    e1.record()
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="conv2d", choices=["conv2d", "c2c4096", "ragged", "mixed3d", "autograd", "copy"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-mixed", action="store_true", help="conv2d: skip the BF16-MFMA 256^3 comparison")
    ap.add_argument("--cached-tune", action="store_true",
                    help="ragged: use the stored autotuner pick for the batch instead of timing the candidates")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "vkdispatch_amd benchmarks one GPU per process; multi-GPU sharding is "
                                   "a design note only and is not run", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "conv2d":
        out = bench_conv(args.steps, args.warmup, args)
        if not args.no_mixed:
            # the metric's second half: BF16-MFMA 256^3 transform vs the FP32 path
            m = bench_mixed3d(args.steps, args.warmup, args)
            d = m["detail"]
            out["bf16_mfma_vs_fp32"] = {
                "value": m["value"], "unit": "x", "workload": "3-D C2C 256^3, bf16 storage, fp32 output",
                "ms_bf16": d["bf16"]["ms"], "ms_fp32": d["fp32"]["ms"],
                "rel_l2_vs_fp64_quantized": d["bf16"]["rel_l2_vs_fp64_quantized"],
                "mfma_utilisation": d["bf16"]["mfma_utilisation"],
                "fp16_storage_speedup": d["fp16"]["speedup_vs_fp32"],
                "timing": "median of 7 interleaved rounds of back-to-back launches per leg"}
    elif args.workload == "ragged":
        out = bench_ragged(args.steps, args.warmup, args)
    elif args.workload == "mixed3d":
        out = bench_mixed3d(args.steps, args.warmup, args)
    elif args.workload == "autograd":
        out = bench_autograd(args.steps, args.warmup, args)
    elif args.workload == "copy":
        out = bench_copy(args.steps, args.warmup, args)
    else:
        out = bench_c2c(args.steps, args.warmup, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
