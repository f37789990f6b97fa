#!/usr/bin/env python3
"""Throughput of the FFT engine on one MI355X.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-cpu] [--no-extras]

Workload (the headline line): `fft` Forward along the innermost axis of an Array DIM2
(Complex Float) of 16384 x 4096, re/im uniform[-1, 1) -- 16384 independent 4096-point
transforms, inputs resident in HBM before the timed region.  One step = one call.

  value      GFLOP/s by the 5 N log2 N convention: 16384 * 5 * 4096 * 12 flop per step.
  roofline   HBM-bound; achieved = algorithmic bytes (one read + one write of the array,
             2 * 16384 * 4096 * 8 B = 1 GiB per launch) / average kernel duration, timed with
             HIP events on the stream the kernel is launched on; peak 8 TB/s.  `traffic` is the
             HBM bytes per launch measured with rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by a
             copy kernel of known size), read from profiles/ when present.
  cpu_baseline  the oracle's recursive FFT in float arithmetic (the CPU path restated; FFTW is
             not installed), over the same 16384 x 4096 array, on every host core available.
  extras     the other configurations (fft1D 1024 Complex Double Forward then Inverse, four-step 2^22
             row, Reverse/Inverse, fft2D 4096^2, batched
             fft2D 64 x 512^2, fft3D 256^3 Complex Double, ragged batch, spectral convolution) with their own
             time and roofline fraction; the ragged batch and the spectral convolution also
             against the plain per-call path on the same GPU.

  Before the W warmup steps, untimed steps run for 50 ms so that a fresh process's clocks have
  ramped up (the first ~30 calls otherwise run 5-8% slower).

  --profile-headline runs only the headline plan (and a plain device copy of the same array,
             the reference for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE) for profiling.
  --profile-workload W  runs only workload W (fft2d, fft3d, four_step, ragged, spectral) for
             warmup + steps calls, for profiling.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "accelerate-fft_amd"))

HBM_PEAK_GBS = 8000.0
RAGGED_SIZES = sorted({b << k for b in (1, 3, 5) for k in range(4, 14)} | {257, 1009, 4099})


def ragged_lengths(count=10000, seed=4):
    """Lengths drawn log-uniformly from [16, 40960] and snapped (in log distance) to the size set
    {2^k, 3*2^k, 5*2^k : k = 4..13} + {257, 1009, 4099}."""
    import numpy as np
    rng = np.random.default_rng(seed)
    u = rng.uniform(math.log(16), math.log(40960), count)
    logs = np.log(np.array(RAGGED_SIZES, dtype=np.float64))
    return [RAGGED_SIZES[int(np.argmin(np.abs(logs - v)))] for v in u]
METRIC = "FFT GFLOP/s (5N log2N) + % HBM roofline; fwd+adjoint time vs plain, same GPU"


def uniform_c(shape, dtype, seed, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    real = torch.float32 if dtype == torch.complex64 else torch.float64
    re = torch.rand(shape, generator=g, device=device, dtype=real) * 2 - 1
    im = torch.rand(shape, generator=g, device=device, dtype=real) * 2 - 1
    return torch.complex(re, im)


def normal_c(shape, dtype, seed, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    real = torch.float32 if dtype == torch.complex64 else torch.float64
    return torch.complex(torch.randn(shape, generator=g, device=device, dtype=real),
                         torch.randn(shape, generator=g, device=device, dtype=real))


def time_calls(fn, steps, warmup):
    """Average seconds per call, HIP events on the current stream (the one the library
    launches on) bracketing `steps` calls after `warmup` untimed ones."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(steps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / 1e3 / steps


def read_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 counter summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(rows, n, min_seconds=12.0, max_reps=400):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    x = O.uniform(2, (rows, n), np.complex64)
    O.fft_rows_f32(x[:64], -1, nthreads=cores)  # warm-up, twiddles
    reps, t0 = 0, time.perf_counter()
    while True:
        O.fft_rows_f32(x, -1, nthreads=cores)
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or reps >= max_reps:
            break
    flops = 5.0 * n * math.log2(n) * rows * reps
    return {"value": round(flops / el / 1e9, 2), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": f"oracle recursive radix-4/2 FFT in float, fft Forward over the full {rows}x{n} "
                      f"Complex Float array, {reps} reps in {el:.1f} s on {cores} threads"}


def ragged_extras(afft, Mode, torch, dev, args):
    """10,000 Complex Float transforms of mixed lengths through one ragged plan (the lengths that fit
    one workgroup in one persistent launch, the longer ones as batched plans over their gathered
    rows; eager and as a replayed hipGraph) against one plain call per transform and one batched
    call per length, the latter both on pre-gathered batches and on the packed buffer (gather,
    call, scatter)."""
    import ctypes
    lengths = ragged_lengths()
    total = sum(lengths)
    x = uniform_c((total,), torch.complex64, 4, dev)
    out = torch.empty_like(x)
    plan = afft.RaggedPlan(lengths, torch.complex64)
    flops = sum(5.0 * n * math.log2(n) for n in lengths)
    res = {"ragged_total_elements": total}
    t = time_calls(lambda: plan.execute(Mode.Forward, x, out=out), max(5, args.steps // 2), 3)
    tg = time_calls(lambda: plan.execute(Mode.Forward, x, out=out, graph=True), max(5, args.steps // 2), 3)
    # plain: one fft1D call per transform, straight through the C ABI (no tensor allocation)
    lib = afft._lib.lib()
    offs = [0]
    for n in lengths[:-1]:
        offs.append(offs[-1] + n)
    xp, op = x.data_ptr(), out.data_ptr()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args_list = [(n, ctypes.c_void_p(xp + 8 * o), ctypes.c_void_p(op + 8 * o)) for n, o in zip(lengths, offs)]

    def per_call():
        for n, a, b in args_list:
            lib.afft_fft1D(0, 0, n, a, b, s)

    tp = time_calls(per_call, 3, 1)
    # per length: the rows of each length gathered contiguously (gather not timed), one call each
    by = {}
    for n, o in zip(lengths, offs):
        by.setdefault(n, []).append(o)
    groups = [(n, torch.stack([x[o:o + n] for o in os_])) for n, os_ in by.items()]

    def per_length():
        for n, g in groups:
            afft.fft(Mode.Forward, g)

    tl = time_calls(per_length, max(3, args.steps // 5), 1)
    del groups
    # the same per-length calls on the packed buffer itself: each length's elements gathered into a
    # contiguous batch (one index_select), transformed, and scattered back (one index_copy_)
    idx = []
    for n, os_ in by.items():
        start = torch.tensor(os_, dtype=torch.int64, device=dev)
        idx.append((n, len(os_), (start[:, None] + torch.arange(n, device=dev)[None, :]).reshape(-1)))

    def per_length_packed():
        for n, cnt, ix in idx:
            y = afft.fft(Mode.Forward, x.index_select(0, ix).view(cnt, n))
            out.index_copy_(0, ix, y.view(-1))

    tlp = time_calls(per_length_packed, max(3, args.steps // 5), 1)
    del idx
    res["ragged_persistent"] = {"us": round(t * 1e6, 1), "gflops": round(flops / t / 1e9, 1),
                                "hbm_frac": round(16.0 * total / t / 1e9 / HBM_PEAK_GBS, 4)}
    res["ragged_persistent_graph"] = {"us": round(tg * 1e6, 1), "gflops": round(flops / tg / 1e9, 1)}
    res["ragged_plain_per_call"] = {"us": round(tp * 1e6, 1), "calls": len(lengths)}
    res["ragged_per_length_batched"] = {"us": round(tl * 1e6, 1), "calls": len(by)}
    res["ragged_per_length_packed"] = {"us": round(tlp * 1e6, 1), "calls": len(by),
                                       "note": "per-length calls on the packed buffer: gather, fft, scatter"}
    res["ragged_speedup_vs_per_call"] = round(tp / min(t, tg), 2)
    res["ragged_speedup_vs_per_length"] = round(tl / min(t, tg), 2)
    res["ragged_speedup_vs_per_length_packed"] = round(tlp / min(t, tg), 2)
    plan.close()
    return res


def spectral_extras(afft, Mode, torch, dev, args):
    """y = Inverse(fft2D(x) * H) over 64 x 512 x 512 Complex Float, forward + VJP: the fused
    kernels (the VJP recomputing F x, and the forward storing it for the VJP) against the plain
    per-call composition (one fft2D per image, separate product, conjugate-product, batch-sum and
    scale kernels)."""
    B, h, w = 64, 512, 512
    x = normal_c((B, h, w), torch.complex64, 5, dev)
    H = normal_c((h, w), torch.complex64, 6, dev)
    t = normal_c((B, h, w), torch.complex64, 7, dev)
    M = h * w

    def fused():
        y = afft.spectral_conv(x, H)
        gy = 2 * (y - t)
        return afft.spectral_conv_vjp(x, H, gy)

    def fused_saved():  # the forward stores F x, the VJP reads it back
        y, spec = afft.spectral_conv(x, H, save_spectrum=True)
        gy = 2 * (y - t)
        return afft.spectral_conv_vjp(None, H, gy, spectrum=spec)

    def plain():
        ys = []
        for b in range(B):
            ys.append(afft.fft2D(Mode.Reverse, afft.fft2D(Mode.Forward, x[b]) * H) * (1.0 / M))
        y = torch.stack(ys)
        gy = 2 * (y - t)
        gxs, gH = [], torch.zeros_like(H)
        for b in range(B):
            Fx = afft.fft2D(Mode.Forward, x[b])
            Fg = afft.fft2D(Mode.Forward, gy[b])
            gxs.append(afft.fft2D(Mode.Reverse, torch.conj(H) * Fg) * (1.0 / M))
            gH += torch.conj(Fx) * Fg
        return torch.stack(gxs), gH * (1.0 / M)

    tf = time_calls(fused, max(3, args.steps // 10), 2)
    tfs = time_calls(fused_saved, max(3, args.steps // 10), 2)
    tp = time_calls(plain, 3, 1)
    tfwd = time_calls(lambda: afft.spectral_conv(x, H), max(3, args.steps // 5), 2)
    return {"spectral_conv_fused_fwd_vjp": {"us": round(tf * 1e6, 1)},
            "spectral_conv_fused_fwd_only": {"us": round(tfwd * 1e6, 1),
                                             "hbm_frac_per_pass_3": round(3 * 2 * 8.0 * B * M / tfwd / 1e9 / HBM_PEAK_GBS, 4)},
            "spectral_conv_fused_fwd_vjp_saved": {"us": round(tfs * 1e6, 1),
                                                  "note": "forward stores F x, VJP reads it"},
            "spectral_conv_plain_fwd_vjp": {"us": round(tp * 1e6, 1)},
            "spectral_conv_speedup_fwd_vjp": round(tp / tf, 2),
            "spectral_conv_speedup_fwd_vjp_saved": round(tp / tfs, 2)}


def _prof_fft2d(afft, Mode, torch, dev):
    x = normal_c((4096, 4096), torch.complex64, 3, dev)
    return lambda: afft.fft2D(Mode.Forward, x)


def _prof_fft3d(afft, Mode, torch, dev):
    x = normal_c((256, 256, 256), torch.complex128, 3, dev)
    return lambda: afft.fft3D(Mode.Forward, x)


def _prof_four_step(afft, Mode, torch, dev):
    x = uniform_c((1, 1 << 22), torch.complex64, 2, dev)
    return lambda: afft.fft(Mode.Forward, x)


def _prof_ragged(afft, Mode, torch, dev):
    lengths = ragged_lengths()
    x = uniform_c((sum(lengths),), torch.complex64, 4, dev)
    out = torch.empty_like(x)
    plan = afft.RaggedPlan(lengths, torch.complex64)
    return lambda: plan.execute(Mode.Forward, x, out=out, graph=True)


def _prof_spectral(afft, Mode, torch, dev):
    x = normal_c((64, 512, 512), torch.complex64, 5, dev)
    H = normal_c((512, 512), torch.complex64, 6, dev)
    t = normal_c((64, 512, 512), torch.complex64, 7, dev)

    def call():
        y = afft.spectral_conv(x, H)
        afft.spectral_conv_vjp(x, H, 2 * (y - t))
    return call


PROFILE_WORKLOADS = {"fft2d": _prof_fft2d, "fft3d": _prof_fft3d, "four_step": _prof_four_step,
                     "ragged": _prof_ragged, "spectral": _prof_spectral}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--profile-headline", action="store_true")
    ap.add_argument("--profile-workload", choices=sorted(PROFILE_WORKLOADS))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (each GPU would run its own "
              "shard of rows or images); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import afft
    from afft import Mode

    dev = torch.device("cuda", 0)
    if args.profile_workload:
        call = PROFILE_WORKLOADS[args.profile_workload](afft, Mode, torch, dev)
        for _ in range(args.warmup + args.steps):
            call()
        torch.cuda.synchronize()
        print(json.dumps({"profile_workload": args.profile_workload, "steps": args.steps}), flush=True)
        return

    rows, n = 16384, 4096
    x = uniform_c((rows, n), torch.complex64, 2, dev)
    out = torch.empty_like(x)
    plan = afft.Plan((rows, n), torch.complex64, "innermost")
    nl = plan.info()["launches"]

    def step():
        plan.execute(x, -1, 1.0, out=out)

    if args.profile_headline:
        for _ in range(args.warmup + args.steps):
            step()
        for _ in range(args.steps):
            out.copy_(x)  # plain copy of the same 512 MiB: counter calibration
        torch.cuda.synchronize()
        print(json.dumps({"profile_headline": True, "steps": args.steps}), flush=True)
        return

    # settle: untimed steps until 50 ms have passed, so that the clocks of a freshly started
    # process have ramped up before the W warmup steps (the first ~30 calls otherwise run 5-8%
    # slower); then K timed steps between barriers (synchronize) on both sides
    settle_end = time.perf_counter() + 0.05
    while time.perf_counter() < settle_end:
        step()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    kern = time_calls(step, args.steps, 2)  # event-timed on the launch stream
    flops = 5.0 * n * math.log2(n) * rows
    alg_bytes = 2.0 * rows * n * 8
    achieved = alg_bytes / kern / 1e9
    line = {
        "metric": METRIC,
        "value": round(flops / wall / 1e9, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: re/im uniform[-1,1) from torch.rand seed 2, resident in HBM",
        "config": {"workload": "fft Forward along the innermost axis of Array DIM2 (Complex Float) 16384x4096",
                   "rows": rows, "n": n, "mode": "Forward", "launches_per_step": nl},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic("afft_rows_f32_4096_fft_16384x4096.txt"),
                     "kernel_avg_us": round(kern * 1e6, 2)},
    }
    del plan

    extras = {}
    if not args.no_extras:
        def add(name, fn, n_total, bytes_alg, flops_call):
            try:
                t = time_calls(fn, max(5, args.steps // 2), 3)
                extras[name] = {"us": round(t * 1e6, 2), "gflops": round(flops_call / t / 1e9, 1),
                                "hbm_frac": round(bytes_alg / t / 1e9 / HBM_PEAK_GBS, 4)}
            except afft.AfftError as e:
                extras[name] = {"unavailable": str(e)}

        for mode in (Mode.Reverse, Mode.Inverse):
            add(f"fft_16384x4096_c64_{mode.name}", lambda m=mode: afft.fft(m, x), rows * n, alg_bytes, flops)
        del x, out
        torch.cuda.empty_cache()
        # the reference's CPU-runnable case: fft1D Complex Double, N = 1024, Forward then Inverse
        # (latency-bound: two launches of one workgroup each)
        v = uniform_c((1024,), torch.complex128, 0, dev)
        add("fft1D_1024_c128_forward_then_inverse",
            lambda: afft.fft1D(Mode.Inverse, afft.fft1D(Mode.Forward, v)), 1024, 2 * 32.0 * 1024,
            2 * 5.0 * 1024 * 10)
        del v
        big = uniform_c((1, 1 << 22), torch.complex64, 2, dev)
        N = 1 << 22
        add("fft_1x4194304_c64_four_step", lambda: afft.fft(Mode.Forward, big), N, 16.0 * N, 5.0 * N * 22)
        del big
        m2 = normal_c((4096, 4096), torch.complex64, 3, dev)
        N = 4096 * 4096
        add("fft2D_4096x4096_c64_Forward", lambda: afft.fft2D(Mode.Forward, m2), N, 16.0 * N, 5.0 * N * 24)
        add("fft2D_4096x4096_c64_Inverse", lambda: afft.fft2D(Mode.Inverse, m2), N, 16.0 * N, 5.0 * N * 24)
        del m2
        # batched 2D (the spectral-convolution batch): 64 images of 512 x 512, two passes
        mb = normal_c((64, 512, 512), torch.complex64, 5, dev)
        N = 64 * 512 * 512
        add("fft2D_many_64x512x512_c64_Forward", lambda: afft.fft2D_many(Mode.Forward, mb), N, 16.0 * N,
            5.0 * N * 18)
        del mb
        c3 = normal_c((256, 256, 256), torch.complex128, 3, dev)
        N = 256 ** 3
        add("fft3D_256x256x256_c128_Forward", lambda: afft.fft3D(Mode.Forward, c3), N, 32.0 * N, 5.0 * N * 24)
        del c3
        torch.cuda.empty_cache()
        extras.update(ragged_extras(afft, Mode, torch, dev, args))
        extras.update(spectral_extras(afft, Mode, torch, dev, args))
        line["extras"] = extras

    if not args.no_cpu:
        cb = cpu_baseline(rows, n)
        line["cpu_baseline"] = cb
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
