"""Benchmark of the DAS/SIR hot paths on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]

Headline step: one TFM beamforming of the C2 workload -- 64 x 64 A-scans of
2048 samples (FP32 data, FP64 delays and sample indices), 256 x 256 pixels --
with the data already resident in HBM.  value = pixel-channel sums per second
(N_pix * N_ch / step time).  The JSON line carries:
  roofline      the DAS forward kernel's algorithmic bytes (in-bounds pairs x 4
                B + pixels x 4 B) per launch over its HIP-event-timed duration,
                against 8 TB/s HBM;
  cpu_baseline  the reference's own DAS<float> (das.cc compiled from source,
                oracle/_ref) or, where that is absent, the oracle's
                restatement, on the same image on the host cores;
  extra         the other measured paths (FP64 / reference-float TFM, C1 SIR,
                C3 batched SIR, C4 SAFT forward/backward, C5 RCA frames).
Only one GPU is used; --gpus N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
METRIC = "DAS pixel·channel sums/s + SIR obs·elem/s vs FP64 CPU; %HBM roofline; bwd/fwd"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_launches(fn, steps, warmup):
    """Average duration (ms) of fn() over `steps` launches on the current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def cpu_threads():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def bench_c2(steps, warmup, want_cpu, want_extra=True):
    import numpy as np
    import torch

    import workloads
    from dream_amd.torch_ops import DasPlan

    w = workloads.c2(device="cuda", dtype=torch.float32)
    Y = w["Y"].contiguous()
    plan = DasPlan("tfm", w["K"], w["Ro"], w["Gt"], w["Gr"], precision="f32")
    n_pix, n_ch = plan.No, plan.n_ascans
    out = torch.empty(n_pix, dtype=torch.float32, device="cuda")
    # Exact in-bounds pair count: beamform all-ones data (integer sums, exact in FP32).
    ones = torch.ones_like(Y)
    n_inb = int(plan.forward(ones, w["dt"], 0.0, w["cp"]).double().sum().item())
    del ones
    step = lambda: plan.forward(Y, w["dt"], 0.0, w["cp"], out=out)  # noqa: E731
    log(f"C2 TFM f32: warmup {warmup}, timing {steps} steps")
    ms = time_launches(step, steps, warmup)
    sums_per_s = n_pix * n_ch / (ms * 1e-3)
    alg_bytes = n_inb * 4 + n_pix * 4
    achieved = alg_bytes / (ms * 1e-3)
    res = dict(ms=ms, value=sums_per_s, n_inb=n_inb, alg_bytes=alg_bytes, achieved=achieved)

    extra = {}
    # FP64 and the reference-float path on the same image.
    for prec, dtype in ((("f64", torch.float64), ("f32_ref", torch.float32)) if want_extra else ()):
        p = DasPlan("tfm", w["K"], w["Ro"], w["Gt"], w["Gr"], precision=prec)
        Yp = Y.to(dtype)
        o = torch.empty(n_pix, dtype=dtype, device="cuda")
        m = time_launches(lambda: p.forward(Yp, w["dt"], 0.0, w["cp"], out=o), steps, warmup)
        extra[f"c2_tfm_{prec}"] = dict(ms=round(m, 4), sums_per_s=n_pix * n_ch / (m * 1e-3))
        del Yp
    cpu = None
    if want_cpu:
        cpu = cpu_baseline_c2(w, n_pix, n_ch)
    return res, extra, cpu


def cpu_baseline_c2(w, n_pix, n_ch):
    import numpy as np

    from oracle import oracle as O
    threads = cpu_threads()
    os.environ["DREAM_NUM_THREADS"] = str(threads)
    use_ref = O.have_ref()
    Y = w["Y"].cpu().numpy().T.copy()
    times = []
    for _ in range(3):
        t = time.perf_counter()
        O.das(Y, w["Gt"], w["Gr"], w["Ro"], w["dt"], 0.0, w["cp"], "tfm", "ignore", "f32",
              nthreads=threads, use_ref=use_ref)
        times.append(time.perf_counter() - t)
    t = sorted(times)[1]
    return {"value": n_pix * n_ch / t, "unit": "pixel·channel sums/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": f"full C2 TFM image ({n_pix} pixels x {n_ch} A-scans), DAS<float> on the "
                      f"host, {threads} threads, median of 3"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: headline kernel only, plus a 256 MiB "
                         "device-to-device copy used to calibrate the HBM byte counters")
    args = ap.parse_args()
    if args.profile:
        args.no_cpu = True
        args.no_extra = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    from dream_amd import _lib
    _lib.require_device()

    res, extra, cpu = bench_c2(args.steps, args.warmup, not args.no_cpu, not args.no_extra)
    if args.profile:
        # Calibration: a plain copy of 256 MiB (256 MiB read + 256 MiB written).
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 268435456 bytes")
    line = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "pixel·channel sums/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "C2 TFM das: 64 x 64 A-scans x 2048 samples at 50 MHz, 256 x 256 "
                               "pixels at 0.125 mm, FP32 data with FP64 delays, nearest-lower "
                               "sample (reference rule)",
                   "n_pixels": 65536, "n_channels": 4096, "in_bounds_pairs": res["n_inb"]},
        "roofline": {"bound": "hbm", "achieved": res["achieved"] / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": res["achieved"] / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
    }
    if not args.no_extra:
        line["extra"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
