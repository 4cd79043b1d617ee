"""Benchmark of the DAS/SIR hot paths on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]

Headline step: one TFM beamforming of the C2 workload -- 64 x 64 A-scans of
2048 samples (FP32 data, FP64 delays and sample indices), 256 x 256 pixels --
with the data already resident in HBM.  value = pixel-channel sums per second
(N_pix * N_ch / step time).  The JSON line carries:
  roofline      the beamforming kernel's algorithmic bytes per launch (the
                in-bounds pairs it gathers x 4 B + pixels x 4 B; with one
                array on both sides it gathers each pair (i, j), (j, i) once
                from A-scans a preparation kernel symmetrizes) over its
                duration from HIP events recorded around each launch on its
                stream, against 8 TB/s HBM; the preparation kernel's time and
                bytes beside it;
  cpu_baseline  the reference's own DAS<float> (das.cc compiled from source,
                oracle/_ref) or, where that is absent, the oracle's
                restatement, on the same image on the host cores (all
                threads, and one thread on an eighth of the image), with the
                CPU model;
  extra         the other measured paths (FP64 / reference-float TFM, the C2
                image through the host-buffer entry point with its PCIe
                copies, a 256 MiB device copy, C1 SIR, C3 batched SIR, C4
                SAFT forward/backward, C5 RCA frames, the C2 image summing
                every pair on its own, and with the out-of-bounds status read
                back each call) and, beside the headline's cpu_baseline, the
                reference's DAS<float> on C4 and a C5 sub-volume and the
                oracle's SIR on the host cores (all of C1, a bounded sample of
                C3).
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
# HBM bytes per launch of the headline kernel, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE and scaled by a device copy of known size.
COUNTERS = os.path.join(ROOT, "profiles", "das_tfm_c2_counters.md")


def measured_traffic():
    """traffic_bytes_per_launch from the committed counter summary, or None."""
    try:
        with open(COUNTERS) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None
METRIC = "DAS pixel·channel sums/s + SIR obs·elem/s vs FP64 CPU; %HBM roofline; bwd/fwd"
# MI355X: 256 CUs at 2.4 GHz; a CU's 4 SIMDs retire 64 VALU lane-operations
# per cycle (a wave instruction takes 4 cycles); one L1 tag lookup per CU per
# cycle.
N_CU, CLOCK_HZ = 256, 2.4e9
C3_COUNTERS = os.path.join(ROOT, "profiles", "sir_kernel_c3_counters.md")
C5_COUNTERS = os.path.join(ROOT, "profiles", "das_rca_c5_counters.md")


def _counter_value(path, key):
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def valu_issue_frac(contributions, ms):
    """C3: VALU busy lane-cycles per second (the committed counters' busy
    lane-cycles per contribution -- an FP64 rsqrt holds its SIMD for four
    passes -- x contributions / time) over the device's rate (256 CUs x 64
    lanes x 2.4 GHz), or None."""
    per = _counter_value(C3_COUNTERS, "valu_busy_lane_cycles_per_contribution")
    if per is None:
        return None
    return per * contributions / (ms * 1e-3) / (N_CU * 64 * CLOCK_HZ)


def c5_bound():
    """C5: the address/tag path of the gathers, as the committed counters
    measured it (TA busy fraction and L1 tag lookups per gather, under the
    profiler), or Nones."""
    return dict(ta_busy_frac_profiled=_counter_value(C5_COUNTERS, "ta_busy_frac_profiled"),
                tag_lookups_per_gather=_counter_value(C5_COUNTERS, "tag_lookups_per_gather"))


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


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def _median_time(fn, runs=3):
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return sorted(ts)[len(ts) // 2]


def bench_c2(steps, warmup, want_cpu, want_extra=True):
    import numpy as np
    import torch

    import workloads
    from dream_amd.torch_ops import DasPlan

    w = workloads.c2(device="cuda", dtype=torch.float32)
    Y = w["Y"].contiguous()
    plan = DasPlan("tfm", w["K"], w["Ro"], w["Gt"], w["Gr"], precision="f32")
    n_pix, n_ch = plan.No, plan.n_ascans
    L = plan.Lt
    out = torch.empty(n_pix, dtype=torch.float32, device="cuda")
    # Exact in-bounds pair counts: beamform all-ones data (integer sums, exact
    # in FP32), and data that is one on the A-scans with transmit == receive.
    ones = torch.ones_like(Y)
    n_inb = int(plan.forward(ones, w["dt"], 0.0, w["cp"]).double().sum().item())
    ones.zero_()
    ones[torch.arange(L, device="cuda") * (L + 1)] = 1.0
    n_diag = int(plan.forward(ones, w["dt"], 0.0, w["cp"]).double().sum().item())
    del ones
    # One array on both sides and a zero delay: the forward gathers each
    # pair (i, j), (j, i) once from the symmetrized A-scans.
    reciprocal = plan.Gr is plan.Gt
    gathered = (n_inb - n_diag) // 2 + n_diag if reciprocal else n_inb
    step = lambda: plan.forward(Y, w["dt"], 0.0, w["cp"], out=out)  # noqa: E731
    log(f"C2 TFM f32: warmup {warmup}, timing {steps} steps")
    ms = time_launches(step, steps, warmup)
    sums_per_s = n_pix * n_ch / (ms * 1e-3)
    # The beamforming kernel's own duration: HIP events recorded on its
    # stream around each launch, over a second pass of the same steps.
    plan.time_kernels(steps)
    for _ in range(steps):
        step()
    n_t, prep_total, fwd_total = plan.kernel_times()
    kernel_ms = fwd_total / n_t
    prep_ms = prep_total / n_t
    alg_bytes = gathered * 4 + n_pix * 4
    prep_bytes = (n_ch * w["K"] + (L * (L + 1) // 2) * w["K"]) * 4 if reciprocal else 0
    achieved = alg_bytes / (kernel_ms * 1e-3)
    res = dict(ms=ms, value=sums_per_s, n_inb=n_inb, n_diag=n_diag, gathered=gathered,
               alg_bytes=alg_bytes, achieved=achieved, kernel_ms=kernel_ms, prep_ms=prep_ms,
               prep_bytes=prep_bytes, reciprocal=reciprocal)

    extra = {}
    if want_extra:
        # The same image summing every (transmit, receive) pair on its own:
        # the receive array passed as a separate copy, so the reciprocal
        # pairing (one gather per pair (i, j), (j, i)) is not used.
        full = DasPlan("tfm", w["K"], w["Ro"], w["Gt"], w["Gr"], precision="f32")
        full.Gr = full.Gt.clone()
        m = time_launches(lambda: full.forward(Y, w["dt"], 0.0, w["cp"], out=out), steps, warmup)
        extra["c2_tfm_f32_full_pair_sum"] = dict(ms=round(m, 4),
                                                 sums_per_s=n_pix * n_ch / (m * 1e-3))
        del full
        # The reference's das() returns the out-of-bounds status
        # (extra_src/das.cc:338-351): the kernel collects it and each call
        # reads it back (one synchronisation per call).
        m = time_launches(lambda: plan.forward(Y, w["dt"], 0.0, w["cp"], out=out,
                                               err_level="ignore", want_status=True),
                          steps, warmup)
        extra["c2_tfm_f32_with_status"] = dict(ms=round(m, 4),
                                               sums_per_s=n_pix * n_ch / (m * 1e-3))
    # FP64 and the reference-float path on the same image.
    for prec, dtype in ((("f64", torch.float64), ("f32_ref", torch.float32)) if want_extra else ()):
        p = DasPlan("tfm", w["K"], w["Ro"], w["Gt"], w["Gr"], precision=prec)
        Yp = Y.to(dtype)
        o = torch.empty(n_pix, dtype=dtype, device="cuda")
        m = time_launches(lambda: p.forward(Yp, w["dt"], 0.0, w["cp"], out=o), steps, warmup)
        extra[f"c2_tfm_{prec}"] = dict(ms=round(m, 4), sums_per_s=n_pix * n_ch / (m * 1e-3))
        del Yp
    if want_extra:
        extra["c2_tfm_f32_host_call"] = c2_host_call(w, Y, n_pix, n_ch)
        extra["device_copy_256MiB"] = device_copy_rate(steps, warmup)
    cpu = None
    if want_cpu:
        cpu = cpu_baseline_c2(w, n_pix, n_ch)
    return res, extra, cpu


def c2_host_call(w, Y, n_pix, n_ch, calls=10):
    """The C2 image through the host-buffer entry point (the reference's
    das() contract: Y, Ro, G in host memory, Im returned to the host), so the
    time includes the copy of the 33.5 MB of A-scans over PCIe and the image
    back.  Median of `calls` calls after one warm-up."""
    import numpy as np
    from dream_amd import das_f

    Yh = Y.cpu().numpy().T                 # K x n_ascans, Fortran order
    args = (Yh, w["Gt"], w["Gr"], w["Ro"], w["dt"], 0.0, w["cp"], "tfm", "ignore")
    das_f.das(*args)
    ts = []
    for _ in range(calls):
        t0 = time.perf_counter()
        das_f.das(*args)
        ts.append(time.perf_counter() - t0)
    ms = float(np.median(ts)) * 1e3
    return dict(ms=round(ms, 3), sums_per_s=n_pix * n_ch / (ms * 1e-3),
                host_bytes_in=int(Yh.nbytes), calls=calls)


def device_copy_rate(steps, warmup):
    """Measured device-to-device copy rate (256 MiB read + 256 MiB written per
    copy), beside the 8 TB/s HBM peak the roofline is priced against."""
    import torch
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    ms = time_launches(lambda: dst.copy_(src), steps, warmup)
    return dict(ms=round(ms, 4), gb_per_s=2 * src.numel() * 4 / (ms * 1e-3) / 1e9)


def _sir_device_problem(p, torch):
    q = dict(p)
    q["Ro"] = torch.tensor(p["Ro"].T.copy(), device="cuda")
    d = p["delay"]
    q["delay"] = torch.tensor(d, device="cuda") if hasattr(d, "__len__") else float(d)
    return q


def _sir_contributions(p, oracle_free=True):
    """Sub-element contributions of one SIR problem (points x elements x subs)."""
    import numpy as np
    nx = len(_subgrid(p["a"], p["dx"]))
    ny = len(_subgrid(p["b"], p["dy"]))
    return p["Ro"].shape[0] * np.asarray(p["G"]).reshape(-1, 3).shape[0] * nx * ny


def _subgrid(w, d):
    out, s = [], -w / 2.0 + d / 2.0
    while s <= w / 2.0:
        out.append(s)
        s += d
    return out


def bench_c1(steps, warmup):
    """C1 alone (FP64 SIR, one problem): the plan replayed as a captured graph
    and launched eagerly, and through dream_arr_rect_dev (plan set-up
    included), per call."""
    import torch

    import workloads
    from dream_amd.torch_ops import SirBatch, arr_rect
    w = workloads.c1()
    q = _sir_device_problem(w, torch)
    batch = SirBatch([q], out_dtype=torch.float64)
    out = {}
    for name, g in (("graph", True), ("eager", False)):
        ms = time_launches(lambda: batch.run(use_graph=g), steps, warmup)
        out[f"c1_plan_{name}"] = dict(ms=round(ms, 4))
    H = torch.empty((w["Ro"].shape[0], w["nt"]), dtype=torch.float64, device="cuda")
    ms = time_launches(lambda: arr_rect(q["Ro"], w["a"], w["b"], w["dx"], w["dy"], w["dt"],
                                        w["nt"], q["delay"], w["v"], w["cp"], w["G"],
                                        err_level="stop", out=H), max(2, steps // 4), 2)
    out["c1_arr_rect_dev_call"] = dict(ms=round(ms, 4))
    return out


def bench_sir(steps, warmup):
    """C1 (FP64, one problem) and C3 (FP32, 32 uneven problems, one persistent launch
    under a captured graph, against one dream_arr_rect_dev call per problem)."""
    import numpy as np
    import torch

    import workloads
    from dream_amd.torch_ops import SirBatch, arr_rect
    out = {}
    # C1
    w = workloads.c1()
    batch = SirBatch([_sir_device_problem(w, torch)], out_dtype=torch.float64)
    # A launch of two small kernels (status and output zeroing, the SIR sum):
    # eager launches beat a graph replay here.
    ms = time_launches(lambda: batch.run(use_graph=False), steps, warmup)
    ms_g = time_launches(lambda: batch.run(use_graph=True), steps, warmup)
    n_oe = w["Ro"].shape[0] * w["G"].shape[0]
    contrib = _sir_contributions(w)
    out["c1_sir_f64"] = dict(ms=round(ms, 4), graph_ms=round(ms_g, 4),
                             obs_elem_per_s=n_oe / (ms * 1e-3),
                             contributions_per_s=contrib / (ms * 1e-3),
                             roofline_frac=(contrib * 16 + w["Ro"].shape[0] * w["nt"] * 8)
                             / (ms * 1e-3) / HBM_PEAK)
    # C3
    probs = workloads.c3()
    dev = [_sir_device_problem(p, torch) for p in probs]
    n_oe = sum(p["Ro"].shape[0] * p["G"].shape[0] for p in probs)
    contrib = sum(_sir_contributions(p) for p in probs)
    out_bytes = sum(p["Ro"].shape[0] * p["nt"] * 4 for p in probs)
    log(f"C3: {len(probs)} problems, {n_oe} obs-elem, {contrib:.3e} contributions")
    batch = SirBatch(dev, out_dtype=torch.float32)
    ms_b = time_launches(lambda: batch.run(use_graph=True), max(2, steps // 10), 1)
    codes, oobs = batch.status()
    # In-bounds contributions exactly: total minus the kernel's out-of-bounds count.
    inb = contrib - sum(oobs)
    Hs = [torch.empty_like(H) for H in batch.H]

    def per_call():
        for p, q, H in zip(probs, dev, Hs):
            arr_rect(q["Ro"], p["a"], p["b"], p["dx"], p["dy"], p["dt"], p["nt"], q["delay"],
                     p["v"], p["cp"], p["G"], foc_met="ud", focal=p["focal"], apod_met="hann",
                     err_level="ignore", out_dtype=torch.float32, out=H)
    ms_p = time_launches(per_call, max(1, steps // 25), 1)
    # The SIR kernel is bound by FP64 VALU issue, not HBM (its histograms
    # live in LDS): the rate is reported against that bound from the
    # committed counter summary, not as an HBM fraction.
    out["c3_sir_batched_f32"] = dict(
        ms=round(ms_b, 3), obs_elem_per_s=n_oe / (ms_b * 1e-3),
        contributions_per_s=contrib / (ms_b * 1e-3), in_bounds_contributions=int(inb),
        bound="VALU issue", valu_issue_frac=valu_issue_frac(contrib, ms_b),
        per_call_ms=round(ms_p, 3), speedup_vs_per_call=ms_p / ms_b)
    out.update(bench_annu(steps, warmup))
    out.update(bench_att(steps, warmup))
    return out


def bench_att(steps, warmup):
    """C1 with attenuation (alpha = 5 dB/(cm MHz)): 64 points x 16 elements x
    3600 sub-elements, nt = 400, FP64, through the host entry point (copies
    included).  Each in-range contribution costs nt/2 complex exponentials."""
    import time

    import numpy as np

    import workloads
    from dream_amd.dream_arr_rect import dream_arr_rect
    w = workloads.c1()

    def run():
        return dream_arr_rect(w["Ro"], [w["a"], w["b"]], w["G"], [w["dx"], w["dy"], w["dt"], w["nt"]],
                              w["delay"], [w["v"], w["cp"], 5.0], "off", 0.0, "off", [0.0, 0.0],
                              "off", None, 0.0, "ignore")
    for _ in range(max(warmup, 1)):
        run()
    n = max(2, steps // 10)
    t = time.perf_counter()
    for _ in range(n):
        run()
    ms = (time.perf_counter() - t) * 1e3 / n
    contrib = _sir_contributions(w)
    return {"c1_sir_attenuated_f64": dict(ms=round(ms, 3), contributions=int(contrib),
                                          exponentials_per_s=contrib * (w["nt"] // 2) / (ms * 1e-3))}


def _disc_subelements(R, d):
    """Sub-elements of the disc row walk (rows ys = -R + d/2 while ys <= R)."""
    import math
    n, ys = 0, -R + d / 2.0
    while ys <= R:
        rx = math.sqrt(max(R * R - ys * ys, 0.0))
        xs = -rx + d / 2.0
        while xs <= rx:
            n += 1
            xs += d
        ys += d
    return n


def bench_annu(steps, warmup):
    """Annular array (dream_arr_annu): the reference's test radii over 4096
    points, FP64 output.  Contributions = points x sum over discs of their
    sub-elements."""
    import torch

    import workloads
    from dream_amd.torch_ops import arr_annu
    w = workloads.annu()
    Ro = torch.tensor(w["Ro"].T.copy(), dtype=torch.float64, device="cuda")
    No = Ro.shape[1]
    H = torch.empty((No, w["nt"]), dtype=torch.float64, device="cuda")

    def run():
        arr_annu(Ro, w["Gr"], w["dx"], w["dy"], w["dt"], w["nt"], w["delay"], w["v"], w["cp"],
                 foc_met=w["foc_met"], focal=w["focal"], err_level=w["err_level"], out=H)
    ms = time_launches(run, max(2, steps // 10), 1)
    contrib = No * sum(_disc_subelements(r, w["dx"]) for r in w["Gr"])
    return {"annu_4096_points_f64": dict(ms=round(ms, 3), contributions=contrib,
                                         contributions_per_s=contrib / (ms * 1e-3))}


def bench_batches(steps, warmup):
    """Uneven batches of small problems, where one persistent launch matters:
    96 TFM images (8-48 elements, A-scans of 512-2048 samples, 16^2-80^2 pixels)
    and 128 SIR problems (16-256 points, 4-16 elements, nt 200-800), each in one
    launch (captured graph) against one plain call per problem."""
    import numpy as np
    import torch

    import workloads
    from dream_amd.torch_ops import DasBatch, DasPlan, SirBatch, arr_rect
    out = {}
    probs = workloads.das_uneven_batch(device="cuda")
    batch = DasBatch("tfm", probs, precision="f32")
    plans = [DasPlan("tfm", p["K"], p["Ro"], p["Gt"], p["Gr"], precision="f32") for p in probs]
    outs = [torch.empty(pl.No, dtype=torch.float32, device="cuda") for pl in plans]

    def per_call():
        for p, pl, o in zip(probs, plans, outs):
            pl.forward(p["Y"], p["dt"], 0.0, p["cp"], out=o)
    ms_b = time_launches(lambda: batch.run(use_graph=True), steps, warmup)
    ms_p = time_launches(per_call, max(5, steps // 5), 2)
    pairs = sum(pl.No * pl.n_ascans for pl in plans)
    out["das_uneven_batch_f32"] = dict(problems=len(probs), ms=round(ms_b, 4),
                                       per_call_ms=round(ms_p, 4),
                                       speedup_vs_per_call=ms_p / ms_b,
                                       sums_per_s=pairs / (ms_b * 1e-3))
    sprobs = workloads.sir_small_uneven_batch()
    dev = [_sir_device_problem(p, torch) for p in sprobs]
    sb = SirBatch(dev, out_dtype=torch.float32)
    Hs = [torch.empty_like(H) for H in sb.H]

    def sir_per_call():
        for p, q, H in zip(sprobs, dev, Hs):
            arr_rect(q["Ro"], p["a"], p["b"], p["dx"], p["dy"], p["dt"], p["nt"], q["delay"],
                     p["v"], p["cp"], p["G"], foc_met="ud", focal=p["focal"], apod_met="hann",
                     err_level="ignore", out_dtype=torch.float32, out=H)
    ms_b = time_launches(lambda: sb.run(use_graph=True), steps, warmup)
    ms_p = time_launches(sir_per_call, max(2, steps // 25), 1)
    n_oe = sum(p["Ro"].shape[0] * p["G"].shape[0] for p in sprobs)
    out["sir_small_uneven_batch_f32"] = dict(problems=len(sprobs), ms=round(ms_b, 4),
                                             per_call_ms=round(ms_p, 4),
                                             speedup_vs_per_call=ms_p / ms_b,
                                             obs_elem_per_s=n_oe / (ms_b * 1e-3))
    return out


def bench_conv(steps, warmup):
    """Data synthesis by convolution (fftconv_p / sum_fftconv path): 4096
    columns (64 x 64 A-scans) of 2000 samples convolved with one 300-sample
    pulse, FP64 and FP32, as multiply-adds (of the direct sum) per second;
    and a long-filter case through both paths."""
    import torch

    from dream_amd.torch_ops import conv_sum
    out = {}
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        g = torch.Generator(device="cuda").manual_seed(7)
        A = torch.randn((4096, 2000), generator=g, device="cuda", dtype=dt)
        b = torch.randn(300, generator=g, device="cuda", dtype=dt)
        y = torch.empty((4096, 2299), device="cuda", dtype=dt)
        ms = time_launches(lambda: conv_sum(A, b, out=y), steps, warmup)
        out[f"conv_4096x2000_by_300_{name}"] = dict(ms=round(ms, 4),
                                                   macs_per_s=4096 * 2000 * 300 / (ms * 1e-3))
    # Long filters (impulse responses convolved with each other, the
    # reference's double-path synthesis): 1024 columns of 2000 samples by a
    # 2000-sample filter, FFT path against the direct sums.
    g = torch.Generator(device="cuda").manual_seed(8)
    A = torch.randn((1024, 2000), generator=g, device="cuda", dtype=torch.float64)
    b = torch.randn(2000, generator=g, device="cuda", dtype=torch.float64)
    y = torch.empty((1024, 3999), device="cuda", dtype=torch.float64)
    entry = {}
    for path in ("fft", "direct"):
        os.environ["DREAM_HIP_CONV"] = path
        try:
            ms = time_launches(lambda: conv_sum(A, b, out=y), max(2, steps // 10), 2)
        finally:
            os.environ.pop("DREAM_HIP_CONV", None)
        entry[f"{path}_ms"] = round(ms, 4)
    entry["macs_per_s_fft"] = 1024 * 2000 * 2000 / (entry["fft_ms"] * 1e-3)
    out["conv_1024x2000_by_2000_f64"] = entry
    return out


def bench_das_diff(steps, warmup):
    """C4 SAFT (linear, apodized) forward and adjoint; C5 RCA 8 frames forward and
    apodization adjoint."""
    import torch
    out = bench_das_diff_c4(steps, warmup)
    torch.cuda.empty_cache()
    out.update(bench_das_diff_c5(steps, warmup))
    return out


def bench_das_diff_c4(steps, warmup):
    """C4: SAFT, 128 A-scans x 4096 samples, 512 x 512 pixels, linear interpolation with
    apodization; forward and the adjoint (gradients for Y, apodization and cp)."""
    import torch

    import workloads
    from dream_amd.torch_ops import DasPlan
    out = {}
    w = workloads.c4(device="cuda", dtype=torch.float32)
    plan = DasPlan("saft", w["K"], w["Ro"], w["Gt"], None, precision="f32")
    Y = w["Y"].contiguous()
    apod = torch.ones(plan.Lt, dtype=torch.float32, device="cuda")
    g = torch.randn(plan.No, dtype=torch.float32, device="cuda")
    fwd = time_launches(lambda: plan.forward(Y, w["dt"], 0.0, w["cp"], apod, None, "linear"),
                        steps, warmup)
    bwd = time_launches(lambda: plan.adjoint(Y, g, w["dt"], 0.0, w["cp"], apod, None, "linear"),
                        steps, warmup)
    pairs = plan.No * plan.n_ascans
    # In-bounds pairs (nearest-lower index in [0, K)): all-ones data, exact sums.
    n_inb = int(plan.forward(torch.ones_like(Y), w["dt"], 0.0, w["cp"]).double().sum().item())
    alg_bytes = n_inb * 2 * 4 + plan.No * 4              # two taps per pair, FP32
    out["c4_saft_linear_f32"] = dict(fwd_ms=round(fwd, 4), bwd_ms=round(bwd, 4),
                                     bwd_over_fwd=bwd / fwd, sums_per_s=pairs / (fwd * 1e-3),
                                     in_bounds_pairs=n_inb,
                                     roofline_frac=alg_bytes / (fwd * 1e-3) / HBM_PEAK)
    return out


def bench_das_diff_c5(steps, warmup):
    """C5: RCA row transmit, 128 x 128 A-scans x 4096 samples, 128 x 128 x 256 voxels,
    8 frames in one launch; forward and the apodization adjoint."""
    import torch

    import workloads
    from dream_amd.torch_ops import DasPlan
    out = {}
    w = workloads.c5(device="cuda", dtype=torch.float32)
    plan = DasPlan("rca_rowtx", w["K"], w["Ro"], w["Gt"], w["Gr"], precision="f32")
    Y = w["Y"].contiguous()
    F = Y.shape[0]
    ones = torch.ones(plan.Lt, dtype=torch.float32, device="cuda")
    g = torch.randn((F, plan.No), dtype=torch.float32, device="cuda")
    fwd = time_launches(lambda: plan.forward(Y, w["dt"], 0.0, w["cp"]), 2, 1)
    bwd = time_launches(lambda: plan.adjoint(Y, g, w["dt"], 0.0, w["cp"], ones, ones, "nearest",
                                             want_gy=False, want_cp=False), 2, 1)
    pairs = plan.No * plan.n_ascans * F
    n_inb = int(plan.forward(torch.ones_like(Y), w["dt"], 0.0, w["cp"]).double().sum().item())
    alg_bytes = n_inb * 4 + F * plan.No * 4
    # Y (268 MB per frame) is re-read from L2/MALL by every voxel tile, so the
    # no-reuse byte count exceeds HBM bandwidth: it is not a bound for C5.
    # The forward is bound by the gathers' address and tag path (TA busy).
    out["c5_rca_8frames_f32"] = dict(fwd_ms=round(fwd, 3), bwd_apod_ms=round(bwd, 3),
                                     bwd_over_fwd=bwd / fwd, sums_per_s=pairs / (fwd * 1e-3),
                                     in_bounds_pairs=n_inb, bound="TA / L1 tag lookups",
                                     no_reuse_bytes_per_s=alg_bytes / (fwd * 1e-3), **c5_bound())
    return out


def _cpu_das(Y, Gt, Gr, Ro, dt, cp, method, threads, use_ref):
    """One DAS<float> call on the host: the reference's das.cc compiled from
    source (use_ref) or the oracle's restatement, `threads` threads."""
    from oracle import oracle as O
    os.environ["DREAM_NUM_THREADS"] = str(threads)
    O.das(Y, Gt, Gr, Ro, dt, 0.0, cp, method, "ignore", "f32", nthreads=threads,
          use_ref=use_ref)


def cpu_baseline_c2(w, n_pix, n_ch):
    """The reference's DAS<float> (das.cc compiled from source; the oracle's
    restatement where that checker is absent) on the whole C2 image with all
    host threads, median of 3; and on one thread over the first 8192 pixels
    (1/8 of the image), median of 3, in the same unit."""
    from oracle import oracle as O
    threads = cpu_threads()
    use_ref = O.have_ref()
    Y = w["Y"].cpu().numpy().T.copy()
    t = _median_time(lambda: _cpu_das(Y, w["Gt"], w["Gr"], w["Ro"], w["dt"], w["cp"], "tfm",
                                      threads, use_ref))
    n1 = 8192
    t1 = _median_time(lambda: _cpu_das(Y, w["Gt"], w["Gr"], w["Ro"][:n1], w["dt"], w["cp"],
                                       "tfm", 1, use_ref))
    return {"value": n_pix * n_ch / t, "unit": "pixel·channel sums/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": f"full C2 TFM image ({n_pix} pixels x {n_ch} A-scans), DAS<float> on the "
                      f"host, {threads} threads, median of 3",
            "cpu_model": cpu_model(),
            "value_1_thread": n1 * n_ch / t1,
            "sample_1_thread": f"first {n1} pixels of the C2 image x {n_ch} A-scans, 1 thread, "
                               "median of 3"}


def cpu_baselines_c4_c5():
    """C4 and C5 on the host cores, the reference's DAS<float> in its own
    (reference) mode: nearest-lower sample, no apodization or interpolation,
    forward only.  C4: the whole 512 x 512 SAFT image.  C5: one frame over a
    16 x 16 x 256-voxel sub-volume of the 128 x 128 x 256 grid (x, y indices
    56..71), extrapolated linearly in voxels and frames to the 8-frame volume
    (x 512), labelled as such."""
    import numpy as np
    import torch

    import workloads
    from oracle import oracle as O
    threads = cpu_threads()
    use_ref = O.have_ref()
    kind = "reference" if use_ref else "port"
    out = {}
    w = workloads.c4(device="cpu", dtype=torch.float32)
    Y = w["Y"].numpy().T.copy()
    n_pc = w["Ro"].shape[0] * w["Gt"].shape[0]
    t = _median_time(lambda: _cpu_das(Y, w["Gt"], None, w["Ro"], w["dt"], w["cp"], "saft",
                                      threads, use_ref))
    out["c4_saft"] = dict(ms=round(t * 1e3, 3), sums_per_s=n_pc / t, cores=threads, kind=kind,
                          sample=f"full C4 image ({w['Ro'].shape[0]} pixels x 128 A-scans), "
                                 f"nearest-lower, {threads} threads, median of 3")
    w = workloads.c5(device="cpu", dtype=torch.float32, n_frames=1)
    nx, ny, nz = 128, 128, 256
    ix, iy, iz = np.meshgrid(np.arange(56, 72), np.arange(56, 72), np.arange(nz), indexing="ij")
    sub = w["Ro"][(ix + nx * (iy + ny * iz)).ravel("F")]
    Y = w["Y"][0].numpy().T.copy()
    n_ch = w["Gt"].shape[0] * w["Gr"].shape[0]
    t = _median_time(lambda: _cpu_das(Y, w["Gt"], w["Gr"], sub, w["dt"], w["cp"], "rca_rowtx",
                                      threads, use_ref), runs=1)
    scale = (nx * ny * nz * 8) / sub.shape[0]
    out["c5_rca_rowtx"] = dict(ms=round(t * 1e3, 3), sums_per_s=sub.shape[0] * n_ch / t,
                               cores=threads, kind=kind,
                               extrapolated_8_frames_s=round(t * scale, 2),
                               sample=f"one frame, {sub.shape[0]} voxels (16 x 16 x 256 of the "
                                      f"grid) x {n_ch} A-scans, {threads} threads, one run; "
                                      f"8 frames x full grid extrapolated x{scale:.0f}")
    return out


def cpu_baseline_c1():
    """C1 SIR on the host cores: the oracle's FP64 restatement of
    dream_arr_rect (the reference's own SIR source needs FFTW, which this
    image lacks), threaded like the reference, median of 3."""
    import numpy as np

    import workloads
    from oracle import oracle as O
    threads = cpu_threads()
    w = workloads.c1()
    times = []
    for _ in range(3):
        t = time.perf_counter()
        O.dream_arr_rect(w["Ro"], w["a"], w["b"], w["dx"], w["dy"], w["dt"], w["nt"],
                         w["delay"], w["v"], w["cp"], w["G"], err_level="stop", nthreads=threads)
        times.append(time.perf_counter() - t)
    t = sorted(times)[1]
    n_oe = w["Ro"].shape[0] * w["G"].shape[0]
    return dict(ms=round(t * 1e3, 3), obs_elem_per_s=n_oe / t, cores=threads, kind="port",
                sample="full C1 problem (64 points x 16 elements x 3600 sub-elements), FP64, "
                       f"{threads} threads, median of 3")


def cpu_baseline_c3_sample(n_points=512):
    """A bounded sample of C3 on the host cores: the first problem's first
    n_points observation points with all its elements, the oracle's FP64
    restatement, threaded; reported as sub-element contributions per second
    (the C3 line's unit), median of 3."""
    import workloads
    from oracle import oracle as O
    threads = cpu_threads()
    p = dict(workloads.c3()[0])
    p["Ro"] = p["Ro"][:n_points]
    p["delay"] = p["delay"][:n_points]
    times = []
    for _ in range(3):
        t = time.perf_counter()
        O.dream_arr_rect(p["Ro"], p["a"], p["b"], p["dx"], p["dy"], p["dt"], p["nt"], p["delay"],
                         p["v"], p["cp"], p["G"], foc_met=p["foc_met"], focal=p["focal"],
                         apod_met=p["apod_met"], err_level="ignore", nthreads=threads)
        times.append(time.perf_counter() - t)
    t = sorted(times)[1]
    contrib = _sir_contributions(p)
    return dict(ms=round(t * 1e3, 3), contributions_per_s=contrib / t, contributions=contrib,
                cores=threads, kind="port",
                sample=f"C3 problem 0, first {p['Ro'].shape[0]} points x {p['G'].shape[0]} "
                       f"elements, FP64, {threads} threads, median of 3")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: headline kernel only, plus a 256 MiB "
                         "device-to-device copy used to calibrate the HBM byte counters")
    ap.add_argument("--part", choices=("c1", "sir", "c4", "c5", "batch", "conv"),
                    help="profile one of the other paths alone (short run, no JSON line): "
                         "c1 = C1 SIR alone, "
                         "sir = C1 + C3 SIR, c4 = SAFT forward/adjoint, c5 = RCA frames, "
                         "batch = uneven batches of small DAS and SIR problems, "
                         "conv = data-synthesis convolution")
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
    if args.part:
        part = {"c1": bench_c1, "sir": bench_sir, "c4": bench_das_diff_c4, "c5": bench_das_diff_c5,
                "batch": bench_batches, "conv": bench_conv}[args.part]
        log(json.dumps(part(args.steps, args.warmup)))
        return

    res, extra, cpu = bench_c2(args.steps, args.warmup, not args.no_cpu, not args.no_extra)
    if args.profile:
        # Calibration: a plain copy of 256 MiB (256 MiB read + 256 MiB written).
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 268435456 bytes")
    if not args.no_extra:
        extra.update(bench_sir(args.steps, args.warmup))
        extra.update(bench_batches(args.steps, args.warmup))
        extra.update(bench_conv(args.steps, args.warmup))
        extra.update(bench_das_diff(args.steps, args.warmup))
        if not args.no_cpu:
            extra["c1_sir_cpu_baseline"] = cpu_baseline_c1()
            extra["c3_sir_cpu_sample"] = cpu_baseline_c3_sample()
            extra["cpu_baselines"] = cpu_baselines_c4_c5()
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
                   "n_pixels": 65536, "n_channels": 4096, "in_bounds_pairs": res["n_inb"],
                   "gathered_in_bounds_pairs": res["gathered"],
                   "reciprocal_pairs": res["reciprocal"]},
        "roofline": {"bound": "hbm", "achieved": res["achieved"] / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": res["achieved"] / HBM_PEAK,
                     "traffic": measured_traffic(),
                     "kernel": "das_fwd_kernel<float, double, TFM, nearest>",
                     "kernel_ms": res["kernel_ms"],
                     "prep_kernel": "das_symmetrize_kernel<float, double>",
                     "prep_kernel_ms": res["prep_ms"], "prep_bytes": res["prep_bytes"]},
        "cpu_baseline": cpu,
    }
    if not args.no_extra:
        line["extra"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
