"""GPU parity of the SIR kernels (dreamrect / dream_arr_rect) against the oracle.

Bars: per-contribution sample indices bit for bit (direct dump, including
contributions placed on rounding boundaries), FP64 output within 1e-12 of
the oracle (an index moved by one sample would change a bin by ~1e-3), FP32
output within 2e-5 relative L2, the reference's error codes.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _gpu_indices(xo, yo, zo, a, b, dx, dy, dt, delay, cp):
    from dream_amd import _lib
    L = _lib.lib()
    cnt = ctypes.c_int64(0)
    _lib.check(L.dream_hip_rect_indices(xo, yo, zo, a, b, dx, dy, dt, delay, cp, None, 0,
                                        ctypes.byref(cnt)))
    out = np.empty(cnt.value, dtype=np.int64)
    _lib.check(L.dream_hip_rect_indices(xo, yo, zo, a, b, dx, dy, dt, delay, cp,
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        cnt.value, ctypes.byref(cnt)))
    return out


def test_contribution_indices_bit_exact(gpu, oracle):
    rng = np.random.default_rng(10)
    n_boundary = 0
    for trial in range(60):
        a, b = rng.uniform(0.2, 1.5), rng.uniform(2.0, 15.0)
        dx = dy = rng.choice([0.05, 0.03, 0.1])
        dt = rng.choice([0.02, 0.025, 0.01])
        cp = rng.uniform(1400, 1600)
        xo, yo, zo = rng.uniform(-8, 8), rng.uniform(-8, 8), rng.uniform(1, 80)
        delay = rng.uniform(-5, 20)
        if trial % 2:
            # Put one contribution on the rounding boundary: (t - delay)/dt = k + 1/2.
            it, _ = oracle.rect_contributions(xo, yo, zo, a, b, dx, dy, dt, 0.0, 1.0, cp)
            t0 = float(it[len(it) // 2]) * dt
            xs, ys = oracle.subgrid(a, dx), oracle.subgrid(b, dy)
            j = len(it) // 2
            rx, ry = xo - xs[j % len(xs)], yo - ys[j // len(xs)]
            t = np.sqrt(rx * rx + ry * ry + zo * zo) * 1000.0 / cp
            delay = t - (np.floor(t / dt) + 0.5) * dt
            n_boundary += 1
        ref, _ = oracle.rect_contributions(xo, yo, zo, a, b, dx, dy, dt, delay, 1.0, cp)
        got = _gpu_indices(xo, yo, zo, a, b, dx, dy, dt, delay, cp)
        assert np.array_equal(got, ref), (trial, np.nonzero(got != ref)[0][:5])
    assert n_boundary >= 20


def _gpu_arr_rect(w, **kw):
    from dream_amd import dream_arr_rect
    args = dict(foc_met="off", focal=0.0, steer_met="off", steer_par=[0.0, 0.0], apod_met="off",
                apod=None, win_par=0.0)
    args.update(kw)
    return dream_arr_rect.dream_arr_rect(
        w["Ro"], [w["a"], w["b"]], w["G"], [w["dx"], w["dy"], w["dt"], w["nt"]], w["delay"],
        [w["v"], w["cp"], 0.0], args["foc_met"], args["focal"], args["steer_met"],
        args["steer_par"], args["apod_met"], args["apod"], args["win_par"],
        w.get("err_level", "stop"))


def test_c1_fp64(gpu, oracle):
    import workloads
    w = workloads.c1()
    ref, err, oob = oracle.dream_arr_rect(w["Ro"], w["a"], w["b"], w["dx"], w["dy"], w["dt"],
                                          w["nt"], w["delay"], w["v"], w["cp"], w["G"],
                                          err_level="stop")
    assert err == 0 and oob == 0
    H = _gpu_arr_rect(w)
    assert H.shape == ref.shape and H.flags.f_contiguous
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(H != 0, ref != 0)


@pytest.mark.parametrize("opt", [
    dict(foc_met="x", focal=12.0), dict(foc_met="y", focal=9.0), dict(foc_met="xy", focal=15.0),
    dict(foc_met="x+y", focal=15.0), dict(foc_met="ud", focal="ud"),
    dict(steer_met="x", steer_par=[8.0, 0.0]), dict(steer_met="y", steer_par=[0.0, -5.0]),
    dict(steer_met="xy", steer_par=[4.0, 3.0]),
    dict(apod_met="ud"), dict(apod_met="triangle"), dict(apod_met="gauss"),
    dict(apod_met="raised"), dict(apod_met="hann"), dict(apod_met="hamming"),
    dict(apod_met="simply"), dict(apod_met="clamped"),
])
def test_options_fp64(gpu, oracle, opt):
    rng = np.random.default_rng(len(str(opt)))
    L = 9
    G = np.c_[np.linspace(-4, 4, L), rng.uniform(-0.5, 0.5, L), np.zeros(L)]
    No = 40
    Ro = np.c_[rng.uniform(-5, 5, No), rng.uniform(-3, 3, No), rng.uniform(8, 30, No)]
    w = dict(Ro=Ro, a=0.6, b=5.0, dx=0.05, dy=0.05, dt=0.02, nt=1600, delay=2.0, v=1.0, cp=1500.0,
             G=G, err_level="ignore")
    kw = dict(opt)
    if kw.get("focal") == "ud":
        kw["focal"] = rng.uniform(0, 1, L)
    if kw.get("apod_met") == "ud":
        kw["apod"] = rng.uniform(0.2, 1.0, L)
    o = dict(foc_met=kw.get("foc_met", "off"), focal=kw.get("focal", 0.0),
             steer_met=kw.get("steer_met", "off"), theta=kw.get("steer_par", [0, 0])[0],
             phi=kw.get("steer_par", [0, 0])[1], apod_met=kw.get("apod_met", "off"),
             apod=kw.get("apod"), apod_par=0.0)
    ref, err, _ = oracle.dream_arr_rect(Ro, w["a"], w["b"], w["dx"], w["dy"], w["dt"], w["nt"],
                                        w["delay"], 1.0, w["cp"], G, err_level="ignore", **o)
    H = _gpu_arr_rect(w, **kw)
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()


def test_fp32_output_and_store_mode(gpu, oracle):
    """Many points (one work item per point, column store) and FP32 sums."""
    from dream_amd.torch_ops import arr_rect
    rng = np.random.default_rng(12)
    L, No = 24, 3000
    G = np.c_[np.linspace(-6, 6, L), np.zeros(L), np.zeros(L)]
    Ro = np.c_[rng.uniform(-8, 8, No), np.zeros(No), rng.uniform(5, 40, No)]
    delay = np.linalg.norm(Ro, axis=1) * 1000 / 1500 - 8.0
    args = (0.45, 8.0, 0.05, 0.05, 0.02, 1800)
    ref, err, oob = oracle.dream_arr_rect(Ro, *args, delay, 1.0, 1500.0, G, apod_met="hann",
                                          err_level="ignore")
    Rt = torch.tensor(Ro.T.copy(), device="cuda")
    dt_ = torch.tensor(delay, device="cuda")
    H64, code, oob64 = arr_rect(Rt, *args, dt_, 1.0, 1500.0, G, apod_met="hann",
                                err_level="ignore", out_dtype=torch.float64)
    assert np.max(np.abs(H64.cpu().numpy().T - ref)) <= 1e-12 * np.abs(ref).max()
    assert oob64 == oob
    H32, _, _ = arr_rect(Rt, *args, dt_, 1.0, 1500.0, G, apod_met="hann", err_level="ignore",
                         out_dtype=torch.float32)
    assert rel_l2(H32.double().cpu().numpy().T, ref) <= 2e-5


def test_long_sir_global_mode(gpu, oracle):
    """nt beyond the LDS histogram: accumulation straight into global memory."""
    from dream_amd import dreamrect
    Ro = np.array([[0.3, 0.2, 10.0], [1.0, -1.0, 60.0]])
    args = (1.0, 6.0, 0.05, 0.05, 0.005)
    nt = 12000
    ref, _, _ = oracle.dreamrect(Ro, *args, nt, 0.0, 1.0, 1500.0, "stop")
    H = dreamrect.dreamrect(Ro, [1.0, 6.0], [0.05, 0.05, 0.005, nt], 0.0, [1.0, 1500.0, 0.0],
                            "stop")
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()


def test_batch_mixing_global_and_lds_problems(gpu, oracle):
    """One plan whose problems use both accumulation modes (nt beyond and
    within the LDS histogram), FP64 and FP32, against the oracle."""
    from dream_amd.torch_ops import SirBatch
    rng = np.random.default_rng(51)
    specs = [(12000, 0.005), (400, 0.025), (9000, 0.005), (700, 0.02)]
    probs, refs = [], []
    for nt, dt in specs:
        L, No = 3, 40
        G = np.c_[np.linspace(-1.5, 1.5, L), np.zeros(L), np.zeros(L)]
        Ro = np.c_[rng.uniform(-2, 2, No), rng.uniform(-1, 1, No), rng.uniform(8, 12, No)]
        d = Ro[:, 2] * 1000.0 / 1500.0 - 1.0
        probs.append(dict(Ro=torch.tensor(Ro.T.copy(), device="cuda"), a=0.6, b=2.0, dx=0.05,
                          dy=0.05, dt=dt, nt=nt, delay=torch.tensor(d, device="cuda"), v=1.0,
                          cp=1500.0, G=G))
        ref, _, _ = oracle.dream_arr_rect(Ro, 0.6, 2.0, 0.05, 0.05, dt, nt, d, 1.0, 1500.0, G,
                                          err_level="ignore", nthreads=4)
        refs.append(ref)
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        for H, ref in zip(SirBatch(probs, out_dtype=dtype).run(), refs):
            torch.cuda.synchronize()
            Hn = H.double().cpu().numpy().T
            if dtype == torch.float64:
                assert np.max(np.abs(Hn - ref)) <= tol * np.abs(ref).max()
            else:
                assert rel_l2(Hn, ref) <= tol


@pytest.mark.parametrize("walk", ["rows", "columns", "auto"])
def test_row_and_column_walks(gpu, oracle, monkeypatch, walk):
    """Both sub-element walks of rectangles (rows: x inner; columns: y inner,
    in row segments) and the per-point choice, forced by DREAM_HIP_SIR_WALK,
    against the oracle: many points (column store), few points (items split
    over work-groups, atomic column adds), a long SIR (global accumulation),
    out-of-range counts and the last-element status rule, FP32 output."""
    from dream_amd.torch_ops import arr_rect
    if walk != "auto":
        monkeypatch.setenv("DREAM_HIP_SIR_WALK", walk)
    rng = np.random.default_rng(77)
    L = 7
    G = np.c_[np.linspace(-3, 3, L), rng.uniform(-1, 1, L), rng.uniform(-0.2, 0.2, L)]
    cases = [(300, 0.4, 7.0, 0.02, 1500, "ignore"), (3, 1.1, 4.0, 0.025, 900, "ignore"),
             (4, 0.5, 3.0, 0.005, 11000, "ignore"), (50, 0.7, 5.0, 0.02, 700, "stop")]
    total_oob = 0
    for No, a, b, dt, nt, lvl in cases:
        Ro = np.c_[rng.uniform(-6, 6, No), rng.uniform(-4, 4, No), rng.uniform(6, 30, No)]
        # Delays that put part of the arrivals before sample 0 and past nt.
        delay = np.linalg.norm(Ro, axis=1) * 1000 / 1500 - rng.uniform(-1.0, 0.5 * nt * dt, No)
        focal = rng.uniform(0, 0.5, L)
        # The oracle under `stop` ends a thread at its first out-of-range
        # contribution, so its sum and count come from `ignore`; the device
        # computes everything at every level.
        ref, _, oob = oracle.dream_arr_rect(Ro, a, b, 0.05, 0.05, dt, nt, delay, 1.0, 1500.0, G,
                                            foc_met="ud", focal=focal, apod_met="hann",
                                            err_level="ignore", nthreads=4)
        _, err, _ = oracle.dream_arr_rect(Ro, a, b, 0.05, 0.05, dt, nt, delay, 1.0, 1500.0, G,
                                          foc_met="ud", focal=focal, apod_met="hann",
                                          err_level=lvl, nthreads=4)
        Rt = torch.tensor(Ro.T.copy(), device="cuda")
        dt_ = torch.tensor(delay, device="cuda")
        H, code, n = arr_rect(Rt, a, b, 0.05, 0.05, dt, nt, dt_, 1.0, 1500.0, G, foc_met="ud",
                              focal=focal, apod_met="hann", err_level=lvl)
        assert n == oob and code == err, (No, n, oob, code, err)
        total_oob += oob
        assert np.max(np.abs(H.cpu().numpy().T - ref)) <= 1e-12 * np.abs(ref).max(), No
        H32, _, n32 = arr_rect(Rt, a, b, 0.05, 0.05, dt, nt, dt_, 1.0, 1500.0, G, foc_met="ud",
                               focal=focal, apod_met="hann", err_level="ignore",
                               out_dtype=torch.float32)
        assert n32 == oob and rel_l2(H32.double().cpu().numpy().T, ref) <= 2e-5
    assert total_oob > 0


def test_error_codes_gpu(gpu):
    from dream_amd import dream_arr_rect, dreamrect
    Ro = np.array([[0.0, 0.0, 20.0], [0.0, 0.0, 21.0], [0.0, 0.0, 22.0]])
    with pytest.raises(RuntimeError):
        dreamrect.dreamrect(Ro, [0.9, 10.0], [0.05, 0.05, 0.025, 10], 0.0, [1.0, 1500.0, 0.0], "stop")
    H = dreamrect.dreamrect(Ro, [0.9, 10.0], [0.05, 0.05, 0.025, 10], 0.0, [1.0, 1500.0, 0.0],
                            "ignore")
    assert np.all(H == 0)
    # The last-element rule: OOB only in the first element -> status none.
    from dream_amd.torch_ops import arr_rect
    Rt = torch.tensor(Ro.T.copy(), device="cuda")
    delay = 20.0 / 1.5 - 1.0
    G_first = np.array([[40.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    _, c1, n1 = arr_rect(Rt, 0.9, 10.0, 0.05, 0.05, 0.025, 200, delay, 1.0, 1500.0, G_first,
                         err_level="ignore")
    _, c2, n2 = arr_rect(Rt, 0.9, 10.0, 0.05, 0.05, 0.025, 200, delay, 1.0, 1500.0,
                         G_first[::-1].copy(), err_level="ignore")
    assert n1 > 0 and n1 == n2 and c1 == 0 and c2 == 1
    _, c3, _ = arr_rect(Rt, 0.9, 10.0, 0.05, 0.05, 0.025, 200, delay, 1.0, 1500.0, G_first,
                        err_level="stop")
    assert c3 == 3


def test_batch_plan_uneven_with_graph_replays(gpu, oracle):
    import workloads
    from dream_amd.torch_ops import SirBatch
    probs = workloads.c3(seed=33, n_problems=5)
    # Shrink for the oracle: at most 300 points each.
    for p in probs:
        keep = min(300, p["Ro"].shape[0])
        p["Ro"] = p["Ro"][:keep]
        p["delay"] = p["delay"][:keep]
    dev = []
    for p in probs:
        q = dict(p)
        q["Ro"] = torch.tensor(p["Ro"].T.copy(), device="cuda")
        q["delay"] = torch.tensor(p["delay"], device="cuda")
        dev.append(q)
    batch = SirBatch(dev, out_dtype=torch.float64)
    refs = []
    for p in probs:
        ref, err, oob = oracle.dream_arr_rect(p["Ro"], p["a"], p["b"], p["dx"], p["dy"], p["dt"],
                                              p["nt"], p["delay"], 1.0, p["cp"], p["G"],
                                              foc_met="ud", focal=p["focal"], apod_met="hann",
                                              err_level="ignore", nthreads=4)
        refs.append((ref, oob))
    for rep in range(3):
        Hs = batch.run(use_graph=True)
        torch.cuda.synchronize()
        codes, oobs = batch.status()
        for H, (ref, oob), o in zip(Hs, refs, oobs):
            assert np.max(np.abs(H.cpu().numpy().T - ref)) <= 1e-12 * np.abs(ref).max(), rep
            assert o == oob
        torch.cuda.synchronize()


def test_autotuned_batch_equals_untuned(gpu):
    """Every SIR batch launch shape (one work-group per item, persistent
    grids) computes each item the same way: tuned and untuned runs agree."""
    import torch

    import workloads
    from dream_amd import _lib
    from dream_amd.torch_ops import SirBatch
    probs = workloads.sir_small_uneven_batch(n_problems=12)
    dev = []
    for p in probs:
        q = dict(p)
        q["Ro"] = torch.tensor(p["Ro"].T.copy(), device="cuda")
        q["delay"] = torch.tensor(p["delay"], device="cuda")
        dev.append(q)
    try:
        _lib.lib().dream_hip_set_autotune(0)
        ref = [h.clone() for h in SirBatch(dev, out_dtype=torch.float64).run(use_graph=False)]
        _lib.lib().dream_hip_set_autotune(1)
        tuned = SirBatch(dev, out_dtype=torch.float64)
        for _ in range(2):
            got = tuned.run(use_graph=True)
            torch.cuda.synchronize()
            for a, b in zip(got, ref):
                assert torch.allclose(a, b, rtol=1e-12, atol=1e-12 * float(b.abs().max()))
    finally:
        _lib.lib().dream_hip_set_autotune(1)


def test_rsqrt_seed_accuracy(gpu):
    """The hardware FP64 rsqrt seed that the SIR and SAFT fast index paths
    start from: its relative error must stay within 2^-23 (the CPU emulation
    in test_certification.py shows both certifications hold up to there; the
    margins were sized on the 2^-24.2 measured on MI355X).  Mantissas dense
    over [1, 4) (one binade of s per binade of the result), exponents spread."""
    from dream_amd import _lib
    m = np.linspace(1.0, 4.0, 1 << 20, endpoint=False)
    s = np.concatenate([m * 2.0 ** e for e in (-20, -4, 0, 6, 12, 30)])
    y = np.empty_like(s)
    _lib.check(_lib.lib().dream_hip_rsq_seed(s.ctypes.data, s.size, y.ctypes.data))
    rel = np.abs(y * np.sqrt(s) - 1.0)
    assert rel.max() <= 2.0 ** -23, np.log2(rel.max())

