"""GPU parity of the DAS kernels against the oracle and the reference's vectors.

Bars: FP64 -- the reference's sample indices bit for bit (ramp data) and
<= 1e-12 relative error; FP32 reference arithmetic (das_f) -- the reference's
float indices bit for bit; FP32 data with FP64 delays -- the FP64 indices bit
for bit and <= 2e-5 relative L2 error against the FP64 oracle.
"""
import os

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
METHODS = ("saft", "tfm", "rca_coltx", "rca_rowtx")


def _golden():
    d = np.load(os.path.join(HERE, "golden", "das_ref.npz"))
    names = sorted({k.split("__")[0] for k in d.files if k.endswith("__delay")})
    return d, names


def _setup(method, K=700, L=9, No=500, seed=0):
    rng = np.random.default_rng(seed)
    x = np.linspace(-4.0, 4.0, L)
    Gt = np.c_[x, np.zeros(L), np.zeros(L)]
    Gr = Gt.copy() if method == "tfm" else np.c_[np.zeros(L), x, np.zeros(L)]
    n_ascans = L if method == "saft" else L * L
    Y = rng.standard_normal((K, n_ascans))
    Ro = np.c_[rng.uniform(-6, 6, No), rng.uniform(-5, 5, No), rng.uniform(0.5, 9, No)]
    return Y, Gt, Gr, Ro, rng


def test_golden_reference_vectors(gpu):
    """das / das_f on the GPU against the reference's own outputs."""
    from dream_amd import das, das_f
    d, names = _golden()
    for n in names:
        m = next(x for x in METHODS if n.startswith(x + "_"))
        level = n.rsplit("_", 1)[1]
        dt, cp = d[f"{n}__scalars"]
        delay = d[f"{n}__delay"]
        dl = delay if delay.size > 1 else float(delay[0])
        Y, Gt, Gr, Ro = d[f"{m}__Y"], d[f"{m}__Gt"], d[f"{m}__Gr"], d[f"{m}__Ro"]
        Im = das.das(Y, Gt, Gr, Ro, dt, dl, cp, m, level, "gpu")
        ref = d[f"{n}__f64__Im"]
        scale = np.abs(Y).max() * Y.shape[1]
        assert np.max(np.abs(Im - ref)) <= 1e-13 * scale, n
        Imf = das_f.das(Y.astype(np.float32), Gt.astype(np.float32), Gr.astype(np.float32),
                        Ro.astype(np.float32), np.float32(dt), np.float32(dl) if np.isscalar(dl)
                        else dl.astype(np.float32), np.float32(cp), m, level, "gpu")
        reff = d[f"{n}__f32__Im"]
        assert np.max(np.abs(Imf.astype(np.float64) - reff)) <= 2e-6 * scale, n


@pytest.mark.parametrize("method", METHODS)
def test_ramp_indices_exact_all_precisions(gpu, oracle, method):
    """Y[a, k] = k: sums are exact integers in any order, so equality means
    equal sample-index multisets per pixel."""
    from dream_amd import das, das_f
    from dream_amd.torch_ops import DasPlan
    _, Gt, Gr, Ro, rng = _setup(method, K=900, No=3000, seed=1)
    # Pixels on elements (a zero squared distance) and 1e-120 from one (a
    # squared distance below the range of the forward's fast square root).
    Ro[0], Ro[1], Ro[2] = Gt[0], Gt[4], Gr[2]
    Ro[3] = Gt[1] + np.array([1e-120, 0.0, 0.0])
    n_ascans = Gt.shape[0] if method == "saft" else Gt.shape[0] * Gr.shape[0]
    ramp = np.repeat(np.arange(900, dtype=np.float64)[:, None], n_ascans, axis=1)
    for delay in (0.0, 0.31, rng.uniform(-0.4, 0.9, Ro.shape[0])):
        ref64, _ = oracle.das(ramp, Gt, Gr, Ro, 0.02, delay, 1540.0, method, "ignore", "f64")
        gpu64 = das.das(ramp, Gt, Gr, Ro, 0.02, delay, 1540.0, method, "ignore", "gpu")
        assert np.array_equal(gpu64, ref64)
        ref32, _ = oracle.das(ramp, Gt, Gr, Ro, 0.02, delay, 1540.0, method, "ignore", "f32")
        gpu32 = das_f.das(ramp, Gt, Gr, Ro, 0.02, delay, 1540.0, method, "ignore", "gpu")
        assert np.array_equal(gpu32, ref32)
        # FP32 data, FP64 delays: the FP64 indices.
        plan = DasPlan(method, 900, Ro, Gt, Gr, precision="f32")
        Yt = torch.tensor(ramp.T.copy(), dtype=torch.float32, device="cuda")
        dl = torch.tensor(delay, device="cuda") if np.ndim(delay) else delay
        Im = plan.forward(Yt, 0.02, dl, 1540.0)
        assert np.array_equal(Im.cpu().numpy().astype(np.float64), ref64)


@pytest.mark.parametrize("delay", [0.0, 0.37])
def test_saft_positions_on_sample_boundaries(gpu, oracle, delay):
    """Pixels straight below one element whose sample position 2 (z/cp +
    delay) fs lies within 1e-13 .. 1e-4 of an integer, on both sides: the
    kernel's fast square root is certified there or replaced by the IEEE one,
    so the FP64 indices (ramp data) must come out bit for bit, for FP64 data
    and for FP32 data with FP64 delays."""
    from dream_amd import das
    from dream_amd.torch_ops import DasPlan
    dt, cp, K = 0.02, 1500.0, 4000
    fs = 1000.0 / dt
    Gt = np.zeros((1, 3))
    z = []
    for n in np.arange(200, 3800, 240):
        z0 = (n / (2.0 * fs) - delay / 1000.0) * cp
        z += [z0 * (1.0 + m * 2.0 ** -30) for m in range(-64, 65)]
        z += [z0 * (1.0 + m * 2.0 ** -40) for m in range(-64, 65)]
        z += [z0 * (1.0 + m * 2.0 ** -52) for m in range(-20, 21)]
    z = np.array(z)
    Ro = np.c_[np.zeros_like(z), np.zeros_like(z), z]
    ramp = np.arange(K, dtype=np.float64)[:, None]
    ref, _ = oracle.das(ramp, Gt, Gt, Ro, dt, delay, cp, "saft", "ignore", "f64")
    assert np.unique(np.floor(ref)).size >= 2 * 15        # both sides of every boundary
    assert np.array_equal(das.das(ramp, Gt, Gt, Ro, dt, delay, cp, "saft", "ignore", "gpu"), ref)
    plan = DasPlan("saft", K, Ro, Gt, None, precision="f32")
    Yt = torch.tensor(ramp.T.copy(), dtype=torch.float32, device="cuda")
    Im = plan.forward(Yt, dt, delay, cp).cpu().numpy().astype(np.float64)
    assert np.array_equal(Im, ref)
    # Linear interpolation of a ramp returns the position itself, here within
    # 1e-4 of an integer (FP32 output: within 1e-3).
    lin = plan.forward(Yt, dt, delay, cp, interp="linear").cpu().numpy().astype(np.float64)
    assert np.max(np.abs(lin - np.rint(lin))) <= 1e-3
    assert np.max(np.abs(lin - ref)) <= 1.0 + 1e-3


@pytest.mark.parametrize("method", METHODS)
def test_random_data_f64_and_error_codes(gpu, oracle, method):
    from dream_amd import das
    Y, Gt, Gr, Ro, _ = _setup(method, seed=2)
    for level, code in (("ignore", 1), ("warn", 2)):
        ref, err = oracle.das(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, method, level, "f64", nthreads=4)
        Im = das.das(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, method, level, "gpu")
        assert err == code
        assert np.max(np.abs(Im - ref)) <= 1e-12 * np.abs(ref).max()
    with pytest.raises(RuntimeError):
        das.das(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, method, "stop", "gpu")
    # In-bounds everywhere: no error even with "stop".
    Ro_in = Ro.copy()
    Ro_in[:, 2] = np.clip(Ro_in[:, 2], 0.5, 2.0)
    Ro_in[:, :2] *= 0.1
    Yl = np.vstack([Y] * 3)
    ref, err = oracle.das(Yl, Gt, Gr, Ro_in, 0.02, 0.0, 1500.0, method, "stop", "f64", nthreads=4)
    assert err == 0
    Im = das.das(Yl, Gt, Gr, Ro_in, 0.02, 0.0, 1500.0, method, "stop", "gpu")
    assert np.max(np.abs(Im - ref)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("method", METHODS)
def test_status_on_device_every_call(gpu, oracle, method):
    """Every forward computes the out-of-bounds status on the device (the
    reference's das always does, extra_src/das.cc:338-351); DasPlan.status
    reads the last call's code later.  One plan, calls with and without
    out-of-range pairs in turn: each reports its own status (a call with
    none after one with some reports none), chunked launches report what
    one launch does, and the synchronous code agrees."""
    import ctypes

    from dream_amd import _lib
    from dream_amd.torch_ops import DasPlan
    Y, Gt, Gr, Ro, _ = _setup(method, seed=2)
    Yl = np.vstack([Y] * 3)
    K = Yl.shape[0]
    saft = method == "saft"
    Yd = torch.tensor(Yl.T.copy(), device="cuda")
    plan = DasPlan(method, K, Ro, Gt, None if saft else Gr, precision="f64")
    seen = set()
    for delay in (30.0, 0.0, 30.0, 0.0):
        _, err = oracle.das(Yl, Gt, Gr, Ro, 0.02, delay, 1500.0, method, "warn", "f64",
                            nthreads=4)
        seen.add(err)
        plan.forward(Yd, 0.02, delay, 1500.0)
        assert plan.status("warn") == err, (delay, err)
        assert plan.status("stop") == (3 if err else 0)
        _, code = plan.forward(Yd, 0.02, delay, 1500.0, err_level="warn", want_status=True)
        assert code == err
    assert seen == {0, 2}
    try:
        _lib.lib().dream_hip_test_chunking(ctypes.c_int64(64 * 8), ctypes.c_int64(0))
        for delay, err in ((30.0, 2), (0.0, 0)):
            plan.forward(Yd, 0.02, delay, 1500.0)
            assert plan.status("warn") == err
    finally:
        _lib.lib().dream_hip_test_chunking(ctypes.c_int64(0), ctypes.c_int64(0))


def test_c2_tfm_f32_vs_f64_oracle(gpu, oracle):
    """The C2 TFM configuration: FP32 data / FP64 delays within 2e-5 rel-L2 of
    the FP64 oracle; FP64 within 1e-12; reference-float path against the
    reference-float oracle."""
    import workloads
    from dream_amd.torch_ops import DasPlan
    w = workloads.c2(device="cuda", dtype=torch.float64)
    Y64 = w["Y"]
    Y_np = Y64.cpu().numpy().T.copy()            # K x n_ascans
    ref, _ = oracle.das(Y_np, w["Gt"], w["Gr"], w["Ro"], w["dt"], 0.0, w["cp"], "tfm", "ignore",
                        "f64")
    p32 = DasPlan("tfm", w["K"], w["Ro"], w["Gt"], w["Gr"], precision="f32")
    Im32 = p32.forward(Y64.float().contiguous(), w["dt"], 0.0, w["cp"]).double().cpu().numpy()
    # Y rounded to FP32 is part of the FP32 path; the bar includes it.
    assert rel_l2(Im32, ref) <= 2e-5
    p64 = DasPlan("tfm", w["K"], w["Ro"], w["Gt"], w["Gr"], precision="f64")
    Im64 = p64.forward(Y64, w["dt"], 0.0, w["cp"]).cpu().numpy()
    assert np.max(np.abs(Im64 - ref)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("method", METHODS)
def test_linear_apod_forward_and_adjoint(gpu, oracle, method):
    from dream_amd.torch_ops import DasPlan
    Y, Gt, Gr, Ro, rng = _setup(method, seed=3, No=700)
    Lr = 0 if method == "saft" else Gr.shape[0]
    at = rng.uniform(0.5, 1.5, Gt.shape[0])
    ar = rng.uniform(0.5, 1.5, Lr) if Lr else None
    g = rng.standard_normal(Ro.shape[0])
    ref, _ = oracle.das_ext(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, method, "linear", at, ar)
    gY, gat, gar, gcp = oracle.das_ext_adjoint(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, g, method,
                                               "linear", at, ar)
    for prec, dt_, tol in (("f64", torch.float64, 1e-12), ("f32", torch.float32, 2e-5)):
        plan = DasPlan(method, Y.shape[0], Ro, Gt, Gr, precision=prec)
        Yt = torch.tensor(Y.T.copy(), dtype=dt_, device="cuda")
        att = torch.tensor(at, dtype=dt_, device="cuda")
        art = torch.tensor(ar, dtype=dt_, device="cuda") if Lr else None
        Im = plan.forward(Yt, 0.02, 0.1, 1500.0, att, art, "linear").double().cpu().numpy()
        assert rel_l2(Im, ref) <= tol
        gt = torch.tensor(g, dtype=dt_, device="cuda")
        rY, rat, rar, rcp = plan.adjoint(Yt, gt, 0.02, 0.1, 1500.0, att, art, "linear")
        atol = 1e-10 if prec == "f64" else 5e-5
        assert rel_l2(rY.double().cpu().numpy().T, gY) <= atol
        assert rel_l2(rat.double().cpu().numpy(), gat) <= atol
        if Lr:
            assert rel_l2(rar.double().cpu().numpy(), gar) <= atol
        assert abs(float(rcp) - gcp) <= (1e-9 if prec == "f64" else 1e-3) * abs(gcp)


@pytest.mark.parametrize("method", METHODS)
def test_dot_product_identity_gpu(gpu, method):
    from dream_amd.torch_ops import DasPlan
    Y, Gt, Gr, Ro, rng = _setup(method, seed=4, No=900)
    plan = DasPlan(method, Y.shape[0], Ro, Gt, Gr, precision="f64")
    y = torch.tensor(rng.standard_normal(Y.T.shape), device="cuda")
    g = torch.tensor(rng.standard_normal(Ro.shape[0]), device="cuda")
    for interp in ("linear", "nearest"):
        Ay = plan.forward(y, 0.02, 0.1, 1500.0, interp=interp)
        ATg, _, _, _ = plan.adjoint(y, g, 0.02, 0.1, 1500.0, interp=interp)
        lhs, rhs = float(Ay @ g), float((y * ATg).sum())
        assert abs(lhs - rhs) <= 1e-12 * float(Ay.norm() * g.norm())


def test_frames_equal_single_calls(gpu):
    from dream_amd.torch_ops import DasPlan
    method = "rca_rowtx"
    Y, Gt, Gr, Ro, rng = _setup(method, seed=5, No=1000)
    F = 3
    Yf = torch.tensor(np.stack([Y.T * (f + 1) for f in range(F)]), dtype=torch.float32,
                      device="cuda").contiguous()
    plan = DasPlan(method, Y.shape[0], Ro, Gt, Gr, precision="f32")
    Im = plan.forward(Yf, 0.02, 0.1, 1500.0)
    for f in range(F):
        single = plan.forward(Yf[f].contiguous(), 0.02, 0.1, 1500.0)
        assert torch.equal(Im[f], single)


@pytest.mark.parametrize("method,want_gy", [("saft", True), ("rca_rowtx", True),
                                            ("rca_rowtx", False)])
def test_adjoint_frames_equal_single_calls(gpu, method, want_gy):
    """Frames in one adjoint launch: each frame's gY is its own single-frame
    gY; the apodization and cp gradients are the sums over the frames."""
    from dream_amd.torch_ops import DasPlan
    Y, Gt, Gr, Ro, rng = _setup(method, seed=9, No=1500)
    F = 3
    Yf = torch.tensor(np.stack([Y.T * (f + 1) for f in range(F)]), device="cuda").contiguous()
    g = torch.tensor(rng.standard_normal((F, Ro.shape[0])), device="cuda").contiguous()
    plan = DasPlan(method, Y.shape[0], Ro, Gt, Gr, precision="f64")
    at = torch.tensor(rng.uniform(0.5, 1.5, Gt.shape[0]), device="cuda")
    ar = None if method == "saft" else torch.tensor(rng.uniform(0.5, 1.5, Gr.shape[0]),
                                                     device="cuda")
    interp = "linear" if want_gy else "nearest"
    gY, gat, gar, gcp = plan.adjoint(Yf, g, 0.02, 0.1, 1500.0, at, ar, interp, want_gy=want_gy,
                                     want_cp=want_gy)
    sat, sar, scp = 0.0, 0.0, 0.0
    for f in range(F):
        y1, a1, r1, c1 = plan.adjoint(Yf[f].contiguous(), g[f].contiguous(), 0.02, 0.1, 1500.0,
                                      at, ar, interp, want_gy=want_gy, want_cp=want_gy)
        if want_gy:
            assert rel_l2(gY[f].cpu().numpy(), y1.cpu().numpy()) <= 1e-12
            scp = scp + float(c1)
        sat = sat + a1.cpu().numpy()
        if ar is not None:
            sar = sar + r1.cpu().numpy()
    assert rel_l2(gat.cpu().numpy(), sat) <= 1e-12
    if ar is not None:
        assert rel_l2(gar.cpu().numpy(), sar) <= 1e-12
    if want_gy:
        assert abs(float(gcp) - scp) <= 1e-10 * abs(scp)


@pytest.mark.parametrize("method", ["tfm", "rca_coltx", "rca_rowtx"])
def test_frame_groups_match_single_frames(gpu, method):
    """FP32 calls of several frames read them in groups of four (one
    gather serving four frames; five frames: a full group and a group with
    three padding frames).  Each frame's image is its single-frame image,
    and the nearest-lower apodization gradients are the sums of the single
    frames' (to FP32 rounding: the frames are summed in another order)."""
    from dream_amd.torch_ops import DasPlan
    Y, Gt, Gr, Ro, rng = _setup(method, seed=11, No=1300)
    if method == "tfm":
        Gr = Gr + np.array([0.05, 0.0, 0.0])        # two arrays: not the reciprocal path
    F = 5
    Yf = torch.tensor(np.stack([Y.T * (f + 1) + f for f in range(F)]), dtype=torch.float32,
                      device="cuda").contiguous()
    g = torch.tensor(rng.standard_normal((F, Ro.shape[0])), dtype=torch.float32,
                     device="cuda").contiguous()
    plan = DasPlan(method, Y.shape[0], Ro, Gt, Gr, precision="f32")
    at = torch.tensor(rng.uniform(0.5, 1.5, Gt.shape[0]), dtype=torch.float32, device="cuda")
    ar = torch.tensor(rng.uniform(0.5, 1.5, Gr.shape[0]), dtype=torch.float32, device="cuda")
    Im = plan.forward(Yf, 0.02, 0.1, 1500.0)
    _, gat, gar, _ = plan.adjoint(Yf, g, 0.02, 0.1, 1500.0, at, ar, "nearest", want_gy=False,
                                  want_cp=False)
    sat, sar = 0.0, 0.0
    for f in range(F):
        single = plan.forward(Yf[f].contiguous(), 0.02, 0.1, 1500.0).cpu().numpy()
        assert rel_l2(Im[f].cpu().numpy(), single) <= 1e-6, f
        _, a1, r1, _ = plan.adjoint(Yf[f].contiguous(), g[f].contiguous(), 0.02, 0.1, 1500.0,
                                    at, ar, "nearest", want_gy=False, want_cp=False)
        sat = sat + a1.double().cpu().numpy()
        sar = sar + r1.double().cpu().numpy()
    assert rel_l2(gat.double().cpu().numpy(), sat) <= 1e-5
    assert rel_l2(gar.double().cpu().numpy(), sar) <= 1e-5


def test_autograd_function(gpu, oracle):
    from dream_amd.torch_ops import DasPlan, das_diff
    method = "saft"
    Y, Gt, Gr, Ro, rng = _setup(method, seed=6, No=800)
    plan = DasPlan(method, Y.shape[0], Ro, Gt, Gr, precision="f64")
    Yt = torch.tensor(Y.T.copy(), device="cuda", requires_grad=True)
    at = torch.tensor(rng.uniform(0.5, 1.5, Gt.shape[0]), device="cuda", requires_grad=True)
    cp = torch.tensor(1500.0, dtype=torch.float64, requires_grad=True)
    target = torch.tensor(rng.standard_normal(Ro.shape[0]), device="cuda")
    Im = das_diff(plan, Yt, at, None, cp, 0.02, 0.1, "linear")
    loss = ((Im - target) ** 2).sum()
    loss.backward()
    g = 2 * (Im - target).detach().cpu().numpy()
    gY, gat, _, gcp = oracle.das_ext_adjoint(Y, Gt, None, Ro, 0.02, 0.1, 1500.0, g, method,
                                             "linear", at.detach().cpu().numpy())
    assert rel_l2(Yt.grad.cpu().numpy().T, gY) <= 1e-10
    assert rel_l2(at.grad.cpu().numpy(), gat) <= 1e-10
    assert abs(float(cp.grad) - gcp) <= 1e-9 * abs(gcp)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_split_launch_small_image(gpu, oracle, prec):
    """Few pixels and many channel blocks: the forward launch splits tiles
    over work-groups and adds partial sums in a second kernel."""
    from dream_amd.torch_ops import DasPlan
    rng = np.random.default_rng(8)
    L, K, No = 64, 900, 256
    x = np.linspace(-6.0, 6.0, L)
    G = np.c_[x, np.zeros(L), np.zeros(L)]
    Y = rng.standard_normal((K, L * L))
    Ro = np.c_[rng.uniform(-6, 6, No), np.zeros(No), rng.uniform(1, 8, No)]
    ref, _ = oracle.das(Y, G, G, Ro, 0.02, 0.05, 1500.0, "tfm", "ignore", "f64")
    dt_ = torch.float64 if prec == "f64" else torch.float32
    plan = DasPlan("tfm", K, Ro, G, G, precision=prec)
    Yt = torch.tensor(Y.T.copy(), dtype=dt_, device="cuda")
    for _ in range(2):
        Im = plan.forward(Yt, 0.02, 0.05, 1500.0).double().cpu().numpy()
        assert rel_l2(Im, ref) <= (1e-13 if prec == "f64" else 2e-6)


def _uneven_problems(method, rng, n, dtype):
    """n problems of one method with different arrays, grids and A-scan lengths."""
    probs = []
    for i in range(n):
        L = int(rng.integers(3, 12))
        K = int(rng.integers(300, 900))
        No = int(rng.integers(1, 700))
        x = np.linspace(-3.0 - i % 3, 3.0 + i % 2, L)
        Gt = np.c_[x, np.zeros(L), np.zeros(L)]
        Gr = Gt.copy() if method == "tfm" else np.c_[np.zeros(L), x, np.zeros(L)]
        n_ascans = L if method == "saft" else L * L
        Ro = np.c_[rng.uniform(-5, 5, No), rng.uniform(-4, 4, No), rng.uniform(0.5, 8, No)]
        Yh = rng.standard_normal((n_ascans, K))
        probs.append(dict(K=K, Ro=Ro, Gt=Gt, Gr=None if method == "saft" else Gr,
                          Yh=Yh, Y=torch.tensor(Yh, dtype=dtype, device="cuda"),
                          dt=0.02, delay=0.05 * (i % 3), cp=1500.0))
    return probs


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("prec", ["f64", "f32", "f32_ref"])
def test_uneven_batch_equals_single_calls(gpu, method, prec):
    """One persistent launch over uneven problems gives each problem's single-call
    result: indices bit for bit (ramp data), values to rounding (random data)."""
    from dream_amd.torch_ops import DasBatch, DasPlan
    rng = np.random.default_rng(11)
    dtype = torch.float64 if prec == "f64" else torch.float32
    probs = _uneven_problems(method, rng, 7, dtype)
    batch = DasBatch(method, probs, precision=prec)
    outs = [o.clone() for o in batch.run(use_graph=False)]
    for p, o in zip(probs, outs):
        plan = DasPlan(method, p["K"], p["Ro"], p["Gt"], p["Gr"], precision=prec)
        ref = plan.forward(p["Y"], p["dt"], p["delay"], p["cp"]).double().cpu().numpy()
        tol = 1e-13 if prec == "f64" else 2e-6
        assert rel_l2(o.double().cpu().numpy(), ref) <= tol
    # Ramp data Y[a, k] = k: every sum is an exact integer in any order.
    for p in probs:
        p["Y"].copy_(torch.arange(p["K"], dtype=dtype, device="cuda").expand_as(p["Y"]))
    ramp = DasBatch(method, probs, precision=prec)
    got = [o.clone() for o in ramp.run(use_graph=False)]
    for p, o in zip(probs, got):
        plan = DasPlan(method, p["K"], p["Ro"], p["Gt"], p["Gr"], precision=prec)
        ref = plan.forward(p["Y"], p["dt"], p["delay"], p["cp"])
        assert torch.equal(o, ref)


def test_uneven_batch_graph_replays_and_oracle(gpu, oracle):
    from dream_amd.torch_ops import DasBatch
    rng = np.random.default_rng(12)
    probs = _uneven_problems("tfm", rng, 9, torch.float64)
    batch = DasBatch("tfm", probs, precision="f64")
    first = [o.clone() for o in batch.run(use_graph=True)]
    for _ in range(3):
        torch.cuda.synchronize()
        scratch = torch.randn(1 << 16, device="cuda")      # eager work between replays
        scratch.mul_(2.0)
        outs = batch.run(use_graph=True)
        torch.cuda.synchronize()
        for a, b in zip(first, outs):
            assert torch.equal(a, b)
    for p, o in zip(probs[:3], first[:3]):
        ref, _ = oracle.das(p["Yh"].T.copy(), p["Gt"], p["Gr"], p["Ro"], p["dt"], p["delay"],
                            p["cp"], "tfm", "ignore", "f64", nthreads=4)
        assert np.max(np.abs(o.cpu().numpy() - ref)) <= 1e-12 * max(np.abs(ref).max(), 1.0)


@pytest.mark.parametrize("pitch,K", [(0.025, 1024), (0.1, 4096)])
def test_saft_adjoint_multi_region_work_groups(gpu, oracle, pitch, K):
    """A SAFT image large enough for adjoint work-groups of 2 and 4 regions
    (one window flush per work-group).  The coarse grid makes the windows of
    4-region groups overflow into direct gY atomics.  Gradients against the
    FP64 oracle with tuning off (the default, 4 regions) and on (the tuned
    pick, after the candidates ran on the same accumulators)."""
    from dream_amd import _lib
    from dream_amd.torch_ops import DasPlan
    rng = np.random.default_rng(41)
    L, n = 128, 512
    x = np.linspace(-6.35, 6.35, L)
    Gt = np.c_[x, np.zeros(L), np.zeros(L)]
    X, Z = np.meshgrid(pitch * (np.arange(n) - n / 2), 2.0 + pitch * np.arange(n))
    Ro = np.c_[X.ravel("F"), np.zeros(n * n), Z.ravel("F")]
    Y = rng.standard_normal((K, L))
    at = rng.uniform(0.5, 1.5, L)
    g = rng.standard_normal(n * n)
    gY, gat, _, gcp = oracle.das_ext_adjoint(Y, Gt, None, Ro, 0.02, 0.1, 1500.0, g, "saft",
                                             "linear", at, None, nthreads=8)
    plan = DasPlan("saft", K, Ro, Gt, None, precision="f64")
    Yt = torch.tensor(Y.T.copy(), device="cuda")
    att = torch.tensor(at, device="cuda")
    gt = torch.tensor(g, device="cuda")
    try:
        for tune in (0, 1, 1):
            _lib.lib().dream_hip_set_autotune(tune)
            rY, rat, _, rcp = plan.adjoint(Yt, gt, 0.02, 0.1, 1500.0, att, None, "linear")
            assert rel_l2(rY.cpu().numpy().T, gY) <= 1e-10
            assert rel_l2(rat.cpu().numpy(), gat) <= 1e-10
            assert abs(float(rcp) - gcp) <= 1e-9 * abs(gcp)
    finally:
        _lib.lib().dream_hip_set_autotune(1)


@pytest.mark.parametrize("method", ["tfm", "rca_rowtx"])
def test_two_leg_adjoint_multi_region_work_groups(gpu, oracle, method):
    """Two-leg adjoints with the Y-gradient on an image large enough for
    work-groups of several regions (tuning off: the largest; on: the pick),
    against the FP64 oracle."""
    from dream_amd import _lib
    from dream_amd.torch_ops import DasPlan
    rng = np.random.default_rng(42)
    L, n, K = 32, 128, 1024
    x = np.linspace(-3.1, 3.1, L)
    Gt = np.c_[x, np.zeros(L), np.zeros(L)]
    Gr = Gt.copy() if method == "tfm" else np.c_[np.zeros(L), x, np.zeros(L)]
    X, Z = np.meshgrid(0.05 * (np.arange(n) - n / 2), 2.0 + 0.05 * np.arange(n))
    Ro = np.c_[X.ravel("F"), np.zeros(n * n), Z.ravel("F")]
    Y = rng.standard_normal((K, L * L))
    at, ar = rng.uniform(0.5, 1.5, L), rng.uniform(0.5, 1.5, L)
    g = rng.standard_normal(n * n)
    gY, gat, gar, gcp = oracle.das_ext_adjoint(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, g, method,
                                               "linear", at, ar, nthreads=8)
    plan = DasPlan(method, K, Ro, Gt, Gr, precision="f64")
    Yt = torch.tensor(Y.T.copy(), device="cuda")
    att, art = torch.tensor(at, device="cuda"), torch.tensor(ar, device="cuda")
    gt = torch.tensor(g, device="cuda")
    try:
        for tune in (0, 1, 1):
            _lib.lib().dream_hip_set_autotune(tune)
            rY, rat, rar, rcp = plan.adjoint(Yt, gt, 0.02, 0.1, 1500.0, att, art, "linear")
            assert rel_l2(rY.cpu().numpy().T, gY) <= 1e-10
            assert rel_l2(rat.cpu().numpy(), gat) <= 1e-10
            assert rel_l2(rar.cpu().numpy(), gar) <= 1e-10
            assert abs(float(rcp) - gcp) <= 1e-9 * abs(gcp)
    finally:
        _lib.lib().dream_hip_set_autotune(1)


def test_autotuned_shapes_give_the_same_image(gpu):
    """The forward launch shape picked by the autotuner (waves per work-group,
    split) changes only the summation order: ramp data give identical images
    with tuning on and off, for a small image that has split candidates."""
    from dream_amd import _lib
    from dream_amd.torch_ops import DasPlan
    rng = np.random.default_rng(21)
    L, K, No = 48, 900, 300
    x = np.linspace(-6.0, 6.0, L)
    G = np.c_[x, np.zeros(L), np.zeros(L)]
    Ro = np.c_[rng.uniform(-6, 6, No), np.zeros(No), rng.uniform(1, 8, No)]
    Y = torch.arange(K, dtype=torch.float32, device="cuda").expand(L * L, K).contiguous()
    try:
        _lib.lib().dream_hip_set_autotune(0)
        off = DasPlan("tfm", K, Ro, G, G, precision="f32").forward(Y, 0.02, 0.05, 1500.0)
        _lib.lib().dream_hip_set_autotune(1)
        plan = DasPlan("tfm", K, Ro, G, G, precision="f32")
        first = plan.forward(Y, 0.02, 0.05, 1500.0).clone()     # tunes
        again = plan.forward(Y, 0.02, 0.05, 1500.0)             # reuses the pick
        assert torch.equal(first, off) and torch.equal(again, off)
    finally:
        _lib.lib().dream_hip_set_autotune(1)


@pytest.mark.parametrize("prec", ["f64", "f32", "f32_ref"])
def test_reciprocal_tfm_equals_full_pair_sum(gpu, oracle, prec):
    """One array on both sides and a zero delay: the forward sums each pair
    (i, j), (j, i) once over the symmetrized data (same sample index, the
    legs being computed by the same expression).  Against the full pair sum
    (the receive array passed as a separate copy): bit for bit on ramp data,
    to rounding on random data, linear interpolation, frames, split launches
    and the out-of-bounds status; a non-zero delay reads Y itself."""
    from dream_amd.torch_ops import DasPlan
    dt_ = torch.float64 if prec == "f64" else torch.float32
    # Split launch; unsplit, ragged blocks; an A-scan length that is not a
    # multiple of four (scalar symmetrizing).  Some pairs fall past K.
    for L, No, K in ((64, 3000, 1200), (24, 20000, 1200), (7, 500, 1201)):
        rng = np.random.default_rng(81 + L)
        x = np.linspace(-8.0, 8.0, L)
        G = np.c_[x, np.zeros(L), np.zeros(L)]
        Ro = np.c_[rng.uniform(-10, 10, No), np.zeros(No), rng.uniform(1, 12, No)]
        rec = DasPlan("tfm", K, Ro, G, G, precision=prec)
        assert rec.Gr is rec.Gt
        full = DasPlan("tfm", K, Ro, G, G, precision=prec)
        full.Gr = full.Gt.clone()
        ramp = torch.arange(K, dtype=dt_, device="cuda").remainder(512).expand(L * L, K)
        Yr = torch.stack([ramp, 2 * ramp]).contiguous()
        a = rec.forward(Yr, 0.02, 0.0, 1500.0)
        b = full.forward(Yr, 0.02, 0.0, 1500.0)
        assert torch.equal(a, b), (L, prec)
        Yn = torch.tensor(rng.standard_normal((2, L * L, K)), dtype=dt_, device="cuda")
        for interp in ("nearest", "linear"):
            a = rec.forward(Yn, 0.02, 0.0, 1500.0, interp=interp).double().cpu().numpy()
            b = full.forward(Yn, 0.02, 0.0, 1500.0, interp=interp).double().cpu().numpy()
            assert rel_l2(a, b) <= (1e-13 if prec == "f64" else 1e-6), (L, prec, interp)
        a, ca = rec.forward(Yn[0].contiguous(), 0.02, 0.0, 1500.0, err_level="warn",
                            want_status=True)
        b, cb = full.forward(Yn[0].contiguous(), 0.02, 0.0, 1500.0, err_level="warn",
                             want_status=True)
        assert ca == cb == 2
        # Non-zero delay: the device-side test sends the forward to Y (every
        # pair gathered; the launch shape, tuned for the paired job list, may
        # sum in another order).
        a = rec.forward(Yr, 0.02, 0.3, 1500.0)
        b = full.forward(Yr, 0.02, 0.3, 1500.0)
        assert torch.equal(a, b), (L, prec)
        a = rec.forward(Yn, 0.02, 0.3, 1500.0).double().cpu().numpy()
        b = full.forward(Yn, 0.02, 0.3, 1500.0).double().cpu().numpy()
        assert rel_l2(a, b) <= (1e-13 if prec == "f64" else 1e-6), (L, prec)
    # And against the oracle on the C2-like case through the host path.
    Y, Gt, Gr, Ro, rng = _setup("tfm", K=800, L=16, No=900, seed=12)
    ref, err = oracle.das(Y, Gt, Gt, Ro, 0.02, 0.0, 1500.0, "tfm", "warn", "f64", nthreads=4)
    from dream_amd import das
    Im = das.das(Y, Gt, Gt.copy(), Ro, 0.02, 0.0, 1500.0, "tfm", "warn", "gpu")
    assert np.max(np.abs(Im - ref)) <= 1e-13 * np.abs(Y).max() * Y.shape[1]
