"""GPU parity on the benchmark configurations themselves, and the contract
checks around them.

- C5 (RCA row transmit, 128 x 128 elements at 0.2 mm, K = 4096, two frames):
  the full array and A-scan length over a 16 x 16 x 256-voxel sub-volume of
  C5's grid -- forward sample indices bit for bit (ramp data), FP32 echo data
  within 2e-5 relative L2 of the FP64 oracle, and the pixel-major adjoint
  (row and column apodization gradients, the backward C5 times) against the
  oracle's FP64 adjoint: FP64 <= 1e-10, FP32 <= 5e-5 (extra_src/das.cc:533-626).
- The pixel-major adjoint on TFM and both RCA methods, nearest and linear.
- C4 (SAFT, 128 elements, K = 4096, 512 x 512 pixels): loss
  ||DAS(y; apod, c) - DAS(y; 1, 1480)||^2 at c = 1500; gradients of
  DASFunction against the oracle's FP64 adjoint, a GPU dot-product test,
  central differences, and the FP32 forward within 2e-5 of the FP64 oracle.
- C3: all 32 problems (256 points each) in one FP32 SirBatch against the FP64
  oracle: <= 2e-5 relative L2 per problem, equal out-of-bounds counts/codes.
- Non-square RCA arrays against the reference's golden vectors.
- Long calls cut into chunks give the single-launch results; a call
  cancelled between chunks returns DREAM_HIP_ECANCELLED.
- Tensors the kernels cannot read are refused before any launch.
"""
import os
import threading
import time

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------------ C5

@pytest.fixture(scope="module")
def c5_sub():
    """C5's arrays, two frames of echo data and a 16 x 16 x 256-voxel
    sub-volume of its grid (x and y indices 56..71, every z)."""
    import workloads
    w = workloads.c5(n_frames=2, device="cuda", dtype=torch.float32)
    nx, ny, nz = 128, 128, 256
    ix, iy, iz = np.meshgrid(np.arange(56, 72), np.arange(56, 72), np.arange(nz), indexing="ij")
    sel = (ix + nx * (iy + ny * iz)).ravel("F")
    w["Ro_sub"] = w["Ro"][sel]
    return w


def _frame_f64(Y, f):
    """Frame f of a (F, n_ascans, K) tensor as the oracle's (K, n_ascans) FP64."""
    return Y[f].double().cpu().numpy().T


def test_c5_subvolume_forward_indices_and_f32(gpu, oracle, c5_sub):
    from dream_amd.torch_ops import DasPlan
    w = c5_sub
    Ro, Gt, Gr, K = w["Ro_sub"], w["Gt"], w["Gr"], w["K"]
    n_ascans = Gt.shape[0] * Gr.shape[0]
    # Ramp data: FP64 sums of sample indices are exact in any order.
    ramp = np.broadcast_to(np.arange(K, dtype=np.float64)[:, None], (K, n_ascans))
    ref, _ = oracle.das(ramp, Gt, Gr, Ro, w["dt"], 0.0, w["cp"], "rca_rowtx", "ignore", "f64")
    p64 = DasPlan("rca_rowtx", K, Ro, Gt, Gr, precision="f64")
    r = torch.arange(K, dtype=torch.float64, device="cuda").expand(n_ascans, K)
    Yr = torch.stack([r, 2 * r]).contiguous()
    Im = p64.forward(Yr, w["dt"], 0.0, w["cp"]).cpu().numpy()
    assert np.array_equal(Im[0], ref) and np.array_equal(Im[1], 2 * ref)
    # FP32 data, FP64 delays: k mod 1024 keeps every partial sum below 2^24.
    ref_m, _ = oracle.das(np.mod(ramp, 1024.0), Gt, Gr, Ro, w["dt"], 0.0, w["cp"], "rca_rowtx",
                          "ignore", "f64")
    p32 = DasPlan("rca_rowtx", K, Ro, Gt, Gr, precision="f32")
    Ym = torch.remainder(r, 1024.0).float().unsqueeze(0).contiguous()
    assert np.array_equal(p32.forward(Ym, w["dt"], 0.0, w["cp"]).double().cpu().numpy()[0], ref_m)
    # Echo data, two frames in one launch.
    Im32 = p32.forward(w["Y"], w["dt"], 0.0, w["cp"]).double().cpu().numpy()
    for f in range(2):
        ref_f, _ = oracle.das(_frame_f64(w["Y"], f), Gt, Gr, Ro, w["dt"], 0.0, w["cp"],
                              "rca_rowtx", "ignore", "f64")
        assert rel_l2(Im32[f], ref_f) <= 2e-5, f


@pytest.mark.parametrize("interp", ["nearest", "linear"])
def test_c5_subvolume_apodization_adjoint(gpu, oracle, c5_sub, interp):
    """The pixel-major adjoint (no Y gradient) over two frames against the
    sum of the oracle's per-frame FP64 adjoints."""
    from dream_amd.torch_ops import DasPlan
    w = c5_sub
    Ro, Gt, Gr, K = w["Ro_sub"], w["Gt"], w["Gr"], w["K"]
    rng = np.random.default_rng(51)
    No = Ro.shape[0]
    at = rng.uniform(0.5, 1.5, Gt.shape[0])
    ar = rng.uniform(0.5, 1.5, Gr.shape[0])
    g = rng.standard_normal((2, No))
    sat, sar, scp = 0.0, 0.0, 0.0
    for f in range(2):
        _, a1, r1, c1 = oracle.das_ext_adjoint(_frame_f64(w["Y"], f), Gt, Gr, Ro, w["dt"], 0.0,
                                               w["cp"], g[f], "rca_rowtx", interp, at, ar,
                                               want_gy=False)
        sat, sar, scp = sat + a1, sar + r1, scp + c1
    for prec, dt_, tol in (("f64", torch.float64, 1e-10), ("f32", torch.float32, 5e-5)):
        plan = DasPlan("rca_rowtx", K, Ro, Gt, Gr, precision=prec)
        Y = w["Y"].to(dt_)
        gat, gar = (torch.tensor(v, dtype=dt_, device="cuda") for v in (at, ar))
        gt = torch.tensor(g, dtype=dt_, device="cuda")
        _, rat, rar, rcp = plan.adjoint(Y, gt, w["dt"], 0.0, w["cp"], gat, gar, interp,
                                        want_gy=False, want_cp=True)
        assert rel_l2(rat.double().cpu().numpy(), sat) <= tol, prec
        assert rel_l2(rar.double().cpu().numpy(), sar) <= tol, prec
        if interp == "linear":
            assert abs(float(rcp) - scp) <= (1e-9 if prec == "f64" else 1e-3) * abs(scp), prec


@pytest.mark.parametrize("method", ["tfm", "rca_coltx", "rca_rowtx"])
@pytest.mark.parametrize("interp", ["nearest", "linear"])
def test_pixel_major_adjoint_against_oracle(gpu, oracle, method, interp):
    from dream_amd.torch_ops import DasPlan
    rng = np.random.default_rng(17)
    L, K, No = 12, 900, 1500
    x = np.linspace(-4.0, 4.0, L)
    Gt = np.c_[x, np.zeros(L), np.zeros(L)]
    Gr = Gt.copy() if method == "tfm" else np.c_[np.zeros(L), x, np.zeros(L)]
    Y = rng.standard_normal((K, L * L)).astype(np.float32).astype(np.float64)
    Ro = np.c_[rng.uniform(-6, 6, No), rng.uniform(-5, 5, No), rng.uniform(0.5, 9, No)]
    at, ar = rng.uniform(0.5, 1.5, L), rng.uniform(0.5, 1.5, L)
    g = rng.standard_normal(No)
    _, ref_t, ref_r, ref_cp = oracle.das_ext_adjoint(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, g, method,
                                                     interp, at, ar, want_gy=False)
    for prec, dt_, tol in (("f64", torch.float64, 1e-10), ("f32", torch.float32, 5e-5)):
        plan = DasPlan(method, K, Ro, Gt, Gr, precision=prec)
        Yt = torch.tensor(Y.T.copy(), dtype=dt_, device="cuda")
        a_t, a_r, gt = (torch.tensor(v, dtype=dt_, device="cuda") for v in (at, ar, g))
        _, rat, rar, rcp = plan.adjoint(Yt, gt, 0.02, 0.1, 1500.0, a_t, a_r, interp,
                                        want_gy=False, want_cp=True)
        assert rel_l2(rat.double().cpu().numpy(), ref_t) <= tol
        assert rel_l2(rar.double().cpu().numpy(), ref_r) <= tol
        if interp == "linear":
            assert abs(float(rcp) - ref_cp) <= (1e-9 if prec == "f64" else 1e-3) * abs(ref_cp)


# ------------------------------------------------------------------------ C4

@pytest.fixture(scope="module")
def c4():
    import workloads
    return workloads.c4(device="cuda", dtype=torch.float64)


def test_c4_loss_gradients_against_fp64_adjoint(gpu, oracle, c4):
    """loss = ||DAS(y; apod, c) - DAS(y; 1, 1480)||^2 at c = 1500, linear
    interpolation; DASFunction's gradients w.r.t. y, apod and c against the
    oracle's FP64 adjoint of the same loss."""
    from dream_amd.torch_ops import DasPlan, das_diff
    w = c4
    G, Ro, K, dt = w["Gt"], w["Ro"], w["K"], w["dt"]
    L = G.shape[0]
    Yn = w["Y"].cpu().numpy().T
    rng = np.random.default_rng(41)
    apod = rng.uniform(0.8, 1.2, L)
    tgt_ref, _ = oracle.das_ext(Yn, G, None, Ro, dt, 0.0, w["cp_target"], "saft", "linear",
                                np.ones(L), None)
    im_ref, _ = oracle.das_ext(Yn, G, None, Ro, dt, 0.0, 1500.0, "saft", "linear", apod, None)
    gY_ref, ga_ref, _, gcp_ref = oracle.das_ext_adjoint(Yn, G, None, Ro, dt, 0.0, 1500.0,
                                                        2.0 * (im_ref - tgt_ref), "saft",
                                                        "linear", apod, None)
    plan = DasPlan("saft", K, Ro, G, None, precision="f64")
    ones = torch.ones(L, dtype=torch.float64, device="cuda")
    target = plan.forward(w["Y"], dt, 0.0, w["cp_target"], ones, None, "linear")
    assert rel_l2(target.cpu().numpy(), tgt_ref) <= 1e-12
    Y = w["Y"].clone().requires_grad_(True)
    a = torch.tensor(apod, device="cuda", requires_grad=True)
    c = torch.tensor(1500.0, dtype=torch.float64, requires_grad=True)
    Im = das_diff(plan, Y, a, None, c, dt, 0.0, "linear")
    assert rel_l2(Im.detach().cpu().numpy(), im_ref) <= 1e-12
    loss = ((Im - target) ** 2).sum()
    loss.backward()
    assert rel_l2(Y.grad.cpu().numpy().T, gY_ref) <= 1e-10
    assert rel_l2(a.grad.cpu().numpy(), ga_ref) <= 1e-10
    assert abs(float(c.grad) - gcp_ref) <= 1e-9 * abs(gcp_ref)


def test_c4_dot_product_identity(gpu, c4):
    from dream_amd.torch_ops import DasPlan
    w = c4
    G, Ro, K, dt = w["Gt"], w["Ro"], w["K"], w["dt"]
    gen = torch.Generator(device="cuda").manual_seed(44)
    for prec, dtype, tol in (("f64", torch.float64, 1e-12), ("f32", torch.float32, 2e-5)):
        plan = DasPlan("saft", K, Ro, G, None, precision=prec)
        y = torch.randn((G.shape[0], K), generator=gen, device="cuda", dtype=dtype)
        g = torch.randn(Ro.shape[0], generator=gen, device="cuda", dtype=dtype)
        apod = torch.rand(G.shape[0], generator=gen, device="cuda", dtype=dtype) + 0.5
        for interp in ("linear", "nearest"):
            Ay = plan.forward(y, dt, 0.0, 1500.0, apod, None, interp)
            ATg, _, _, _ = plan.adjoint(y, g, dt, 0.0, 1500.0, apod, None, interp,
                                        want_apod=False, want_cp=False)
            lhs = float((Ay.double() * g.double()).sum())
            rhs = float((y.double() * ATg.double()).sum())
            assert abs(lhs - rhs) <= tol * float(Ay.double().norm() * g.double().norm()), \
                (prec, interp)


def test_c4_central_differences(gpu, c4):
    """Central differences at h = 1e-3 relative for the apodization weights and
    samples of y (the loss is quadratic in each, so the difference is exact up
    to rounding).  In c the image is piecewise smooth (linear interpolation
    between samples) and a 1e-3 relative step moves the C4 travel times by
    about 4 samples, most of a 5 MHz period at 50 MHz: there the difference of
    the linear functional <DAS(y; apod, c), g> at h = 1e-8 c is used."""
    from dream_amd.torch_ops import DasPlan, das_diff
    w = c4
    G, Ro, K, dt = w["Gt"], w["Ro"], w["K"], w["dt"]
    L = G.shape[0]
    plan = DasPlan("saft", K, Ro, G, None, precision="f64")
    ones = torch.ones(L, dtype=torch.float64, device="cuda")
    target = plan.forward(w["Y"], dt, 0.0, w["cp_target"], ones, None, "linear")
    rng = np.random.default_rng(42)
    apod0 = torch.tensor(rng.uniform(0.8, 1.2, L), device="cuda")

    def loss(Y, apod, c):
        Im = plan.forward(Y, dt, 0.0, c, apod, None, "linear")
        return float(((Im - target) ** 2).sum())

    Y = w["Y"].clone().requires_grad_(True)
    a = apod0.clone().requires_grad_(True)
    c = torch.tensor(1500.0, dtype=torch.float64, requires_grad=True)
    Im = das_diff(plan, Y, a, None, c, dt, 0.0, "linear")
    ((Im - target) ** 2).sum().backward()
    for n in (5, 64, 120):
        h = 1e-3 * float(apod0[n])
        ap, am = apod0.clone(), apod0.clone()
        ap[n] += h
        am[n] -= h
        fd = (loss(w["Y"], ap, 1500.0) - loss(w["Y"], am, 1500.0)) / (2 * h)
        assert abs(fd - float(a.grad[n])) <= 1e-6 * abs(float(a.grad[n])), n
    gY = Y.grad
    for flat in torch.topk((gY.abs() * w["Y"].abs()).flatten(), 3).indices.tolist():
        e, k = divmod(flat, K)
        h = 1e-3 * abs(float(w["Y"][e, k]))
        yp, ym = w["Y"].clone(), w["Y"].clone()
        yp[e, k] += h
        ym[e, k] -= h
        fd = (loss(yp, apod0, 1500.0) - loss(ym, apod0, 1500.0)) / (2 * h)
        assert abs(fd - float(gY[e, k])) <= 1e-6 * abs(float(gY[e, k])), (e, k)
    gfix = (2.0 * (Im - target)).detach()
    h = 1e-8 * 1500.0
    fp = float(plan.forward(w["Y"], dt, 0.0, 1500.0 + h, apod0, None, "linear") @ gfix)
    fm = float(plan.forward(w["Y"], dt, 0.0, 1500.0 - h, apod0, None, "linear") @ gfix)
    fd = (fp - fm) / (2 * h)
    assert abs(fd - float(c.grad)) <= 1e-6 * abs(float(c.grad))


def test_c4_forward_f32_against_fp64_oracle(gpu, oracle, c4):
    from dream_amd.torch_ops import DasPlan
    w = c4
    G, Ro, K, dt = w["Gt"], w["Ro"], w["K"], w["dt"]
    apod = np.random.default_rng(43).uniform(0.8, 1.2, G.shape[0])
    Y32 = w["Y"].float()
    ref, _ = oracle.das_ext(Y32.double().cpu().numpy().T, G, None, Ro, dt, 0.0, 1500.0, "saft",
                            "linear", apod, None)
    plan = DasPlan("saft", K, Ro, G, None, precision="f32")
    Im = plan.forward(Y32, dt, 0.0, 1500.0, torch.tensor(apod, dtype=torch.float32,
                                                          device="cuda"), None, "linear")
    assert rel_l2(Im.double().cpu().numpy(), ref) <= 2e-5


# ------------------------------------------------------------------------ C3

def test_c3_all_problems_f32_against_oracle(gpu, oracle):
    """Every one of C3's 32 problems ('ud' focusing, Hann apodization), 256
    points of each, in one FP32 SirBatch."""
    import workloads
    from dream_amd.torch_ops import SirBatch
    probs = workloads.c3()
    assert len(probs) == 32
    rng = np.random.default_rng(31)
    dev, host = [], []
    for p in probs:
        keep = np.sort(rng.choice(p["Ro"].shape[0], size=min(256, p["Ro"].shape[0]),
                                  replace=False))
        q = dict(p)
        q["Ro"], q["delay"] = p["Ro"][keep], p["delay"][keep]
        host.append(q)
        d = dict(q)
        d["Ro"] = torch.tensor(q["Ro"].T.copy(), device="cuda")
        d["delay"] = torch.tensor(q["delay"], device="cuda")
        dev.append(d)
    batch = SirBatch(dev, out_dtype=torch.float32)
    H = [h.double().cpu().numpy() for h in batch.run(use_graph=False)]
    codes, oobs = batch.status()
    for i, q in enumerate(host):
        ref, e, oob = oracle.dream_arr_rect(q["Ro"], q["a"], q["b"], q["dx"], q["dy"], q["dt"],
                                            q["nt"], q["delay"], q["v"], q["cp"], q["G"], "ud",
                                            q["focal"], apod_met="hann", err_level="ignore",
                                            nthreads=8)
        assert rel_l2(H[i], ref.T) <= 2e-5, i
        assert oobs[i] == oob and codes[i] == e, i


# ------------------------------------------------------------- non-square RCA

def test_nonsquare_rca_against_reference_vectors(gpu):
    """The reference reads the stripe extents at rows[Lc], rows[2 Lc - 1] and
    columns[Lr - 1] (extra_src/das.cc:464-467, :560-563); das/das_f on the GPU
    must give its vectors for arrays of 8 x 11 and 12 x 9 elements."""
    from dream_amd import das, das_f
    d = np.load(os.path.join(HERE, "golden", "das_rca_nonsquare_ref.npz"))
    for key in ("c8r11", "c12r9"):
        Y, Gt, Gr, Ro = (d[f"{key}__{n}"] for n in ("Y", "Gt", "Gr", "Ro"))
        scale = np.abs(Y).max() * Y.shape[1]
        for m in ("rca_coltx", "rca_rowtx"):
            Im = das.das(Y, Gt, Gr, Ro, 0.025, 0.3, 1510.0, m, "ignore", "gpu")
            assert np.max(np.abs(Im - d[f"{key}__{m}__f64__Im"])) <= 1e-13 * scale, (key, m)
            f = np.float32
            Imf = das_f.das(Y.astype(f), Gt.astype(f), Gr.astype(f), Ro.astype(f), f(0.025),
                            f(0.3), f(1510.0), m, "ignore", "gpu")
            ref = d[f"{key}__{m}__f32__Im"].astype(np.float64)
            assert np.max(np.abs(Imf.astype(np.float64) - ref)) <= 2e-6 * scale, (key, m)


def test_rca_extents_outside_the_tables_are_refused(gpu):
    from dream_amd.torch_ops import DasPlan
    Gt = np.c_[np.linspace(-1, 1, 4), np.zeros(4), np.zeros(4)]
    Gr = np.c_[np.zeros(20), np.linspace(-1, 1, 20), np.zeros(20)]
    with pytest.raises(RuntimeError, match="stripe extents"):
        DasPlan("rca_coltx", 64, np.array([[0.0, 0.0, 1.0]]), Gt, Gr, precision="f64")


# ------------------------------------------------------ chunks and cancellation

@pytest.fixture
def chunking(gpu):
    """Set the chunking test hook; always restore the defaults."""
    yield gpu.dream_hip_test_chunking
    gpu.dream_hip_test_chunking(0, 0)
    gpu.dream_hip_set_running()


def _tfm_case(No=3000, L=16, K=1024, seed=61):
    rng = np.random.default_rng(seed)
    x = np.linspace(-4.0, 4.0, L)
    G = np.c_[x, np.zeros(L), np.zeros(L)]
    Ro = np.c_[rng.uniform(-6, 6, No), np.zeros(No), rng.uniform(1, 9, No)]
    Y = rng.standard_normal((L * L, K))
    return G, Ro, Y


# 131072 pixels (2048 tiles of 64: a grid that fills the device, so the
# forward is not split) of an 8-element TFM array: 64 pairs per pixel.
BIG = dict(No=131072, L=8, K=512)
TILE_WORK = 64 * 64


def test_chunked_calls_equal_single_launches(gpu, chunking):
    from dream_amd.torch_ops import DasPlan, SirBatch
    G, Ro, Y = _tfm_case(**BIG)
    Yt = torch.tensor(np.stack([Y, 2 * Y]), device="cuda")
    g = torch.tensor(np.random.default_rng(62).standard_normal((2, Ro.shape[0])), device="cuda")
    plan = DasPlan("tfm", Y.shape[1], Ro, G, G, precision="f64")
    one = plan.forward(Yt, 0.02, 0.1, 1500.0, interp="linear")
    adj1 = plan.adjoint(Yt, g, 0.02, 0.1, 1500.0, interp="linear")
    pix1 = plan.adjoint(Yt, g, 0.02, 0.1, 1500.0, interp="linear", want_gy=False)
    # Chunks of 3 and 1000 tiles (block ranges) and of one frame.
    for work in (3 * TILE_WORK, 1000 * TILE_WORK, 2048 * TILE_WORK):
        chunking(work, 0)
        assert torch.equal(plan.forward(Yt, 0.02, 0.1, 1500.0, interp="linear"), one)
        adj = plan.adjoint(Yt, g, 0.02, 0.1, 1500.0, interp="linear")
        pix = plan.adjoint(Yt, g, 0.02, 0.1, 1500.0, interp="linear", want_gy=False)
        for x, y in zip(adj + pix[1:], adj1 + pix1[1:]):
            if x is not None:
                assert rel_l2(x.cpu().numpy(), y.cpu().numpy()) <= 1e-12, work
        chunking(0, 0)
    # A small image (heuristic shape: split partial sums) in chunks of 3 tiles,
    # each chunk reduced on its own: the same fixed-order sums.
    # (64 elements: 16 jobs per tile, so the heuristic splits each tile in two.)
    G, Ro, Y = _tfm_case(No=3000, L=64, K=512)
    Yt = torch.tensor(Y, device="cuda")
    gpu.dream_hip_set_autotune(0)
    try:
        small = DasPlan("tfm", Y.shape[1], Ro, G, G, precision="f64")
        one = small.forward(Yt, 0.02, 0.1, 1500.0)
        chunking(3 * 64 * 4096, 0)
        assert torch.equal(small.forward(Yt, 0.02, 0.1, 1500.0), one)
        chunking(0, 0)
    finally:
        gpu.dream_hip_set_autotune(1)
    # SIR plan: items in chunks.
    import workloads
    probs = workloads.sir_small_uneven_batch(n_problems=12)
    dev = []
    for p in probs:
        q = dict(p)
        q["Ro"] = torch.tensor(p["Ro"].T.copy(), device="cuda")
        q["delay"] = torch.tensor(p["delay"], device="cuda")
        dev.append(q)
    batch = SirBatch(dev, out_dtype=torch.float64)
    ref = [h.clone() for h in batch.run(use_graph=False)]
    ref_status = batch.status()
    chunking(200000, 0)
    got = batch.run(use_graph=False)
    for a, b in zip(got, ref):
        assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) <= 1e-13
    assert batch.status() == ref_status


def test_cancel_between_chunks(gpu, chunking):
    from dream_amd import _lib
    from dream_amd.torch_ops import DasPlan, SirBatch, arr_rect
    G, Ro, Y = _tfm_case(**BIG)
    Yt = torch.tensor(Y, device="cuda")
    plan = DasPlan("tfm", Y.shape[1], Ro, G, G, precision="f64")
    full = plan.forward(Yt, 0.02, 0.1, 1500.0)
    tiles_per_chunk = 3
    chunking(TILE_WORK * tiles_per_chunk, 2)            # abort after the second chunk
    out = torch.full((Ro.shape[0],), float("nan"), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.Cancelled):
        plan.forward(Yt, 0.02, 0.1, 1500.0, out=out)
    torch.cuda.synchronize()
    assert gpu.dream_hip_is_running() == 0
    done = ~torch.isnan(out)
    # Exactly the first two chunks' pixels (in visiting order) were written.
    assert int(done.sum()) == 2 * tiles_per_chunk * 64
    assert torch.equal(out[done], full[done])
    # A cancelled call leaves the library usable once running is set again.
    chunking(0, 0)
    gpu.dream_hip_set_running()
    assert torch.equal(plan.forward(Yt, 0.02, 0.1, 1500.0), full)
    # The same with split partial sums (a small image, heuristic shape).
    Gs, Rs, Ys = _tfm_case(No=3000, L=64, K=512)
    Yst = torch.tensor(Ys, device="cuda")
    gpu.dream_hip_set_autotune(0)
    try:
        small = DasPlan("tfm", Ys.shape[1], Rs, Gs, Gs, precision="f64")
        sfull = small.forward(Yst, 0.02, 0.1, 1500.0)
        chunking(3 * 64 * 4096, 2)
        sout = torch.full((Rs.shape[0],), float("nan"), dtype=torch.float64, device="cuda")
        with pytest.raises(_lib.Cancelled):
            small.forward(Yst, 0.02, 0.1, 1500.0, out=sout)
        torch.cuda.synchronize()
        sdone = ~torch.isnan(sout)
        assert int(sdone.sum()) == 2 * 3 * 64 and torch.equal(sout[sdone], sfull[sdone])
    finally:
        gpu.dream_hip_set_autotune(1)
        chunking(0, 0)
        gpu.dream_hip_set_running()
    # The adjoint and a SIR plan stop the same way.
    g = torch.ones(Ro.shape[0], dtype=torch.float64, device="cuda")
    chunking(TILE_WORK * tiles_per_chunk, 1)
    with pytest.raises(_lib.Cancelled):
        plan.adjoint(Yt, g, 0.02, 0.1, 1500.0, interp="linear")
    with pytest.raises(_lib.Cancelled):
        chunking(TILE_WORK * tiles_per_chunk, 1)
        plan.adjoint(Yt, g, 0.02, 0.1, 1500.0, interp="linear", want_gy=False)
    chunking(0, 0)
    gpu.dream_hip_set_running()
    import workloads
    p = workloads.c1()
    Rt = torch.tensor(p["Ro"].T.copy(), device="cuda")
    dl = torch.tensor(p["delay"], device="cuda")
    chunking(100000, 1)
    with pytest.raises(_lib.Cancelled):
        arr_rect(Rt, p["a"], p["b"], p["dx"], p["dy"], p["dt"], p["nt"], dl, p["v"], p["cp"],
                 p["G"])
    chunking(0, 0)
    gpu.dream_hip_set_running()
    H, code, _ = arr_rect(Rt, p["a"], p["b"], p["dx"], p["dy"], p["dt"], p["nt"], dl, p["v"],
                          p["cp"], p["G"], err_level="stop")
    assert code == 0 and float(H.abs().max()) > 0


def test_abort_from_another_thread_stops_a_long_call(gpu, chunking):
    """A long forward (8 frames of C5's array over 128 x 128 x 64 voxels,
    1.4e11 pairs, about 50 ms in 32 chunks) is stopped by dream_hip_abort from
    another thread a quarter of the way through."""
    from dream_amd import _lib
    from dream_amd.torch_ops import DasPlan
    import workloads
    Gc, Gw = workloads.rca_arrays()
    x = np.linspace(-12.8, 12.6, 128)
    X, Yg, Z = np.meshgrid(x, x, np.linspace(5, 45, 64), indexing="ij")
    Ro = np.c_[X.ravel("F"), Yg.ravel("F"), Z.ravel("F")]
    plan = DasPlan("rca_rowtx", 4096, Ro, Gc, Gw, precision="f32")
    Y = torch.randn((8, 128 * 128, 4096), device="cuda", dtype=torch.float32)
    plan.forward(Y, 0.025, 0.0, 1540.0)                    # tune first
    torch.cuda.synchronize()
    chunking(2 ** 32, 0)                                    # ~4.4e9 pairs per chunk
    t0 = time.perf_counter()
    full = plan.forward(Y, 0.025, 0.0, 1540.0)
    torch.cuda.synchronize()
    t_full = time.perf_counter() - t0
    timer = threading.Timer(t_full / 4, gpu.dream_hip_abort, [0])
    timer.start()
    t0 = time.perf_counter()
    with pytest.raises(_lib.Cancelled):
        plan.forward(Y, 0.025, 0.0, 1540.0)
    torch.cuda.synchronize()
    t_cancel = time.perf_counter() - t0
    timer.join()
    assert t_cancel < 0.8 * t_full, (t_cancel, t_full)
    del full


# ------------------------------------------------------------ input checks

def test_unreadable_tensors_are_refused_before_launch(gpu):
    from dream_amd.torch_ops import DasPlan, arr_rect, conv_sum
    G, Ro, Y = _tfm_case(No=200, L=4, K=128)
    plan = DasPlan("tfm", 128, Ro, G, G, precision="f32")
    Yc = torch.tensor(Y, dtype=torch.float32)
    with pytest.raises(ValueError, match="must be on"):
        plan.forward(Yc, 0.02, 0.0, 1500.0)                                  # host memory
    Yg = Yc.cuda()
    with pytest.raises(ValueError, match="float32"):
        plan.forward(Yg.double(), 0.02, 0.0, 1500.0)                         # dtype
    with pytest.raises(ValueError, match="contiguous"):
        plan.forward(Yg.t().contiguous().t(), 0.02, 0.0, 1500.0)             # strided view
    with pytest.raises(ValueError, match="elements"):
        plan.forward(Yg, 0.02, 0.0, 1500.0, torch.ones(3, device="cuda"),
                     torch.ones(4, device="cuda"))                            # apod too short
    with pytest.raises(ValueError, match="apod"):
        plan.forward(Yg, 0.02, 0.0, 1500.0, torch.ones(4, device="cuda", dtype=torch.float64),
                     torch.ones(4, device="cuda", dtype=torch.float64))      # apod dtype
    with pytest.raises(ValueError, match="delay"):
        plan.forward(Yg, 0.02, torch.zeros(7, dtype=torch.float64), 1500.0)  # delay length
    with pytest.raises(ValueError, match="g"):
        plan.adjoint(Yg, torch.ones(199, device="cuda"), 0.02, 0.0, 1500.0)  # g length
    with pytest.raises(ValueError, match="Ro"):
        arr_rect(torch.zeros((3, 4), dtype=torch.float64), 1.0, 1.0, 0.1, 0.1, 0.02, 100, 0.0,
                 1.0, 1500.0, np.zeros((1, 3)))                              # host points
    with pytest.raises(ValueError, match="focal"):
        arr_rect(torch.zeros((3, 4), dtype=torch.float64, device="cuda"), 1.0, 1.0, 0.1, 0.1,
                 0.02, 100, 0.0, 1.0, 1500.0, np.zeros((3, 3)), foc_met="ud", focal=[0.0])
    with pytest.raises(ValueError, match="must be on"):
        conv_sum(torch.ones((2, 8), dtype=torch.float64, device="cuda"),
                 torch.ones(3, dtype=torch.float64))                         # host filter
    torch.cuda.synchronize()


def test_overlapping_frames_are_refused(gpu):
    """The forward reads each frame through a view of K x n_ascans samples
    from its start: a frame stride below that (frames overlapping) is refused
    before any launch, and a stride above it is honoured."""
    import ctypes
    from dream_amd import _lib
    from dream_amd.torch_ops import DasPlan, _ptr, _stream
    L, K, No = 8, 300, 64
    G = np.c_[np.linspace(-4.0, 4.0, L), np.zeros(L), np.zeros(L)]
    Ro = np.c_[np.zeros(No), np.zeros(No), np.linspace(2.0, 6.0, No)]
    plan = DasPlan("tfm", K, Ro, G, G, precision="f32")
    d, dtype = plan._delay(0.0)
    stride = L * L * K
    Y = torch.randn(2 * stride + 7, dtype=torch.float32, device="cuda")
    Im = torch.zeros(2 * No, dtype=torch.float32, device="cuda")

    def call(y_stride):
        return _lib.lib().dream_hip_das_dev(
            plan.h, _ptr(Y), _ptr(plan.Ro), _ptr(plan.Gt), _ptr(plan.Gr), 0.02, dtype, _ptr(d),
            1500.0, None, None, 0, 2, y_stride, No, _ptr(Im), _lib.ERR_LEVEL["ignore"], None,
            _stream())

    assert call(stride - 1) != 0
    assert "overlap" in _lib.lib().dream_hip_last_error().decode()
    assert call(stride + 7) == 0
    torch.cuda.synchronize()
    ref0 = plan.forward(Y[:stride].view(L * L, K), 0.02, 0.0, 1500.0)
    ref1 = plan.forward(Y[stride + 7:].view(L * L, K), 0.02, 0.0, 1500.0)
    # (The one-frame and two-frame launches may be tuned to different
    # shapes, which sum in another order.)
    for got, ref in ((Im[:No], ref0), (Im[No:], ref1)):
        assert torch.allclose(got, ref, rtol=1e-5, atol=1e-5 * float(ref.abs().max()))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_split_sums_of_flat_and_strided_images_agree(gpu, prec):
    """A split forward launch (work-groups sharing a tile, partial sums added
    after it) adds four pixels at a time when the image rows are contiguous
    and pixel by pixel, in visiting order, when they are not: the two give
    the same two-frame image bit for bit.  128 x 128 pairs make 64 jobs per
    tile, so the untuned shape splits each tile over eight work-groups."""
    from dream_amd import _lib
    from dream_amd.torch_ops import DasPlan, _ptr, _stream
    L, K, F = 128, 256, 2
    x = np.linspace(-8.0, 8.0, L)
    Gt = np.c_[x, np.zeros(L), np.zeros(L)]
    Gr = np.c_[x, np.zeros(L), np.full(L, 0.01)]
    X, Z = np.meshgrid(np.linspace(-6.0, 6.0, 64), np.linspace(2.0, 5.0, 64), indexing="ij")
    Ro = np.c_[X.ravel(), np.zeros(X.size), Z.ravel()]
    No = Ro.shape[0]
    dtype = torch.float64 if prec == "f64" else torch.float32
    plan = DasPlan("tfm", K, Ro, Gt, Gr, precision=prec)
    d, dtype_code = plan._delay(0.0)
    torch.manual_seed(3)
    Y = torch.randn(F * L * L * K, dtype=dtype, device="cuda")

    def run(im_stride):
        Im = torch.full(((F - 1) * im_stride + No,), 7.0, dtype=dtype, device="cuda")
        rc = _lib.lib().dream_hip_das_dev(
            plan.h, _ptr(Y), _ptr(plan.Ro), _ptr(plan.Gt), _ptr(plan.Gr), 0.02, dtype_code,
            _ptr(d), 1500.0, None, None, 0, F, 0, im_stride, _ptr(Im),
            _lib.ERR_LEVEL["ignore"], None, _stream())
        assert rc == 0, _lib.lib().dream_hip_last_error().decode()
        torch.cuda.synchronize()
        return [Im[f * im_stride:f * im_stride + No] for f in range(F)]

    gpu.dream_hip_set_autotune(0)
    try:
        flat = run(No)
        strided = run(No + 4)
    finally:
        gpu.dream_hip_set_autotune(1)
    for a, b in zip(flat, strided):
        assert torch.equal(a, b)
        assert float(a.abs().max()) > 0.0
    # Against single-frame calls of the plan (own launch shape: to rounding).
    for f in range(F):
        ref = plan.forward(Y[f * L * L * K:(f + 1) * L * L * K].view(L * L, K), 0.02, 0.0,
                           1500.0)
        assert rel_l2(flat[f].double().cpu().numpy(), ref.double().cpu().numpy()) <= (
            1e-13 if prec == "f64" else 1e-5)


def test_frames_of_4gib_are_refused(gpu):
    """Gather offsets are 32-bit byte offsets inside one frame: a frame of
    4 GiB or more (FP64 TFM, 256 x 256 elements, K = 8192) is refused by the
    forward and the adjoint before any launch, not wrapped."""
    from dream_amd.torch_ops import DasPlan
    L, K = 256, 8192
    G = np.c_[np.linspace(-32.0, 32.0, L), np.zeros(L), np.zeros(L)]
    Ro = np.c_[np.zeros(64), np.zeros(64), np.linspace(5.0, 20.0, 64)]
    plan = DasPlan("tfm", K, Ro, G, G, precision="f64")
    Y = torch.empty((L * L, K), dtype=torch.float64, device="cuda")      # 4.29 GB, not read
    with pytest.raises(RuntimeError, match="4 GiB"):
        plan.forward(Y, 0.02, 0.0, 1500.0)
    with pytest.raises(RuntimeError, match="4 GiB"):
        plan.adjoint(Y, torch.ones(64, dtype=torch.float64, device="cuda"), 0.02, 0.0, 1500.0,
                     want_gy=False)
    del Y
    torch.cuda.synchronize()
