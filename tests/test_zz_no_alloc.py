"""dream_hip_das_adjoint_dev_noalloc, called through the C ABI, against
dream_hip_das_adjoint_dev on the same inputs: two calls in a row, the second
on a smaller problem (the shared fixed workspace keeps the first call's
contents, so this checks that it is cleared), both kernels of the adjoint
(region-major with the Y gradient, pixel-major without), FP64 within 1e-12
and FP32 within 1e-6 relative L2; a problem beyond the workspace is refused
with DREAM_HIP_EINVAL; and the call leaves the device's free memory
unchanged."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
EINVAL = -1


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _case(method, L, K, No, F, seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(-4.0, 4.0, L)
    Gt = np.c_[x, np.zeros(L), np.zeros(L)]
    Gr = None if method == "saft" else (Gt.copy() if method == "tfm"
                                        else np.c_[np.zeros(L), x, np.zeros(L)])
    Ro = np.c_[rng.uniform(-6, 6, No), rng.uniform(-3, 3, No), rng.uniform(0.5, 9, No)]
    n_ascans = L if method == "saft" else L * L
    Y = rng.standard_normal((F, n_ascans, K))
    g = rng.standard_normal((F, No))
    at = rng.uniform(0.5, 1.5, L)
    ar = None if method == "saft" else rng.uniform(0.5, 1.5, L)
    return Gt, Gr, Ro, Y, g, at, ar


def _adjoint(fn, plan, Y, g, at, ar, interp, want_gy):
    """One adjoint call through the C ABI entry point `fn`; returns
    (rc, gY, g_apod_t, g_apod_r, g_cp)."""
    dt_ = plan.data_dtype
    F = Y.shape[0]
    gY = torch.full_like(Y, float("nan")) if want_gy else None
    gat = torch.full((plan.Lt,), float("nan"), dtype=dt_, device="cuda")
    gar = (torch.full((plan.Lr,), float("nan"), dtype=dt_, device="cuda")
           if plan.type != 0 else None)
    gcp = torch.full((1,), float("nan"), dtype=dt_, device="cuda")
    rc = fn(plan.h, _vp(Y), _vp(plan.Ro), _vp(plan.Gt), _vp(plan.Gr), 0.02, 0.1, 1500.0,
            _vp(at), _vp(ar), 1 if interp == "linear" else 0, F, plan.n_ascans * plan.K,
            plan.No, _vp(g), _vp(gY), _vp(gat), _vp(gar), _vp(gcp),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, gY, gat, gar, gcp


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_adjoint_equals_allocating_adjoint(gpu, prec):
    from dream_amd.torch_ops import DasPlan
    dt_ = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-12 if prec == "f64" else 1e-6
    # Larger call first, then a smaller one, for both adjoint kernels.
    cases = [("tfm", 16, 1024, 5000, 2, "linear", True),
             ("rca_rowtx", 8, 600, 700, 1, "nearest", True),
             ("rca_coltx", 16, 1024, 4000, 2, "linear", False),
             ("tfm", 8, 512, 600, 1, "linear", False),
             ("saft", 32, 2048, 6000, 2, "linear", True),
             ("saft", 8, 700, 500, 1, "nearest", True)]
    for i, (method, L, K, No, F, interp, want_gy) in enumerate(cases):
        print(f"noalloc case {i}: {method} L={L} K={K} No={No} F={F} {interp} gy={want_gy}",
              flush=True)
        Gt, Gr, Ro, Y, g, at, ar = _case(method, L, K, No, F, seed=70 + i)
        plan = DasPlan(method, K, Ro, Gt, Gr, precision=prec)
        Yt = torch.tensor(Y, dtype=dt_, device="cuda")
        gt = torch.tensor(g, dtype=dt_, device="cuda")
        att = torch.tensor(at, dtype=dt_, device="cuda")
        art = None if ar is None else torch.tensor(ar, dtype=dt_, device="cuda")
        ref = _adjoint(gpu.dream_hip_das_adjoint_dev, plan, Yt, gt, att, art, interp, want_gy)
        got = _adjoint(gpu.dream_hip_das_adjoint_dev_noalloc, plan, Yt, gt, att, art, interp,
                       want_gy)
        assert ref[0] == 0 and got[0] == 0, gpu.dream_hip_last_error()
        for r, x in zip(ref[1:], got[1:]):
            if r is None:
                assert x is None
                continue
            assert not torch.isnan(x).any(), i
            assert rel_l2(x.double().cpu().numpy(), r.double().cpu().numpy()) <= tol, i


def test_noalloc_adjoint_refuses_problems_beyond_the_workspace(gpu):
    from dream_amd.torch_ops import DasPlan
    L, K, No = 512, 8, 64                     # 512 x 512 A-scans + 1024 + 1 > 262144 doubles
    rng = np.random.default_rng(5)
    x = np.linspace(-4.0, 4.0, L)
    G = np.c_[x, np.zeros(L), np.zeros(L)]
    Ro = np.c_[rng.uniform(-6, 6, No), np.zeros(No), rng.uniform(0.5, 9, No)]
    plan = DasPlan("tfm", K, Ro, G, G, precision="f32")
    Y = torch.zeros((1, L * L, K), dtype=torch.float32, device="cuda")
    g = torch.ones((1, No), dtype=torch.float32, device="cuda")
    rc, _, gat, _, _ = _adjoint(gpu.dream_hip_das_adjoint_dev_noalloc, plan, Y, g, None, None,
                                "nearest", False)
    assert rc == EINVAL
    assert b"fixed workspace" in gpu.dream_hip_last_error()
    assert torch.isnan(gat).all()                        # nothing was written


def test_noalloc_adjoint_leaves_free_memory_unchanged(gpu):
    from dream_amd.torch_ops import DasPlan
    Gt, Gr, Ro, Y, g, at, ar = _case("rca_rowtx", 16, 1024, 3000, 2, seed=90)
    plan = DasPlan("rca_rowtx", 1024, Ro, Gt, Gr, precision="f32")
    Yt = torch.tensor(Y, dtype=torch.float32, device="cuda")
    gt = torch.tensor(g, dtype=torch.float32, device="cuda")
    gY = torch.empty_like(Yt)
    gat = torch.empty(plan.Lt, dtype=torch.float32, device="cuda")
    gar = torch.empty(plan.Lr, dtype=torch.float32, device="cuda")
    gcp = torch.empty(1, dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def call(want_gy):
        return gpu.dream_hip_das_adjoint_dev_noalloc(
            plan.h, _vp(Yt), _vp(plan.Ro), _vp(plan.Gt), _vp(plan.Gr), 0.02, 0.1, 1500.0, None,
            None, 1, 2, plan.n_ascans * plan.K, plan.No, _vp(gt), _vp(gY) if want_gy else None,
            _vp(gat), _vp(gar), _vp(gcp), stream)
    assert call(True) == 0 and call(False) == 0           # first calls on the device
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for w in (True, False, True):
        assert call(w) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
