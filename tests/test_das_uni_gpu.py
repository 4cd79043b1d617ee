"""GPU parity of das_uni (DAS over implicit uniform grids) against the oracle's
restatement of the das_uni kernels.  Bars: sample indices bit for bit (ramp
data, whose sums are exact integers in any order); random data 1e-12 (FP64)
and 2e-6 relative L2 (reference float arithmetic)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    # method, gt, gr, delay -- maxima off their grids and non-square RCA arrays
    ("saft", [-2.0, 0.5, 2.2], None, 0.3),
    ("tfm", [-2.0, 0.5, 2.0], [-1.5, 0.25, 1.6], 0.2),
    ("rca_coltx", [-2.0, 0.5, 2.2], [-1.5, 0.5, 1.5], 0.25),
    ("rca_rowtx", [-1.8, 0.4, 1.8], [-2.0, 0.5, 2.3], 0.35),
]
RO3 = np.array([[-3.0, 0.25, 3.1], [-1.2, 0.6, 1.2], [1.5, 0.2, 6.0]])


@pytest.mark.parametrize("method,gt,gr,delay", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_das_uni_matches_oracle(gpu, oracle, method, gt, gr, delay, dtype):
    from dream_amd.das_uni import das_uni
    prec = "f64" if dtype == np.float64 else "f32"
    Lt = oracle.uni_count(*gt, dtype=dtype)
    Lr = 1 if gr is None else oracle.uni_count(*gr, dtype=dtype)
    n_ascans = Lt if method == "saft" else Lt * Lr
    K = 700
    ramp = np.asfortranarray(np.repeat(np.arange(K, dtype=dtype)[:, None], n_ascans, axis=1))
    got = das_uni(ramp, gt, gr, RO3, 0.02, delay, 1500.0, method)
    ref = oracle.das_uni(ramp, gt, gr, RO3, 0.02, delay, 1500.0, method, prec)
    assert got.dtype == dtype and got.shape == ref.shape
    assert np.array_equal(got, ref)
    rng = np.random.default_rng(9)
    Y = np.asfortranarray(rng.standard_normal((K, n_ascans)).astype(dtype))
    got = das_uni(Y, gt, gr, RO3, 0.02, delay, 1500.0, method)
    ref = oracle.das_uni(Y, gt, gr, RO3, 0.02, delay, 1500.0, method, prec)
    if dtype == np.float64:
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(Y).max() * n_ascans
    else:
        assert rel_l2(got, ref) <= 2e-6


def test_das_uni_device_linear_and_reuse(gpu, oracle):
    """The device entry point with the reference rule equals the host one;
    the handle is reused while the setup is unchanged."""
    import ctypes

    import torch

    from dream_amd import _lib
    from dream_amd.das_uni import das_uni
    gt, gr = [-2.0, 0.5, 2.0], [-2.0, 0.5, 2.0]
    Lt = Lr = 9
    K = 512
    rng = np.random.default_rng(10)
    Y = np.asfortranarray(rng.standard_normal((K, Lt * Lr)))
    host = das_uni(Y, gt, gr, RO3, 0.02, 0.1, 1500.0, "tfm")
    L = _lib.lib()
    h = ctypes.c_void_p()
    g = np.array(gt)
    r = np.array(gr)
    R = np.ascontiguousarray(RO3.ravel(order="F"))
    _lib.check(L.dream_hip_das_uni_create(1, K, g.ctypes.data_as(_lib._dp),
                                          r.ctypes.data_as(_lib._dp), R.ctypes.data_as(_lib._dp),
                                          _lib.F64, ctypes.byref(h)))
    try:
        assert L.dream_hip_das_uni_setup_has_changed(h, 1, K, g.ctypes.data_as(_lib._dp),
                                                     r.ctypes.data_as(_lib._dp),
                                                     R.ctypes.data_as(_lib._dp), _lib.F64) == 0
        Yt = torch.tensor(Y.T.copy(), device="cuda")
        Im = torch.empty(host.size, dtype=torch.float64, device="cuda")
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(L.dream_hip_das_uni_dev(h, ctypes.c_void_p(Yt.data_ptr()), 0.02, 0.1, 1500.0,
                                           _lib.NEAREST, ctypes.c_void_p(Im.data_ptr()), s))
        assert np.array_equal(Im.cpu().numpy(), host)
        _lib.check(L.dream_hip_das_uni_dev(h, ctypes.c_void_p(Yt.data_ptr()), 0.02, 0.1, 1500.0,
                                           _lib.LINEAR, ctypes.c_void_p(Im.data_ptr()), s))
        assert np.isfinite(Im.cpu().numpy()).all()
    finally:
        L.dream_hip_das_uni_destroy(h)
