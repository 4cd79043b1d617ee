"""GPU parity of the annular-array SIR (dream_arr_annu) against the oracle.
Bars: FP64 output within 1e-12 of the largest sample (a moved sample index
moves a whole contribution, far above that bar); FP32 output 2e-5 relative
L2; error codes as in the reference."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu

CP = 1000.0
# Radii of the reference's own annular test (tests/m_files/test_dream_arr_annu.m).
GR = [3.5, 3.6, 5.0, 5.1, 6.0, 6.1, 7.0, 7.1, 8.0]


def _points(seed, No):
    rng = np.random.default_rng(seed)
    return np.c_[rng.uniform(-6, 6, No), rng.uniform(-6, 6, No), rng.uniform(8, 30, No)]


@pytest.mark.parametrize("foc,apod", [("x", "off"), ("off", "off"), ("ud", "ud"),
                                      ("xy", "triangle"), ("x+y", "clamped"), ("on", "simply"),
                                      ("y", "hann")])
def test_dream_arr_annu_module(gpu, oracle, foc, apod):
    from dream_amd.dream_arr_annu import dream_arr_annu
    rng = np.random.default_rng(11)
    No = 36
    Ro = _points(5, No)
    d = Ro[:, 2] * 1000.0 / CP - 2.0
    nr = len(GR)
    focal = rng.uniform(0.0, 1.5, nr) if foc == "ud" else 25.0
    w = rng.uniform(0.2, 1.0, nr)
    G = np.array(GR).reshape(nr, 1)
    H = dream_arr_annu(Ro, G, [0.05, 0.05, 0.05, 900], d, [1.0, CP, 0.0], foc, focal, apod, w,
                       1.0, "stop")
    ref, e, _ = oracle.dream_arr_annu(Ro, GR, 0.05, 0.05, 0.05, 900, d, 1.0, CP, foc_met=foc,
                                      focal=focal, apod_met=apod, apod=w, apod_par=0.0,
                                      err_level="stop", nthreads=2)
    assert e == 0
    assert H.shape == (900, No) and H.flags.f_contiguous
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()


def test_zero_centre_even_count_and_fp32_device_path(gpu, oracle):
    """Gr[0] = 0 (centre disc left zero), an even radius count (the last
    radius is computed for the status only), 'ud' focusing with a negative
    and an out-of-range delay; FP64 and FP32 device outputs."""
    from dream_amd.torch_ops import arr_annu
    gr = [0.0, 1.0, 2.0, 3.0, 4.0, 5.5]
    No = 50
    Ro = _points(7, No)
    d = Ro[:, 2] * 1000.0 / CP - 1.0
    focal = [0.0, -0.3, 1e7, 0.4, 0.0, 0.0]
    ref, e, _ = oracle.dream_arr_annu(Ro, gr, 0.04, 0.04, 0.02, 700, d, 1.0, CP, foc_met="ud",
                                      focal=focal, err_level="ignore", nthreads=2)
    Ro_d = torch.tensor(Ro.T.copy(), device="cuda")
    d_d = torch.tensor(d, device="cuda")
    H, code, _ = arr_annu(Ro_d, gr, 0.04, 0.04, 0.02, 700, d_d, 1.0, CP, foc_met="ud",
                          focal=focal, err_level="ignore")
    torch.cuda.synchronize()
    assert code == e
    Hn = H.cpu().numpy().T
    assert np.max(np.abs(Hn - ref)) <= 1e-12 * np.abs(ref).max()
    H32, _, _ = arr_annu(Ro_d, gr, 0.04, 0.04, 0.02, 700, d_d, 1.0, CP, foc_met="ud",
                         focal=focal, err_level="ignore", out_dtype=torch.float32)
    torch.cuda.synchronize()
    assert rel_l2(H32.cpu().numpy().T, ref) <= 2e-5


def test_annu_out_of_bounds_codes(gpu, oracle):
    from dream_amd.dream_arr_annu import dream_arr_annu
    from dream_amd.torch_ops import arr_annu
    # Two points, so the oracle runs two threads: the reference reads no
    # status back from a single-thread run (src/dream_arr_annu.cc:612-624).
    Ro = np.array([[0.0, 0.0, 10.0], [2.0, 0.0, 10.0]])
    args = (Ro, np.array(GR).reshape(-1, 1), [0.05, 0.05, 0.25, 45], 0.0, [1.0, CP, 0.0],
            "off", 0.0, "off", None, 0.0)
    with pytest.raises(RuntimeError):
        dream_arr_annu(*args, "stop")
    H = dream_arr_annu(*args, "warn")
    ref, e, _ = oracle.dream_arr_annu(Ro, GR, 0.05, 0.05, 0.25, 45, 0.0, 1.0, CP,
                                      err_level="warn", nthreads=2)
    assert e == 2 and np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()
    # Only an earlier (larger) disc runs out of range: the status is the last
    # disc's, none, although samples were dropped.
    Ro_d = torch.tensor(Ro.T.copy(), device="cuda")
    _, code, oob = arr_annu(Ro_d, [8.0, 1.0, 1.5], 0.05, 0.05, 0.25, 45, 0.0, 1.0, CP,
                            err_level="warn")
    _, e, oob_ref = oracle.dream_arr_annu(Ro, [8.0, 1.0, 1.5], 0.05, 0.05, 0.25, 45, 0.0, 1.0,
                                          CP, err_level="warn", nthreads=2)
    assert code == e == 0 and oob == oob_ref > 0


def test_row_vector_uses_first_radius_only(gpu, oracle):
    """As the reference reads G.shape[0] radii, a 1 x n row is one disc."""
    from dream_amd.dream_arr_annu import dream_arr_annu
    Ro = _points(9, 12)
    H = dream_arr_annu(Ro, np.array([GR]), [0.05, 0.05, 0.05, 900], 0.0, [1.0, CP, 0.0], "off",
                       0.0, "off", None, 0.0, "ignore")
    ref, _, _ = oracle.dream_arr_annu(Ro, GR[:1], 0.05, 0.05, 0.05, 900, 0.0, 1.0, CP,
                                      err_level="ignore", nthreads=2)
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()
