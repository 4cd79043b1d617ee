"""GPU parity of the attenuated SIR (alpha > 0; src/attenuation.cc:72-140)
against the oracle.  The GPU sums every point's contributions in the
frequency domain and transforms once; the oracle transforms each
contribution (as the reference does), so the two agree to rounding: the bar
is 1e-10 of the largest sample (FP64 output).  Error codes as without
attenuation."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CP = 1500.0


def _close(H, ref, tol=1e-10):
    return np.max(np.abs(H - ref)) <= tol * np.abs(ref).max()


@pytest.mark.parametrize("nt", [128, 129])
def test_dreamrect_attenuated(gpu, oracle, nt):
    from dream_amd.dreamrect import dreamrect
    rng = np.random.default_rng(31)
    No = 12
    Ro = np.c_[rng.uniform(-2, 2, No), rng.uniform(-2, 2, No), rng.uniform(8, 15, No)]
    d = Ro[:, 2] * 1000.0 / CP - 1.0
    H = dreamrect(Ro, [1.0, 2.0], [0.1, 0.1, 0.05, nt], d, [1.0, CP, 5.0], "ignore")
    ref, _, _ = oracle.dreamrect(Ro, 1.0, 2.0, 0.1, 0.1, 0.05, nt, d, 1.0, CP,
                                 err_level="ignore", nthreads=2, alpha=5.0)
    assert _close(H, ref)


def test_dream_arr_rect_attenuated_focused_apodized(gpu, oracle):
    from dream_amd.dream_arr_rect import dream_arr_rect
    rng = np.random.default_rng(32)
    L, No = 4, 8
    G = np.c_[np.linspace(-1.5, 1.5, L), np.zeros(L), np.zeros(L)]
    Ro = np.c_[rng.uniform(-3, 3, No), rng.uniform(-1, 1, No), rng.uniform(10, 20, No)]
    d = Ro[:, 2] * 1000.0 / CP - 2.0
    H = dream_arr_rect(Ro, [0.5, 1.5], G, [0.1, 0.1, 0.04, 200], d, [1.0, CP, 2.5], "xy", 15.0,
                       "x", [5.0, 0.0], "hann", None, 0.0, "ignore")
    ref, _, _ = oracle.dream_arr_rect(Ro, 0.5, 1.5, 0.1, 0.1, 0.04, 200, d, 1.0, CP, G,
                                      foc_met="xy", focal=15.0, steer_met="x", theta=5.0,
                                      apod_met="hann", err_level="ignore", nthreads=2,
                                      alpha=2.5)
    assert _close(H, ref)


def test_disc_cylinder_and_annulus_attenuated(gpu, oracle):
    from dream_amd.dream_arr_annu import dream_arr_annu
    from dream_amd.dreamcirc import dreamcirc
    from dream_amd.dreamcylind import dreamcylind
    Ro = np.array([[0.0, 0.0, 10.0], [0.5, 0.2, 12.0], [-1.0, 0.3, 9.0]])
    H = dreamcirc(Ro, [0.6], [0.1, 0.1, 0.05, 128], 5.0, [1.0, CP, 5.0], "ignore")
    ref, _, _ = oracle.dreamcirc(Ro, 0.6, 0.1, 0.1, 0.05, 128, 5.0, 1.0, CP,
                                 err_level="ignore", nthreads=2, alpha=5.0)
    assert _close(H, ref)
    H = dreamcylind(Ro, [0.5, 1.2, -6.0], [0.1, 0.1, 0.05, 128], 5.0, [1.0, CP, 5.0], "ignore")
    ref, _, _ = oracle.dreamcylind(Ro, 0.5, 1.2, -6.0, 0.1, 0.1, 0.05, 128, 5.0, 1.0, CP,
                                   err_level="ignore", nthreads=2, alpha=5.0)
    assert _close(H, ref)
    gr = [0.4, 0.5, 0.8, 0.9, 1.1]
    H = dream_arr_annu(Ro, np.array(gr).reshape(-1, 1), [0.1, 0.1, 0.05, 128], 5.0,
                       [1.0, CP, 5.0], "x", 12.0, "off", None, 0.0, "ignore")
    ref, _, _ = oracle.dream_arr_annu(Ro, gr, 0.1, 0.1, 0.05, 128, 5.0, 1.0, CP, foc_met="x",
                                      focal=12.0, err_level="ignore", nthreads=2, alpha=5.0)
    assert _close(H, ref)


def test_attenuated_single_point_many_chunks_and_codes(gpu, oracle):
    """One point: its contributions are split over many work-groups whose
    spectra add atomically; out-of-range samples give the usual codes."""
    from dream_amd.dreamrect import dreamrect
    Ro = np.array([[0.2, 0.1, 10.0]])
    H = dreamrect(Ro, [2.0, 4.0], [0.05, 0.05, 0.05, 160], 5.5, [1.0, CP, 4.0], "ignore")
    ref, _, _ = oracle.dreamrect(Ro, 2.0, 4.0, 0.05, 0.05, 0.05, 160, 5.5, 1.0, CP,
                                 err_level="ignore", nthreads=1, alpha=4.0)
    assert _close(H, ref)
    Ro2 = np.array([[0.0, 0.0, 10.0], [2.0, 0.0, 10.0]])
    with pytest.raises(RuntimeError):
        dreamrect(Ro2, [2.0, 4.0], [0.05, 0.05, 0.05, 28], 5.5, [1.0, CP, 4.0], "stop")
    H = dreamrect(Ro2, [2.0, 4.0], [0.05, 0.05, 0.05, 28], 5.5, [1.0, CP, 4.0], "ignore")
    ref, e, _ = oracle.dreamrect(Ro2, 2.0, 4.0, 0.05, 0.05, 0.05, 28, 5.5, 1.0, CP,
                                 err_level="ignore", nthreads=2, alpha=4.0)
    assert e == 1 and _close(H, ref)


@pytest.mark.parametrize("nt", [9000, 9001])
def test_long_responses_split_their_spectra(gpu, oracle, nt):
    """nt above one work-group's share of LDS (3584 bins): each point's
    spectrum is accumulated in bin blocks by several work-groups."""
    from dream_amd.dreamrect import dreamrect
    Ro = np.array([[0.0, 0.0, 10.0], [0.4, 0.1, 12.0], [-0.7, 0.0, 11.0]])
    d = Ro[:, 2] * 1000.0 / CP - 1.0
    H = dreamrect(Ro, [0.3, 0.3], [0.1, 0.1, 0.05, nt], d, [1.0, CP, 3.0], "ignore")
    ref, _, _ = oracle.dreamrect(Ro, 0.3, 0.3, 0.1, 0.1, 0.05, nt, d, 1.0, CP,
                                 err_level="ignore", nthreads=2, alpha=3.0)
    assert _close(H, ref)


def test_device_entry_point_with_attenuation(gpu, oracle):
    """dream_hip_sir_att_dev on device tensors (arrays of rectangles and of
    discs, 'ud' delays, FP64 and FP32 output) against the oracle."""
    import torch
    from dream_amd.torch_ops import sir_attenuated
    rng = np.random.default_rng(35)
    L, No, nt = 3, 6, 300
    G = np.c_[np.linspace(-1.0, 1.0, L), np.zeros(L), np.zeros(L)]
    Ro = np.c_[rng.uniform(-2, 2, No), rng.uniform(-1, 1, No), rng.uniform(8, 14, No)]
    d = Ro[:, 2] * 1000.0 / CP - 1.5
    focal = rng.uniform(0.0, 0.3, L)
    Rt = torch.tensor(Ro.T.copy(), device="cuda")
    dt_ = torch.tensor(d, device="cuda")
    base = dict(Ro=Rt, dx=0.1, dy=0.1, dt=0.05, nt=nt, delay=dt_, v=1.0, cp=CP, G=G,
                foc_met="ud", focal=focal, apod_met="hann", err_level="ignore")
    ref, e, oob = oracle.dream_arr_rect(Ro, 0.4, 0.8, 0.1, 0.1, 0.05, nt, d, 1.0, CP, G,
                                        foc_met="ud", focal=focal, apod_met="hann",
                                        err_level="ignore", nthreads=2, alpha=4.0)
    for out_dtype, tol in ((torch.float64, 1e-10), (torch.float32, 1e-6)):
        H, code, n_oob = sir_attenuated(dict(base, a=0.4, b=0.8), 4.0, out_dtype)
        assert _close(H.double().cpu().numpy().T, ref, tol)
        assert code == e and n_oob == oob
    ref, e, oob = oracle.dream_arr_circ(Ro, 0.5, 0.1, 0.1, 0.05, nt, d, 1.0, CP, G, foc_met="ud",
                                        focal=focal, apod_met="hann", err_level="ignore",
                                        nthreads=2, alpha=4.0)
    H, code, n_oob = sir_attenuated(dict(base, shape="circ", R=0.5), 4.0)
    assert _close(H.cpu().numpy().T, ref) and code == e and n_oob == oob
