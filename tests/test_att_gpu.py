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
