"""Edge cases on the GPU: no observation points or pixels, elements narrower
than one sub-element step (no rows: zero response, as the reference's empty
walk gives), a centre disc of zero radius, nt = 1."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CP = 1500.0


def test_no_points(gpu, oracle):
    from dream_amd.das import das
    from dream_amd.dream_arr_annu import dream_arr_annu
    from dream_amd.dream_arr_rect import dream_arr_rect
    Ro = np.zeros((0, 3))
    H = dream_arr_rect(Ro, [0.5, 2.0], np.zeros((1, 3)), [0.05, 0.05, 0.02, 100], 0.0,
                       [1.0, CP, 0.0], "off", 0.0, "off", [0.0, 0.0], "off", None, 0.0, "stop")
    assert H.shape == (100, 0)
    Ha = dream_arr_annu(Ro, np.array([[1.0], [1.5], [2.0]]), [0.05, 0.05, 0.02, 100], 0.0,
                        [1.0, CP, 0.0], "off", 0.0, "off", None, 0.0, "stop")
    assert Ha.shape == (100, 0)
    G = np.c_[np.linspace(-1, 1, 4), np.zeros(4), np.zeros(4)]
    Im = das(np.ones((64, 16)), G, G, Ro, 0.02, 0.0, CP, "tfm", "stop", "gpu")
    assert Im.size == 0


def test_elements_without_rows_give_zeros(gpu, oracle):
    from dream_amd.dreamcirc import dreamcirc
    from dream_amd.dreamrect import dreamrect
    Ro = np.array([[0.0, 0.0, 10.0], [1.0, 0.5, 12.0]])
    # b < dy: the reference's row walk is empty.
    H = dreamrect(Ro, [1.0, 0.02], [0.05, 0.05, 0.02, 200], 0.0, [1.0, CP, 0.0], "stop")
    ref, e, _ = oracle.dreamrect(Ro, 1.0, 0.02, 0.05, 0.05, 0.02, 200, 0.0, 1.0, CP, nthreads=2)
    assert e == 0 and np.all(ref == 0.0) and np.all(H == 0.0)
    # A disc below half a step.
    H = dreamcirc(Ro, [0.01], [0.05, 0.05, 0.02, 200], 0.0, [1.0, CP, 0.0], "stop")
    ref, _, _ = oracle.dreamcirc(Ro, 0.01, 0.05, 0.05, 0.02, 200, 0.0, 1.0, CP, nthreads=2)
    assert np.array_equal(H, ref)


def test_zero_centre_disc_only_and_nt_one(gpu, oracle):
    from dream_amd.dream_arr_annu import dream_arr_annu
    from dream_amd.dreamrect import dreamrect
    Ro = np.array([[0.0, 0.0, 10.0], [1.0, 0.5, 12.0]])
    H = dream_arr_annu(Ro, np.array([[0.0]]), [0.05, 0.05, 0.02, 300], 0.0, [1.0, CP, 0.0],
                       "off", 0.0, "off", None, 0.0, "stop")
    assert H.shape == (300, 2) and np.all(H == 0.0)
    # nt = 1: every contribution but those landing on sample 0 is out of range.
    d = 10.0 * 1000.0 / CP
    H = dreamrect(Ro[:1], [0.3, 0.3], [0.05, 0.05, 0.02, 1], d, [1.0, CP, 0.0], "ignore")
    ref, _, _ = oracle.dreamrect(Ro[:1], 0.3, 0.3, 0.05, 0.05, 0.02, 1, d, 1.0, CP,
                                 err_level="ignore", nthreads=1)
    assert H.shape == (1, 1) and np.max(np.abs(H - ref)) <= 1e-12 * max(np.abs(ref).max(), 1.0)
