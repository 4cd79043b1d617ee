"""The oracle's attenuated SIR (Attenuation::att, src/attenuation.cc:72-140,
inside the DR sums) pinned by limits and by an independent FFT.  FFTW (the
reference's transform) is absent here and the reference ships no golden
vectors for attenuation; the transform convention (complex-to-real inverse,
divided by nt) and the magnitude law are pinned through numpy's FFT below.
The dispersion phase -w r a b/pi^2, b = log(1/(a1 w)), is pinned by the
Kramers-Kronig (minimum-phase) relation to the magnitude law, measured on the
oracle's own samples."""
import numpy as np

from oracle import oracle as O

CP, DT = 1500.0, 0.05


def test_vanishing_alpha_is_the_plain_sir():
    Ro = np.array([[0.0, 0.0, 10.0], [1.0, 0.5, 12.0]])
    H0, _, _ = O.dreamrect(Ro, 0.5, 0.5, 0.1, 0.1, DT, 128, 5.0, 1.0, CP, nthreads=2)
    H1, _, _ = O.dreamrect(Ro, 0.5, 0.5, 0.1, 0.1, DT, 128, 5.0, 1.0, CP, nthreads=2,
                           alpha=1e-9)
    assert np.abs(H1 - H0).max() <= 1e-8 * np.abs(H0).max()


def test_each_pulse_keeps_its_mass():
    """X[0] = 1: every attenuated pulse sums to 1, so the time integral equals
    the plain SIR's (same in-range contributions)."""
    Ro = np.array([[0.0, 0.0, 10.0], [1.0, 0.5, 12.0], [-2.0, 0.0, 15.0]])
    for nt in (128, 131):
        H0, _, _ = O.dream_arr_rect(Ro, 0.4, 1.0, 0.1, 0.1, DT, nt, 5.0, 1.0, CP,
                                    np.array([[-0.5, 0, 0], [0.5, 0, 0]]), nthreads=2)
        H5, _, _ = O.dream_arr_rect(Ro, 0.4, 1.0, 0.1, 0.1, DT, nt, 5.0, 1.0, CP,
                                    np.array([[-0.5, 0, 0], [0.5, 0, 0]]), nthreads=2,
                                    alpha=5.0)
        assert np.allclose(H5.sum(axis=0), H0.sum(axis=0), rtol=1e-12, atol=0)
        assert np.abs(H5 - H0).max() > 0.1 * np.abs(H0).max()


def _one_sub_element(xo, yo, zo, delay, nt, alpha):
    """A single sub-element at the origin: a = dx, b = dy."""
    H, e, _ = O.dreamrect(np.array([[xo, yo, zo], [xo, yo, zo]]), 0.1, 0.1, 0.1, 0.1, DT, nt,
                          delay, 1.0, CP, nthreads=2, alpha=alpha)
    r = np.sqrt(xo * xo + yo * yo + zo * zo)
    ai = 0.1 * 0.1 / (2 * np.pi * r) / DT * 1000.0
    it = int(np.rint((r * 1000.0 / CP - delay) / DT))
    return H[:, 0], r, ai, it, e


def test_spectrum_magnitude_is_the_absorption_law():
    """|X_k| = exp(-r a w_k/(2 pi)) with a = alpha / 8.686 per cm MHz in Np per
    m Hz (1 Np = 8.686 dB; 1/(cm MHz) = 10^-4/(m Hz)) and w_k = 2 pi k/(nt dt)."""
    for nt in (128, 129):
        alpha = 3.0
        h, r, ai, it, e = _one_sub_element(0.0, 0.0, 20.0, 10.0, nt, alpha)
        assert e == 0
        X = np.fft.rfft(h) / ai
        k = np.arange(X.size)
        a = alpha / (8.686 * 1e4)
        w = 2 * np.pi * k / (nt * DT * 1e-6)
        want = np.exp(-(r * 1e-3) * a * w / (2 * np.pi))
        if nt % 2 == 0:                  # only Re X[n/2] reaches the samples
            assert np.allclose(np.abs(X[:-1]), want[:-1], rtol=1e-9, atol=1e-12)
        else:
            assert np.allclose(np.abs(X), want, rtol=1e-9, atol=1e-12)
        assert abs(X[0] - 1.0) < 1e-12


def test_a_delay_of_m_samples_is_a_circular_shift():
    """Moving the sample index by m leaves r, and so the pulse, unchanged and
    rolls it circularly by m (the t w phase term)."""
    h1, _, ai, it, _ = _one_sub_element(0.3, 0.0, 20.0, 10.0, 128, 4.0)
    h2, _, _, it2, _ = _one_sub_element(0.3, 0.0, 20.0, 10.0 - 7 * DT, 128, 4.0)
    assert it2 == it + 7
    assert np.allclose(h2, np.roll(h1, 7), rtol=0, atol=1e-12 * np.abs(h1).max())
    # The pulse peaks just after its sample (absorption delays and spreads it)
    # and is larger after its sample than before it.
    assert it <= np.argmax(h1) <= it + 5
    assert h1[it:it + 10].sum() > 2 * h1[it - 10:it].sum()


def test_attenuation_reaches_discs_cylinders_and_annuli():
    Ro = np.array([[0.0, 0.0, 10.0], [0.5, 0.2, 12.0]])
    for H0, H5 in (
        (O.dreamcirc(Ro, 0.5, 0.1, 0.1, DT, 128, 5.0, 1.0, CP, nthreads=2)[0],
         O.dreamcirc(Ro, 0.5, 0.1, 0.1, DT, 128, 5.0, 1.0, CP, nthreads=2, alpha=5.0)[0]),
        (O.dreamcylind(Ro, 0.5, 1.0, 8.0, 0.1, 0.1, DT, 128, 5.0, 1.0, CP, nthreads=2)[0],
         O.dreamcylind(Ro, 0.5, 1.0, 8.0, 0.1, 0.1, DT, 128, 5.0, 1.0, CP, nthreads=2,
                       alpha=5.0)[0]),
):
        assert np.allclose(H5.sum(axis=0), H0.sum(axis=0), rtol=1e-12, atol=1e-9)
        assert np.abs(H5 - H0).max() > 0.05 * np.abs(H0).max()
    # Annuli: every disc is attenuated, then the rings are formed and moved
    # one sample later (which cuts the last sample of the spread pulses).
    D = {R: O.dreamcirc(Ro, R, 0.1, 0.1, DT, 128, 5.0, 1.0, CP, nthreads=2, alpha=5.0)[0]
         for R in (0.4, 0.5, 0.8)}
    Ha, _, _ = O.dream_arr_annu(Ro, [0.4, 0.5, 0.8], 0.1, 0.1, DT, 128, 5.0, 1.0, CP,
                                nthreads=2, alpha=5.0)
    assert np.array_equal(Ha[1:], (D[0.4] + (D[0.8] - D[0.5]))[:-1])


def test_dispersion_phase_is_the_minimum_phase_of_linear_absorption():
    """Kramers-Kronig for absorption linear in frequency (Futterman 1962; Aki
    and Richards 1980, ch. 5): a causal, minimum-phase filter with
    ln|X(w)| = -c w has the phase (2 c / pi) w ln(w / w0) up to a pure delay.
    Measured on one attenuated sub-element's spectrum (numpy's FFT of the
    oracle's samples, the t w delay removed): the magnitude decay c and the
    slope s of phase/w against ln w must satisfy s = (2/pi) c.  A dropped
    pi, a sign error or a missing dispersion term fails.  The reference
    frequency w0 = pi / (0.95 dt) is the reference's causality constant
    (src/attenuation.cc:96-98)."""
    for nt, alpha, zo in ((257, 3.0, 20.0), (401, 8.0, 35.0)):
        h, r, ai, it, e = _one_sub_element(0.0, 0.0, zo, zo * 1000.0 / CP - 5.0, nt, alpha)
        assert e == 0
        X = np.fft.rfft(h) / ai
        k = np.arange(1, X.size)
        w = 2 * np.pi * k / (nt * DT * 1e-6)                 # rad/s
        t = it * DT * 1e-6                                    # s
        Xk = X[1:] * np.exp(1j * t * w)                       # remove the sample delay
        phase = np.unwrap(np.angle(Xk))
        keep = np.abs(Xk) > 1e-6                              # above the FFT's rounding
        assert keep.sum() > 40
        c = -np.polyfit(w[keep], np.log(np.abs(Xk[keep])), 1)[0]        # ln|X| = -c w
        s, icpt = np.polyfit(np.log(w[keep]), phase[keep] / w[keep], 1)  # phase/w = s ln w + b
        assert abs(s - 2.0 * c / np.pi) <= 1e-6 * abs(s)
        assert s > 0
        a1 = DT * 1e-6 * 0.95 / np.pi
        assert abs(icpt - s * np.log(a1)) <= 1e-6 * abs(s * np.log(a1))
