"""The oracle's annular-array SIR (restatement of ArrAnnu::dream_arr_annu,
src/dream_arr_annu.cc:186-254, 361-505) pinned by closed forms and
invariants.  The reference's SIR sources need FFTW, absent here, so there is
no executable reference and the reference ships no golden vectors for this
function; the disc sum it builds on is pinned in test_oracle_circ.py."""
import numpy as np
import pytest

from oracle import oracle as O

CP, DT, DX = 1000.0, 0.25, 0.05
# Radii of the reference's own annular test (tests/m_files/test_dream_arr_annu.m).
GR = [3.5, 3.6, 5.0, 5.1, 6.0, 6.1, 7.0, 7.1, 8.0]


def test_ring_radii_focusing_and_apodization_closed_forms():
    assert np.array_equal(O.annu_ring_radii([1.0, 2.0, 4.0, 5.0, 7.0]), [0.0, 3.0, 6.0])
    # An even count: the last radius has no ring.
    assert np.array_equal(O.annu_ring_radii([1.0, 2.0, 4.0, 9.0]), [0.0, 3.0])
    # Spherical focusing law: the outermost ring gets no delay, the centre the most.
    F, rmax = 25.0, 6.0
    assert O.annu_focusing("x", F, rmax, rmax, CP) == 0.0
    want = (np.hypot(rmax, F) - F) * 1000.0 / CP
    assert abs(O.annu_focusing("xy", F, 0.0, rmax, CP) - want) < 1e-15
    for m in ("y", "x+y"):
        assert O.annu_focusing(m, F, 0.0, rmax, CP) == O.annu_focusing("x", F, 0.0, rmax, CP)
    assert O.annu_focusing("ud", 1.25, 3.0, rmax, CP) == 1.25
    assert O.annu_focusing("off", F, 0.0, rmax, CP) == 0.0
    # Window values at half the outer radius.
    h = 0.5 * rmax
    assert O.annu_apodization("triangle", 1, None, h, rmax) == 0.5
    assert O.annu_apodization("simply", 1, None, h, rmax) == 0.75
    assert O.annu_apodization("clamped", 1, None, h, rmax) == 0.5625
    assert O.annu_apodization("gauss", 1, None, h, rmax, 2.0) == pytest.approx(np.exp(-0.5))
    assert O.annu_apodization("raised", 1, None, h, rmax, 0.0) == pytest.approx(0.0, abs=1e-15)
    assert O.annu_apodization("hann", 1, None, h, rmax) == 1.0
    assert O.annu_apodization("ud", 2, [0.1, 0.2, 0.3], h, rmax) == 0.3


def test_single_disc_is_dreamcirc_one_sample_later():
    """One radius: the array is the centre disc, moved one sample later."""
    Ro = np.array([[0.0, 0.0, 10.0], [2.0, -1.0, 12.0], [5.0, 0.5, 8.0]])
    Hd, _, _ = O.dreamcirc(Ro, 2.5, DX, DX, DT, 300, 0.0, 1.0, CP, nthreads=2)
    Ha, e, _ = O.dream_arr_annu(Ro, [2.5], DX, DX, DT, 300, 0.0, 1.0, CP, nthreads=2)
    assert e == 0
    assert np.all(Ha[0] == 0.0)
    assert np.array_equal(Ha[1:], Hd[:-1])


def test_ring_is_outer_disc_minus_inner_disc():
    """Gr = [0, a, b]: the centre disc has no area (left zero) and the one ring
    is disc(b) - disc(a), one sample later."""
    Ro = np.array([[0.0, 0.0, 10.0], [3.0, 1.0, 9.0]])
    a, b = 2.0, 3.2
    Da, _, _ = O.dreamcirc(Ro, a, DX, DX, DT, 300, 0.0, 1.0, CP, nthreads=2)
    Db, _, _ = O.dreamcirc(Ro, b, DX, DX, DT, 300, 0.0, 1.0, CP, nthreads=2)
    Ha, e, _ = O.dream_arr_annu(Ro, [0.0, a, b], DX, DX, DT, 300, 0.0, 1.0, CP, nthreads=2)
    assert e == 0
    assert np.all(Ha[0] == 0.0)
    assert np.array_equal(Ha[1:], (Db - Da)[:-1])


def _window(z, a, b):
    """Arrival interval [us] of a ring of radii a < b on its axis at depth z."""
    return np.hypot(z, a) * 1000.0 / CP, np.hypot(z, b) * 1000.0 / CP


@pytest.mark.parametrize("ring", [1, 2, 3])
def test_isolated_ring_on_axis_matches_the_analytic_ring_sir(ring):
    """On the axis the SIR of a ring is v*cp between its inner and outer arrival
    times (the analytic circular-piston result, differenced).  A one-hot 'ud'
    apodization isolates ring n; 'ud' focusing moves it later by
    rint(focal[n]/dt) + 1 samples, and the weight scales the plateau."""
    z = 10.0
    gr = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0]
    w = np.zeros(len(gr))
    w[ring] = 0.7
    focal = np.zeros(len(gr))
    focal[:4] = [0.0, 1.3, 2.6, 3.9]                # us, per ring
    H, e, _ = O.dream_arr_annu(np.array([[0.0, 0.0, z]]), gr, 0.02, 0.02, 0.02, 1000, 0.0, 1.0,
                               CP, foc_met="ud", focal=focal, apod_met="ud", apod=w, nthreads=1)
    h = H[:, 0]
    dt = 0.02
    t0, t1 = _window(z, gr[2 * ring - 1], gr[2 * ring])
    s = np.rint(focal[ring] / dt) + 1
    i0, i1 = int(np.rint(t0 / dt) + s), int(np.rint(t1 / dt) + s)
    # Outside the window the discrete outer and inner discs differ only by
    # their sub-element grids (a few m/s here).
    nz = np.nonzero(np.abs(h) > 0.05 * 0.7 * CP)[0]
    assert abs(nz[0] - i0) <= 1 and abs(nz[-1] - i1) <= 1
    plateau = h[i0 + 3:i1 - 3]
    assert abs(plateau.mean() - 0.7 * CP) < 0.03 * CP


def test_x_focusing_aligns_the_rings_at_the_focal_depth():
    """Focused at F, the ring pulses of an on-axis point at depth F all arrive
    about sqrt(F^2 + r_max^2)/cp + dt; unfocused they spread over the ring
    arrival times."""
    F = 25.0
    Ro = np.array([[0.0, 0.0, F]])
    Hf, _, _ = O.dream_arr_annu(Ro, GR, DX, DX, 0.05, 1200, 0.0, 1.0, CP, foc_met="x", focal=F,
                                nthreads=1)
    Hn, _, _ = O.dream_arr_annu(Ro, GR, DX, DX, 0.05, 1200, 0.0, 1.0, CP, nthreads=1)
    rr = O.annu_ring_radii(GR)
    t_star = np.hypot(F, rr[-1]) * 1000.0 / CP + 0.05
    t = np.arange(1200) * 0.05
    stats = []
    for H in (Hf, Hn):
        h = np.abs(H[:, 0])
        hr = h[t > 20.0]                 # the rings, after the centre disc's pulse
        tr = t[t > 20.0]
        centre = (hr * tr).sum() / hr.sum()
        stats.append((centre, np.sqrt((hr * (tr - centre) ** 2).sum() / hr.sum())))
    (cf, wf), (cn, wn) = stats
    assert abs(cf - t_star) < 0.05          # within a sample of the aligned arrival
    assert wf < 0.1 and wn > 4 * wf


def test_points_are_independent_and_delays_shift():
    Ro = np.array([[0.0, 0.0, 10.0], [1.0, 2.0, 11.0], [-3.0, 0.5, 9.0]])
    H, e, _ = O.dream_arr_annu(Ro, GR, DX, DX, DT, 300, [0.0, 0.5, 1.0], 1.0, CP, nthreads=3)
    assert e == 0
    for p in range(3):
        Hp, _, _ = O.dream_arr_annu(Ro[p:p + 1], GR, DX, DX, DT, 300, [0.0, 0.5, 1.0][p], 1.0,
                                    CP, nthreads=1)
        assert np.array_equal(H[:, p], Hp[:, 0])
    # A delay of k samples moves the response k samples earlier.
    H2, _, _ = O.dream_arr_annu(Ro[:1], GR, DX, DX, DT, 300, 4 * DT, 1.0, CP, nthreads=1)
    assert np.allclose(H2[:-4, 0], H[4:, 0], rtol=0, atol=1e-9 * CP)


def test_shift_beyond_nt_or_negative():
    Ro = np.array([[0.0, 0.0, 10.0], [1.0, 0.0, 10.0]])
    base, _, _ = O.dream_arr_annu(Ro, [2.0], DX, DX, DT, 200, 0.0, 1.0, CP, nthreads=2)
    late, _, _ = O.dream_arr_annu(Ro, [2.0], DX, DX, DT, 200, 0.0, 1.0, CP, foc_met="ud",
                                  focal=[1e6], nthreads=2)
    assert np.all(late == 0.0)
    # focal = -3 dt: shift -2, so sample i holds the ring's sample i + 2 (i.e.
    # the unshifted array's sample i + 3).
    early, _, _ = O.dream_arr_annu(Ro, [2.0], DX, DX, DT, 200, 0.0, 1.0, CP, foc_met="ud",
                                   focal=[-3 * DT], nthreads=2)
    assert np.array_equal(early[:-3], base[3:])
    assert np.all(early[-2:] == 0.0)


def test_error_codes():
    Ro = np.array([[0.0, 0.0, 10.0], [2.0, 0.0, 10.0]])
    # nt too short for the outer discs.
    _, e, oob = O.dream_arr_annu(Ro, GR, DX, DX, DT, 45, 0.0, 1.0, CP, err_level="warn",
                                 nthreads=2)
    assert e == 2 and oob > 0
    H, e, _ = O.dream_arr_annu(Ro, GR, DX, DX, DT, 45, 0.0, 1.0, CP, err_level="stop",
                               nthreads=2)
    assert e == 3 and np.all(H == 0.0)
    # A point's status is its last disc's: when only an earlier (larger) disc
    # runs out of range, 'warn' reports none although samples were dropped.
    _, e, oob = O.dream_arr_annu(Ro, [8.0, 1.0, 1.5], DX, DX, DT, 45, 0.0, 1.0, CP,
                                 err_level="warn", nthreads=2)
    assert e == 0 and oob > 0
