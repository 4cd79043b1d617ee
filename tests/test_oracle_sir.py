"""Pins of the SIR oracle (dreamrect / dream_arr_rect restatement).

The reference's SIR sources need FFTW and cannot be compiled in this image,
so the restatement is checked against what the mathematics fixes:
  * the analytic on-axis plateau h = cp for a piston (Lockwood-Willette; the
    discrete sum over an annulus of width cp*dt gives cp exactly in the limit);
  * the time integral sum(h)*dt = 1000 v * integral of dA/(2 pi r), against 2-D
    adaptive quadrature;
  * the first-arrival sample index from the nearest sub-element;
  * mirror symmetry, superposition over elements, delay shifts, 'ud'
    focusing / apodization equivalences, linearity in v;
  * the error-code rules of src/dream_error.cc:40-76 and
    src/dream_arr_rect.cc:227-239 (the last element's status per point).
A dropped factor (2 pi, dt, 1000, ds, v), a wrong sign of the delay or of an
element offset, or swapped coordinates fails at least one of these.
"""
import numpy as np
import pytest
from scipy import integrate


def test_subgrid_accumulated_sequence(oracle):
    # The count and coordinates are those of the floating-point walk
    # s = -w/2 + d/2; while s <= w/2: s += d (src/dreamrect.cc:293-331).
    for w, d in ((0.9, 0.05), (10.0, 0.05), (0.2, 0.05), (1.0, 0.05), (7.3, 0.11)):
        xs = oracle.subgrid(w, d)
        ref = []
        s = -w / 2.0 + d / 2.0
        while s <= w / 2.0:
            ref.append(s)
            s += d
        assert np.array_equal(xs, ref)
    assert oracle.subgrid(0.9, 0.05).size == 18
    assert oracle.subgrid(10.0, 0.05).size == 200


def test_plateau_equals_cp(oracle):
    # On-axis point over a 4 x 4 mm piston: h = cp from z/cp to the first edge.
    cp, z, dt = 1500.0, 10.0, 0.01
    H, err, oob = oracle.dreamrect([[0.0, 0.0, z]], 4.0, 4.0, 0.01, 0.01, dt, 1000, 0.0, 1.0, cp,
                                   "stop", nthreads=1)
    h = H[:, 0]
    t0 = z / cp * 1000.0                          # us
    t_edge = np.hypot(z, 2.0) / cp * 1000.0
    i0, i1 = int(np.ceil(t0 / dt)) + 1, int(np.floor(t_edge / dt)) - 1
    plateau = h[i0:i1]
    assert plateau.size >= 8
    assert abs(plateau.mean() / cp - 1.0) < 0.01
    assert np.all(np.abs(plateau / cp - 1.0) < 0.05)
    assert np.all(h[: int(t0 / dt) - 1] == 0.0)     # nothing before the first arrival


def test_time_integral_matches_quadrature(oracle):
    # sum(h) dt = 1000 v * iint dA / (2 pi r) (no out-of-bounds samples).
    a, b, dx, dy, dt, v, cp = 1.2, 3.0, 0.01, 0.01, 0.005, 0.7, 1480.0
    for (xo, yo, zo) in ((0.3, -0.4, 5.0), (2.5, 1.0, 3.0), (0.0, 0.0, 8.0)):
        H, err, oob = oracle.dreamrect([[xo, yo, zo]], a, b, dx, dy, dt, 3000, 0.0, v, cp,
                                       "stop", nthreads=1)
        assert err == 0 and oob == 0
        f = lambda y, x: 1.0 / (2 * np.pi * np.sqrt((xo - x) ** 2 + (yo - y) ** 2 + zo ** 2))  # noqa: E731
        val, _ = integrate.dblquad(f, -a / 2, a / 2, -b / 2, b / 2, epsabs=1e-13, epsrel=1e-12)
        assert abs(H.sum() * dt / (1000.0 * v * val) - 1.0) < 2e-5


def test_first_arrival_index(oracle):
    a, b, dx, dy, dt, cp = 0.9, 10.0, 0.05, 0.05, 0.025, 1500.0
    rng = np.random.default_rng(3)
    for _ in range(5):
        xo, yo, zo = rng.uniform(-3, 3), rng.uniform(-6, 6), rng.uniform(5, 20)
        delay = rng.uniform(0, 3)
        H, err, _ = oracle.dreamrect([[xo, yo, zo]], a, b, dx, dy, dt, 2000, delay, 1.0, cp,
                                     "stop", nthreads=1)
        xs, ys = oracle.subgrid(a, dx), oracle.subgrid(b, dy)
        X, Yg = np.meshgrid(xs, ys)
        r = np.sqrt((xo - X) ** 2 + (yo - Yg) ** 2 + zo ** 2)
        first = int(np.rint((r.min() * 1000.0 / cp - delay) / dt))
        last = int(np.rint((r.max() * 1000.0 / cp - delay) / dt))
        nz = np.nonzero(H[:, 0])[0]
        assert nz[0] == first and nz[-1] == last


def test_mirror_symmetry(oracle):
    a, b, dx, dy, dt, cp = 0.8, 6.0, 0.02, 0.02, 0.02, 1500.0
    pts = np.array([[1.3, 0.7, 9.0], [-1.3, 0.7, 9.0], [1.3, -0.7, 9.0], [-1.3, -0.7, 9.0]])
    H, err, _ = oracle.dreamrect(pts, a, b, dx, dy, dt, 1000, 5.0, 1.0, cp, "stop", nthreads=1)
    for j in range(1, 4):
        assert np.linalg.norm(H[:, j] - H[:, 0]) / np.linalg.norm(H[:, 0]) < 2e-3
    # Swapping the roles of x and y (a <-> b, xo <-> yo) gives the same SIR.
    H2, _, _ = oracle.dreamrect([[0.7, 1.3, 9.0]], b, a, dy, dx, dt, 1000, 5.0, 1.0, cp, "stop",
                                nthreads=1)
    assert np.linalg.norm(H2[:, 0] - H[:, 0]) / np.linalg.norm(H[:, 0]) < 2e-3


def test_superposition_over_elements(oracle):
    a, b, dx, dy, dt, cp, nt = 0.5, 4.0, 0.05, 0.05, 0.02, 1500.0, 800
    G = np.array([[-2.0, 0.5, 0.0], [1.0, -0.25, 0.3], [3.5, 0.0, -0.2]])
    Ro = np.array([[0.4, 0.1, 7.0], [-1.0, 2.0, 9.0]])
    H, err, _ = oracle.dream_arr_rect(Ro, a, b, dx, dy, dt, nt, 2.0, 1.0, cp, G, nthreads=1)
    S = np.zeros_like(H)
    for g in G:
        Hg, _, _ = oracle.dreamrect(Ro - g, a, b, dx, dy, dt, nt, 2.0, 1.0, cp, "stop", nthreads=1)
        S += Hg
    assert np.max(np.abs(H - S)) <= 1e-13 * np.max(np.abs(S))
    assert np.count_nonzero(H) == np.count_nonzero(S)


def test_delay_shift_moves_samples(oracle):
    a, b, dx, dy, dt, cp = 0.9, 10.0, 0.05, 0.05, 0.025, 1500.0
    Ro = [[0.5, 0.0, 30.0]]
    H0, _, _ = oracle.dreamrect(Ro, a, b, dx, dy, dt, 400, 19.0, 1.0, cp, "stop", nthreads=1)
    H1, _, _ = oracle.dreamrect(Ro, a, b, dx, dy, dt, 400, 19.0 + 7 * dt, 1.0, cp, "stop",
                                nthreads=1)
    # A later start time moves the response 7 samples earlier.
    assert np.linalg.norm(H1[:-7, 0] - H0[7:, 0]) / np.linalg.norm(H0) < 1e-2
    c0 = np.sum(np.arange(400) * H0[:, 0]) / H0.sum()
    c1 = np.sum(np.arange(400) * H1[:, 0]) / H1.sum()
    assert abs((c0 - c1) - 7.0) < 0.05


def test_ud_focus_and_apodization_equivalences(oracle):
    a, b, dx, dy, dt, cp, nt = 0.9, 10.0, 0.05, 0.05, 0.025, 1500.0, 400
    Ro = np.array([[0.2, 0.1, 20.0], [-3.0, 0.0, 25.0]])
    G = np.array([[0.0, 0.0, 0.0]])
    d = np.array([12.0, 15.0])
    f = 1.75
    H, _, _ = oracle.dream_arr_rect(Ro, a, b, dx, dy, dt, nt, d, 1.0, cp, G, foc_met="ud",
                                    focal=[f], nthreads=1)
    R, _, _ = oracle.dreamrect(Ro, a, b, dx, dy, dt, nt, d - f, 1.0, cp, "stop", nthreads=1)
    assert np.array_equal(H, R)     # d = (delay - foc) - steer, bit for bit
    w = 0.37
    Ha, _, _ = oracle.dream_arr_rect(Ro, a, b, dx, dy, dt, nt, d, 1.0, cp, G, apod_met="ud",
                                     apod=[w], nthreads=1)
    Rw, _, _ = oracle.dreamrect(Ro, a, b, dx, dy, dt, nt, d, 1.0, cp, "stop", nthreads=1)
    assert np.max(np.abs(Ha - w * Rw)) <= 1e-14 * np.max(np.abs(Rw))
    Hv, _, _ = oracle.dreamrect(Ro, a, b, dx, dy, dt, nt, d, 2.5, cp, "stop", nthreads=1)
    assert np.max(np.abs(Hv - 2.5 * Rw)) <= 1e-14 * np.max(np.abs(Rw))


def test_error_codes(oracle):
    a, b, dx, dy, dt, cp = 0.9, 10.0, 0.05, 0.05, 0.025, 1500.0
    Ro = [[0.0, 0.0, 20.0], [0.0, 0.0, 21.0], [0.0, 0.0, 22.0]]
    # Everything arrives after nt*dt: out of bounds everywhere.
    for level, code in (("ignore", 1), ("warn", 2), ("stop", 3)):
        _, err, _ = oracle.dreamrect(Ro, a, b, dx, dy, dt, 10, 0.0, 1.0, cp, level, nthreads=2)
        assert err == code
    # One thread: the reference never reads its status back -> none.
    _, err, _ = oracle.dreamrect(Ro, a, b, dx, dy, dt, 10, 0.0, 1.0, cp, "ignore", nthreads=1)
    assert err == 0
    # dream_arr_rect keeps, per point, the status of the LAST element only.
    G_first_oob = np.array([[40.0, 0.0, 0.0], [0.0, 0.0, 0.0]])   # far element first
    G_last_oob = G_first_oob[::-1].copy()
    delay = 20.0 / 1.5 - 1.0
    _, e1, n1 = oracle.dream_arr_rect(Ro, a, b, dx, dy, dt, 200, delay, 1.0, cp, G_first_oob,
                                      err_level="ignore", nthreads=2)
    _, e2, n2 = oracle.dream_arr_rect(Ro, a, b, dx, dy, dt, 200, delay, 1.0, cp, G_last_oob,
                                      err_level="ignore", nthreads=2)
    assert n1 > 0 and n2 == n1
    assert e1 == 0 and e2 == 1
