"""The oracle's disc SIR (dreamcirc / dream_arr_circ restatement of
src/dreamcirc.cc:286-352 and src/dream_arr_circ.cc) pinned by closed forms
and invariants; the reference's SIR sources need FFTW, absent here, so there
is no executable reference for it (its operation order is pinned by reading
src/dreamcirc.cc:300-348 only)."""
import numpy as np

from oracle import oracle as O


def test_on_axis_plateau_is_cp():
    """On the axis of a circular piston the SIR is cp (v = 1) between z/cp and
    sqrt(z^2 + R^2)/cp (the analytic circular-piston SIR)."""
    R, z, cp, dt = 2.0, 10.0, 1500.0, 0.01
    t0 = z / cp * 1000.0
    t1 = np.sqrt(z * z + R * R) / cp * 1000.0
    H, e, _ = O.dreamcirc(np.array([[0.0, 0.0, z]]), R, 0.01, 0.01, dt, 400, t0 - 0.5, 1.0, cp,
                          nthreads=2)
    h = H[:, 0]
    i0, i1 = int(round(0.5 / dt)), int(round((t1 - t0 + 0.5) / dt))
    assert e == 0
    assert abs(h[i0 + 3:i1 - 3].mean() - cp) < 2e-3 * cp
    assert np.all(h[:i0 - 2] == 0) and np.all(h[i1 + 3:] == 0)


def test_mass_equals_sum_of_amplitudes():
    """Every contribution lands in range, so sum(h) = sum over sub-elements of
    v ds / (2 pi r) / dt * 1000, and the sub-element count is that of the
    reference's accumulated row walk (about pi R^2 / ds)."""
    R, dx, dt, cp = 1.3, 0.05, 0.02, 1500.0
    xo, yo, zo = 0.7, -0.4, 8.0
    H, e, _ = O.dreamcirc(np.array([[xo, yo, zo]]), R, dx, dx, dt, 600, 4.0, 1.0, cp, nthreads=1)
    total, n = 0.0, 0
    ys = -R + dx / 2.0
    while ys <= R:
        rxs = np.sqrt(R * R - ys * ys)
        xs = -rxs + dx / 2.0
        while xs <= rxs:
            r = np.sqrt((xo - xs) ** 2 + (yo - ys) ** 2 + zo ** 2)
            total += dx * dx / (2 * np.pi * r) / dt * 1000.0
            n += 1
            xs += dx
        ys += dx
    assert e == 0
    assert abs(H.sum() - total) <= 1e-12 * total
    assert abs(n - np.pi * R * R / (dx * dx)) < 0.02 * n


def test_array_of_one_disc_equals_dreamcirc_and_mirror_symmetry():
    R, dx, dt, cp = 0.8, 0.04, 0.02, 1500.0
    Ro = np.array([[0.3, 0.2, 6.0], [-0.3, 0.2, 6.0], [0.3, -0.2, 6.0]])
    d = 3.0
    H1, _, _ = O.dreamcirc(Ro, R, dx, dx, dt, 300, d, 1.0, cp, nthreads=2)
    H2, _, _ = O.dream_arr_circ(Ro, R, dx, dx, dt, 300, d, 1.0, cp, np.zeros((1, 3)),
                                nthreads=2)
    assert np.array_equal(H1, H2)
    # The disc's sub-element grid is symmetric only to within a sub-element
    # (a row's walk starts at -sqrt(R^2 - ys^2) + dx/2), so mirrored points
    # agree in their time integral, not sample by sample.
    s = H1.sum(axis=0)
    assert abs(s[0] - s[1]) < 0.02 * s[0] and abs(s[0] - s[2]) < 0.02 * s[0]
