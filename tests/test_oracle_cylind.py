"""The oracle's cylindrically-curved-element SIR (dreamcylind /
dream_arr_cylind, restating src/dreamcylind.cc:275-353, 442-517 and
src/dream_arr_cylind.cc:197-250) pinned by closed forms and invariants.  The
reference's SIR sources need FFTW, absent here, so there is no executable
reference, and the reference ships no golden vectors for these functions."""
import numpy as np

from oracle import oracle as O

CP = 1500.0


def _rows(b, Rc, dy):
    """Row angles and positions of the reference's walk (the test's own
    statement of src/dreamcylind.cc:292-310, 349)."""
    R = abs(Rc)
    hb = b / 2.0
    span = 2.0 * np.arctan(hb / np.sqrt(R * R - hb * hb))
    dphi = np.arcsin(dy / R)
    out, ph = [], -span / 2.0 + dphi / 2.0
    while ph <= span / 2.0:
        out.append(ph)
        ph += dphi
    return np.array(out), dphi


def test_flat_limit_is_the_rectangle():
    """With a huge curvature radius the element is the flat a x b rectangle:
    the same time integral to 1e-6 and nearly every sample equal."""
    a, b, dx = 1.0, 4.0, 0.05
    Ro = np.array([[0.3, 0.8, 12.0], [-1.0, 2.5, 20.0], [2.0, -1.0, 8.0]])
    d = 4.0
    Hr, _, _ = O.dreamrect(Ro, a, b, dx, dx, 0.02, 900, d, 1.0, CP, nthreads=2)
    for Rc in (1e7, -1e7):
        Hc, e, _ = O.dreamcylind(Ro, a, b, Rc, dx, dx, 0.02, 900, d, 1.0, CP, nthreads=2)
        assert e == 0
        assert np.allclose(Hc.sum(axis=0), Hr.sum(axis=0), rtol=1e-6, atol=0)
        assert np.linalg.norm(Hc - Hr) <= 1e-3 * np.linalg.norm(Hr)


def test_point_at_the_centre_of_curvature_sees_one_distance_per_column():
    """Every surface point of a column xs is sqrt(xs^2 + Rc^2) from the centre
    of curvature (0, 0, Rc) (focused) or (0, 0, -|Rc|) (defocused), so a
    narrow strip answers there in a single sample, with the summed amplitude
    of all its sub-elements, ds = |Rc| dx dphi."""
    a, b, dx, dt = 0.1, 6.0, 0.05, 0.01
    for Rc, zo in ((20.0, 20.0), (-20.0, -20.0)):
        phis, dphi = _rows(b, Rc, dx)
        H, e, _ = O.dreamcylind(np.array([[0.0, 0.0, zo], [0.0, 0.0, zo]]), a, b, Rc, dx, dx,
                                dt, 3000, 0.0, 1.0, CP, nthreads=2)
        h = H[:, 0]
        assert e == 0
        xs = np.array([-a / 2 + dx / 2, -a / 2 + 3 * dx / 2])
        r = np.sqrt(xs ** 2 + Rc ** 2)
        bins = np.rint(r * 1000.0 / CP / dt).astype(int)
        assert set(np.nonzero(h)[0]) <= set(bins)
        ds = abs(Rc) * dx * dphi
        want = (len(phis) * ds / (2 * np.pi * r) / dt * 1000.0).sum()
        assert abs(h.sum() - want) <= 1e-12 * want


def test_focused_and_defocused_differ_as_mirror_images():
    """Focused at +Rc seen from z equals defocused at -Rc seen from -z with the
    z offsets mirrored: rz = z - zs versus z + zs, so (z) <-> (-z) swaps them."""
    a, b, dx = 2.0, 3.0, 0.05
    Ro = np.array([[0.4, 0.3, 15.0], [1.5, -0.7, 9.0]])
    Rm = Ro * np.array([1.0, 1.0, -1.0])
    Hf, _, _ = O.dreamcylind(Ro, a, b, 10.0, dx, dx, 0.02, 800, 2.0, 1.0, CP, nthreads=2)
    Hd, _, _ = O.dreamcylind(Rm, a, b, -10.0, dx, dx, 0.02, 800, 2.0, 1.0, CP, nthreads=2)
    assert np.array_equal(Hf, Hd)
    # ... and they differ from each other seen from the same side.
    Hd2, _, _ = O.dreamcylind(Ro, a, b, -10.0, dx, dx, 0.02, 800, 2.0, 1.0, CP, nthreads=2)
    assert np.linalg.norm(Hd2 - Hf) > 0.1 * np.linalg.norm(Hf)


def test_b_beyond_the_curvature_diameter_gives_zero():
    H, e, oob = O.dreamcylind(np.array([[0.0, 0.0, 10.0], [1.0, 0.0, 10.0]]), 1.0, 9.0, 4.0,
                              0.05, 0.05, 0.02, 400, 0.0, 1.0, CP, nthreads=2)
    assert e == 0 and oob == 0 and np.all(H == 0.0)


def test_array_of_one_element_equals_dreamcylind_and_superposition():
    a, b, Rc, dx = 0.8, 5.0, 30.0, 0.05
    Ro = np.array([[0.3, 0.2, 16.0], [-0.3, 0.9, 12.0], [2.0, -0.2, 20.0]])
    d = 5.0
    H1, _, _ = O.dreamcylind(Ro, a, b, Rc, dx, dx, 0.02, 900, d, 1.0, CP, nthreads=2)
    H2, _, _ = O.dream_arr_cylind(Ro, a, b, Rc, dx, dx, 0.02, 900, d, 1.0, CP, np.zeros((1, 3)),
                                  nthreads=2)
    assert np.array_equal(H1, H2)
    # Two elements: the sum of each alone (element offset = moved point).
    G = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    H, _, _ = O.dream_arr_cylind(Ro, a, b, Rc, dx, dx, 0.02, 900, d, 1.0, CP, G, nthreads=2)
    Ha, _, _ = O.dreamcylind(Ro - G[0], a, b, Rc, dx, dx, 0.02, 900, d, 1.0, CP, nthreads=2)
    Hb, _, _ = O.dreamcylind(Ro - G[1], a, b, Rc, dx, dx, 0.02, 900, d, 1.0, CP, nthreads=2)
    assert np.allclose(H, Ha + Hb, rtol=0, atol=1e-12 * np.abs(H).max())


def test_error_codes():
    Ro = np.array([[0.0, 0.0, 10.0], [2.0, 0.0, 10.0]])
    _, e, oob = O.dreamcylind(Ro, 2.0, 4.0, 15.0, 0.05, 0.05, 0.02, 90, 5.0, 1.0, CP,
                              err_level="warn", nthreads=2)
    assert e == 2 and oob > 0
    H, e, _ = O.dreamcylind(Ro, 2.0, 4.0, 15.0, 0.05, 0.05, 0.02, 90, 5.0, 1.0, CP,
                            err_level="stop", nthreads=2)
    assert e == 3
