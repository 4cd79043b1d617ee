"""Pins of the oracle's extensions (linear interpolation, apodization, adjoint).

No reference exists for these, so they are fixed by:
  * reduction to the reference path: the nearest rule without apodization
    equals the reference-pinned das (bit for bit);
  * the linear rule against numpy.interp for single pairs;
  * linearity: <A y, g> = <y, A^T g> (dot-product test) and the apodization
    gradients equal directional derivatives of a function linear in them;
  * central finite differences in cp.
"""
import numpy as np
import pytest

METHODS = ("saft", "tfm", "rca_coltx", "rca_rowtx")


def _setup(method, K=600, L=8, No=150, seed=0):
    rng = np.random.default_rng(seed)
    x = np.linspace(-3.0, 3.0, L)
    Gt = np.c_[x, np.zeros(L), np.zeros(L)]
    Gr = Gt.copy() if method == "tfm" else np.c_[np.zeros(L), x, np.zeros(L)]
    n_ascans = L if method == "saft" else L * L
    t = np.arange(K)
    # Smooth band-limited data: a sum of Gaussian-windowed cosines per A-scan.
    Y = np.zeros((K, n_ascans))
    for a in range(n_ascans):
        c = rng.uniform(100, 500)
        Y[:, a] = np.exp(-((t - c) / 25.0) ** 2) * np.cos(2 * np.pi * 0.1 * (t - c))
    Y += 0.01 * rng.standard_normal(Y.shape)
    Ro = np.c_[rng.uniform(-3, 3, No), rng.uniform(-2, 2, No), rng.uniform(2, 7, No)]
    return Y, Gt, Gr, Ro, rng


@pytest.mark.parametrize("method", METHODS)
def test_nearest_reduces_to_reference_path(oracle, method):
    Y, Gt, Gr, Ro, _ = _setup(method)
    ref, _ = oracle.das(Y, Gt, Gr, Ro, 0.02, 0.15, 1500.0, method, "ignore", "f64", nthreads=2)
    ext, _ = oracle.das_ext(Y, Gt, Gr, Ro, 0.02, 0.15, 1500.0, method, interp="nearest")
    # Same sample indices; rca_rowtx sums in another order, hence the bound.
    assert np.max(np.abs(ext - ref)) <= 1e-13 * np.max(np.abs(ref))
    # Ramp data (Y[k] = k): every partial sum is an exact integer, so equal
    # sums mean equal index multisets per pixel, whatever the order.
    ramp = np.repeat(np.arange(Y.shape[0], dtype=np.float64)[:, None], Y.shape[1], axis=1)
    ref_r, _ = oracle.das(ramp, Gt, Gr, Ro, 0.02, 0.15, 1500.0, method, "ignore", "f64", nthreads=2)
    ext_r, _ = oracle.das_ext(ramp, Gt, Gr, Ro, 0.02, 0.15, 1500.0, method, interp="nearest")
    assert np.array_equal(ext_r, ref_r)
    Lr = 0 if method == "saft" else Gr.shape[0]
    ext1, _ = oracle.das_ext(Y, Gt, Gr, Ro, 0.02, 0.15, 1500.0, method, interp="nearest",
                             apod_t=np.ones(Gt.shape[0]), apod_r=np.ones(Lr) if Lr else None)
    assert np.max(np.abs(ext1 - ext)) == 0.0


def test_linear_matches_numpy_interp(oracle):
    # One element, one pixel straight below it: u = 2 (z/cp + delay) fs.
    K = 300
    rng = np.random.default_rng(5)
    y = rng.standard_normal(K)
    Gt = np.array([[0.0, 0.0, 0.0]])
    for z in (1.234, 2.0, 3.3333):
        dt, cp, delay = 0.02, 1500.0, 0.05
        u = 2.0 * (z * (1.0 / cp) + delay / 1000.0) * (1.0 / dt) * 1000.0
        Im, _ = oracle.das_ext(y[:, None], Gt, None, [[0.0, 0.0, z]], dt, delay, cp, "saft",
                               interp="linear")
        assert abs(Im[0] - np.interp(u, np.arange(K), y)) < 1e-12


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_dot_product_identity(oracle, method, interp):
    Y, Gt, Gr, Ro, rng = _setup(method, seed=11)
    Lr = 0 if method == "saft" else Gr.shape[0]
    at = rng.uniform(0.5, 1.5, Gt.shape[0])
    ar = rng.uniform(0.5, 1.5, Lr) if Lr else None
    y = rng.standard_normal(Y.shape)
    g = rng.standard_normal(Ro.shape[0])
    Ay, _ = oracle.das_ext(y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, method, interp, at, ar)
    gY, _, _, _ = oracle.das_ext_adjoint(y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, g, method, interp, at, ar)
    lhs, rhs = float(Ay @ g), float(np.sum(y * gY))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.linalg.norm(Ay) * np.linalg.norm(g))


@pytest.mark.parametrize("method", METHODS)
def test_apodization_gradients_are_directional_derivatives(oracle, method):
    Y, Gt, Gr, Ro, rng = _setup(method, seed=3)
    Lr = 0 if method == "saft" else Gr.shape[0]
    at = rng.uniform(0.5, 1.5, Gt.shape[0])
    ar = rng.uniform(0.5, 1.5, Lr) if Lr else None
    g = rng.standard_normal(Ro.shape[0])
    _, gat, gar, _ = oracle.das_ext_adjoint(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, g, method, "linear",
                                            at, ar, want_gy=False)
    base, _ = oracle.das_ext(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, method, "linear", at, ar)
    dat = rng.standard_normal(at.size)
    # Im is linear in apod_t (for fixed apod_r): exact directional derivative.
    moved, _ = oracle.das_ext(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, method, "linear", at + dat, ar)
    assert abs((moved - base) @ g - gat @ dat) <= 1e-10 * (abs(gat @ dat) + 1.0)
    if Lr:
        dar = rng.standard_normal(Lr)
        moved, _ = oracle.das_ext(Y, Gt, Gr, Ro, 0.02, 0.1, 1500.0, method, "linear", at, ar + dar)
        assert abs((moved - base) @ g - gar @ dar) <= 1e-10 * (abs(gar @ dar) + 1.0)


@pytest.mark.parametrize("method", METHODS)
def test_cp_gradient_central_differences(oracle, method):
    Y, Gt, Gr, Ro, rng = _setup(method, seed=4, No=400)
    g = rng.standard_normal(Ro.shape[0])
    cp = 1500.0
    _, _, _, gcp = oracle.das_ext_adjoint(Y, Gt, Gr, Ro, 0.02, 0.1, cp, g, method, "linear",
                                          want_gy=False)
    # Piecewise-linear in cp: with a step of 1e-9 cp almost no pair crosses a
    # sample boundary, and the central difference converges to the gradient.
    h = 1e-9 * cp
    fp, _ = oracle.das_ext(Y, Gt, Gr, Ro, 0.02, 0.1, cp + h, method, "linear")
    fm, _ = oracle.das_ext(Y, Gt, Gr, Ro, 0.02, 0.1, cp - h, method, "linear")
    fd = float((fp - fm) @ g / (2 * h))
    assert abs(fd - gcp) <= 1e-5 * abs(gcp)
