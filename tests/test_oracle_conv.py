"""The oracle's FFT convolutions (fftconv_p, sum_fftconv) pinned against the
direct definition of linear convolution and its identities.

FFTW, which the reference's FFT class wraps (include/fft.h), is absent here,
so the reference's exact FFT rounding is unpinned; the sums themselves are
fixed by the direct definition below, which shares no code with the FFT
path."""
import numpy as np
import pytest

from oracle import oracle as O


def _direct(a, b):
    out = np.zeros(len(a) + len(b) - 1)
    for i, bi in enumerate(b):
        out[i:i + len(a)] += bi * a
    return out


@pytest.mark.parametrize("ma,mb,n", [(1000, 400, 5), (7, 1, 3), (1, 9, 2), (300, 300, 1)])
def test_fftconv_p_matches_direct(ma, mb, n):
    rng = np.random.default_rng(ma + mb)
    A = rng.standard_normal((ma, n))
    B = rng.standard_normal((mb, n))
    if mb > 1 and n > 1:                               # a filter per column
        Z = O.fftconv_p(A, B)
        for c in range(n):
            d = _direct(A[:, c], B[:, c])
            assert np.max(np.abs(Z[:, c] - d)) <= 1e-12 * max(np.abs(d).max(), 1.0)
    else:
        # A 1 x n or mb x 1 B is ONE filter of its n (mb) samples, as in
        # extra_src/fftconv_p.cc:170-173.
        Z = O.fftconv_p(A, B)
        assert Z.shape == (ma + B.size - 1, n)
    b = rng.standard_normal(mb)                        # one filter for every column
    Zv = O.fftconv_p(A, b)
    for c in range(n):
        d = _direct(A[:, c], b)
        assert np.max(np.abs(Zv[:, c] - d)) <= 1e-12 * max(np.abs(d).max(), 1.0)


def test_fftconv_identities_and_modes():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((64, 4))
    delta = np.zeros(5)
    delta[2] = 1.0                                      # shift by two samples
    Z = O.fftconv_p(A, delta)
    assert np.allclose(Z[2:66], A, atol=1e-13) and np.allclose(Z[:2], 0, atol=1e-13)
    b = rng.standard_normal(9)
    Y0 = rng.standard_normal((72, 4))
    assert np.allclose(O.fftconv_p(A, b, Y0, "+="), Y0 + O.fftconv_p(A, b), atol=1e-13)
    assert np.allclose(O.fftconv_p(A, b, Y0, "-="), Y0 - O.fftconv_p(A, b), atol=1e-13)
    # Commutativity on one column.
    assert np.allclose(O.fftconv_p(A[:, :1], b)[:, 0], O.fftconv_p(b, A[:, 0])[:, 0], atol=1e-13)


def test_sum_fftconv_matches_direct():
    rng = np.random.default_rng(2)
    M, N, L, K = 120, 6, 5, 33
    H = rng.standard_normal((M, N, L))
    U = rng.standard_normal((K, L))
    S = O.sum_fftconv(H, U)
    for n in range(N):
        d = sum(_direct(H[:, n, l], U[:, l]) for l in range(L))
        assert np.max(np.abs(S[:, n] - d)) <= 1e-12 * np.abs(d).max()
