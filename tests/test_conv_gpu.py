"""GPU parity of the batched convolution (fftconv_p, sum_fftconv, conv_sum)
against the oracle.  Bars: FP64 1e-12 relative to the column's largest
magnitude; FP32 1e-5 relative L2."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ma,mb,n", [(1000, 400, 50), (7, 1, 3), (1, 9, 2), (2500, 3000, 4),
                                     (64, 300, 0)])
def test_fftconv_p_matrix_and_vector(gpu, oracle, ma, mb, n):
    from dream_amd.fftconv_p import fftconv_p
    rng = np.random.default_rng(ma * 7 + mb)
    A = np.asfortranarray(rng.standard_normal((ma, n)))
    B = np.asfortranarray(rng.standard_normal((mb, n)))
    b = rng.standard_normal(mb)
    # A 1 x n (or mb x 1) B is one filter of its n (mb) samples, as in the
    # reference (extra_src/fftconv_p.cc:170-173).
    for filt in ((B, b) if n else (b,)):
        Y = fftconv_p(A, filt)
        ref = oracle.fftconv_p(A, filt)
        assert Y.shape == ref.shape and Y.flags.f_contiguous
        if n:
            assert np.max(np.abs(Y - ref)) <= 1e-12 * np.abs(ref).max()


def test_fftconv_p_argument_check(gpu):
    from dream_amd.fftconv_p import fftconv_p
    with pytest.raises(RuntimeError, match="same number of columns"):
        fftconv_p(np.ones((10, 3)), np.ones((4, 2)))


def test_sum_fftconv_normal_and_in_place(gpu, oracle):
    from dream_amd.sum_fftconv import sum_fftconv
    rng = np.random.default_rng(5)
    M, N, L, K = 700, 33, 9, 120
    H = np.asfortranarray(rng.standard_normal((M, N, L)))
    U = np.asfortranarray(rng.standard_normal((K, L)))
    ref = oracle.sum_fftconv(H, U)
    Y = sum_fftconv(H, U)
    assert np.max(np.abs(Y - ref)) <= 1e-12 * np.abs(ref).max()
    Y0 = np.asfortranarray(rng.standard_normal(ref.shape))
    Y1 = Y0.copy(order="F")
    assert sum_fftconv(H, U, Y1) is None
    assert np.max(np.abs(Y1 - (Y0 + ref))) <= 1e-12 * np.abs(Y0 + ref).max()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_conv_sum_device_modes(gpu, oracle, dtype):
    from dream_amd.torch_ops import conv_sum
    rng = np.random.default_rng(6)
    L, N, MA, MB = 3, 40, 513, 257
    A = rng.standard_normal((L, N, MA))
    Bl = rng.standard_normal((L, MB))
    ref = sum(oracle.fftconv_p(A[l].T, Bl[l]) for l in range(L)).T
    At = torch.tensor(A, dtype=dtype, device="cuda")
    Bt = torch.tensor(Bl, dtype=dtype, device="cuda")
    out = conv_sum(At, Bt)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert rel_l2(out.double().cpu().numpy(), ref) <= tol
    base = torch.tensor(rng.standard_normal(ref.shape), dtype=dtype, device="cuda")
    acc = base.clone()
    conv_sum(At, Bt, out=acc, mode="+=")
    assert rel_l2(acc.double().cpu().numpy(), base.double().cpu().numpy() + ref) <= tol
    conv_sum(At, Bt, out=acc, mode="-=")
    assert rel_l2(acc.double().cpu().numpy(), base.double().cpu().numpy()) <= 10 * tol
    # One filter per column (two-dimensional A).
    Bc = rng.standard_normal((N, MB))
    out2 = conv_sum(At[0].contiguous(), torch.tensor(Bc, dtype=dtype, device="cuda"))
    ref2 = oracle.fftconv_p(A[0].T, Bc.T).T
    assert rel_l2(out2.double().cpu().numpy(), ref2) <= tol


@pytest.fixture
def conv_path():
    """Force the direct or the FFT path (DREAM_HIP_CONV); restore after."""
    import os
    old = os.environ.get("DREAM_HIP_CONV")

    def set_path(p):
        if p is None:
            os.environ.pop("DREAM_HIP_CONV", None)
        else:
            os.environ["DREAM_HIP_CONV"] = p
    yield set_path
    set_path(old)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_fft_path_equals_direct_sums(gpu, oracle, conv_path, dtype, tol):
    """The FFT path (power-of-two transforms in LDS, two columns per complex
    transform) against the oracle and the direct path: a shared filter (odd
    column count), per-column filters (split by conjugate symmetry), layers
    with their own filters summed in the frequency domain, output lengths
    just under and at a power of two, the += and -= modes."""
    from dream_amd.torch_ops import conv_sum
    rng = np.random.default_rng(91)
    cases = [(3000, 1000, 7, 1, "vec"), (600, 400, 9, 1, "col"), (1500, 500, 6, 3, "layer"),
             (700, 325, 5, 1, "vec"), (513, 512, 4, 1, "col"), (1, 900, 3, 1, "vec")]
    for MA, MB, N, L, kind in cases:
        A = rng.standard_normal((L, N, MA))
        if kind == "vec":
            B = rng.standard_normal(MB)
        elif kind == "col":
            B = rng.standard_normal((N, MB))
        else:
            B = rng.standard_normal((L, MB))
        At = torch.tensor(A if L > 1 else A[0], dtype=dtype, device="cuda")
        Bt = torch.tensor(B, dtype=dtype, device="cuda")
        # Reference: the direct definition, summed over layers.
        ref = np.zeros((N, MA + MB - 1))
        for l in range(L):
            for n in range(N):
                b = B if kind == "vec" else (B[n] if kind == "col" else B[l])
                ref[n] += np.convolve(A[l, n], b)
        outs = {}
        for path in ("fft", "direct"):
            conv_path(path)
            Y = conv_sum(At, Bt).double().cpu().numpy()
            scale = np.abs(ref).max()
            assert np.max(np.abs(Y - ref)) <= tol * scale * (10 if path == "fft" else 1), \
                (MA, MB, N, L, kind, path)
            outs[path] = Y
            Y0 = torch.tensor(rng.standard_normal(ref.shape), dtype=dtype, device="cuda")
            for mode, sign in (("+=", 1.0), ("-=", -1.0)):
                Yi = Y0.clone()
                conv_sum(At, Bt, out=Yi, mode=mode)
                want = Y0.double().cpu().numpy() + sign * ref
                assert np.max(np.abs(Yi.double().cpu().numpy() - want)) <= \
                    tol * 10 * np.abs(want).max(), (mode, path)
        conv_path(None)


def test_fftconv_p_long_filters_take_the_fft_path(gpu, oracle, conv_path):
    """The reference's double-path synthesis convolves impulse responses with
    each other (python/tests/test_das_tfm.py:110-115): long filters, where
    the direct sum costs M_A M_B per sample; the automatic choice goes through
    the FFT path and agrees with the oracle."""
    from dream_amd.fftconv_p import fftconv_p
    rng = np.random.default_rng(92)
    H = np.asfortranarray(rng.standard_normal((1800, 16)))
    h = rng.standard_normal(1800)
    conv_path(None)
    Y = fftconv_p(H, h)
    ref = oracle.fftconv_p(H, h)
    assert np.max(np.abs(Y - ref)) <= 1e-11 * np.abs(ref).max()
