"""``Y = fftconv_p(A, B)`` on the GPU: drop-in for the reference's Python
module ``fftconv_p`` (python/src/py_fftconv_p.cc, python/include/
py_fftconv_p.hpp:34-82).

Each column of A (M_A x N, Fortran order) is convolved with the matching
column of B, or with B when B is a vector; Y is (M_A + M_B - 1) x N in
float64.  The reference multiplies zero-padded FFTs; the GPU evaluates the
same linear convolution directly (``dream_hip_fftconv_p``), so results agree
to rounding.
"""
import numpy as np

from . import _lib


def _fortran(x):
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


def fftconv_p(A, B):
    A = _fortran(A)
    B = _fortran(B)
    A_M = A.shape[0]
    A_N = A.shape[1] if A.ndim == 2 else 1
    B_M = B.shape[0]
    B_N = B.shape[1] if B.ndim == 2 else 1
    if B_M != 1 and B_N != 1 and B_N != A_N:
        raise RuntimeError(
            "Argument 2 must be a vector or a matrix with the same number of columns as arg 1!")
    if B_M == 1 or B_N == 1:
        B_M, B_N = B_M * B_N, 1
    Y = np.empty((A_M + B_M - 1, A_N), dtype=np.float64, order="F")
    _lib.check(_lib.lib().dream_hip_fftconv_p(Y.ctypes.data, A.ctypes.data, A_M, A_N,
                                              B.ctypes.data, B_M, B_N, _lib.CONV_EQU))
    return Y
