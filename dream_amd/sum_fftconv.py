"""``Y = sum_fftconv(H, U)`` and in-place ``sum_fftconv(H, U, Y)`` on the GPU,
following the reference's Octave/MATLAB gateway (extra_src/oct_sum_fftconv.cc,
extra_src/mex_sum_fftconv.cc; the reference has no Python binding for it).

H is an M x N x L array and U a K x L matrix (Fortran order); Y(:,n) =
sum_l H(:,n,l) (*) U(:,l), an (M + K - 1) x N float64 matrix.  In in-place
mode the sum is added to the given Y (a Fortran-ordered float64 array of
that shape), as the reference's in-place mode does.
"""
import numpy as np

from . import _lib


def sum_fftconv(H, U, Y=None):
    H = np.asfortranarray(np.asarray(H, dtype=np.float64))
    if H.ndim != 3:
        raise RuntimeError("Argument 1 should be a 3D Matrix")
    U = np.asfortranarray(np.asarray(U, dtype=np.float64))
    if U.ndim == 1:
        U = U.reshape(-1, 1, order="F")
    M, N, L = H.shape
    if U.shape[1] != L:
        raise RuntimeError("3rd dimension of arg 1 must match the number of columns in arg 2")
    K = U.shape[0]
    if Y is None:
        out = np.empty((M + K - 1, N), dtype=np.float64, order="F")
        mode = _lib.CONV_EQU
    else:
        if Y.shape != (M + K - 1, N) or Y.dtype != np.float64 or not Y.flags.f_contiguous:
            raise RuntimeError("In-place mode needs a Fortran-ordered float64 Y of size "
                               "(M+K-1) x N")
        out = Y
        mode = _lib.CONV_SUM
    _lib.check(_lib.lib().dream_hip_sum_fftconv(out.ctypes.data, H.ctypes.data, M, N, L,
                                                U.ctypes.data, K, mode))
    return None if Y is not None else out
