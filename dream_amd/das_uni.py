"""``Im = das_uni(Y, gt, gr, Ro, dt, delay, cp, method)`` on the GPU: DAS over
implicit uniform grids, following the reference's Octave/MATLAB gateway
(extra_src/oct_das_uni.cc, extra_src/mex_das_uni.cc; the reference has no
Python binding for it) and DAS_UNI<T> (include/das_uni.h).

gt = [min_t, pitch_t, max_t], gr = [min_r, pitch_r, max_r] (ignored for
SAFT), Ro = 3 x 3 [min_Rx dx max_Rx; min_Ry dy max_Ry; min_Rz dz max_Rz] [mm];
Y is K x n_ascans.  float32 Y selects the reference's float arithmetic,
anything else float64.  Only the scalar (first) delay is used, as in
DAS_UNI<T>::cl_das_uni.  The result has one value per grid pixel, ny fastest,
then nz, then nx.
"""
import ctypes

import numpy as np

from . import _lib

_handles = {}


def _handle(prec, das_type, K, gt, gr, Ro):
    h = _handles.get(prec)
    L = _lib.lib()
    if h is not None and not L.dream_hip_das_uni_setup_has_changed(
            h, das_type, K, gt.ctypes.data_as(_lib._dp), gr.ctypes.data_as(_lib._dp),
            Ro.ctypes.data_as(_lib._dp), prec):
        return h
    if h is not None:
        L.dream_hip_das_uni_destroy(h)
        _handles.pop(prec)
    h = ctypes.c_void_p()
    _lib.check(L.dream_hip_das_uni_create(das_type, K, gt.ctypes.data_as(_lib._dp),
                                          gr.ctypes.data_as(_lib._dp),
                                          Ro.ctypes.data_as(_lib._dp), prec, ctypes.byref(h)))
    _handles[prec] = h
    return h


def das_uni(Y, gt, gr, Ro, dt, delay, cp, method="tfm"):
    use_float = np.asarray(Y).dtype == np.float32
    T = np.float32 if use_float else np.float64
    prec = _lib.F32_REF if use_float else _lib.F64
    Y = np.asfortranarray(np.asarray(Y, dtype=T))
    if Y.ndim == 1:
        Y = Y.reshape(-1, 1, order="F")
    gt = np.ascontiguousarray(np.asarray(gt, dtype=np.float64).ravel()[:3])
    gr = np.ascontiguousarray(np.asarray(gr, dtype=np.float64).ravel()[:3]) \
        if gr is not None and np.size(gr) >= 3 else gt.copy()
    Ro = np.asfortranarray(np.asarray(Ro, dtype=np.float64).reshape(3, 3))
    Ro = np.ascontiguousarray(Ro.ravel(order="F"))
    if gt.size != 3 or Ro.size != 9:
        raise RuntimeError("das_uni: gt must have 3 elements and Ro must be 3 x 3")
    das_type = _lib.DAS_TYPE[method]
    h = _handle(prec, das_type, Y.shape[0], gt, gr, Ro)
    No, Lt, Lr = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.lib().dream_hip_das_uni_shape(h, ctypes.byref(No), ctypes.byref(Lt),
                                                  ctypes.byref(Lr)))
    n_ascans = Lt.value if das_type == 0 else Lt.value * Lr.value
    if Y.shape[1] != n_ascans:
        raise RuntimeError(f"das_uni: Y must have {n_ascans} A-scans for this array")
    Im = np.empty(No.value, dtype=T)
    # The scalars are values of T (the float gateway reads them as float).
    d = float(T(np.asarray(delay, dtype=np.float64).ravel()[0]))
    _lib.check(_lib.lib().dream_hip_das_uni(h, Y.ctypes.data, float(T(dt)), d, float(T(cp)),
                                            Im.ctypes.data))
    return Im
