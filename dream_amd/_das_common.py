"""Shared body of the ``das`` (float64) and ``das_f`` (float32) modules.

Mirrors ``py_das<T>`` (python/include/py_das.hpp:34-247): argument checks and
messages, a persistent DAS object per precision that is rebuilt only when
``das_setup_has_changed`` (include/das.h:77-94), and the error behaviour.
Both ``device="cpu"`` and ``device="gpu"`` run on the MI355X (this package
has no CPU path; the reference's Python gateway rejects "gpu",
python/include/py_das.hpp:146-148).  ``verbose`` is optional here.
"""
import ctypes

import numpy as np

from . import _args, _lib

FUNC = "das"


class _Handle:
    def __init__(self):
        self.h = ctypes.c_void_p(None)

    def ensure(self, das_type, K, No, Lt, Lr, prec):
        L = _lib.lib()
        if self.h.value is not None and not L.dream_hip_das_setup_has_changed(
                self.h, das_type, K, No, Lt, Lr, prec):
            return self.h
        self.release()
        _lib.check(L.dream_hip_das_create(das_type, K, No, Lt, Lr, prec, ctypes.byref(self.h)))
        return self.h

    def release(self):
        if self.h.value is not None and _lib._lib is not None:
            _lib._lib.dream_hip_das_destroy(self.h)
        self.h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def run(handle, dtype, prec, Y, Gt, Gr, Ro, dt, delay, cp, method, err_level, device, verbose):
    Y = np.asfortranarray(np.asarray(Y, dtype=dtype))
    if Y.ndim == 1:
        Y = Y.reshape(-1, 1, order="F")
    K = Y.shape[0]
    Gt = _args.as_matrix(Gt, dtype)
    if Gt.shape[1] != 3:
        raise RuntimeError("Argument 2 must be a (num transmit elements) x 3 matrix!")
    Lt = Gt.shape[0]
    das_type = _args.das_method(FUNC, method)
    Lr = 0
    Gr_c = None
    if das_type != 0:
        Gr_c = _args.as_matrix(Gr, dtype)
        if Gr_c.shape[1] != 3:
            raise RuntimeError("Argument 3 must be a (num recieve elements) x 3 matrix!")
        Lr = Gr_c.shape[0]
    Ro = _args.obs_points(FUNC, Ro, dtype)
    No = Ro.shape[0]
    d, delay_type = _args.delay(FUNC, delay, No, dtype)
    level = _args.err_level(FUNC, err_level)
    if device not in ("cpu", "gpu"):
        raise RuntimeError("Error in das - unkown compute device string!")
    n_ascans = Lt if das_type == 0 else Lt * Lr
    if Y.shape[1] != n_ascans:
        raise RuntimeError(f"das: Y has {Y.shape[1]} A-scans, the array needs {n_ascans}")

    _lib.require_device()
    h = handle.ensure(das_type, K, No, Lt, Lr, prec)
    Im = np.zeros(No, dtype=dtype)
    code = ctypes.c_int(0)
    scalar = (lambda x: float(np.float32(x))) if dtype == np.float32 else float
    _lib.check(_lib.lib().dream_hip_das(
        h, Y.ctypes.data, Ro.ctypes.data, Gt.ctypes.data,
        Gr_c.ctypes.data if Gr_c is not None else None, scalar(dt), delay_type,
        np.ascontiguousarray(d).ctypes.data, scalar(cp), Im.ctypes.data, level,
        ctypes.byref(code)))
    if code.value == 3:
        print("DAS out-of-bounds!")
        raise RuntimeError("DAS out-of-bounds!")
    if code.value == 2:
        print("Warning: DAS out-of-bounds!")
    return Im
