"""``Im = das(Y, Gt, Gr, Ro, dt, delay, cp, method, err_level, device, verbose)``
in float32 on the GPU: drop-in for the reference's Python module ``das_f``
(python/src/py_das.cc:40-64, built with USE_FLOAT).

All inputs are taken as float32 and every delay is computed with the
reference's float arithmetic (constants are double expressions rounded to
float, extra_src/das.cc:316-318), so the sample indices equal the
reference's float ones bit for bit.  For FP32 data with FP64 delays, use
``dream_amd.torch_ops`` with precision "f32".
"""
import numpy as np

from . import _das_common, _lib

_handle = _das_common._Handle()


def das(Y, Gt, Gr, Ro, dt, delay, cp, method="tfm", err_level="stop", device="gpu",
        verbose=""):
    return _das_common.run(_handle, np.float32, _lib.F32_REF, Y, Gt, Gr, Ro, dt, delay, cp,
                           method, err_level, device, verbose)
