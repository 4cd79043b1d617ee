"""``Im = das(Y, Gt, Gr, Ro, dt, delay, cp, method, err_level, device, verbose)``
in float64 on the GPU: drop-in for the reference's Python module ``das``
(python/src/py_das.cc:66-90 with python/include/py_das.hpp:34-247).

method: "saft", "tfm", "rca" / "rca_coltx", "rca_rowtx"; err_level:
"ignore", "warn", "stop".  Y is K x n_ascans (Fortran order); the result is
a length-No float64 vector.  Sample indices are the reference's FP64 ones bit
for bit.
"""
import numpy as np

from . import _das_common, _lib

_handle = _das_common._Handle()


def das(Y, Gt, Gr, Ro, dt, delay, cp, method="tfm", err_level="stop", device="gpu",
        verbose=""):
    return _das_common.run(_handle, np.float64, _lib.F64, Y, Gt, Gr, Ro, dt, delay, cp, method,
                           err_level, device, verbose)
