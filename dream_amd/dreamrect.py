"""``H = dreamrect(Ro, geom_par, s_par, delay, m_par, err_level)`` on the GPU.

Drop-in for the reference's Python module ``dreamrect``
(python/src/py_dreamrect.cc:55-157, 230-237): the SIR of one rectangular
element centred at the origin; same checks, output (nt x No float64, Fortran
order) and error behaviour.  Attenuation (alpha > 0) runs the
attenuated sum of src/attenuation.cc:72-140.
"""
import ctypes

import numpy as np

from . import _args, _lib

FUNC = "dreamrect"


def dreamrect(Ro, geom_par, s_par, delay, m_par, err_level):
    Ro = _args.obs_points(FUNC, Ro, np.float64)
    No = Ro.shape[0]
    a, b = _args.vector(FUNC, geom_par, 2, "the geometry")
    dx, dy, dt, nt_f = _args.vector(FUNC, s_par, 4, "sampling par")
    nt = int(nt_f)
    d, delay_type = _args.delay(FUNC, delay, No, np.float64)
    v, cp, alpha = _args.vector(FUNC, m_par, 3, "material parameters par")
    level = _args.err_level(FUNC, err_level)

    _lib.require_device()
    H = np.zeros((nt, No), dtype=np.float64, order="F")
    code = ctypes.c_int(0)
    Ro_c = np.asfortranarray(Ro)
    d_c = np.ascontiguousarray(d, dtype=np.float64)
    _lib.check(_lib.lib().dream_hip_rect(
        float(alpha), Ro_c.ctypes.data, No, float(a), float(b), float(dx), float(dy), float(dt),
        nt, delay_type, d_c.ctypes.data, float(v), float(cp), H.ctypes.data, level,
        ctypes.byref(code)))
    if code.value == 3:
        print("SIR out-of-bounds!")
        raise RuntimeError(f"Error in {FUNC}!")
    if code.value == 2:
        print("Warning: SIR out-of-bounds!")
    return H
