"""``H = dreamcylind(Ro, geom_par, s_par, delay, m_par, err_level)`` on the GPU.

Drop-in for the reference's Python module ``dreamcylind``
(python/src/py_dreamcylind.cc:42-161, 169-239): the SIR of one cylindrically
curved element centred at the origin, geom_par = [a, b, Rcurv] (Rcurv > 0
focused/concave, < 0 defocused/convex); same checks, output (nt x No
float64, Fortran order) and error behaviour.  With b > 2|Rcurv| the
reference prints an error message per observation point and yields a zero
response; this module prints it once and yields the same zeros.
Attenuation (alpha > 0) runs the
attenuated sum of src/attenuation.cc:72-140.
"""
import ctypes

import numpy as np

from . import _args, _lib

FUNC = "dreamcylind"


def dreamcylind(Ro, geom_par, s_par, delay, m_par, err_level):
    Ro = _args.obs_points(FUNC, Ro, np.float64)
    No = Ro.shape[0]
    a, b, Rcurv = _args.vector(FUNC, geom_par, 3, "the geometry")
    dx, dy, dt, nt_f = _args.vector(FUNC, s_par, 4, "sampling par")
    nt = int(nt_f)
    d, delay_type = _args.delay(FUNC, delay, No, np.float64)
    v, cp, alpha = _args.vector(FUNC, m_par, 3, "material parameters par")
    level = _args.err_level(FUNC, err_level)

    _lib.require_device()
    if b > 2.0 * abs(Rcurv):
        print("Error in dreamcylind: the y-size, b, must be less than the curvature diameter "
              "2*Rcurv!\n")
    H = np.zeros((nt, No), dtype=np.float64, order="F")
    code = ctypes.c_int(0)
    Ro_c = np.asfortranarray(Ro)
    d_c = np.ascontiguousarray(d, dtype=np.float64)
    _lib.check(_lib.lib().dream_hip_cylind(
        float(alpha), Ro_c.ctypes.data, No, float(a), float(b), float(Rcurv), float(dx),
        float(dy), float(dt), nt, delay_type, d_c.ctypes.data, float(v), float(cp),
        H.ctypes.data, level, ctypes.byref(code)))
    if code.value == 3:
        print("SIR out-of-bounds!")
        raise RuntimeError(f"Error in {FUNC}!")
    if code.value == 2:
        print("Warning: SIR out-of-bounds!")
    return H
