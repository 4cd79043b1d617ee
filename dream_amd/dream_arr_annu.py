"""``H = dream_arr_annu(Ro, G, s_par, delay, m_par, foc_met, focal, apod_met, apod,
win_par, err_level)`` on the GPU.

Drop-in for the reference's Python module ``dream_arr_annu``
(python/src/py_dream_arr_annu.cc:34-201, 209-317): same arguments, same
checks and messages, same output (an nt x No float64 array in Fortran
order), same error behaviour.  G is the vector of radii: the centre disc's,
then an (inner, outer) pair per ring; as in the reference, the radii are the
first ``G.shape[0]`` entries (a 1-D array counts as a column).  The focusing
delays ('ud') and apodization weights ('ud') are vectors of the same length
as G, of which one value per ring is used.  ``win_par`` is accepted and not
used (include/arg_parser_python.h:536-540).  Attenuation (alpha > 0) runs
the attenuated sum of src/attenuation.cc:72-140.
"""
import ctypes

import numpy as np

from . import _args, _lib

FUNC = "dream_arr_annu"


def dream_arr_annu(Ro, G, s_par, delay, m_par, foc_met, focal, apod_met, apod, win_par,
                   err_level):
    Ro = _args.obs_points(FUNC, Ro, np.float64)
    No = Ro.shape[0]
    Gr = _args.annu_radii(FUNC, G)
    nr = Gr.shape[0]
    dx, dy, dt, nt_f = _args.vector(FUNC, s_par, 4, "sampling par")
    nt = int(nt_f)
    d, delay_type = _args.delay(FUNC, delay, No, np.float64)
    v, cp, alpha = _args.vector(FUNC, m_par, 3, "material parameters par")
    foc, focal_v = _args.focus(FUNC, foc_met, focal, nr)
    do_apod, apod_m, apod_v, apod_par = _args.apodization(FUNC, apod_met, apod, nr)
    level = _args.err_level(FUNC, err_level)

    _lib.require_device()
    H = np.zeros((nt, No), dtype=np.float64, order="F")
    code = ctypes.c_int(0)
    Ro_c = np.asfortranarray(Ro)
    focal_c = np.ascontiguousarray(focal_v, dtype=np.float64)
    apod_c = np.ascontiguousarray(apod_v, dtype=np.float64)
    d_c = np.ascontiguousarray(d, dtype=np.float64)
    _lib.check(_lib.lib().dream_hip_arr_annu(
        float(alpha), Ro_c.ctypes.data, No, float(dx), float(dy), float(dt), nt, delay_type,
        d_c.ctypes.data, float(v), float(cp), nr, Gr.ctypes.data, foc, focal_c.ctypes.data,
        apod_c.ctypes.data, do_apod, apod_m, apod_par, H.ctypes.data, level, ctypes.byref(code)))
    if code.value == 3:
        print("SIR out-of-bounds!")
        raise RuntimeError(f"Error in {FUNC}!")
    if code.value == 2:
        print("Warning: SIR out-of-bounds!")
    return H
