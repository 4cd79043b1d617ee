"""``H = dream_arr_circ(Ro, geom_par, G, s_par, delay, m_par, foc_met, focal,
steer_met, steer_par, apod_met, apod, win_par, err_level)`` on the GPU.

Drop-in for the reference's Python module ``dream_arr_circ``
(python/src/py_dream_arr_circ.cc:34-235, 237-366): same arguments, same
checks and messages, same output (an nt x No float64 array in Fortran
order), same error behaviour (``"stop"`` with an out-of-bounds sample prints
``SIR out-of-bounds!`` and raises ``RuntimeError``; ``"warn"`` prints a
warning).  As in the reference's Python gateway, ``win_par`` is accepted and
not used (include/arg_parser_python.h:536-540).  Attenuation (alpha > 0)
runs the attenuated sum of src/attenuation.cc:72-140.
"""
import ctypes

import numpy as np

from . import _args, _lib

FUNC = "dream_arr_circ"


def dream_arr_circ(Ro, geom_par, G, s_par, delay, m_par, foc_met, focal, steer_met, steer_par,
                   apod_met, apod, win_par, err_level):
    Ro = _args.obs_points(FUNC, Ro, np.float64)
    No = Ro.shape[0]
    (R,) = _args.vector(FUNC, geom_par, 1, "the geometry")
    G = _args.grid(FUNC, G, np.float64)
    L = G.shape[0]
    dx, dy, dt, nt_f = _args.vector(FUNC, s_par, 4, "sampling par")
    nt = int(nt_f)
    d, delay_type = _args.delay(FUNC, delay, No, np.float64)
    v, cp, alpha = _args.vector(FUNC, m_par, 3, "material parameters par")
    foc, focal_v = _args.focus(FUNC, foc_met, focal, L)
    steer, theta, phi = _args.steer(FUNC, steer_met, steer_par)
    do_apod, apod_m, apod_v, apod_par = _args.apodization(FUNC, apod_met, apod, L)
    level = _args.err_level(FUNC, err_level)

    _lib.require_device()
    H = np.zeros((nt, No), dtype=np.float64, order="F")
    code = ctypes.c_int(0)
    Ro_c = np.asfortranarray(Ro)
    G_c = np.asfortranarray(G)
    focal_c = np.ascontiguousarray(focal_v, dtype=np.float64)
    apod_c = np.ascontiguousarray(apod_v, dtype=np.float64)
    d_c = np.ascontiguousarray(d, dtype=np.float64)
    _lib.check(_lib.lib().dream_hip_arr_circ(
        float(alpha), Ro_c.ctypes.data, No, float(R), float(dx), float(dy), float(dt),
        nt, delay_type, d_c.ctypes.data, float(v), float(cp), L, G_c.ctypes.data, foc,
        focal_c.ctypes.data, steer, theta, phi, apod_c.ctypes.data, do_apod, apod_m, apod_par,
        H.ctypes.data, level, ctypes.byref(code)))
    if code.value == 3:
        print("SIR out-of-bounds!")
        raise RuntimeError(f"Error in {FUNC}!")
    if code.value == 2:
        print("Warning: SIR out-of-bounds!")
    return H
