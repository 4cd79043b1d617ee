"""Array helpers of the oracle (focusing, steering, apodization, aperture
extent) against the reference's own arr_functions.cc (golden vectors), plus
closed-form spot checks."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_helpers_match_reference(oracle):
    g = np.load(os.path.join(HERE, "golden", "arr_functions_ref.npz"))
    L = oracle.lib()
    gx, gy = g["gx"], g["gy"]
    G = np.c_[gx, gy, g["gz"]]
    assert tuple(oracle.max_dim_arr(G)) == tuple(g["dims"])
    xm, ym, rm = g["dims"]
    cp, focal = float(g["cp"]), float(g["focal"])
    for m in range(7):
        mine = [L.dro_focusing(m, focal, gx[i], gy[i], xm, ym, rm, cp) for i in range(gx.size)]
        assert np.array_equal(mine, g["foc"][m]), m
    for m in range(4):
        mine = [L.dro_beamsteering(m, float(g["theta"]), float(g["phi"]), gx[i], gy[i], xm, ym, cp)
                for i in range(gx.size)]
        assert np.array_equal(mine, g["steer"][m]), m
    av = np.ascontiguousarray(g["apod_vec"])
    p = av.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for m in range(8):
        mine = [L.dro_apodization(m, i, p, -99.0, gx[i], gy[i], rm, float(g["apod_par"]))
                for i in range(gx.size)]
        assert np.array_equal(mine, g["apod"][m]), m


def test_helper_closed_forms(oracle):
    L = oracle.lib()
    # hann: 1 at the centre, 0 at the aperture edge; triangle likewise.
    assert L.dro_apodization(3, 0, None, 0.0, 0.0, 0.0, 5.0, 0.0) == 1.0
    assert abs(L.dro_apodization(3, 0, None, 0.0, 5.0, 0.0, 5.0, 0.0)) < 1e-15
    assert L.dro_apodization(0, 0, None, 0.0, 0.0, 0.0, 5.0, 0.0) == 1.0
    # x focusing: the outermost element has zero delay, the centre the largest
    # (sqrt(x_max^2+F^2) - F) * 1000 / cp.
    assert L.dro_focusing(1, 20.0, 7.5, 0.0, 7.5, 0.0, 7.5, 1500.0) == 0.0
    assert abs(L.dro_focusing(1, 20.0, 0.0, 0.0, 7.5, 0.0, 7.5, 1500.0)
               - (np.hypot(7.5, 20.0) - 20.0) * 1000 / 1500) < 1e-12
    # x steering at theta: linear in gx, zero at gx = -x_max.
    assert abs(L.dro_beamsteering(1, 10.0, 0.0, -7.5, 0.0, 7.5, 0.0, 1500.0)) < 1e-15
