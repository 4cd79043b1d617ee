"""Argument checks of the reference's Python gateways (include/arg_parser_python.h).

The reference prints a message and raises ``RuntimeError`` when an argument
is malformed; these helpers do the same with the same messages.  Matrices are
taken column-major (Fortran order), as the reference's ``py::array_t<T,
f_style>`` arguments convert them.
"""
import numpy as np

from . import _lib


def fail(func, msg, exc_msg=None):
    print(f"{func} {msg}" if msg else func)
    raise RuntimeError(exc_msg or f"Error in {func}!")


def as_matrix(a, dtype):
    """2-D Fortran-ordered array of `dtype` (scalars and vectors become 1 x n)."""
    a = np.asarray(a, dtype=dtype)
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(1, -1)
    return np.asfortranarray(a)


def obs_points(func, Ro, dtype):
    # check_obs_points (include/arg_parser_python.h:83-97)
    Ro = as_matrix(Ro, dtype)
    if Ro.shape[1] != 3:
        fail(func, "requires that observation point arg must be a (number of observation "
                   "points) x 3 matrix!")
    return Ro


def grid(func, G, dtype):
    # check_array (:99-113)
    G = as_matrix(G, dtype)
    if G.shape[1] != 3:
        fail(func, "requires that array grid arg must be a (number of array elements) x 3 matrix!")
    return G


def annu_radii(func, G):
    """check_array_annu (include/arg_parser_python.h:115-127) and the radii the
    reference reads: the first G.shape[0] values of the column-major data
    (python/src/py_dream_arr_annu.cc:73-81).  A 1-D array is a column of radii."""
    a = np.asarray(G, dtype=np.float64)
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(-1, 1)
    m, n = a.shape[0], a.shape[1]
    if m != 1 and n != 1 and (n * n) % 2 == 0:
        fail(func, "requires that array grid arg must be a (number of radii) vector with an odd "
                   "number of elements!")
    return np.ascontiguousarray(np.asfortranarray(a).ravel(order="F")[:m])


def vector(func, a, n, what, dtype=np.float64):
    """An n-element row or column vector (parse_geometry/parse_sampling/parse_material)."""
    m = as_matrix(a, dtype)
    if not ((m.shape[0] == n and m.shape[1] == 1) or (m.shape[0] == 1 and m.shape[1] == n)):
        if n == 1:
            fail(func, f"requires that the {what} arg must be a scalar!")
        fail(func, f"requires that {what} arg must be a {n} element vector!")
    return m.ravel(order="F")


def delay(func, d, No, dtype):
    # check_delay (:257-275): a scalar, or one value per observation point.
    m = as_matrix(d, dtype)
    n = m.shape[0] * m.shape[1]
    if n > 1 and n != No:
        fail(func, f"requires that delay arg must be a {n} element vector (with a length equal "
                   "to the number of observation points) or a scalar!")
    flat = np.ascontiguousarray(m.ravel(order="F"))
    return flat, (0 if n == 1 else 1)


def err_level(func, s):
    # parse_error_arg (:277-310)
    if s not in _lib.ERR_LEVEL:
        fail(func, "Unknown error level!")
    return _lib.ERR_LEVEL[s]


def focus(func, foc_str, focal, L):
    # parse_focus_args (:316-390)
    if foc_str not in _lib.FOC_MET:
        fail(func, "Unknown focusing method!")
    met = _lib.FOC_MET[foc_str]
    f = as_matrix(focal, np.float64)
    n = f.shape[0] * f.shape[1]
    if foc_str == "ud":
        if n != L:
            fail(func, "the time delay vector arg for user defined ('ud') focusing\n!delays must "
                       "have the same length as the number of array elements!")
        return met, np.ascontiguousarray(f.ravel(order="F"))
    if n != 1:
        fail(func, "arg must be a scalar for non-user defined focusing!")
    return met, np.ascontiguousarray(f.ravel(order="F"))


def steer(func, steer_str, steer_par):
    # parse_steer_args (:394-442)
    if steer_str not in _lib.STEER_MET:
        fail(func, "Unknown beem steering method!")
    p = as_matrix(steer_par, np.float64)
    if not ((p.shape[0] == 2 and p.shape[1] == 1) or (p.shape[0] == 1 and p.shape[1] == 2)):
        fail(func, "steer_par must be a 2 element vector!")
    v = p.ravel(order="F")
    return _lib.STEER_MET[steer_str], float(v[0]), float(v[1])


def apodization(func, apod_str, apod, L):
    # parse_apod_args (:446-545).  The Python gateway never reads win_par, so
    # apod_par stays 0.0 (:536-540).
    if apod_str == "off":
        return 0, 0, np.ones(L), 0.0
    if apod_str not in _lib.APOD_MET:
        fail(func, "Unknown apodization method!")
    w = np.ones(L)
    if apod_str == "ud":
        a = as_matrix(apod, np.float64)
        if a.shape[0] * a.shape[1] != L:
            fail(func, "the length of apodization vector argument must be the same as the number "
                       "of array elements!")
        w = np.ascontiguousarray(a.ravel(order="F"))
    return 1, _lib.APOD_MET[apod_str], w, 0.0


def das_method(func, s):
    # parse_das_arg (:548-582)
    if s not in _lib.DAS_TYPE:
        fail(func, "Unknown delay-and-sum method!")
    return _lib.DAS_TYPE[s]
