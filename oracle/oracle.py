"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker or the timed CPU baseline; the
product package ``dream_amd`` never imports it.

Two libraries are wrapped:

* ``oracle/libdream_oracle.so`` -- this project's C++ restatement of the
  reference's CPU paths (``oracle/dream_oracle.cc``).
* ``oracle/_ref/libdream_ref.so`` -- the reference's own ``das.cc`` and
  ``arr_functions.cc`` compiled from source (``make -C oracle ref``), present
  only where the reference tree was available at build time.

Array conventions follow the reference: column-major (Fortran) matrices, ``Ro``
is No x 3, ``G`` is L x 3, ``H`` is nt x No, ``Y`` is K x n_ascans.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdream_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdream_ref.so")

# Enum numbering of include/dream.h:42-102.
ERR_LEVEL = {"none": 0, "ignore": 1, "warn": 2, "stop": 3, "parallel_stop": 4}
FOC = {"off": 0, "x": 1, "y": 3, "xy": 5, "on": 5, "x+y": 6, "ud": 7}
STEER = {"off": 0, "x": 1, "y": 2, "xy": 3}
APOD = {"triangle": 0, "gauss": 1, "raised": 2, "hann": 3, "hamming": 4,
        "simply": 5, "clamped": 6, "ud": 7}
DAS_TYPE = {"saft": 0, "tfm": 1, "rca": 2, "rca_coltx": 2, "rca_rowtx": 3}

_c_i64 = ctypes.c_int64
_c_dp = ctypes.POINTER(ctypes.c_double)
_c_fp = ctypes.POINTER(ctypes.c_float)
_c_lp = ctypes.POINTER(ctypes.c_longlong)
_c_i64p = ctypes.POINTER(ctypes.c_int64)

_lib = None
_ref = None


def build():
    """Compile the oracle (and the reference checker when the tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "libdream_oracle.so"], check=True)
    if os.path.isdir(os.environ.get("DREAM_REFERENCE", "/root/reference")):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "libdream_oracle.so"], check=True)
        L = ctypes.CDLL(LIB_PATH)
        L.dro_subgrid.restype = _c_i64
        L.dro_subgrid.argtypes = [ctypes.c_double, ctypes.c_double, _c_dp, _c_i64]
        L.dro_max_dim_arr.argtypes = [_c_dp, _c_dp, _c_i64, _c_dp, _c_dp, _c_dp]
        L.dro_focusing.restype = ctypes.c_double
        L.dro_focusing.argtypes = [ctypes.c_int] + [ctypes.c_double] * 7
        L.dro_beamsteering.restype = ctypes.c_double
        L.dro_beamsteering.argtypes = [ctypes.c_int] + [ctypes.c_double] * 7
        L.dro_apodization.restype = ctypes.c_double
        L.dro_apodization.argtypes = [ctypes.c_int, _c_i64, _c_dp] + [ctypes.c_double] * 5
        L.dro_rect_contributions.restype = _c_i64
        L.dro_rect_contributions.argtypes = [ctypes.c_double] * 12 + [_c_i64p, _c_dp, _c_i64]
        L.dro_dreamrect.restype = ctypes.c_int
        L.dro_dreamrect.argtypes = [_c_dp, _c_i64] + [ctypes.c_double] * 5 + [
            _c_i64, ctypes.c_int, _c_dp, ctypes.c_double, ctypes.c_double, _c_dp,
            ctypes.c_int, ctypes.c_int, _c_lp]
        L.dro_dream_arr_rect.restype = ctypes.c_int
        L.dro_dream_arr_rect.argtypes = [_c_dp, _c_i64] + [ctypes.c_double] * 5 + [
            _c_i64, ctypes.c_int, _c_dp, ctypes.c_double, ctypes.c_double, _c_i64, _c_dp,
            ctypes.c_int, _c_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _c_dp,
            ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_dp, ctypes.c_int, ctypes.c_int,
            _c_lp]
        L.dro_dreamcirc.restype = ctypes.c_int
        L.dro_dreamcirc.argtypes = [_c_dp, _c_i64] + [ctypes.c_double] * 4 + [
            _c_i64, ctypes.c_int, _c_dp, ctypes.c_double, ctypes.c_double, _c_dp,
            ctypes.c_int, ctypes.c_int, _c_lp]
        L.dro_dream_arr_circ.restype = ctypes.c_int
        L.dro_dream_arr_circ.argtypes = [_c_dp, _c_i64] + [ctypes.c_double] * 4 + [
            _c_i64, ctypes.c_int, _c_dp, ctypes.c_double, ctypes.c_double, _c_i64, _c_dp,
            ctypes.c_int, _c_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _c_dp,
            ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_dp, ctypes.c_int, ctypes.c_int,
            _c_lp]
        L.dro_dreamcylind.restype = ctypes.c_int
        L.dro_dreamcylind.argtypes = [_c_dp, _c_i64] + [ctypes.c_double] * 6 + [
            _c_i64, ctypes.c_int, _c_dp, ctypes.c_double, ctypes.c_double, _c_dp,
            ctypes.c_int, ctypes.c_int, _c_lp]
        L.dro_dream_arr_cylind.restype = ctypes.c_int
        L.dro_dream_arr_cylind.argtypes = [_c_dp, _c_i64] + [ctypes.c_double] * 6 + [
            _c_i64, ctypes.c_int, _c_dp, ctypes.c_double, ctypes.c_double, _c_i64, _c_dp,
            ctypes.c_int, _c_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _c_dp,
            ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_dp, ctypes.c_int, ctypes.c_int,
            _c_lp]
        L.dro_set_alpha.argtypes = [ctypes.c_double]
        L.dro_set_alpha.restype = None
        L.dro_dream_arr_annu.restype = ctypes.c_int
        L.dro_dream_arr_annu.argtypes = [_c_dp, _c_i64] + [ctypes.c_double] * 3 + [
            _c_i64, ctypes.c_int, _c_dp, ctypes.c_double, ctypes.c_double, _c_i64, _c_dp,
            ctypes.c_int, _c_dp, _c_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_dp,
            ctypes.c_int, ctypes.c_int, _c_lp]
        L.dro_annu_ring_radii.argtypes = [_c_dp, _c_i64, _c_dp]
        L.dro_annu_focusing.restype = ctypes.c_double
        L.dro_annu_focusing.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4
        L.dro_annu_apodization.restype = ctypes.c_double
        L.dro_annu_apodization.argtypes = [ctypes.c_int, _c_i64, _c_dp] + [ctypes.c_double] * 3
        for name, ptr, scal in (("dro_das_f64", _c_dp, ctypes.c_double),
                                ("dro_das_f32", _c_fp, ctypes.c_float)):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, ptr, _c_i64, ptr, _c_i64, ptr, _c_i64, ptr, _c_i64,
                          scal, ctypes.c_int, ptr, scal, ptr, ctypes.c_int, ctypes.c_int]
        L.dro_das_ext_f64.restype = ctypes.c_int
        L.dro_das_ext_f64.argtypes = [ctypes.c_int, ctypes.c_int, _c_dp, _c_i64, _c_dp, _c_i64,
                                      _c_dp, _c_i64, _c_dp, _c_i64, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _c_dp, _c_dp, _c_dp,
                                      ctypes.c_int]
        L.dro_das_ext_adj_f64.restype = None
        L.dro_das_ext_adj_f64.argtypes = [ctypes.c_int, ctypes.c_int, _c_dp, _c_i64, _c_dp,
                                          _c_i64, _c_dp, _c_i64, _c_dp, _c_i64,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          _c_dp, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp,
                                          ctypes.c_int]
        _lib = L
    return _lib


def have_ref():
    return os.path.exists(REF_PATH)


def ref():
    """The reference-backed checker (das.cc + arr_functions.cc from source)."""
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        for name, ptr, scal in (("ref_das_f64", _c_dp, ctypes.c_double),
                                ("ref_das_f32", _c_fp, ctypes.c_float)):
            f = getattr(R, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, ptr, _c_i64, ptr, _c_i64, ptr, _c_i64, ptr, _c_i64,
                          scal, ctypes.c_int, ptr, scal, ptr, ctypes.c_int]
        R.ref_max_dim_arr.argtypes = [_c_dp, _c_dp, _c_dp, _c_i64, _c_dp, _c_dp, _c_dp]
        R.ref_focusing.restype = ctypes.c_double
        R.ref_focusing.argtypes = [ctypes.c_int] + [ctypes.c_double] * 7
        R.ref_beamsteering.restype = ctypes.c_double
        R.ref_beamsteering.argtypes = [ctypes.c_int] + [ctypes.c_double] * 8
        R.ref_apodization.restype = ctypes.c_double
        R.ref_apodization.argtypes = [ctypes.c_int, _c_i64, _c_dp] + [ctypes.c_double] * 5
        _ref = R
    return _ref


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _f32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _p(a, ctype=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctype))


# ----------------------------------------------------------------------------- helpers

def subgrid(w, d):
    """Sub-element centres of one rectangle side (src/dreamrect.cc:286-331)."""
    n = lib().dro_subgrid(float(w), float(d), None, 0)
    out = np.empty(n, dtype=np.float64)
    lib().dro_subgrid(float(w), float(d), _p(out), n)
    return out


def max_dim_arr(G):
    G = _f64(G)
    xm, ym, rm = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().dro_max_dim_arr(_p(G[:, 0].copy()), _p(G[:, 1].copy()), G.shape[0],
                          ctypes.byref(xm), ctypes.byref(ym), ctypes.byref(rm))
    return xm.value, ym.value, rm.value


def rect_contributions(xo, yo, zo, a, b, dx, dy, dt, delay, v, cp, weight=1.0):
    """Per-contribution (sample index, amplitude) of one rectangle sum."""
    L = lib()
    args = [float(x) for x in (xo, yo, zo, a, b, dx, dy, dt, delay, v, cp, weight)]
    n = L.dro_rect_contributions(*args, None, None, 0)
    it = np.empty(n, dtype=np.int64)
    ai = np.empty(n, dtype=np.float64)
    L.dro_rect_contributions(*args, it.ctypes.data_as(_c_i64p), _p(ai), n)
    return it, ai


# ----------------------------------------------------------------------------- SIR


class _alpha:
    """Sets the oracle's attenuation coefficient [dB/(cm MHz)] for one call."""

    def __init__(self, alpha):
        self.alpha = float(alpha)

    def __enter__(self):
        lib().dro_set_alpha(self.alpha)

    def __exit__(self, *exc):
        lib().dro_set_alpha(0.0)


def dreamrect(Ro, a, b, dx, dy, dt, nt, delay, v, cp, err_level="stop", nthreads=0, alpha=0.0):
    """Rect::dreamrect (no attenuation).  Returns (H nt x No, err code, oob count)."""
    Ro = _f64(np.atleast_2d(Ro))
    No = Ro.shape[0]
    delay = _f64(np.atleast_1d(np.asarray(delay, dtype=np.float64).ravel()))
    dtype = 0 if delay.size == 1 else 1
    H = np.zeros((int(nt), No), dtype=np.float64, order="F")
    oob = ctypes.c_longlong(0)
    with _alpha(alpha):
        err = lib().dro_dreamrect(_p(Ro), No, a, b, dx, dy, dt, int(nt), dtype, _p(delay), v, cp,
                                  _p(H), ERR_LEVEL[err_level], int(nthreads), ctypes.byref(oob))
    return H, err, oob.value


def dream_arr_rect(Ro, a, b, dx, dy, dt, nt, delay, v, cp, G, foc_met="off", focal=0.0,
                   steer_met="off", theta=0.0, phi=0.0, apod_met="off", apod=None,
                   apod_par=0.0, err_level="stop", nthreads=0, alpha=0.0):
    """ArrRect::dream_arr_rect (no attenuation).  Returns (H, err code, oob count)."""
    Ro = _f64(np.atleast_2d(Ro))
    G = _f64(np.atleast_2d(G))
    No, L = Ro.shape[0], G.shape[0]
    delay = _f64(np.atleast_1d(np.asarray(delay, dtype=np.float64).ravel()))
    dtype = 0 if delay.size == 1 else 1
    focal = _f64(np.atleast_1d(np.asarray(focal, dtype=np.float64).ravel()))
    if focal.size < L:
        focal = _f64(np.resize(focal, L))
    do_apod = apod_met != "off"
    apod_v = _f64(np.ones(L) if apod is None else np.asarray(apod, dtype=np.float64).ravel())
    H = np.zeros((int(nt), No), dtype=np.float64, order="F")
    oob = ctypes.c_longlong(0)
    with _alpha(alpha):
        err = lib().dro_dream_arr_rect(
            _p(Ro), No, a, b, dx, dy, dt, int(nt), dtype, _p(delay), v, cp, L, _p(G),
            FOC[foc_met], _p(focal), STEER[steer_met], theta, phi, _p(apod_v), int(do_apod),
            APOD.get(apod_met, 0), apod_par, _p(H), ERR_LEVEL[err_level], int(nthreads),
            ctypes.byref(oob))
    return H, err, oob.value


def dreamcirc(Ro, R, dx, dy, dt, nt, delay, v, cp, err_level="stop", nthreads=0, alpha=0.0):
    """Circ::dreamcirc (no attenuation).  Returns (H, err code, oob count)."""
    Ro = _f64(np.atleast_2d(Ro))
    No = Ro.shape[0]
    delay = _f64(np.atleast_1d(np.asarray(delay, dtype=np.float64).ravel()))
    H = np.zeros((int(nt), No), dtype=np.float64, order="F")
    oob = ctypes.c_longlong(0)
    with _alpha(alpha):
        err = lib().dro_dreamcirc(_p(Ro), No, R, dx, dy, dt, int(nt), 0 if delay.size == 1 else 1,
                                  _p(delay), v, cp, _p(H), ERR_LEVEL[err_level], int(nthreads),
                                  ctypes.byref(oob))
    return H, err, oob.value


def dream_arr_circ(Ro, R, dx, dy, dt, nt, delay, v, cp, G, foc_met="off", focal=0.0,
                   steer_met="off", theta=0.0, phi=0.0, apod_met="off", apod=None,
                   apod_par=0.0, err_level="stop", nthreads=0, alpha=0.0):
    """ArrCirc::dream_arr_circ (no attenuation).  Returns (H, err code, oob count)."""
    Ro = _f64(np.atleast_2d(Ro))
    G = _f64(np.atleast_2d(G))
    No, L = Ro.shape[0], G.shape[0]
    delay = _f64(np.atleast_1d(np.asarray(delay, dtype=np.float64).ravel()))
    focal = _f64(np.atleast_1d(np.asarray(focal, dtype=np.float64).ravel()))
    if focal.size < L:
        focal = _f64(np.resize(focal, L))
    do_apod = apod_met != "off"
    apod_v = _f64(np.ones(L) if apod is None else np.asarray(apod, dtype=np.float64).ravel())
    H = np.zeros((int(nt), No), dtype=np.float64, order="F")
    oob = ctypes.c_longlong(0)
    with _alpha(alpha):
        err = lib().dro_dream_arr_circ(
            _p(Ro), No, R, dx, dy, dt, int(nt), 0 if delay.size == 1 else 1, _p(delay), v, cp, L,
            _p(G), FOC[foc_met], _p(focal), STEER[steer_met], theta, phi, _p(apod_v), int(do_apod),
            APOD.get(apod_met, 0), apod_par, _p(H), ERR_LEVEL[err_level], int(nthreads),
            ctypes.byref(oob))
    return H, err, oob.value


def dreamcylind(Ro, a, b, Rcurv, dx, dy, dt, nt, delay, v, cp, err_level="stop", nthreads=0, alpha=0.0):
    """Cylind::dreamcylind (no attenuation).  Returns (H, err code, oob count)."""
    Ro = _f64(np.atleast_2d(Ro))
    No = Ro.shape[0]
    delay = _f64(np.atleast_1d(np.asarray(delay, dtype=np.float64).ravel()))
    H = np.zeros((int(nt), No), dtype=np.float64, order="F")
    oob = ctypes.c_longlong(0)
    with _alpha(alpha):
        err = lib().dro_dreamcylind(_p(Ro), No, a, b, Rcurv, dx, dy, dt, int(nt),
                                    0 if delay.size == 1 else 1, _p(delay), v, cp, _p(H),
                                    ERR_LEVEL[err_level], int(nthreads), ctypes.byref(oob))
    return H, err, oob.value


def dream_arr_cylind(Ro, a, b, Rcurv, dx, dy, dt, nt, delay, v, cp, G, foc_met="off",
                     focal=0.0, steer_met="off", theta=0.0, phi=0.0, apod_met="off", apod=None,
                     apod_par=0.0, err_level="stop", nthreads=0, alpha=0.0):
    """ArrCylind::dream_arr_cylind (no attenuation).  Returns (H, err code, oob count)."""
    Ro = _f64(np.atleast_2d(Ro))
    G = _f64(np.atleast_2d(G))
    No, L = Ro.shape[0], G.shape[0]
    delay = _f64(np.atleast_1d(np.asarray(delay, dtype=np.float64).ravel()))
    focal = _f64(np.atleast_1d(np.asarray(focal, dtype=np.float64).ravel()))
    if focal.size < L:
        focal = _f64(np.resize(focal, L))
    do_apod = apod_met != "off"
    apod_v = _f64(np.ones(L) if apod is None else np.asarray(apod, dtype=np.float64).ravel())
    H = np.zeros((int(nt), No), dtype=np.float64, order="F")
    oob = ctypes.c_longlong(0)
    with _alpha(alpha):
        err = lib().dro_dream_arr_cylind(
            _p(Ro), No, a, b, Rcurv, dx, dy, dt, int(nt), 0 if delay.size == 1 else 1, _p(delay), v,
            cp, L, _p(G), FOC[foc_met], _p(focal), STEER[steer_met], theta, phi, _p(apod_v),
            int(do_apod), APOD.get(apod_met, 0), apod_par, _p(H), ERR_LEVEL[err_level],
            int(nthreads), ctypes.byref(oob))
    return H, err, oob.value


def dream_arr_annu(Ro, Gr, dx, dy, dt, nt, delay, v, cp, foc_met="off", focal=0.0,
                   apod_met="off", apod=None, apod_par=0.0, err_level="stop", nthreads=0, alpha=0.0):
    """ArrAnnu::dream_arr_annu (no attenuation).  Gr is the vector of radii (centre
    disc, then inner/outer per ring).  Returns (H, err code, oob count)."""
    Ro = _f64(np.atleast_2d(Ro))
    Gr = _f64(np.asarray(Gr, dtype=np.float64).ravel())
    No, nr = Ro.shape[0], Gr.size
    delay = _f64(np.atleast_1d(np.asarray(delay, dtype=np.float64).ravel()))
    focal = _f64(np.atleast_1d(np.asarray(focal, dtype=np.float64).ravel()))
    if focal.size < nr:
        focal = _f64(np.resize(focal, max(nr, 1)))
    do_apod = apod_met != "off"
    apod_v = _f64(np.ones(max(nr, 1)) if apod is None
                  else np.asarray(apod, dtype=np.float64).ravel())
    H = np.zeros((int(nt), No), dtype=np.float64, order="F")
    oob = ctypes.c_longlong(0)
    with _alpha(alpha):
        err = lib().dro_dream_arr_annu(
            _p(Ro), No, dx, dy, dt, int(nt), 0 if delay.size == 1 else 1, _p(delay), v, cp, nr,
            _p(Gr), FOC[foc_met], _p(focal), _p(apod_v), int(do_apod), APOD.get(apod_met, 0),
            apod_par, _p(H), ERR_LEVEL[err_level], int(nthreads), ctypes.byref(oob))
    return H, err, oob.value


def annu_ring_radii(Gr):
    Gr = _f64(np.asarray(Gr, dtype=np.float64).ravel())
    out = np.zeros((Gr.size + 1) // 2)
    lib().dro_annu_ring_radii(_p(Gr), Gr.size, _p(out))
    return out


def annu_focusing(foc_met, focal, ring_r, ring_r_max, cp):
    return lib().dro_annu_focusing(FOC[foc_met], focal, ring_r, ring_r_max, cp)


def annu_apodization(apod_met, n, apod, ring_r, ring_r_max, apod_par=0.0):
    apod = _f64(np.atleast_1d(np.asarray(apod if apod is not None else [1.0], dtype=np.float64)))
    return lib().dro_annu_apodization(APOD[apod_met], n, _p(apod), ring_r, ring_r_max, apod_par)


# ----------------------------------------------------------------------------- DAS

def _das_args(Y, Gt, Gr, Ro, delay, conv):
    Y = conv(Y)
    Gt = conv(np.atleast_2d(Gt))
    Gr = conv(np.atleast_2d(Gr)) if Gr is not None and np.size(Gr) else None
    Ro = conv(np.atleast_2d(Ro))
    delay = conv(np.atleast_1d(np.asarray(delay).ravel()))
    return Y, Gt, Gr, Ro, delay


def _rca_shapes_ok(method, Lt, Lr):
    """The reference reads the RCA stripe extents at rows[Lt], rows[2 Lt - 1]
    and columns[Lr - 1] (extra_src/das.cc:464-467); outside 2 Lt <= 3 Lr and
    Lr <= 3 Lt those reads leave the element tables (undefined behaviour)."""
    if DAS_TYPE[method] >= 2 and Lr > 0 and not (2 * Lt <= 3 * Lr and Lr <= 3 * Lt):
        raise ValueError("RCA stripe extents out of the element tables")


def das(Y, Gt, Gr, Ro, dt, delay, cp, method="tfm", err_level="ignore", precision="f64",
        nthreads=0, use_ref=False):
    """DAS<T>::das.  precision 'f64' or 'f32' selects T.  Returns (Im, err code).

    use_ref=True runs the reference's own das.cc (oracle/_ref) instead of the
    restatement.
    """
    f32 = precision == "f32"
    conv, ptr = (_f32, ctypes.c_float) if f32 else (_f64, ctypes.c_double)
    Y, Gt, Gr, Ro, delay = _das_args(Y, Gt, Gr, Ro, delay, conv)
    K = Y.shape[0]
    Lt = Gt.shape[0]
    Lr = 0 if Gr is None or DAS_TYPE[method] == 0 else Gr.shape[0]
    No = Ro.shape[0]
    _rca_shapes_ok(method, Lt, Lr)
    dtype = 0 if delay.size == 1 else 1
    Im = np.zeros(No, dtype=np.float32 if f32 else np.float64)
    gr_p = _p(Gr, ptr) if Gr is not None else None
    if use_ref:
        fn = ref().ref_das_f32 if f32 else ref().ref_das_f64
        err = fn(DAS_TYPE[method], _p(Y, ptr), K, _p(Ro, ptr), No, _p(Gt, ptr), Lt, gr_p, Lr,
                 dt, dtype, _p(delay, ptr), cp, _p(Im, ptr), ERR_LEVEL[err_level])
    else:
        fn = lib().dro_das_f32 if f32 else lib().dro_das_f64
        err = fn(DAS_TYPE[method], _p(Y, ptr), K, _p(Ro, ptr), No, _p(Gt, ptr), Lt, gr_p, Lr,
                 dt, dtype, _p(delay, ptr), cp, _p(Im, ptr), ERR_LEVEL[err_level],
                 int(nthreads))
    return Im, err


def das_ext(Y, Gt, Gr, Ro, dt, delay, cp, method="saft", interp="linear", apod_t=None,
            apod_r=None, nthreads=0):
    """Extension forward (FP64): linear interpolation + apodization.

    Returns (Im, any_oob).
    """
    Y, Gt, Gr, Ro, _ = _das_args(Y, Gt, Gr, Ro, [0.0], _f64)
    K, Lt, No = Y.shape[0], Gt.shape[0], Ro.shape[0]
    Lr = 0 if Gr is None or DAS_TYPE[method] == 0 else Gr.shape[0]
    _rca_shapes_ok(method, Lt, Lr)
    at = None if apod_t is None else _f64(np.asarray(apod_t).ravel())
    ar = None if apod_r is None else _f64(np.asarray(apod_r).ravel())
    Im = np.zeros(No, dtype=np.float64)
    oob = lib().dro_das_ext_f64(DAS_TYPE[method], 1 if interp == "linear" else 0, _p(Y), K,
                                _p(Ro), No, _p(Gt), Lt, _p(Gr), Lr, float(dt), float(delay),
                                float(cp), _p(at), _p(ar), _p(Im), int(nthreads))
    return Im, oob


def das_ext_adjoint(Y, Gt, Gr, Ro, dt, delay, cp, g, method="saft", interp="linear",
                    apod_t=None, apod_r=None, want_gy=True, nthreads=0):
    """Adjoint of das_ext for cotangent g.  Returns (gY, g_apod_t, g_apod_r, g_cp)."""
    Y, Gt, Gr, Ro, _ = _das_args(Y, Gt, Gr, Ro, [0.0], _f64)
    K, Lt, No = Y.shape[0], Gt.shape[0], Ro.shape[0]
    Lr = 0 if Gr is None or DAS_TYPE[method] == 0 else Gr.shape[0]
    _rca_shapes_ok(method, Lt, Lr)
    at = None if apod_t is None else _f64(np.asarray(apod_t).ravel())
    ar = None if apod_r is None else _f64(np.asarray(apod_r).ravel())
    g = _f64(np.asarray(g).ravel())
    n_ascans = Lt if Lr == 0 else Lt * Lr
    gY = np.zeros((K, n_ascans), dtype=np.float64, order="F") if want_gy else None
    gat = np.zeros(Lt)
    gar = np.zeros(max(Lr, 1))
    gcp = ctypes.c_double(0.0)
    lib().dro_das_ext_adj_f64(DAS_TYPE[method], 1 if interp == "linear" else 0, _p(Y), K,
                              _p(Ro), No, _p(Gt), Lt, _p(Gr), Lr, float(dt), float(delay),
                              float(cp), _p(at), _p(ar), _p(g), _p(gY), _p(gat), _p(gar),
                              ctypes.byref(gcp), int(nthreads))
    return gY, gat, (gar[:Lr] if Lr else None), gcp.value


# ----------------------------------------------------------------- convolution

def fftconv_p(A, B, Y=None, mode="="):
    """Column-wise linear convolution by zero-padded FFTs of length
    M_A + M_B - 1, the algorithm of fftconv (extra_src/fftconv.cc:32-106)
    driven by FFTConvP::run (extra_src/fftconv_p.cc:154-403): a vector B is
    one filter for every column of A.  mode '=', '+=' or '-=' as ConvMode."""
    A = np.asarray(A, dtype=np.float64)
    A = A.reshape(A.shape[0], -1)
    B = np.asarray(B, dtype=np.float64)
    vec = B.ndim == 1 or B.shape[0] == 1 or B.shape[1] == 1
    B = B.reshape(-1, 1) if vec else B
    n = A.shape[0] + B.shape[0] - 1
    Z = np.fft.irfft(np.fft.rfft(A, n, axis=0) * np.fft.rfft(B, n, axis=0), n, axis=0)
    if Y is None or mode == "=":
        return Z
    return Y + Z if mode == "+=" else Y - Z


def sum_fftconv(H, U):
    """sum_l H(:,n,l) (*) U(:,l) by FFTs, accumulated in the frequency domain
    as add_fftconv does (extra_src/fftconv.cc:119-209)."""
    H = np.asarray(H, dtype=np.float64)
    U = np.asarray(U, dtype=np.float64).reshape(U.shape[0], -1)
    M, N, L = H.shape
    n = M + U.shape[0] - 1
    F = np.zeros((n // 2 + 1, N), dtype=np.complex128)
    for l in range(L):
        F += np.fft.rfft(H[:, :, l], n, axis=0) * np.fft.rfft(U[:, l], n)[:, None]
    return np.fft.irfft(F, n, axis=0)


# ----------------------------------------------------------------- das_uni

def uni_count(mn, pitch, mx, dtype=np.float64):
    """(max - min)/pitch + 1 truncated, with the subtraction and division in T
    (include/das_uni.h:70-78)."""
    T = dtype
    return int(float((T(mx) - T(mn)) / T(pitch)) + 1.0)


def uni_pixels(Ro3, dtype=np.float64):
    """Pixels of das_uni's grid in its linear order: ny fastest, then nz, then
    nx; x = min + ((T) n) * d (opencl/das_uni_double.cl:32-47)."""
    T = dtype
    R = np.asarray(Ro3, dtype=np.float64).reshape(3, 3, order="F")
    N = [uni_count(R[c, 0], R[c, 1], R[c, 2], T) for c in range(3)]
    no = np.arange(N[0] * N[1] * N[2])
    ny = no % N[1]
    nz = (no // N[1]) % N[2]
    nx = (no // N[1] // N[2]) % N[0]
    x = T(R[0, 0]) + nx.astype(T) * T(R[0, 1])
    y = T(R[1, 0]) + ny.astype(T) * T(R[1, 1])
    z = T(R[2, 0]) + nz.astype(T) * T(R[2, 1])
    return np.asarray(x, T), np.asarray(y, T), np.asarray(z, T)


def das_uni(Y, gt, gr, Ro3, dt, delay, cp, method="tfm", precision="f64"):
    """Restatement of the das_uni kernels (opencl/das_uni_double.cl:9-394) with
    IEEE sqrt in place of native_sqrt, vectorised over the pixels.  Y is
    K x n_ascans (Fortran order); returns Im (No,)."""
    T = np.float64 if precision == "f64" else np.float32
    Y = np.asarray(Y, dtype=T)
    K = Y.shape[0]
    xo, yo, zo = uni_pixels(Ro3, T)
    # The scalars arrive as T (extra_src/oct_das_uni.cc:300-340); the
    # constants are double expressions of them rounded to T.
    dt, delay, cp = float(T(dt)), float(T(delay)), float(T(cp))
    Fs = T((1.0 / dt) * 1000.0)
    icp = T(1.0 / cp)
    dms = T(delay / 1000.0)
    gt = [float(v) for v in gt]
    gr = gt if method == "saft" else [float(v) for v in gr]
    Lt = uni_count(*gt, dtype=T)
    Lr = uni_count(*gr, dtype=T)
    et = T(gt[0]) + np.arange(Lt).astype(T) * T(gt[1])
    er = T(gr[0]) + np.arange(Lr).astype(T) * T(gr[1])
    im = np.zeros(xo.shape, dtype=T)

    def gather(a, t_dp):
        k = (t_dp * Fs).astype(np.int64)                  # (int) truncation
        ok = (k >= 0) & (k < K)
        return np.where(ok, Y[np.clip(k, 0, K - 1), a], T(0))

    def leg(gx, gy, gz):
        return np.sqrt((gx * gx + gy * gy) + gz * gz) * icp

    def stripe(p, lo, hi):
        lo, hi = T(lo), T(hi)
        return np.where((p >= lo) & (p <= hi), T(0), np.where(p < lo, lo - p, hi - p)).astype(T)

    if method == "saft":
        for n in range(Lt):
            t = leg(et[n] - xo, yo, zo) + dms
            im = im + gather(n, T(2.0) * t)
    elif method == "tfm":
        for i in range(Lt):
            tt = leg(et[i] - xo, yo, zo) + dms
            for j in range(Lr):
                im = im + gather(i * Lr + j, tt + leg(er[j] - xo, yo, zo))
    elif method in ("rca", "rca_coltx"):
        gy_c = stripe(yo, gr[0], gr[2])                   # rows' [min, max] (:197-200)
        gx_r = stripe(xo, gt[0], gt[2])
        for c in range(Lt):
            tc = leg(et[c] - xo, gy_c, zo) + dms
            for r in range(Lr):
                im = im + gather(c * Lr + r, tc + leg(gx_r, er[r] - yo, zo))
    else:                                                 # rca_rowtx
        gy_c = stripe(yo, gr[0], gr[2])
        gx_r = stripe(xo, gt[0], gt[2])
        for r in range(Lr):
            tr = leg(gx_r, er[r] - yo, zo) + dms          # delay on the row leg (:344-346)
            for c in range(Lt):
                im = im + gather(r * Lt + c, leg(et[c] - xo, gy_c, zo) + tr)
    return im
