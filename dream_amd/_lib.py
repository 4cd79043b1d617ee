"""ctypes binding of libdream_hip.so (the C ABI declared in include/dream_hip.h).

The library is built in-tree by ``python -m dream_amd.build``.  There is no
CPU fallback: if the library is missing or no HIP device is visible the
calls raise.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdream_hip.so")

OK = 0
F64, F32_REF, F32 = 0, 1, 2
NEAREST, LINEAR = 0, 1
CONV_EQU, CONV_SUM, CONV_NEG = 0, 1, 2

# Enum numbering of the reference (include/dream.h:42-102).
ERR_LEVEL = {"ignore": 1, "warn": 2, "stop": 3}
SIR_ERROR = {0: "none", 1: "ignore_out_of_bounds", 2: "warn_out_of_bounds", 3: "out_of_bounds"}
FOC_MET = {"off": 0, "x": 1, "y": 3, "xy": 5, "on": 5, "x+y": 6, "ud": 7}
STEER_MET = {"off": 0, "x": 1, "y": 2, "xy": 3}
APOD_MET = {"triangle": 0, "gauss": 1, "raised": 2, "hann": 3, "hamming": 4,
            "simply": 5, "clamped": 6, "ud": 7}
DAS_TYPE = {"saft": 0, "tfm": 1, "rca": 2, "rca_coltx": 2, "rca_rowtx": 3}

# Every symbol include/dream_hip.h declares.
EXPORTS = (
    "dream_hip_last_error", "dream_hip_version", "dream_hip_device_count",
    "dream_hip_set_autotune",
    "dream_hip_abort", "dream_hip_set_running", "dream_hip_is_running",
    "dream_hip_test_chunking",
    "dream_hip_arr_rect", "dream_hip_rect", "dream_hip_arr_rect_dev",
    "dream_hip_rect_indices", "dream_hip_rsq_seed",
    "dream_hip_sir_att_dev",
    "dream_hip_sir_plan_create", "dream_hip_sir_plan_run",
    "dream_hip_sir_plan_status", "dream_hip_sir_plan_destroy",
    "dream_hip_das_create", "dream_hip_das_destroy", "dream_hip_das_setup_has_changed",
    "dream_hip_das_set_pixel_order", "dream_hip_pixel_order",
    "dream_hip_das", "dream_hip_das_dev", "dream_hip_das_status", "dream_hip_das_adjoint_dev",
    "dream_hip_das_adjoint_dev_noalloc", "dream_hip_das_time_kernels",
    "dream_hip_das_kernel_times",
    "dream_hip_das_batch_create", "dream_hip_das_batch_run", "dream_hip_das_batch_destroy",
    "dream_hip_fftconv_p", "dream_hip_sum_fftconv", "dream_hip_conv_sum_dev",
    "dream_hip_das_uni_create", "dream_hip_das_uni_destroy", "dream_hip_das_uni_shape",
    "dream_hip_das_uni_setup_has_changed", "dream_hip_das_uni", "dream_hip_das_uni_dev",
    "dream_hip_arr_circ", "dream_hip_circ", "dream_hip_arr_circ_dev",
    "dream_hip_arr_annu", "dream_hip_arr_annu_dev",
    "dream_hip_arr_cylind", "dream_hip_cylind", "dream_hip_arr_cylind_dev",
)

_i32, _i64, _dbl = ctypes.c_int, ctypes.c_int64, ctypes.c_double
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class DreamSirProblem(ctypes.Structure):
    """Mirror of the DreamSirProblem struct of include/dream_hip.h."""
    _fields_ = [
        ("Ro", _vp), ("No", _i64),
        ("a", _dbl), ("b", _dbl), ("dx", _dbl), ("dy", _dbl), ("dt", _dbl), ("nt", _i64),
        ("delay_type", ctypes.c_int32), ("delay", _vp),
        ("v", _dbl), ("cp", _dbl),
        ("num_elements", _i64), ("G", _vp),
        ("foc_met", ctypes.c_int32), ("focal", _vp),
        ("steer_met", ctypes.c_int32), ("theta", _dbl), ("phi", _dbl),
        ("apod", _vp), ("do_apod", ctypes.c_int32), ("apod_met", ctypes.c_int32),
        ("apod_par", _dbl),
        ("err_level", ctypes.c_int32),
        ("h", _vp),
        ("shape", ctypes.c_int32), ("R", _dbl),
    ]


class DreamDasProblem(ctypes.Structure):
    """Mirror of the DreamDasProblem struct of include/dream_hip.h."""
    _fields_ = [
        ("K", _i64), ("No", _i64), ("Lt", _i64), ("Lr", _i64),
        ("Y", _vp), ("Ro", _vp), ("Gt", _vp), ("Gr", _vp),
        ("dt", _dbl), ("delay", _dbl), ("cp", _dbl),
        ("apod_t", _vp), ("apod_r", _vp),
        ("perm", _vp),
        ("Im", _vp),
    ]


_lib = None


def _declare(L):
    L.dream_hip_last_error.restype = ctypes.c_char_p
    L.dream_hip_last_error.argtypes = []
    L.dream_hip_version.restype = ctypes.c_char_p
    L.dream_hip_version.argtypes = []
    L.dream_hip_device_count.restype = _i32
    L.dream_hip_device_count.argtypes = []
    L.dream_hip_set_autotune.restype = None
    L.dream_hip_set_autotune.argtypes = [_i32]
    L.dream_hip_abort.restype = None
    L.dream_hip_abort.argtypes = [_i32]
    L.dream_hip_set_running.restype = None
    L.dream_hip_set_running.argtypes = []
    L.dream_hip_is_running.restype = _i32
    L.dream_hip_is_running.argtypes = []
    L.dream_hip_test_chunking.restype = None
    L.dream_hip_test_chunking.argtypes = [_i64, _i64]
    sir_common = [_vp, _i64, _dbl, _dbl, _dbl, _dbl, _dbl, _i64, _i32, _vp, _dbl, _dbl]
    L.dream_hip_arr_rect.restype = _i32
    L.dream_hip_arr_rect.argtypes = [_dbl] + sir_common + [
        _i64, _vp, _i32, _vp, _i32, _dbl, _dbl, _vp, _i32, _i32, _dbl, _vp, _i32, _ip]
    L.dream_hip_rect.restype = _i32
    L.dream_hip_rect.argtypes = [_dbl] + sir_common + [_vp, _i32, _ip]
    L.dream_hip_arr_rect_dev.restype = _i32
    L.dream_hip_arr_rect_dev.argtypes = sir_common + [
        _i64, _vp, _i32, _vp, _i32, _dbl, _dbl, _vp, _i32, _i32, _dbl, _vp, _i32, _i32, _ip,
        _i64p, _vp]
    L.dream_hip_arr_circ.restype = _i32
    L.dream_hip_arr_circ.argtypes = [_dbl, _vp, _i64] + [_dbl] * 4 + [_i64, _i32, _vp, _dbl, _dbl] + [
        _i64, _vp, _i32, _vp, _i32, _dbl, _dbl, _vp, _i32, _i32, _dbl, _vp, _i32, _ip]
    L.dream_hip_circ.restype = _i32
    L.dream_hip_circ.argtypes = [_dbl, _vp, _i64] + [_dbl] * 4 + [_i64, _i32, _vp, _dbl, _dbl,
                                                                _vp, _i32, _ip]
    L.dream_hip_arr_circ_dev.restype = _i32
    L.dream_hip_arr_circ_dev.argtypes = [_vp, _i64] + [_dbl] * 4 + [_i64, _i32, _vp, _dbl, _dbl] + [
        _i64, _vp, _i32, _vp, _i32, _dbl, _dbl, _vp, _i32, _i32, _dbl, _vp, _i32, _i32, _ip,
        _i64p, _vp]
    L.dream_hip_arr_cylind.restype = _i32
    L.dream_hip_arr_cylind.argtypes = [_dbl, _vp, _i64] + [_dbl] * 6 + [_i64, _i32, _vp, _dbl,
                                                                       _dbl] + [
        _i64, _vp, _i32, _vp, _i32, _dbl, _dbl, _vp, _i32, _i32, _dbl, _vp, _i32, _ip]
    L.dream_hip_cylind.restype = _i32
    L.dream_hip_cylind.argtypes = [_dbl, _vp, _i64] + [_dbl] * 6 + [_i64, _i32, _vp, _dbl, _dbl,
                                                                   _vp, _i32, _ip]
    L.dream_hip_arr_cylind_dev.restype = _i32
    L.dream_hip_arr_cylind_dev.argtypes = [_vp, _i64] + [_dbl] * 6 + [_i64, _i32, _vp, _dbl,
                                                                     _dbl] + [
        _i64, _vp, _i32, _vp, _i32, _dbl, _dbl, _vp, _i32, _i32, _dbl, _vp, _i32, _i32, _ip,
        _i64p, _vp]
    L.dream_hip_arr_annu.restype = _i32
    L.dream_hip_arr_annu.argtypes = [_dbl, _vp, _i64] + [_dbl] * 3 + [
        _i64, _i32, _vp, _dbl, _dbl, _i64, _vp, _i32, _vp, _vp, _i32, _i32, _dbl, _vp, _i32, _ip]
    L.dream_hip_arr_annu_dev.restype = _i32
    L.dream_hip_arr_annu_dev.argtypes = [_vp, _i64] + [_dbl] * 3 + [
        _i64, _i32, _vp, _dbl, _dbl, _i64, _vp, _i32, _vp, _vp, _i32, _i32, _dbl, _vp, _i32,
        _i32, _ip, _i64p, _vp]
    L.dream_hip_rect_indices.restype = _i32
    L.dream_hip_rect_indices.argtypes = [_dbl] * 10 + [_i64p, _i64, _i64p]
    L.dream_hip_rsq_seed.restype = _i32
    L.dream_hip_rsq_seed.argtypes = [_vp, _i64, _vp]
    L.dream_hip_sir_att_dev.restype = _i32
    L.dream_hip_sir_att_dev.argtypes = [ctypes.POINTER(DreamSirProblem), ctypes.c_int32, _dbl,
                                        _i32, _vp, _vp, _vp]
    L.dream_hip_sir_plan_create.restype = _i32
    L.dream_hip_sir_plan_create.argtypes = [ctypes.POINTER(DreamSirProblem), ctypes.c_int32, _i32,
                                            ctypes.POINTER(_vp)]
    L.dream_hip_sir_plan_run.restype = _i32
    L.dream_hip_sir_plan_run.argtypes = [_vp, _vp, _i32]
    L.dream_hip_sir_plan_status.restype = _i32
    L.dream_hip_sir_plan_status.argtypes = [_vp, _i32p, _i64p]
    L.dream_hip_sir_plan_destroy.restype = _i32
    L.dream_hip_sir_plan_destroy.argtypes = [_vp]
    L.dream_hip_das_create.restype = _i32
    L.dream_hip_das_create.argtypes = [_i32, _i64, _i64, _i64, _i64, _i32, ctypes.POINTER(_vp)]
    L.dream_hip_das_destroy.restype = _i32
    L.dream_hip_das_destroy.argtypes = [_vp]
    L.dream_hip_das_setup_has_changed.restype = _i32
    L.dream_hip_das_setup_has_changed.argtypes = [_vp, _i32, _i64, _i64, _i64, _i64, _i32]
    L.dream_hip_das_set_pixel_order.restype = _i32
    L.dream_hip_das_set_pixel_order.argtypes = [_vp, _i32p, _i64]
    L.dream_hip_pixel_order.restype = _i32
    L.dream_hip_pixel_order.argtypes = [_vp, _i64, _i32p]
    L.dream_hip_das.restype = _i32
    L.dream_hip_das.argtypes = [_vp, _vp, _vp, _vp, _vp, _dbl, _i32, _vp, _dbl, _vp, _i32, _ip]
    L.dream_hip_das_dev.restype = _i32
    L.dream_hip_das_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _dbl, _i32, _vp, _dbl, _vp, _vp,
                                    _i32, _i64, _i64, _i64, _vp, _i32, _ip, _vp]
    L.dream_hip_das_status.restype = _i32
    L.dream_hip_das_status.argtypes = [_vp, _i32, _ip]
    L.dream_hip_das_uni_create.restype = _i32
    L.dream_hip_das_uni_create.argtypes = [_i32, _i64, _dp, _dp, _dp, _i32, ctypes.POINTER(_vp)]
    L.dream_hip_das_uni_destroy.restype = _i32
    L.dream_hip_das_uni_destroy.argtypes = [_vp]
    L.dream_hip_das_uni_shape.restype = _i32
    L.dream_hip_das_uni_shape.argtypes = [_vp, _i64p, _i64p, _i64p]
    L.dream_hip_das_uni_setup_has_changed.restype = _i32
    L.dream_hip_das_uni_setup_has_changed.argtypes = [_vp, _i32, _i64, _dp, _dp, _dp, _i32]
    L.dream_hip_das_uni.restype = _i32
    L.dream_hip_das_uni.argtypes = [_vp, _vp, _dbl, _dbl, _dbl, _vp]
    L.dream_hip_das_uni_dev.restype = _i32
    L.dream_hip_das_uni_dev.argtypes = [_vp, _vp, _dbl, _dbl, _dbl, _i32, _vp, _vp]
    L.dream_hip_fftconv_p.restype = _i32
    L.dream_hip_fftconv_p.argtypes = [_vp, _vp, _i64, _i64, _vp, _i64, _i64, _i32]
    L.dream_hip_sum_fftconv.restype = _i32
    L.dream_hip_sum_fftconv.argtypes = [_vp, _vp, _i64, _i64, _i64, _vp, _i64, _i32]
    L.dream_hip_conv_sum_dev.restype = _i32
    L.dream_hip_conv_sum_dev.argtypes = [_vp, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _i64, _i64,
                                         _i64, _i32, _i32, _vp]
    L.dream_hip_das_batch_create.restype = _i32
    L.dream_hip_das_batch_create.argtypes = [ctypes.POINTER(DreamDasProblem), ctypes.c_int32,
                                             _i32, _i32, _i32, ctypes.POINTER(_vp)]
    L.dream_hip_das_batch_run.restype = _i32
    L.dream_hip_das_batch_run.argtypes = [_vp, _vp, _i32]
    L.dream_hip_das_batch_destroy.restype = _i32
    L.dream_hip_das_batch_destroy.argtypes = [_vp]
    L.dream_hip_das_adjoint_dev.restype = _i32
    L.dream_hip_das_adjoint_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _dbl, _vp, _vp,
                                            _i32, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]
    L.dream_hip_das_time_kernels.restype = _i32
    L.dream_hip_das_time_kernels.argtypes = [_vp, _i32]
    L.dream_hip_das_kernel_times.restype = _i32
    L.dream_hip_das_kernel_times.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), _dp, _dp]
    L.dream_hip_das_adjoint_dev_noalloc.restype = _i32
    L.dream_hip_das_adjoint_dev_noalloc.argtypes = L.dream_hip_das_adjoint_dev.argtypes


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library with `python -m dream_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


ECANCELLED = -5


class Cancelled(RuntimeError):
    """dream_hip_abort() stopped the call (DREAM_HIP_ECANCELLED)."""


def check(rc):
    """Raise RuntimeError carrying the library's message for a non-zero code
    (Cancelled for a call stopped by dream_hip_abort)."""
    if rc != OK:
        msg = lib().dream_hip_last_error().decode(errors="replace")
        raise (Cancelled if rc == ECANCELLED else RuntimeError)(f"dream_hip error {rc}: {msg}")
    return rc


def require_device():
    """Raise unless a HIP device is visible (the library has no CPU path)."""
    if lib().dream_hip_device_count() < 1:
        raise RuntimeError("no HIP device visible: dream_amd runs on the GPU only")
