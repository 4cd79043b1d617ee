"""Device-tensor API over libdream_hip.so (PyTorch-ROCm tensors as device memory).

Layouts (row-major torch views of the reference's column-major matrices, so
the bytes are the same):
  Ro  (3, No)        rows x, y, z of the observation points / pixels
  G   (3, L)         element positions (Gt, Gr likewise)
  Y   (n_ascans, K) or (F, n_ascans, K)  -- A-scan a at Y[a, :]
  Im  (No,) or (F, No)
  H   (No, nt)       SIR of point n at H[n, :]

Precisions: "f64" (FP64 data and delays, the reference's double path), "f32"
(FP32 data and sums, FP64 geometry, delays and sample indices), "f32_ref"
(the reference's float path, FP32 geometry).
"""
import ctypes

import numpy as np
import torch

from . import _lib

PREC = {"f64": _lib.F64, "f32_ref": _lib.F32_REF, "f32": _lib.F32}
INTERP = {"nearest": _lib.NEAREST, "linear": _lib.LINEAR}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _dev_tensor(x, dtype, device):
    t = torch.as_tensor(x, dtype=dtype, device=device)
    return t.contiguous()


def _same_device(a, b):
    a, b = torch.device(a), torch.device(b)
    if a.type != b.type:
        return False
    ia = a.index if a.index is not None else torch.cuda.current_device()
    ib = b.index if b.index is not None else torch.cuda.current_device()
    return ia == ib


def _check(t, name, device, dtype=None, numel=None, optional=False):
    """Refuse, before any launch, a tensor the kernels cannot read as the
    plan expects: host memory, another GPU, another dtype, a strided view or
    too few elements would be read as raw device bytes."""
    if t is None:
        if optional:
            return
        raise ValueError(f"{name}: a tensor is required")
    if not torch.is_tensor(t):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.device.type != "cuda" or not _same_device(t.device, device):
        raise ValueError(f"{name}: must be on {device}, is on {t.device}")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name}: must be {dtype}, is {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: must have {numel} elements, has {t.numel()}")


def pixel_order_of(Ro):
    """Spatially compact visiting order of the points Ro ((No, 3))."""
    R = np.asfortranarray(np.asarray(Ro.cpu() if torch.is_tensor(Ro) else Ro,
                                     dtype=np.float64).reshape(-1, 3))
    perm = np.empty(R.shape[0], dtype=np.int32)
    _lib.check(_lib.lib().dream_hip_pixel_order(
        R.ctypes.data, R.shape[0], perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return perm


class DasPlan:
    """A DAS object on the GPU (DAS<T> with cached state) for fixed geometry.

    Gt, Gr, Ro: (L, 3) / (No, 3) arrays or tensors in the reference's
    row-per-element convention; stored on the device coordinate-major.
    """

    def __init__(self, method, K, Ro, Gt, Gr=None, precision="f32", device="cuda",
                 pixel_order=True):
        _lib.require_device()
        self.method = method
        self.type = _lib.DAS_TYPE[method]
        self.precision = precision
        self.prec = PREC[precision]
        self.data_dtype = torch.float64 if precision == "f64" else torch.float32
        self.geom_dtype = torch.float32 if precision == "f32_ref" else torch.float64
        self.device = torch.device(device)
        gd = self.geom_dtype
        self.Ro = _dev_tensor(torch.as_tensor(Ro, dtype=gd).reshape(-1, 3).t(), gd, self.device)
        self.Gt = _dev_tensor(torch.as_tensor(Gt, dtype=gd).reshape(-1, 3).t(), gd, self.device)
        self.Lt = self.Gt.shape[1]
        if self.type != 0:
            self.Gr = _dev_tensor(torch.as_tensor(Gr, dtype=gd).reshape(-1, 3).t(), gd,
                                  self.device)
            if self.type == 1 and self.Gr.shape == self.Gt.shape and torch.equal(self.Gr,
                                                                                  self.Gt):
                # One array for both sides: passing the same device array lets
                # the library pair (i, j) with (j, i) (reciprocal TFM).
                self.Gr = self.Gt
            self.Lr = self.Gr.shape[1]
        else:
            self.Gr = None
            self.Lr = 0
        self.No = self.Ro.shape[1]
        self.K = int(K)
        self.n_ascans = self.Lt if self.type == 0 else self.Lt * self.Lr
        self.h = ctypes.c_void_p(None)
        _lib.check(_lib.lib().dream_hip_das_create(self.type, self.K, self.No, self.Lt, self.Lr,
                                                   self.prec, ctypes.byref(self.h)))
        self._delay_cache = {}
        if pixel_order:
            self.set_pixel_order(pixel_order_of(Ro))

    def time_kernels(self, n_max):
        """Record HIP events around the kernels of the next n_max forward calls."""
        _lib.check(_lib.lib().dream_hip_das_time_kernels(self.h, int(n_max)))

    def kernel_times(self):
        """(timed calls, summed preparation ms, summed beamforming ms); stops timing."""
        n = ctypes.c_int(0)
        prep, fwd = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _lib.check(_lib.lib().dream_hip_das_kernel_times(self.h, ctypes.byref(n),
                                                         ctypes.byref(prep), ctypes.byref(fwd)))
        return n.value, prep.value, fwd.value

    def set_pixel_order(self, perm):
        """Visit the pixels in `perm` order (None: 0..No-1)."""
        if perm is None:
            _lib.check(_lib.lib().dream_hip_das_set_pixel_order(self.h, None, 0))
            return
        perm = np.ascontiguousarray(perm, dtype=np.int32)
        _lib.check(_lib.lib().dream_hip_das_set_pixel_order(
            self.h, perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), perm.size))

    def __del__(self):
        try:
            if self.h.value is not None and _lib._lib is not None:
                _lib._lib.dream_hip_das_destroy(self.h)
        except Exception:
            pass

    def _delay(self, delay):
        if torch.is_tensor(delay) and delay.numel() > 1:
            if delay.numel() != self.No:
                raise ValueError(f"delay: needs 1 or {self.No} elements, has {delay.numel()}")
            d = delay.to(device=self.device, dtype=self.geom_dtype).contiguous()
            return d, 1
        val = float(delay)
        if val not in self._delay_cache:
            self._delay_cache[val] = torch.tensor([val], dtype=self.geom_dtype, device=self.device)
        return self._delay_cache[val], 0

    def _frames(self, Y):
        _check(Y, "Y", self.device, self.data_dtype)
        if Y.dim() not in (2, 3) or tuple(Y.shape[-2:]) != (self.n_ascans, self.K):
            raise ValueError(f"Y must be ({self.n_ascans}, {self.K}) or "
                             f"(F, {self.n_ascans}, {self.K}), is {tuple(Y.shape)}")
        return 1 if Y.dim() == 2 else Y.shape[0]

    def _apod(self, apod_t, apod_r):
        _check(apod_t, "apod_t", self.device, self.data_dtype, self.Lt, optional=True)
        if self.type != 0:
            _check(apod_r, "apod_r", self.device, self.data_dtype, self.Lr, optional=True)
            if (apod_t is None) != (apod_r is None):
                raise ValueError("TFM/RCA apodization needs both apod_t and apod_r")

    def forward(self, Y, dt, delay, cp, apod_t=None, apod_r=None, interp="nearest",
                err_level="ignore", want_status=False, out=None):
        """Beamform; returns Im (and the SIRError code when want_status)."""
        F = self._frames(Y)
        self._apod(apod_t, apod_r)
        d, dtype = self._delay(delay)
        if out is not None:
            _check(out, "out", self.device, self.data_dtype, F * self.No)
        Im = out if out is not None else torch.empty(
            (F, self.No) if Y.dim() == 3 else (self.No,), dtype=self.data_dtype, device=Y.device)
        code = ctypes.c_int(0)
        _lib.check(_lib.lib().dream_hip_das_dev(
            self.h, _ptr(Y), _ptr(self.Ro), _ptr(self.Gt), _ptr(self.Gr), float(dt), dtype,
            _ptr(d), float(cp), _ptr(apod_t), _ptr(apod_r), INTERP[interp], F,
            self.n_ascans * self.K, self.No, _ptr(Im), _lib.ERR_LEVEL[err_level],
            ctypes.byref(code) if want_status else None, _stream()))
        return (Im, code.value) if want_status else Im

    def status(self, err_level="ignore"):
        """SIRError code of the last forward call (the device computes the
        out-of-bounds status in every call; this waits for it)."""
        code = ctypes.c_int(0)
        _lib.check(_lib.lib().dream_hip_das_status(self.h, _lib.ERR_LEVEL[err_level],
                                                   ctypes.byref(code)))
        return code.value

    def adjoint(self, Y, g, dt, delay, cp, apod_t=None, apod_r=None, interp="linear",
                want_gy=True, want_apod=True, want_cp=True):
        """Aᵀg: returns (gY or None, g_apod_t, g_apod_r, g_cp)."""
        F = self._frames(Y)
        self._apod(apod_t, apod_r)
        g = g.contiguous() if torch.is_tensor(g) else g
        _check(g, "g", self.device, self.data_dtype, F * self.No)
        dt_ = self.data_dtype
        gY = torch.empty_like(Y) if want_gy else None
        gat = torch.empty(self.Lt, dtype=dt_, device=Y.device) if want_apod else None
        gar = (torch.empty(self.Lr, dtype=dt_, device=Y.device)
               if want_apod and self.type != 0 else None)
        gcp = torch.empty(1, dtype=dt_, device=Y.device) if want_cp else None
        _lib.check(_lib.lib().dream_hip_das_adjoint_dev(
            self.h, _ptr(Y), _ptr(self.Ro), _ptr(self.Gt), _ptr(self.Gr), float(dt), float(delay),
            float(cp), _ptr(apod_t), _ptr(apod_r), INTERP[interp], F, self.n_ascans * self.K,
            self.No, _ptr(g), _ptr(gY), _ptr(gat), _ptr(gar), _ptr(gcp), _stream()))
        return gY, gat, gar, gcp


class DASFunction(torch.autograd.Function):
    """Differentiable DAS: Im = DAS(Y; apod_t, apod_r, cp).

    Gradients with respect to Y, the apodization weights and the scalar cp
    come from the hand-written adjoint kernels.  cp is a 0-d / 1-element
    tensor (its value is read on the host for the forward launch).
    """

    @staticmethod
    def forward(ctx, Y, apod_t, apod_r, cp, plan, dt, delay, interp):
        cpv = float(cp)
        Im = plan.forward(Y.contiguous(), dt, delay, cpv, apod_t, apod_r, interp)
        ctx.save_for_backward(Y, apod_t, apod_r)
        ctx.plan, ctx.dt, ctx.delay, ctx.interp, ctx.cp = plan, dt, delay, interp, cpv
        ctx.cp_shape = cp.shape if torch.is_tensor(cp) else None
        return Im

    @staticmethod
    def backward(ctx, gIm):
        Y, apod_t, apod_r = ctx.saved_tensors
        want_gy = ctx.needs_input_grad[0]
        want_apod = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        want_cp = ctx.needs_input_grad[3]
        gY, gat, gar, gcp = ctx.plan.adjoint(Y.contiguous(), gIm, ctx.dt, ctx.delay, ctx.cp,
                                             apod_t, apod_r, ctx.interp, want_gy, want_apod,
                                             want_cp)
        gcp_out = None
        if want_cp and ctx.cp_shape is not None:
            gcp_out = gcp.reshape(ctx.cp_shape)
        return (gY, gat if ctx.needs_input_grad[1] else None,
                gar if ctx.needs_input_grad[2] else None, gcp_out, None, None, None, None)


def das_diff(plan, Y, apod_t, apod_r, cp, dt, delay, interp="linear"):
    """Differentiable DAS through DASFunction."""
    return DASFunction.apply(Y, apod_t, apod_r, cp, plan, dt, delay, interp)


# ----------------------------------------------------------------------------- SIR

def _host_f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _check_points(Ro):
    """Ro: (3, No) contiguous float64 tensor on a GPU; returns No."""
    if not torch.is_tensor(Ro) or Ro.device.type != "cuda":
        raise ValueError("Ro must be a (3, No) float64 tensor on the GPU")
    _check(Ro, "Ro", Ro.device, torch.float64)
    if Ro.dim() != 2 or Ro.shape[0] != 3:
        raise ValueError(f"Ro must be (3, No), is {tuple(Ro.shape)}")
    return Ro.shape[1]


def _sir_delay(delay, Ro, No):
    """A scalar delay, or one per point moved to Ro's device as float64."""
    if torch.is_tensor(delay) and delay.numel() > 1:
        if delay.numel() != No:
            raise ValueError(f"delay: needs 1 or {No} elements, has {delay.numel()}")
        return delay.to(device=Ro.device, dtype=torch.float64).contiguous(), 1
    return torch.tensor([float(delay)], dtype=torch.float64, device=Ro.device), 0


def _host_lengths(foc_met, focal_v, apod_met, apod_v, L):
    """'ud' focusing and apodization read one host value per element."""
    if foc_met == "ud" and focal_v.size < L:
        raise ValueError(f"focal: 'ud' focusing needs {L} delays, has {focal_v.size}")
    if apod_met == "ud" and apod_v.size < L:
        raise ValueError(f"apod: 'ud' apodization needs {L} weights, has {apod_v.size}")


def _sir_out(out, No, nt, out_dtype, device):
    if out is None:
        return torch.empty((No, int(nt)), dtype=out_dtype, device=device)
    _check(out, "out", device, out_dtype, No * int(nt))
    return out


def arr_rect(Ro, a, b, dx, dy, dt, nt, delay, v, cp, G, foc_met="off", focal=0.0,
             steer_met="off", theta=0.0, phi=0.0, apod_met="off", apod=None, apod_par=0.0,
             err_level="ignore", out_dtype=torch.float64, out=None):
    """dream_arr_rect on device tensors.  Ro: (3, No) float64 device tensor,
    delay: float or (No,) device tensor.  Returns (H (No, nt), SIRError code,
    out-of-bounds contribution count)."""
    _lib.require_device()
    No = _check_points(Ro)
    G = np.asarray(G, dtype=np.float64).reshape(-1, 3)
    L = G.shape[0]
    Gf = np.asfortranarray(G)
    d, dtype = _sir_delay(delay, Ro, No)
    focal_v = _host_f64(np.atleast_1d(focal))
    apod_v = _host_f64(np.ones(L) if apod is None else apod)
    _host_lengths(foc_met, focal_v, apod_met, apod_v, L)
    do_apod = 0 if apod_met == "off" else 1
    H = _sir_out(out, No, nt, out_dtype, Ro.device)
    code = ctypes.c_int(0)
    oob = ctypes.c_int64(0)
    _lib.check(_lib.lib().dream_hip_arr_rect_dev(
        _ptr(Ro), No, a, b, dx, dy, dt, int(nt), dtype, _ptr(d), v, cp, L, Gf.ctypes.data,
        _lib.FOC_MET[foc_met], focal_v.ctypes.data, _lib.STEER_MET[steer_met], theta, phi,
        apod_v.ctypes.data, do_apod, _lib.APOD_MET.get(apod_met, 0), apod_par, _ptr(H),
        _lib.F64 if out_dtype == torch.float64 else _lib.F32, _lib.ERR_LEVEL[err_level],
        ctypes.byref(code), ctypes.byref(oob), _stream()))
    return H, code.value, oob.value


def arr_circ(Ro, R, dx, dy, dt, nt, delay, v, cp, G, foc_met="off", focal=0.0,
             steer_met="off", theta=0.0, phi=0.0, apod_met="off", apod=None, apod_par=0.0,
             err_level="ignore", out_dtype=torch.float64, out=None):
    """dream_arr_circ on device tensors (circular elements of radius R); the
    conventions of arr_rect.  Returns (H (No, nt), SIRError code, out-of-bounds
    contribution count)."""
    _lib.require_device()
    No = _check_points(Ro)
    Gf = np.asfortranarray(np.asarray(G, dtype=np.float64).reshape(-1, 3))
    L = Gf.shape[0]
    d, dtype = _sir_delay(delay, Ro, No)
    focal_v = _host_f64(np.atleast_1d(focal))
    apod_v = _host_f64(np.ones(L) if apod is None else apod)
    _host_lengths(foc_met, focal_v, apod_met, apod_v, L)
    H = _sir_out(out, No, nt, out_dtype, Ro.device)
    code = ctypes.c_int(0)
    oob = ctypes.c_int64(0)
    _lib.check(_lib.lib().dream_hip_arr_circ_dev(
        _ptr(Ro), No, R, dx, dy, dt, int(nt), dtype, _ptr(d), v, cp, L, Gf.ctypes.data,
        _lib.FOC_MET[foc_met], focal_v.ctypes.data, _lib.STEER_MET[steer_met], theta, phi,
        apod_v.ctypes.data, 0 if apod_met == "off" else 1, _lib.APOD_MET.get(apod_met, 0),
        apod_par, _ptr(H), _lib.F64 if out_dtype == torch.float64 else _lib.F32,
        _lib.ERR_LEVEL[err_level], ctypes.byref(code), ctypes.byref(oob), _stream()))
    return H, code.value, oob.value


def arr_cylind(Ro, a, b, Rcurv, dx, dy, dt, nt, delay, v, cp, G, foc_met="off", focal=0.0,
               steer_met="off", theta=0.0, phi=0.0, apod_met="off", apod=None, apod_par=0.0,
               err_level="ignore", out_dtype=torch.float64, out=None):
    """dream_arr_cylind on device tensors (cylindrically curved elements a x b,
    curvature radius Rcurv); the conventions of arr_rect.  Returns (H (No, nt),
    SIRError code, out-of-bounds contribution count)."""
    _lib.require_device()
    No = _check_points(Ro)
    Gf = np.asfortranarray(np.asarray(G, dtype=np.float64).reshape(-1, 3))
    L = Gf.shape[0]
    d, dtype = _sir_delay(delay, Ro, No)
    focal_v = _host_f64(np.atleast_1d(focal))
    apod_v = _host_f64(np.ones(L) if apod is None else apod)
    _host_lengths(foc_met, focal_v, apod_met, apod_v, L)
    H = _sir_out(out, No, nt, out_dtype, Ro.device)
    code = ctypes.c_int(0)
    oob = ctypes.c_int64(0)
    _lib.check(_lib.lib().dream_hip_arr_cylind_dev(
        _ptr(Ro), No, a, b, Rcurv, dx, dy, dt, int(nt), dtype, _ptr(d), v, cp, L, Gf.ctypes.data,
        _lib.FOC_MET[foc_met], focal_v.ctypes.data, _lib.STEER_MET[steer_met], theta, phi,
        apod_v.ctypes.data, 0 if apod_met == "off" else 1, _lib.APOD_MET.get(apod_met, 0),
        apod_par, _ptr(H), _lib.F64 if out_dtype == torch.float64 else _lib.F32,
        _lib.ERR_LEVEL[err_level], ctypes.byref(code), ctypes.byref(oob), _stream()))
    return H, code.value, oob.value


def arr_annu(Ro, Gr, dx, dy, dt, nt, delay, v, cp, foc_met="off", focal=0.0,
             apod_met="off", apod=None, apod_par=0.0, err_level="ignore",
             out_dtype=torch.float64, out=None):
    """dream_arr_annu on device tensors: Ro is (3, No) float64, Gr the host
    vector of radii (centre disc, then inner/outer per ring).  Returns
    (H (No, nt), SIRError code, out-of-bounds contribution count)."""
    _lib.require_device()
    No = _check_points(Ro)
    Gr_v = _host_f64(np.asarray(Gr, dtype=np.float64).ravel())
    nr = Gr_v.shape[0]
    d, dtype = _sir_delay(delay, Ro, No)
    focal_v = _host_f64(np.atleast_1d(focal))
    if focal_v.shape[0] < nr:
        focal_v = _host_f64(np.resize(focal_v, nr))
    apod_v = _host_f64(np.ones(nr) if apod is None else apod)
    H = _sir_out(out, No, nt, out_dtype, Ro.device)
    code = ctypes.c_int(0)
    oob = ctypes.c_int64(0)
    _lib.check(_lib.lib().dream_hip_arr_annu_dev(
        _ptr(Ro), No, dx, dy, dt, int(nt), dtype, _ptr(d), v, cp, nr, Gr_v.ctypes.data,
        _lib.FOC_MET[foc_met], focal_v.ctypes.data, apod_v.ctypes.data,
        0 if apod_met == "off" else 1, _lib.APOD_MET.get(apod_met, 0), apod_par, _ptr(H),
        _lib.F64 if out_dtype == torch.float64 else _lib.F32, _lib.ERR_LEVEL[err_level],
        ctypes.byref(code), ctypes.byref(oob), _stream()))
    return H, code.value, oob.value


def _fill_sir_problem(q, p, H):
    """Fill DreamSirProblem q from the problem dict p (SirBatch's keys) with
    output H; returns the objects whose memory q points into."""
    Ro = p["Ro"]
    No = _check_points(Ro)
    G = np.asfortranarray(np.asarray(p["G"], dtype=np.float64).reshape(-1, 3))
    L = G.shape[0]
    d, dtype = _sir_delay(p.get("delay", 0.0), Ro, No)
    focal = _host_f64(np.atleast_1d(p.get("focal", 0.0)))
    apod = _host_f64(p["apod"] if p.get("apod") is not None else np.ones(L))
    _host_lengths(p.get("foc_met", "off"), focal, p.get("apod_met", "off"), apod, L)
    _check(H, "out", Ro.device, numel=No * int(p["nt"]))
    q.Ro, q.No = Ro.data_ptr(), No
    q.a, q.b = p.get("a", 0.0), p.get("b", 0.0)
    q.dx, q.dy, q.dt, q.nt = p["dx"], p["dy"], p["dt"], int(p["nt"])
    q.delay_type, q.delay = dtype, d.data_ptr()
    q.v, q.cp = p.get("v", 1.0), p["cp"]
    q.num_elements, q.G = L, G.ctypes.data
    q.foc_met, q.focal = _lib.FOC_MET[p.get("foc_met", "off")], focal.ctypes.data
    q.steer_met = _lib.STEER_MET[p.get("steer_met", "off")]
    q.theta, q.phi = p.get("theta", 0.0), p.get("phi", 0.0)
    apod_met = p.get("apod_met", "off")
    q.apod, q.do_apod = apod.ctypes.data, 0 if apod_met == "off" else 1
    q.apod_met, q.apod_par = _lib.APOD_MET.get(apod_met, 0), p.get("apod_par", 0.0)
    q.err_level = _lib.ERR_LEVEL[p.get("err_level", "ignore")]
    q.h = H.data_ptr()
    q.shape = {"rect": 0, "circ": 1, "cylind": 2}[p.get("shape", "rect")]
    q.R = float(p.get("Rcurv", 0.0) if q.shape == 2 else p.get("R", 0.0))
    return [Ro, G, d, focal, apod, H]


def sir_attenuated(problem, alpha, out_dtype=torch.float64, out=None):
    """The SIR of one problem (SirBatch's dict keys: rectangles, discs or
    cylindrically curved elements) with frequency-dependent attenuation
    alpha [dB/(cm MHz)] on device tensors.  Returns (H (No, nt), SIRError
    code, out-of-bounds contribution count)."""
    _lib.require_device()
    Ro = problem["Ro"]
    No = _check_points(Ro)
    H = _sir_out(out, No, problem["nt"], out_dtype, Ro.device)
    q = (_lib.DreamSirProblem * 1)()
    keep = _fill_sir_problem(q[0], problem, H)
    code = ctypes.c_int32(0)
    oob = ctypes.c_int64(0)
    _lib.check(_lib.lib().dream_hip_sir_att_dev(
        q, 1, float(alpha), _lib.F64 if out_dtype == torch.float64 else _lib.F32,
        ctypes.byref(code), ctypes.byref(oob), _stream()))
    del keep
    return H, code.value, oob.value


class SirBatch:
    """Many uneven dream_arr_rect problems in one persistent launch.

    problems: list of dicts with keys Ro ((3, No) float64 device tensor), a, b
    (rectangles), shape="circ" and R (discs) or shape="cylind", a, b and Rcurv
    (cylindrically curved elements), dx, dy, dt, nt, delay (float
    or (No,) device tensor), v, cp, G ((L, 3) host array), and optionally
    foc_met, focal, steer_met, theta, phi, apod_met, apod, apod_par,
    err_level.  Outputs H[i] are (No_i, nt_i)
    tensors of out_dtype allocated by the batch.
    """

    def __init__(self, problems, out_dtype=torch.float32, device="cuda"):
        _lib.require_device()
        self.out_dtype = out_dtype
        self.n = len(problems)
        self._keep = []
        self.H = []
        arr = (_lib.DreamSirProblem * max(self.n, 1))()
        for i, p in enumerate(problems):
            H = torch.zeros((_check_points(p["Ro"]), int(p["nt"])), dtype=out_dtype,
                            device=p["Ro"].device)
            self.H.append(H)
            self._keep += _fill_sir_problem(arr[i], p, H)
        self.plan = ctypes.c_void_p(None)
        _lib.check(_lib.lib().dream_hip_sir_plan_create(arr, self.n,
                                                        _lib.F64 if out_dtype == torch.float64
                                                        else _lib.F32, ctypes.byref(self.plan)))

    def run(self, use_graph=True):
        _lib.check(_lib.lib().dream_hip_sir_plan_run(self.plan, _stream(), 1 if use_graph else 0))
        return self.H

    def status(self):
        codes = (ctypes.c_int32 * max(self.n, 1))()
        oob = (ctypes.c_int64 * max(self.n, 1))()
        _lib.check(_lib.lib().dream_hip_sir_plan_status(self.plan, codes, oob))
        return list(codes)[:self.n], list(oob)[:self.n]

    def __del__(self):
        try:
            if self.plan.value is not None and _lib._lib is not None:
                _lib._lib.dream_hip_sir_plan_destroy(self.plan)
        except Exception:
            pass


class DasBatch:
    """Many uneven DAS problems of one method, precision and sampling rule in
    one persistent forward launch.

    problems: list of dicts with keys K, Ro ((No, 3) array or tensor), Gt,
    Gr (not for SAFT), Y ((n_ascans, K) device tensor of the batch's data
    type), dt, cp, and optionally delay (scalar), apod_t and apod_r (device
    tensors; all problems or none).  Outputs Im[i] ((No_i,) tensors) are
    allocated by the batch; run() overwrites them.
    """

    def __init__(self, method, problems, precision="f32", interp="nearest", device="cuda",
                 pixel_order=True):
        _lib.require_device()
        self.type = _lib.DAS_TYPE[method]
        self.prec = PREC[precision]
        self.data_dtype = torch.float64 if precision == "f64" else torch.float32
        gd = torch.float32 if precision == "f32_ref" else torch.float64
        dev = torch.device(device)
        self.n = len(problems)
        self._keep = []
        self.Im = []
        arr = (_lib.DreamDasProblem * max(self.n, 1))()
        for i, p in enumerate(problems):
            Ro = _dev_tensor(torch.as_tensor(p["Ro"], dtype=gd).reshape(-1, 3).t(), gd, dev)
            Gt = _dev_tensor(torch.as_tensor(p["Gt"], dtype=gd).reshape(-1, 3).t(), gd, dev)
            Gr = (None if self.type == 0 else
                  _dev_tensor(torch.as_tensor(p["Gr"], dtype=gd).reshape(-1, 3).t(), gd, dev))
            Y = p["Y"]
            No, Lt = Ro.shape[1], Gt.shape[1]
            Lr = 0 if Gr is None else Gr.shape[1]
            n_ascans = Lt if self.type == 0 else Lt * Lr
            _check(Y, f"problem {i}: Y", dev, self.data_dtype, n_ascans * int(p["K"]))
            _check(p.get("apod_t"), f"problem {i}: apod_t", dev, self.data_dtype, Lt,
                   optional=True)
            if self.type != 0:
                _check(p.get("apod_r"), f"problem {i}: apod_r", dev, self.data_dtype, Lr,
                       optional=True)
            perm = None
            if pixel_order:
                perm = torch.as_tensor(pixel_order_of(p["Ro"]), device=dev)
            Im = torch.empty(No, dtype=self.data_dtype, device=dev)
            self.Im.append(Im)
            self._keep += [Ro, Gt, Gr, Y, perm, Im, p.get("apod_t"), p.get("apod_r")]
            q = arr[i]
            q.K, q.No, q.Lt, q.Lr = int(p["K"]), No, Lt, Lr
            q.Y, q.Ro, q.Gt = Y.data_ptr(), Ro.data_ptr(), Gt.data_ptr()
            q.Gr = None if Gr is None else Gr.data_ptr()
            q.dt, q.delay, q.cp = float(p["dt"]), float(p.get("delay", 0.0)), float(p["cp"])
            q.apod_t = None if p.get("apod_t") is None else p["apod_t"].data_ptr()
            q.apod_r = None if p.get("apod_r") is None else p["apod_r"].data_ptr()
            q.perm = None if perm is None else perm.data_ptr()
            q.Im = Im.data_ptr()
        self.batch = ctypes.c_void_p(None)
        _lib.check(_lib.lib().dream_hip_das_batch_create(arr, self.n, self.type, self.prec,
                                                         INTERP[interp], ctypes.byref(self.batch)))

    def run(self, use_graph=True):
        _lib.check(_lib.lib().dream_hip_das_batch_run(self.batch, _stream(),
                                                      1 if use_graph else 0))
        return self.Im

    def __del__(self):
        try:
            if self.batch.value is not None and _lib._lib is not None:
                _lib._lib.dream_hip_das_batch_destroy(self.batch)
        except Exception:
            pass


# ----------------------------------------------------------------------- convolution

CONV_MODE = {"=": _lib.CONV_EQU, "+=": _lib.CONV_SUM, "-=": _lib.CONV_NEG}


def conv_sum(A, B, out=None, mode="="):
    """Column-wise linear convolution on device tensors, summed over layers.

    A: (N, M_A) or (L, N, M_A) -- rows are signals (row-major views of the
    reference's column-major M_A x N / M_A x N x L arrays); B: (M_B,) one
    filter, (N, M_B) one per column, or (L, M_B) one per layer.  Returns or
    updates out (N, M_A + M_B - 1): out[n] (mode) sum_l A[l, n] (*) B[...].
    float64 or float32.
    """
    if not torch.is_tensor(A) or A.dtype not in (torch.float64, torch.float32):
        raise ValueError("A must be a float64 or float32 tensor")
    _check(A, "A", A.device, A.dtype)
    _check(B, "B", A.device, A.dtype)
    if out is not None:
        _check(out, "out", A.device, A.dtype)
    A3 = A if A.dim() == 3 else A.unsqueeze(0)
    L, N, MA = A3.shape
    MB = B.shape[-1]
    if B.dim() == 1:
        b_col, b_layer = 0, 0
    elif A.dim() == 2 and B.dim() == 2:
        if B.shape[0] != N:
            raise ValueError("B needs one row per column of A")
        b_col, b_layer = MB, 0
    elif B.dim() == 2:
        if B.shape[0] != L:
            raise ValueError("B needs one row per layer of A")
        b_col, b_layer = 0, MB
    else:
        raise ValueError("B must be (M_B,), (N, M_B) or (L, M_B)")
    if out is None:
        if mode != "=":
            raise ValueError("in-place modes need `out`")
        out = torch.empty((N, MA + MB - 1), dtype=A.dtype, device=A.device)
    if tuple(out.shape) != (N, MA + MB - 1):
        raise ValueError(f"out must be ({N}, {MA + MB - 1}), is {tuple(out.shape)}")
    dt = _lib.F64 if A.dtype == torch.float64 else _lib.F32
    _lib.check(_lib.lib().dream_hip_conv_sum_dev(_ptr(out), _ptr(A3), MA, N, L, MA, N * MA,
                                                 _ptr(B), MB, b_col, b_layer, CONV_MODE[mode],
                                                 dt, _stream()))
    return out
