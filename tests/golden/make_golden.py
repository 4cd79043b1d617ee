"""Write the golden vectors under tests/golden/ from the reference itself.

Runs the reference's own das.cc and arr_functions.cc, compiled from the
DREAM sources into oracle/_ref/libdream_ref.so (``make -C oracle ref``), on
small seeded inputs and stores inputs and outputs as .npz fixtures:

  das_ref.npz             DAS<double>/DAS<float>::das for saft, tfm,
                          rca_coltx, rca_rowtx; scalar and per-pixel delays;
                          ignore / warn error levels (outputs and codes)
  das_rca_nonsquare_ref.npz  both RCA methods on arrays of 8 x 11 and 12 x 9
                          elements (the stripe extents' indexing)
  arr_functions_ref.npz   max_dim_arr, focusing (all methods), beamsteering
                          (all methods), apodization (all methods)

The reference's SIR sources (src/dreamrect.cc, src/dream_arr_rect.cc) need
FFTW, which this image lacks, so no SIR vectors come from the reference.

Usage (in a container that has the DREAM tree):
    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

SEED = 20251202


def das_cases():
    rng = np.random.default_rng(SEED)
    cases = []
    # Linear array on x (SAFT/TFM) and a crossed row/column array (RCA).
    L = 10
    x = np.linspace(-3.3, 3.3, L)
    G = np.c_[x, np.zeros(L), np.zeros(L)]
    G_row = np.c_[np.zeros(L), x, np.zeros(L)]
    K = 320
    # Pixels: an x-z grid (plus a few points outside the stripes for RCA).
    X, Z = np.meshgrid(np.linspace(-5.0, 5.0, 9), np.linspace(1.0, 6.0, 8))
    Ro = np.c_[X.ravel("F"), rng.uniform(-4.5, 4.5, X.size), Z.ravel("F")]
    for method in ("saft", "tfm", "rca_coltx", "rca_rowtx"):
        Gr = G if method == "tfm" else G_row
        n_ascans = L if method == "saft" else L * L
        Y = rng.standard_normal((K, n_ascans))
        vec_delay = rng.uniform(-0.5, 0.8, Ro.shape[0])
        for delay_kind in ("scalar", "vector"):
            delay = 0.37 if delay_kind == "scalar" else vec_delay
            for level in ("ignore", "warn"):
                cases.append(dict(method=method, Y=Y, Gt=G, Gr=Gr, Ro=Ro, dt=0.025, delay=delay,
                                  cp=1480.0, level=level, name=f"{method}_{delay_kind}_{level}"))
    return cases


def make_das():
    out = {}
    for c in das_cases():
        n = c["name"]
        for prec in ("f64", "f32"):
            Im, err = O.das(c["Y"], c["Gt"], c["Gr"], c["Ro"], c["dt"], c["delay"], c["cp"],
                            c["method"], c["level"], prec, use_ref=True)
            out[f"{n}__{prec}__Im"] = Im
            out[f"{n}__{prec}__err"] = np.int32(err)
        # Inputs: Y and the geometry once per method, the delay per case.
        m = c["method"]
        out[f"{m}__Y"] = c["Y"]
        out[f"{m}__Gt"] = c["Gt"]
        out[f"{m}__Gr"] = c["Gr"]
        out[f"{m}__Ro"] = c["Ro"]
        out[f"{n}__delay"] = np.atleast_1d(c["delay"])
        out[f"{n}__scalars"] = np.array([c["dt"], c["cp"]])
    np.savez_compressed(os.path.join(HERE, "das_ref.npz"), **out)
    return len(out)


def make_das_rca_nonsquare():
    """RCA arrays with different column and row counts: the reference reads
    the stripe extents at rows[Lc], rows[2 Lc - 1] and columns[Lr - 1]
    (extra_src/das.cc:464-467, :560-563), so these vectors fix which entries
    it uses."""
    rng = np.random.default_rng(SEED + 2)
    out = {}
    K = 400
    for Lc, Lr in ((8, 11), (12, 9)):
        xc = np.sort(rng.uniform(-3.0, 3.0, Lc))
        yr = np.sort(rng.uniform(-3.0, 3.0, Lr))
        G_col = np.c_[xc, rng.uniform(-0.2, 0.2, Lc), rng.uniform(0.0, 0.1, Lc)]
        G_row = np.c_[rng.uniform(-0.2, 0.2, Lr), yr, rng.uniform(0.0, 0.1, Lr)]
        Ro = np.c_[rng.uniform(-5, 5, 160), rng.uniform(-5, 5, 160), rng.uniform(0.5, 6, 160)]
        Y = rng.standard_normal((K, Lc * Lr))
        key = f"c{Lc}r{Lr}"
        out[f"{key}__Y"], out[f"{key}__Gt"], out[f"{key}__Gr"] = Y, G_col, G_row
        out[f"{key}__Ro"] = Ro
        for method in ("rca_coltx", "rca_rowtx"):
            for prec in ("f64", "f32"):
                Im, err = O.das(Y, G_col, G_row, Ro, 0.025, 0.3, 1510.0, method, "warn", prec,
                                use_ref=True)
                out[f"{key}__{method}__{prec}__Im"] = Im
                out[f"{key}__{method}__{prec}__err"] = np.int32(err)
    np.savez_compressed(os.path.join(HERE, "das_rca_nonsquare_ref.npz"), **out)
    return len(out)


def make_arr_functions():
    import ctypes
    R = O.ref()
    rng = np.random.default_rng(SEED + 1)
    L = 17
    gx = rng.uniform(-9, 9, L)
    gy = rng.uniform(-4, 4, L)
    gz = np.zeros(L)
    xm, ym, rm = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    R.ref_max_dim_arr(p(gx), p(gy), p(gz), L, ctypes.byref(xm), ctypes.byref(ym), ctypes.byref(rm))
    cp = 1487.5
    focal = 23.7
    theta, phi = 11.5, -7.25
    apod_vec = rng.uniform(0.1, 1.0, L)
    apod_par = 0.8
    foc = np.array([[R.ref_focusing(m, focal, gx[i], gy[i], xm.value, ym.value, rm.value, cp)
                     for i in range(L)] for m in range(7)])
    steer = np.array([[R.ref_beamsteering(m, theta, phi, gx[i], gy[i], xm.value, ym.value,
                                          rm.value, cp) for i in range(L)] for m in range(4)])
    apod = np.array([[R.ref_apodization(m, i, p(apod_vec), -99.0, gx[i], gy[i], rm.value,
                                        apod_par) for i in range(L)] for m in range(8)])
    np.savez_compressed(os.path.join(HERE, "arr_functions_ref.npz"), gx=gx, gy=gy, gz=gz,
                        dims=np.array([xm.value, ym.value, rm.value]), cp=cp, focal=focal,
                        theta=theta, phi=phi, apod_vec=apod_vec, apod_par=apod_par, foc=foc,
                        steer=steer, apod=apod)


if __name__ == "__main__":
    if not O.have_ref():
        sys.exit("oracle/_ref/libdream_ref.so is missing: run `make -C oracle ref` first")
    os.environ.setdefault("DREAM_NUM_THREADS", "4")
    print("das_ref.npz entries:", make_das())
    print("das_rca_nonsquare_ref.npz entries:", make_das_rca_nonsquare())
    make_arr_functions()
    print("arr_functions_ref.npz written")
