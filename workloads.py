"""Seeded synthetic workloads of the five benchmark configurations.

C1  dream_arr_rect, 16 elements 0.9 x 10 mm at 1 mm pitch, 64 on-axis points
    z = 20..60 mm, dx = dy = 0.05 mm, dt = 0.025 us, nt = 400, cp = 1500,
    per-point delay 1000 z / cp - 0.5 us.
C2  TFM, 64 elements at 0.5 mm pitch, 64 x 64 A-scans of 2048 samples at 50 MHz,
    cp = 1500, 20 point scatterers (x in [-16, 16], z in [10, 40] mm, amplitude
    U(0.5, 1)), Gaussian-windowed 5 MHz pulses (60 % bandwidth) at the exact
    round-trip time plus N(0, 0.01^2) noise; 256 x 256 pixels at 0.125 mm.
C3  32 uneven dream_arr_rect problems: L ~ U{8..128}, a ~ U(0.2, 1), b ~ U(5, 15),
    pitch a + 0.1, No log-uniform in [1k, 64k] on an x-z plane, nt ~ U{500..4000},
    dt = 0.02, 'ud' focusing from a random focal point, Hann apodization.
C4  SAFT, 128 elements at 0.3 mm pitch, 128 pulse-echo A-scans of 4096 samples
    at 50 MHz, 40 scatterers, 512 x 512 pixels at 0.06 mm.
C5  RCA, 128 x 128 row-column array at 0.2 mm pitch, 128 row transmits x 128
    column receives x 4096 samples at 40 MHz, cp = 1540, 100 scatterers, 8
    frames with N(0, 0.05 mm) jitter; 128 x 128 x 256 voxels.

Echo synthesis runs in torch on the given device.  Matrices are returned in
the reference's row-per-point convention ((No, 3), (L, 3)); Y as
(n_ascans, K) float tensors (A-scan a at Y[a]).
"""
import math

import numpy as np
import torch

F0 = 5.0                      # MHz
BW = 0.6
SIGMA_T = math.sqrt(2.0 * math.log(2.0)) / (math.pi * BW * F0)   # us


def linear_array(L, pitch):
    x = (np.arange(L) - (L - 1) / 2.0) * pitch
    return np.c_[x, np.zeros(L), np.zeros(L)]


def _scatterers(rng, n, xr, zr, yr=(0.0, 0.0)):
    x = rng.uniform(*xr, n)
    y = rng.uniform(*yr, n) if yr != (0.0, 0.0) else np.zeros(n)
    z = rng.uniform(*zr, n)
    amp = rng.uniform(0.5, 1.0, n)
    return np.c_[x, y, z], amp


def _echoes(tau_us, amp, K, fs_mhz, noise, gen, device, dtype):
    """Y[a, k] = sum_s amp_s g(k/fs - tau[a, s]) + noise, g a Gaussian-windowed
    cosine; tau (n_ascans, n_scat) in us.  Evaluated on a +-6 sigma window."""
    tau = torch.as_tensor(tau_us, dtype=torch.float64, device=device)
    A = torch.as_tensor(amp, dtype=torch.float64, device=device)
    n_a, n_s = tau.shape
    half = int(math.ceil(6 * SIGMA_T * fs_mhz)) + 1
    offs = torch.arange(-half, half + 1, device=device, dtype=torch.float64)
    Y = torch.zeros((n_a, K), dtype=torch.float64, device=device)
    chunk = max(1, (1 << 22) // (n_s * offs.numel()))
    for a0 in range(0, n_a, chunk):
        t = tau[a0:a0 + chunk]                                        # (c, s)
        k0 = torch.floor(t * fs_mhz)
        k = k0[..., None] + offs                                      # (c, s, w)
        dtk = k / fs_mhz - t[..., None]
        val = A[None, :, None] * torch.exp(-0.5 * (dtk / SIGMA_T) ** 2) * torch.cos(
            2 * math.pi * F0 * dtk)
        ok = (k >= 0) & (k < K)
        idx = torch.where(ok, k, torch.zeros_like(k)).long()
        val = torch.where(ok, val, torch.zeros_like(val))
        Y[a0:a0 + t.shape[0]].scatter_add_(1, idx.reshape(t.shape[0], -1), val.reshape(t.shape[0], -1))
    Y += noise * torch.randn(Y.shape, generator=gen, dtype=torch.float64, device=gen.device)
    return Y.to(dtype)


def _grid_xz(nx, nz, dx, x0, z0):
    x = x0 + dx * np.arange(nx)
    z = z0 + dx * np.arange(nz)
    X, Z = np.meshgrid(x, z)                       # z fastest in Fortran order
    return np.c_[X.ravel("F"), np.zeros(X.size), Z.ravel("F")]


# --------------------------------------------------------------------- configs

def c1():
    L = 16
    G = linear_array(L, 1.0)
    z = np.linspace(20.0, 60.0, 64)
    Ro = np.c_[np.zeros(64), np.zeros(64), z]
    delay = 1000.0 * z / 1500.0 - 0.5
    return dict(Ro=Ro, a=0.9, b=10.0, dx=0.05, dy=0.05, dt=0.025, nt=400, delay=delay, v=1.0,
                cp=1500.0, G=G, err_level="stop")


def c2(seed=2, device="cpu", dtype=torch.float32):
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device).manual_seed(seed)
    L, pitch, K, fs, cp = 64, 0.5, 2048, 50.0, 1500.0
    G = linear_array(L, pitch)
    S, amp = _scatterers(rng, 20, (-16.0, 16.0), (10.0, 40.0))
    d = np.linalg.norm(G[:, None, :] - S[None, :, :], axis=2)          # (L, s) mm
    tau = (d[:, None, :] + d[None, :, :]).reshape(L * L, -1) / cp * 1000.0   # us, a = t*L + r
    Y = _echoes(tau, amp, K, fs, 0.01, gen, device, dtype)
    Ro = _grid_xz(256, 256, 0.125, -16.0, 10.0)
    return dict(method="tfm", Y=Y, Gt=G, Gr=G, Ro=Ro, dt=1.0 / fs, delay=0.0, cp=cp, K=K)


def c3(seed=3, n_problems=32):
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(n_problems):
        L = int(rng.integers(8, 129))
        a = rng.uniform(0.2, 1.0)
        b = rng.uniform(5.0, 15.0)
        G = linear_array(L, a + 0.1)
        No = int(round(math.exp(rng.uniform(math.log(1e3), math.log(6.4e4)))))
        nt = int(rng.integers(500, 4001))
        half = G[-1, 0] + a / 2
        Ro = np.c_[rng.uniform(-half, half, No), np.zeros(No), rng.uniform(5.0, 80.0, No)]
        cp, dt = 1500.0, 0.02
        # 'ud' focusing delays from a random focal point in the field
        # (the rule of m_files/arr_focus_delays.m:20-31: r_max - r_n in us).
        fx, fz = rng.uniform(-half, half), rng.uniform(5.0, 80.0)
        r = np.hypot(G[:, 0] - fx, fz)
        focal = (r.max() - r) * 1000.0 / cp
        # Start each SIR 0.5 us before the earliest possible arrival.
        dmin = np.sqrt(np.maximum(np.abs(Ro[:, 0:1] - G[:, 0][None, :]) - a / 2, 0.0) ** 2
                       + Ro[:, 2:3] ** 2).min(axis=1)
        delay = dmin * 1000.0 / cp - focal.max() - 0.5
        probs.append(dict(Ro=Ro, a=a, b=b, dx=0.05, dy=0.05, dt=dt, nt=nt, delay=delay, v=1.0,
                          cp=cp, G=G, foc_met="ud", focal=focal, apod_met="hann",
                          err_level="ignore"))
    return probs


def annu():
    """Annular array of the reference's own test (radii 3.5 .. 8 mm, four
    rings around a centre disc, x-focused at 25 mm; 4 MHz sampling, nt 300,
    dx = dy = 0.05 mm) over a 64 x 64 grid of points, x in [-10, 10] mm,
    z in [10, 50] mm."""
    gr = [3.5, 3.6, 5.0, 5.1, 6.0, 6.1, 7.0, 7.1, 8.0]
    x = np.linspace(-10.0, 10.0, 64)
    z = np.linspace(10.0, 50.0, 64)
    X, Z = np.meshgrid(x, z)
    Ro = np.c_[X.ravel("F"), np.zeros(X.size), Z.ravel("F")]
    return dict(Ro=Ro, Gr=gr, dx=0.05, dy=0.05, dt=0.25, nt=300, delay=0.0, v=1.0, cp=1000.0,
                foc_met="x", focal=25.0, err_level="ignore")


def sir_small_uneven_batch(seed=13, n_problems=128):
    """128 small dream_arr_rect problems of uneven size: 16-256 points, 4-16
    elements of 0.2-1.0 x 2-6 mm, nt 200-800, the rest as in c3()."""
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(n_problems):
        L = int(rng.integers(4, 17))
        a = rng.uniform(0.2, 1.0)
        b = rng.uniform(2.0, 6.0)
        G = linear_array(L, a + 0.1)
        No = int(rng.integers(16, 257))
        nt = int(rng.integers(200, 801))
        half = G[-1, 0] + a / 2
        Ro = np.c_[rng.uniform(-half, half, No), np.zeros(No), rng.uniform(5.0, 40.0, No)]
        cp, dt = 1500.0, 0.02
        fx, fz = rng.uniform(-half, half), rng.uniform(5.0, 40.0)
        r = np.hypot(G[:, 0] - fx, fz)
        focal = (r.max() - r) * 1000.0 / cp
        dmin = np.sqrt(np.maximum(np.abs(Ro[:, 0:1] - G[:, 0][None, :]) - a / 2, 0.0) ** 2
                       + Ro[:, 2:3] ** 2).min(axis=1)
        delay = dmin * 1000.0 / cp - focal.max() - 0.5
        probs.append(dict(Ro=Ro, a=a, b=b, dx=0.05, dy=0.05, dt=dt, nt=nt, delay=delay, v=1.0,
                          cp=cp, G=G, foc_met="ud", focal=focal, apod_met="hann",
                          err_level="ignore"))
    return probs


def das_uneven_batch(seed=14, n_problems=96, device="cpu", dtype=torch.float32):
    """96 small TFM problems of uneven size: linear arrays of 8-48 elements at
    0.5 mm pitch, A-scans of 512-2048 samples at 50 MHz, square pixel grids of
    16-80 pixels a side at 0.125 mm; random data (throughput only)."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device).manual_seed(seed)
    probs = []
    for _ in range(n_problems):
        L = int(rng.integers(8, 49))
        K = int(rng.integers(512, 2049))
        n = int(rng.integers(16, 81))
        G = linear_array(L, 0.5)
        Ro = _grid_xz(n, n, 0.125, -n * 0.0625, 10.0)
        Y = torch.randn((L * L, K), generator=gen, device=device, dtype=dtype)
        probs.append(dict(K=K, Ro=Ro, Gt=G, Gr=G, Y=Y, dt=0.02, delay=0.0, cp=1500.0))
    return probs


def c4(seed=4, device="cpu", dtype=torch.float32):
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device).manual_seed(seed)
    L, pitch, K, fs, cp = 128, 0.3, 4096, 50.0, 1500.0
    G = linear_array(L, pitch)
    S, amp = _scatterers(rng, 40, (-16.0, 16.0), (10.0, 40.0))
    d = np.linalg.norm(G[:, None, :] - S[None, :, :], axis=2)
    tau = 2.0 * d / cp * 1000.0
    Y = _echoes(tau, amp, K, fs, 0.01, gen, device, dtype)
    Ro = _grid_xz(512, 512, 0.06, -15.36, 10.0)
    return dict(method="saft", Y=Y, Gt=G, Gr=None, Ro=Ro, dt=1.0 / fs, delay=0.0, cp=cp, K=K,
                cp_target=1480.0)


def rca_arrays(L=128, pitch=0.2):
    x = (np.arange(L) - (L - 1) / 2.0) * pitch
    G_col = np.c_[x, np.zeros(L), np.zeros(L)]       # column electrodes along x
    G_row = np.c_[np.zeros(L), x, np.zeros(L)]       # row electrodes along y
    return G_col, G_row


def c5(seed=5, device="cpu", dtype=torch.float32, n_frames=8, L=128, K=4096, grid=(128, 128, 256)):
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device).manual_seed(seed)
    pitch, fs, cp = 0.2, 40.0, 1540.0
    G_col, G_row = rca_arrays(L, pitch)
    ext = L * pitch / 2
    S0, amp = _scatterers(rng, 100, (-ext, ext), (5.0, 45.0), yr=(-ext, ext))
    frames = []
    x_lo, x_hi, y_lo, y_hi = G_col[0, 0], G_col[-1, 0], G_row[0, 1], G_row[-1, 1]

    def gap(p, lo, hi):
        return np.where((p >= lo) & (p <= hi), 0.0, np.where(p < lo, lo - p, hi - p))

    for f in range(n_frames):
        S = S0 + (rng.normal(0.0, 0.05, S0.shape) if f > 0 else 0.0)
        # Column leg: x distance to the column, y clamped to the stripe; row leg
        # likewise (the geometry of extra_src/das.cc:464-505).
        dc = np.sqrt((G_col[:, None, 0] - S[None, :, 0]) ** 2 + gap(S[None, :, 1], y_lo, y_hi) ** 2
                     + S[None, :, 2] ** 2)                               # (Lc, s)
        dr = np.sqrt(gap(S[None, :, 0], x_lo, x_hi) ** 2 + (G_row[:, None, 1] - S[None, :, 1]) ** 2
                     + S[None, :, 2] ** 2)                               # (Lr, s)
        # Row transmit: A-scan r*Lc + c.
        tau = (dr[:, None, :] + dc[None, :, :]).reshape(L * L, -1) / cp * 1000.0
        frames.append(_echoes(tau, amp, K, fs, 0.02, gen, device, dtype))
    Y = torch.stack(frames)
    nx, ny, nz = grid
    x = -ext + (2 * ext / nx) * np.arange(nx)
    y = -ext + (2 * ext / ny) * np.arange(ny)
    z = 5.0 + (40.0 / nz) * np.arange(nz)
    X, Yg, Z = np.meshgrid(x, y, z, indexing="ij")
    Ro = np.c_[X.ravel("F"), Yg.ravel("F"), Z.ravel("F")]
    return dict(method="rca_rowtx", Y=Y, Gt=G_col, Gr=G_row, Ro=Ro, dt=1.0 / fs, delay=0.0, cp=cp,
                K=K)
