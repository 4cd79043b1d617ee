"""The das_uni restatement pinned against das.

das_uni (opencl/das_uni_double.cl) is das on generated uniform geometry:
for SAFT, TFM and column-transmit RCA (any delay) and for row-transmit RCA
with zero delay its result must equal the reference's compiled DAS<T>::das
(or, without the reference tree, the oracle's das, itself pinned to it) on
the explicitly generated pixel and element arrays, when every max parameter
lies on its grid (then the stripe extents coincide).  Row transmit with a
non-zero delay puts the delay on the other leg (opencl/das_uni_double.cl:
344-346): it equals das there except, possibly, where a pair's exact
position lies within rounding distance of a sample boundary (checked in
extended precision); which of the two roundings the reference's OpenCL
kernel makes there is implementation-defined (native_sqrt)."""
import numpy as np
import pytest

from oracle import oracle as O


def _geometry(method, gt, gr, Ro3, T):
    x, y, z = O.uni_pixels(Ro3, T)
    Ro = np.c_[x, y, z].astype(np.float64)
    Lt = O.uni_count(*gt, dtype=T)
    et = (T(gt[0]) + np.arange(Lt).astype(T) * T(gt[1])).astype(np.float64)
    Gt = np.c_[et, np.zeros(Lt), np.zeros(Lt)]
    if method == "saft":
        return Ro, Gt, None
    Lr = O.uni_count(*gr, dtype=T)
    er = (T(gr[0]) + np.arange(Lr).astype(T) * T(gr[1])).astype(np.float64)
    Gr = (np.c_[er, np.zeros(Lr), np.zeros(Lr)] if method == "tfm"
          else np.c_[np.zeros(Lr), er, np.zeros(Lr)])
    return Ro, Gt, Gr


@pytest.mark.parametrize("method,delay", [("saft", 0.3), ("tfm", 0.2), ("rca_coltx", 0.25),
                                          ("rca_rowtx", 0.0)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_das_uni_equals_das_on_generated_geometry(method, delay, prec):
    T = np.float64 if prec == "f64" else np.float32
    gt = [-2.0, 0.5, 2.0]
    # das reads the RCA stripe extents with the other array's length
    # (extra_src/das.cc:464-467), right only for square arrays.
    gr = [-2.0, 0.5, 2.0] if method.startswith("rca") else [-1.5, 0.5, 1.5]
    Ro3 = np.array([[-3.0, 0.25, 3.0], [-1.0, 0.5, 1.0], [2.0, 0.25, 6.0]])
    Ro, Gt, Gr = _geometry(method, gt, gr, Ro3, T)
    n_ascans = Gt.shape[0] * (1 if Gr is None else Gr.shape[0])
    K = 600
    rng = np.random.default_rng(3)
    Y = rng.standard_normal((K, n_ascans)).astype(T)
    got = O.das_uni(Y, gt, gr, Ro3, 0.02, delay, 1500.0, method, prec)
    ref, _ = O.das(Y, Gt, Gr, Ro, 0.02, delay, 1500.0, method, "ignore",
                   "f64" if prec == "f64" else "f32", nthreads=4, use_ref=O.have_ref())
    assert np.array_equal(got.astype(np.float64), np.asarray(ref, dtype=np.float64))


def test_uniform_counts_truncate():
    assert O.uni_count(0.0, 0.3, 1.0) == 4          # 0, 0.3, 0.6, 0.9
    assert O.uni_count(-1.0, 0.5, 1.0) == 5
    x, y, z = O.uni_pixels(np.array([[0.0, 1.0, 2.0], [0.0, 1.0, 1.0], [5.0, 1.0, 6.0]]))
    # ny fastest, then nz, then nx.
    assert list(y[:4]) == [0.0, 1.0, 0.0, 1.0] and list(z[:4]) == [5.0, 5.0, 6.0, 6.0]
    assert list(x[:5]) == [0.0, 0.0, 0.0, 0.0, 1.0]


def test_row_transmit_delay_agrees_with_das_up_to_rounding():
    """das_uni's row transmit adds the delay to the row leg
    (opencl/das_uni_double.cl:344-346), das's to the column leg
    (extra_src/das.cc:602): the same travel time t_c + t_r + delay rounded in
    another order.  On ramp data (Y[k, a] = k, so a pixel's sum changes with
    any index) the two agree at every pixel except those with a pair whose
    exact position (t_c + t_r + delay) * fs lies within rounding distance of
    a sample boundary -- checked in extended precision.  A delay applied
    twice, with the wrong sign or not at all moves positions by ~15 samples
    and fails at most pixels."""
    T = np.float64
    gt = [-2.0, 0.5, 2.0]
    Ro3 = np.array([[-3.0, 0.25, 3.0], [-1.0, 0.5, 1.0], [2.0, 0.25, 6.0]])
    delay, dt, cp, K = 0.37, 0.02, 1500.0, 600
    Ro, Gt, Gr = _geometry("rca_rowtx", gt, gt, Ro3, T)
    n_ascans = Gt.shape[0] * Gr.shape[0]
    Y = np.repeat(np.arange(K, dtype=T)[:, None], n_ascans, axis=1)
    got = O.das_uni(Y, gt, gt, Ro3, dt, delay, cp, "rca_rowtx", "f64")
    ref, _ = O.das(Y, Gt, Gr, Ro, dt, delay, cp, "rca_rowtx", "ignore", "f64", nthreads=4,
                   use_ref=O.have_ref())
    differ = got != ref
    # Pairs whose exact position lies within 1e-9 samples of an integer.
    L = np.longdouble
    x, y, z = (Ro[:, c].astype(L) for c in range(3))
    lo, hi = L(gt[0]), L(gt[2])

    def gap(p):
        return np.where((p >= lo) & (p <= hi), L(0), np.where(p < lo, lo - p, hi - p))

    fs = L(1000) / L(dt)
    near = np.zeros(Ro.shape[0], dtype=bool)
    for c in Gt[:, 0].astype(L):
        tc = np.sqrt((c - x) ** 2 + gap(y) ** 2 + z ** 2) / L(cp)
        for r in Gr[:, 1].astype(L):
            tr = np.sqrt(gap(x) ** 2 + (r - y) ** 2 + z ** 2) / L(cp)
            u = (tc + tr + L(delay) / L(1000)) * fs
            near |= np.abs(u - np.rint(u)) < L(1e-9)
    assert near.mean() < 0.05
    assert not np.any(differ & ~near), np.flatnonzero(differ & ~near)[:10]
    # The wrong delay rules are not rounding-level.
    for wrong in (2 * delay, -delay, 0.0):
        bad = O.das_uni(Y, gt, gt, Ro3, dt, wrong, cp, "rca_rowtx", "f64")
        assert np.mean(bad != ref) > 0.5
