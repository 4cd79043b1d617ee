"""The das_uni restatement pinned against das.

das_uni (opencl/das_uni_double.cl) is das on generated uniform geometry:
for SAFT, TFM and column-transmit RCA (any delay) and for row-transmit RCA
with zero delay its result must equal the reference's compiled DAS<T>::das
(or, without the reference tree, the oracle's das, itself pinned to it) on
the explicitly generated pixel and element arrays, when every max parameter
lies on its grid (then the stripe extents coincide).  Row transmit with a
non-zero delay puts the delay on the other leg; that rounding difference is
pinned only by reading the kernel (opencl/das_uni_double.cl:344-346)."""
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
