"""The DAS oracle against the reference's own das.cc (golden vectors, and the
compiled reference directly where oracle/_ref was built)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "das_ref.npz")

METHODS = ("saft", "tfm", "rca_coltx", "rca_rowtx")


def _cases():
    d = np.load(GOLD)
    names = sorted({k.split("__")[0] for k in d.files if k.endswith("__delay")})
    return d, names


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_oracle_matches_reference_golden(oracle, prec):
    d, names = _cases()
    assert len(names) == 16
    for n in names:
        m = next(x for x in METHODS if n.startswith(x + "_"))
        level = n.rsplit("_", 1)[1]
        dt, cp = d[f"{n}__scalars"]
        delay = d[f"{n}__delay"]
        Im, err = oracle.das(d[f"{m}__Y"], d[f"{m}__Gt"], d[f"{m}__Gr"], d[f"{m}__Ro"], dt,
                             delay if delay.size > 1 else float(delay[0]), cp, m, level, prec,
                             nthreads=3)
        ref = d[f"{n}__{prec}__Im"]
        assert Im.dtype == ref.dtype
        assert np.array_equal(Im, ref), n            # bit for bit, values and indices
        assert err == int(d[f"{n}__{prec}__err"]), n


def test_oracle_single_thread_drops_error_code(oracle):
    """The reference returns 'none' when it runs one thread (its status is
    never read back, extra_src/das.cc:270-290); the oracle keeps that."""
    d, _ = _cases()
    m = "tfm"
    Im, err = oracle.das(d[f"{m}__Y"], d[f"{m}__Gt"], d[f"{m}__Gr"], d[f"{m}__Ro"], 0.025, 0.37,
                         1480.0, m, "ignore", "f64", nthreads=1)
    assert err == 0
    Im2, err2 = oracle.das(d[f"{m}__Y"], d[f"{m}__Gt"], d[f"{m}__Gr"], d[f"{m}__Ro"], 0.025, 0.37,
                           1480.0, m, "ignore", "f64", nthreads=2)
    assert err2 == 1 and np.array_equal(Im, Im2)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref",
                                                    "libdream_ref.so")),
                    reason="reference checker (oracle/_ref) not built in this tree")
@pytest.mark.parametrize("method", METHODS)
def test_oracle_matches_reference_random(oracle, method):
    rng = np.random.default_rng(7)
    L = 7
    x = np.sort(rng.uniform(-4, 4, L))
    Gt = np.c_[x, rng.uniform(-0.1, 0.1, L), np.zeros(L)]
    Gr = Gt if method == "tfm" else np.c_[np.zeros(L), np.sort(rng.uniform(-4, 4, L)), np.zeros(L)]
    n_ascans = L if method == "saft" else L * L
    K = 700
    Y = rng.standard_normal((K, n_ascans))
    Ro = np.c_[rng.uniform(-6, 6, 300), rng.uniform(-6, 6, 300), rng.uniform(0.5, 9, 300)]
    for prec in ("f64", "f32"):
        for delay in (0.0, -0.2, rng.uniform(-1, 1, 300)):
            a, ea = oracle.das(Y, Gt, Gr, Ro, 0.02, delay, 1540.0, method, "warn", prec, nthreads=4)
            b, eb = oracle.das(Y, Gt, Gr, Ro, 0.02, delay, 1540.0, method, "warn", prec,
                               use_ref=True)
            assert np.array_equal(a, b) and ea == eb


def test_oracle_matches_reference_nonsquare_rca(oracle):
    """Non-square row-column arrays: the reference reads the stripe extents at
    rows[Lc], rows[2 Lc - 1] and columns[Lr - 1] (extra_src/das.cc:464-467,
    :560-563); the oracle must reproduce its vectors bit for bit."""
    d = np.load(os.path.join(HERE, "golden", "das_rca_nonsquare_ref.npz"))
    for key in ("c8r11", "c12r9"):
        Y, Gt, Gr, Ro = (d[f"{key}__{n}"] for n in ("Y", "Gt", "Gr", "Ro"))
        for method in ("rca_coltx", "rca_rowtx"):
            for prec in ("f64", "f32"):
                Im, err = oracle.das(Y, Gt, Gr, Ro, 0.025, 0.3, 1510.0, method, "warn", prec,
                                     nthreads=3)
                assert np.array_equal(Im, d[f"{key}__{method}__{prec}__Im"]), (key, method, prec)
                assert err == int(d[f"{key}__{method}__{prec}__err"])
    # The vectors are sensitive to that indexing: with 8 columns and 11 rows
    # the lower y extent is rows[8], the x of row 8, which no leg reads
    # (rows' legs use the stripe gap, y and z), so moving the rows' x
    # coordinates changes the reference's image only through the extent.
    Y, Gt, Gr, Ro = (d[f"c8r11__{n}"] for n in ("Y", "Gt", "Gr", "Ro"))
    Gr2 = Gr.copy()
    Gr2[:, 0] += 1.5
    Im2, _ = oracle.das(Y, Gt, Gr2, Ro, 0.025, 0.3, 1510.0, "rca_coltx", "warn", "f64",
                        nthreads=3)
    assert not np.array_equal(Im2, d["c8r11__rca_coltx__f64__Im"])


def test_oracle_refuses_rca_extents_outside_the_tables(oracle):
    rng = np.random.default_rng(1)
    Y = rng.standard_normal((50, 4 * 20))
    Gt = np.c_[np.linspace(-1, 1, 4), np.zeros(4), np.zeros(4)]
    Gr = np.c_[np.zeros(20), np.linspace(-1, 1, 20), np.zeros(20)]
    with pytest.raises(ValueError):
        oracle.das(Y, Gt, Gr, np.array([[0.0, 0.0, 1.0]]), 0.02, 0.0, 1500.0, "rca_coltx")
