"""include/dream_hip.hpp compiles against the reference's own dream.h (the
header its gateways include) and links against libdream_hip.so; on the GPU
every drop-in entry point (DAS<T>::das for all methods, cl_das,
ArrRect::dream_arr_rect, Rect::dreamrect, Rect::cl_dreamrect, the `stop`
status rule and cancellation) gives the oracle's results.

The consumer tests/cxx/dropin_check.cc is built by `make -C tests/cxx` (run
by the build entry point where the DREAM tree is present); the GPU test runs
that binary and restates its inputs here."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = os.path.join(os.environ.get("DREAM_REFERENCE", "/root/reference"), "include")
SRC = os.path.join(ROOT, "tests", "cxx", "dropin_check.cc")
EXE = os.path.join(ROOT, "tests", "cxx", "_build", "dropin_check")
METHODS = ("saft", "tfm", "rca_coltx", "rca_rowtx")


def _build(tmp_path):
    exe = str(tmp_path / "dropin_check")
    cmd = ["g++", "-std=c++14", "-O1", "-I" + REF_INC, "-I" + os.path.join(ROOT, "include"),
           "-I/opt/rocm/include", SRC, "-o", exe, "-L" + os.path.join(ROOT, "dream_amd"),
           "-ldream_hip", "-Wl,-rpath," + os.path.join(ROOT, "dream_amd")]
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
def test_dropin_header_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    assert out.startswith("created") or out.startswith("no-device")


# --------------------------------------------------------------- GPU run

def _data(K, n_ascans, T):
    k = np.arange(K)[:, None]
    a = np.arange(n_ascans)[None, :]
    return ((((k + 3 * a) * 37) % 101 - 50).astype(T) / T(8))


def _array(L, dim, T):
    G = np.zeros((L, 3), dtype=T)
    G[:, dim] = (np.arange(L).astype(T) - T(L - 1) / T(2)) * T(0.5)
    return G


def _pixels(T):
    n = np.arange(64)
    return np.c_[((n % 8).astype(T) - T(3.5)) * T(0.375),
                 ((n // 8) % 3).astype(T) * T(0.25) - T(0.25),
                 T(2) + (n // 8).astype(T) * T(0.5)].astype(T)


def _records(out):
    rec = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) >= 2 and parts[0] in ("das64", "das32", "cl_das64", "cl_das32"):
            rec[(parts[0], parts[1])] = (int(parts[2]), np.array([float(v) for v in parts[3:]]))
        elif parts:
            rec[parts[0]] = [float(v) for v in parts[1:]]
    return rec


@pytest.mark.gpu
def test_dropin_classes_compute_the_oracle_results(gpu, oracle):
    if not os.path.exists(EXE):
        pytest.skip("tests/cxx/_build/dropin_check not built (make -C tests/cxx)")
    proc = subprocess.run([EXE, "run"], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr
    rec = _records(proc.stdout)
    K, L = 256, 4
    for tag, prec, T, tol in (("das64", "f64", np.float64, 1e-13), ("das32", "f32", np.float32,
                                                                     2e-6)):
        Ro = _pixels(T)
        for m in METHODS:
            saft = m == "saft"
            Gt = _array(L, 0, T)
            Gr = _array(L, 0 if m == "tfm" else 1, T)
            Y = _data(K, L if saft else L * L, T)
            ref, err = oracle.das(Y, Gt, None if saft else Gr, Ro, T(0.02), T(0.125), T(1500), m,
                                  "ignore", prec, nthreads=4)
            code, Im = rec[(tag, m)]
            scale = np.abs(Y).max() * (L if saft else L * L)
            assert np.max(np.abs(Im - ref)) <= tol * scale, (tag, m)
            assert code == err, (tag, m)
            if m == "tfm":
                rc, Im2 = rec[("cl_" + tag, m)]
                assert rc == 0 and np.max(np.abs(Im2 - ref)) <= tol * scale
    # dream_arr_rect: x focusing at 20 mm, Hann apodization, stop level.
    Ro = np.array([[0.0, 0.0, 3.0], [0.5, 0.25, 3.5], [-0.75, 0.0, 4.0]])
    G = np.array([[-1.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0]])
    H, e, _ = oracle.dream_arr_rect(Ro, 0.5, 2.0, 0.05, 0.05, 0.025, 200, 1.5, 1.0, 1500.0, G,
                                    "x", 20.0, apod_met="hann", err_level="stop", nthreads=2)
    got = np.array(rec["arr_rect"][1:]).reshape(3, 200).T
    assert int(rec["arr_rect"][0]) == e
    assert np.max(np.abs(got - H)) <= 1e-12 * np.abs(H).max()
    assert rec["arr_rect_running"] == [1.0]
    H1, e1, _ = oracle.dreamrect(Ro, 0.5, 2.0, 0.05, 0.05, 0.025, 200, 1.5, 1.0, 1500.0,
                                 err_level="stop", nthreads=2)
    for key, code in (("rect", e1), ("cl_rect", 0)):
        got = np.array(rec[key][1:]).reshape(3, 200).T
        assert int(rec[key][0]) == code
        assert np.max(np.abs(got - H1)) <= 1e-12 * np.abs(H1).max()
    # stop: out_of_bounds and running cleared; the next call sets it again.
    assert rec["stop"] == [3.0, 0.0]
    assert rec["after_stop"] == [0.0, 1.0]
    # Cancelled between chunks: das returns, is_running() reports it.
    assert rec["cancel"] == [0.0, 0.0]
    assert rec["set_running"] == [1.0]
