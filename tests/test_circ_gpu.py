"""GPU parity of the disc SIR (dreamcirc, dream_arr_circ) against the oracle.
Bars: FP64 output within 1e-12 of the largest sample (an index that moves
moves a whole contribution, far above that bar); FP32 2e-5 relative L2;
error codes as in the reference."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_dreamcirc_module(gpu, oracle):
    from dream_amd.dreamcirc import dreamcirc
    rng = np.random.default_rng(1)
    No = 40
    Ro = np.c_[rng.uniform(-3, 3, No), rng.uniform(-3, 3, No), rng.uniform(5, 15, No)]
    d = Ro[:, 2] * 1000.0 / 1500.0 - 0.5
    H = dreamcirc(Ro, [1.5], [0.05, 0.05, 0.02, 500], d, [1.0, 1500.0, 0.0], "stop")
    ref, e, _ = oracle.dreamcirc(Ro, 1.5, 0.05, 0.05, 0.02, 500, d, 1.0, 1500.0, nthreads=2)
    assert e == 0
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("foc,apod", [("off", "off"), ("xy", "hann"), ("ud", "gauss")])
def test_dream_arr_circ_module(gpu, oracle, foc, apod):
    from dream_amd.dream_arr_circ import dream_arr_circ
    rng = np.random.default_rng(2)
    L, No = 7, 30
    G = np.c_[np.linspace(-4, 4, L), np.linspace(-1, 1, L), np.zeros(L)]
    Ro = np.c_[rng.uniform(-5, 5, No), rng.uniform(-2, 2, No), rng.uniform(8, 20, No)]
    d = Ro[:, 2] * 1000.0 / 1500.0 - 1.5
    focal = rng.uniform(0.0, 0.3, L) if foc == "ud" else 25.0
    H = dream_arr_circ(Ro, [0.6], G, [0.04, 0.04, 0.025, 600], d, [1.0, 1500.0, 0.0], foc,
                       focal, "off", [0.0, 0.0], apod, None, 1.5, "ignore")
    # The Python gateway never reads win_par: apod_par stays 0
    # (include/arg_parser_python.h:536-540).
    ref, _, _ = oracle.dream_arr_circ(Ro, 0.6, 0.04, 0.04, 0.025, 600, d, 1.0, 1500.0, G,
                                      foc_met=foc, focal=focal, apod_met=apod, apod_par=0.0,
                                      err_level="ignore", nthreads=2)
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()


def test_arr_circ_out_of_bounds_codes(gpu, oracle):
    from dream_amd.dream_arr_circ import dream_arr_circ
    G = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    # Two points, so the oracle runs two threads: the reference reads no
    # status back from a single-thread run (src/dream_arr_rect.cc:405-425).
    Ro = np.array([[0.0, 0.0, 10.0], [0.1, 0.0, 10.0]])
    args = (Ro, [1.0], G, [0.05, 0.05, 0.02, 50], 6.0, [1.0, 1500.0, 0.0], "off", 0.0, "off",
            [0.0, 0.0], "off", None, 0.0)
    with pytest.raises(RuntimeError):
        dream_arr_circ(*args, "stop")
    H = dream_arr_circ(*args, "ignore")
    ref, e, _ = oracle.dream_arr_circ(Ro, 1.0, 0.05, 0.05, 0.02, 50, 6.0, 1.0, 1500.0, G,
                                      err_level="ignore", nthreads=2)
    assert e == 1 and np.max(np.abs(H - ref)) <= 1e-12 * max(np.abs(ref).max(), 1.0)


def test_mixed_rect_and_disc_batch(gpu, oracle):
    """Rectangles and discs in one persistent launch, FP32 output."""
    from dream_amd.torch_ops import SirBatch
    rng = np.random.default_rng(3)
    probs, refs = [], []
    for i in range(6):
        L, No = int(rng.integers(2, 6)), int(rng.integers(20, 80))
        G = np.c_[np.linspace(-2, 2, L), np.zeros(L), np.zeros(L)]
        Ro = np.c_[rng.uniform(-3, 3, No), np.zeros(No), rng.uniform(6, 12, No)]
        d = Ro[:, 2] * 1000.0 / 1500.0 - 1.0
        p = dict(Ro=torch.tensor(Ro.T.copy(), device="cuda"), dx=0.05, dy=0.05, dt=0.02, nt=400,
                 delay=torch.tensor(d, device="cuda"), v=1.0, cp=1500.0, G=G)
        if i % 2:
            p.update(shape="circ", R=0.5 + 0.1 * i)
            ref, _, _ = oracle.dream_arr_circ(Ro, p["R"], 0.05, 0.05, 0.02, 400, d, 1.0, 1500.0,
                                              G, err_level="ignore", nthreads=2)
        else:
            p.update(a=0.6, b=2.0)
            ref, _, _ = oracle.dream_arr_rect(Ro, 0.6, 2.0, 0.05, 0.05, 0.02, 400, d, 1.0,
                                              1500.0, G, err_level="ignore", nthreads=2)
        probs.append(p)
        refs.append(ref)
    batch = SirBatch(probs, out_dtype=torch.float32)
    for H, ref in zip(batch.run(use_graph=True), refs):
        torch.cuda.synchronize()
        assert rel_l2(H.cpu().numpy().T, ref) <= 2e-5
