"""GPU parity of the cylindrically-curved-element SIR (dreamcylind,
dream_arr_cylind) against the oracle.  Bars: FP64 output within 1e-12 of the
largest sample (a moved sample index moves a whole contribution, far above
that bar); FP32 2e-5 relative L2; error codes as in the reference."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu

CP = 1500.0


@pytest.mark.parametrize("Rc", [25.0, -25.0])
def test_dreamcylind_module(gpu, oracle, Rc):
    from dream_amd.dreamcylind import dreamcylind
    rng = np.random.default_rng(21)
    No = 40
    Ro = np.c_[rng.uniform(-3, 3, No), rng.uniform(-3, 3, No), rng.uniform(5, 30, No)]
    d = Ro[:, 2] * 1000.0 / CP - 1.0
    H = dreamcylind(Ro, [1.5, 8.0, Rc], [0.05, 0.05, 0.02, 700], d, [1.0, CP, 0.0], "stop")
    ref, e, _ = oracle.dreamcylind(Ro, 1.5, 8.0, Rc, 0.05, 0.05, 0.02, 700, d, 1.0, CP,
                                   nthreads=2)
    assert e == 0
    assert H.shape == (700, No) and H.flags.f_contiguous
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("foc,steer,apod", [("off", "off", "off"), ("x", "x", "hann"),
                                            ("ud", "off", "ud"), ("xy", "y", "gauss")])
def test_dream_arr_cylind_module(gpu, oracle, foc, steer, apod):
    from dream_amd.dream_arr_cylind import dream_arr_cylind
    rng = np.random.default_rng(22)
    L, No = 8, 30
    G = np.c_[np.linspace(-3.5, 3.5, L), np.zeros(L), np.zeros(L)]
    Ro = np.c_[rng.uniform(-5, 5, No), rng.uniform(-2, 2, No), rng.uniform(10, 40, No)]
    d = Ro[:, 2] * 1000.0 / CP - 3.0
    focal = rng.uniform(0.0, 0.5, L) if foc == "ud" else 30.0
    w = rng.uniform(0.3, 1.0, L)
    H = dream_arr_cylind(Ro, [0.9, 6.0, 40.0], G, [0.05, 0.05, 0.025, 800], d, [1.0, CP, 0.0],
                         foc, focal, steer, [10.0, 5.0], apod, w, 1.0, "ignore")
    ref, _, _ = oracle.dream_arr_cylind(Ro, 0.9, 6.0, 40.0, 0.05, 0.05, 0.025, 800, d, 1.0, CP,
                                        G, foc_met=foc, focal=focal, steer_met=steer,
                                        theta=10.0, phi=5.0, apod_met=apod, apod=w,
                                        apod_par=0.0, err_level="ignore", nthreads=2)
    assert np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()


def test_cylind_fp32_device_path_and_zero_geometry(gpu, oracle):
    from dream_amd.torch_ops import arr_cylind
    rng = np.random.default_rng(23)
    No = 64
    Ro = np.c_[rng.uniform(-4, 4, No), rng.uniform(-2, 2, No), rng.uniform(8, 25, No)]
    d = Ro[:, 2] * 1000.0 / CP - 2.0
    G = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    ref, _, _ = oracle.dream_arr_cylind(Ro, 1.0, 5.0, -12.0, 0.05, 0.05, 0.02, 900, d, 1.0, CP,
                                        G, err_level="ignore", nthreads=2)
    Ro_d = torch.tensor(Ro.T.copy(), device="cuda")
    d_d = torch.tensor(d, device="cuda")
    H, code, _ = arr_cylind(Ro_d, 1.0, 5.0, -12.0, 0.05, 0.05, 0.02, 900, d_d, 1.0, CP, G,
                            out_dtype=torch.float32)
    torch.cuda.synchronize()
    assert code == 0 and rel_l2(H.cpu().numpy().T, ref) <= 2e-5
    # b > 2|Rcurv|: the reference's row angles are NaN and the response is zero.
    Z, code, oob = arr_cylind(Ro_d, 1.0, 9.0, 4.0, 0.05, 0.05, 0.02, 900, d_d, 1.0, CP, G)
    torch.cuda.synchronize()
    assert code == 0 and oob == 0 and torch.count_nonzero(Z).item() == 0


def test_cylind_in_a_mixed_batch(gpu, oracle):
    from dream_amd.torch_ops import SirBatch
    rng = np.random.default_rng(24)
    probs, refs = [], []
    for i in range(6):
        L, No = int(rng.integers(1, 5)), int(rng.integers(20, 70))
        G = np.c_[np.linspace(-2, 2, L), np.zeros(L), np.zeros(L)]
        Ro = np.c_[rng.uniform(-3, 3, No), np.zeros(No), rng.uniform(6, 20, No)]
        d = Ro[:, 2] * 1000.0 / CP - 1.0
        p = dict(Ro=torch.tensor(Ro.T.copy(), device="cuda"), dx=0.05, dy=0.05, dt=0.02, nt=600,
                 delay=torch.tensor(d, device="cuda"), v=1.0, cp=CP, G=G)
        if i % 2:
            Rc = 15.0 * (1 if i % 4 == 1 else -1)
            p.update(shape="cylind", a=0.7, b=4.0, Rcurv=Rc)
            ref, _, _ = oracle.dream_arr_cylind(Ro, 0.7, 4.0, Rc, 0.05, 0.05, 0.02, 600, d, 1.0,
                                                CP, G, err_level="ignore", nthreads=2)
        else:
            p.update(a=0.6, b=2.0)
            ref, _, _ = oracle.dream_arr_rect(Ro, 0.6, 2.0, 0.05, 0.05, 0.02, 600, d, 1.0, CP,
                                              G, err_level="ignore", nthreads=2)
        probs.append(p)
        refs.append(ref)
    batch = SirBatch(probs, out_dtype=torch.float64)
    for H, ref in zip(batch.run(use_graph=True), refs):
        torch.cuda.synchronize()
        assert np.max(np.abs(H.cpu().numpy().T - ref)) <= 1e-12 * np.abs(ref).max()


def test_cylind_out_of_bounds_codes(gpu, oracle):
    from dream_amd.dreamcylind import dreamcylind
    Ro = np.array([[0.0, 0.0, 10.0], [2.0, 0.0, 10.0]])
    args = (Ro, [2.0, 4.0, 15.0], [0.05, 0.05, 0.02, 90], 5.0, [1.0, CP, 0.0])
    with pytest.raises(RuntimeError):
        dreamcylind(*args, "stop")
    H = dreamcylind(*args, "warn")
    ref, e, _ = oracle.dreamcylind(Ro, 2.0, 4.0, 15.0, 0.05, 0.05, 0.02, 90, 5.0, 1.0, CP,
                                   err_level="warn", nthreads=2)
    assert e == 2 and np.max(np.abs(H - ref)) <= 1e-12 * np.abs(ref).max()
