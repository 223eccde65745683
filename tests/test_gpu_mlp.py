"""GPU kernels of the learned MGFM guess against the numpy restatement
(oracle/mlp_ref.py, BF16 roundings emulated, FP64 accumulation), and what
the trained network gains over the reference's initial guess."""
import numpy as np
import pytest

from oracle import mlp_ref as R
from pgdg_p2_amd import workloads


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, R.NPAD), dtype=np.float32)
    X[:, :R.NIN] = rng.normal(size=(n, R.NIN))
    T = rng.normal(size=(n, R.NOUT)).astype(np.float32)
    return X, T


@pytest.mark.gpu
def test_mlp_forward_matches_restatement(cuda):
    from pgdg_p2_amd.mlp import Mlp
    m = Mlp()
    P = R.init_params(7).astype(np.float32)
    m.set_params(P)
    X, _ = _data(1000, 1)  # not a multiple of 64: ragged last tile
    print("mlp forward n=1000", flush=True)
    y = m.forward(X)
    yr, _ = R.forward(P.astype(np.float64), X.astype(np.float64), emulate=True)
    err = np.abs(y - yr)
    print("forward max err", err.max(), "mean", err.mean())
    assert err.mean() <= 1e-4 and err.max() <= 2e-2
    m.close()


@pytest.mark.gpu
def test_mlp_gradient_matches_restatement(cuda):
    from pgdg_p2_amd.mlp import Mlp
    m = Mlp()
    P = R.init_params(8).astype(np.float32)
    m.set_params(P)
    X, T = _data(4096, 2)
    print("mlp loss_grad n=4096", flush=True)
    L, g = m.loss_grad(X, T)
    Lr, gr = R.loss_grad(P.astype(np.float64), X.astype(np.float64), T.astype(np.float64),
                         emulate=True)
    assert abs(L - Lr) <= 1e-4 * abs(Lr)
    for name, a, b in (("W0", R.OW0, R.OB0), ("b0", R.OB0, R.OW1), ("W1", R.OW1, R.OB1),
                       ("b1", R.OB1, R.OW2), ("W2", R.OW2, R.OB2), ("b2", R.OB2, R.OW3),
                       ("W3", R.OW3, R.OB3), ("b3", R.OB3, R.NPARAM)):
        rel = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        print(f"grad {name}: rel err {rel:.2e}")
        assert rel <= 5e-3, name
    assert np.all(g[~R.live_mask()] == 0)
    m.close()


@pytest.mark.gpu
def test_learned_guess_reduces_newton_iterations(cuda):
    import pgdg_p2_amd as P
    from pgdg_p2_amd.mlp import Mlp
    eos = workloads.MGFM_EOS
    ntr, nte = 1 << 17, 100_000
    d_tr = P.DeviceArray.from_numpy(workloads.mgfm_faces(ntr, seed=0x5EED0004))
    m = Mlp()
    print(f"mlp train n={ntr}", flush=True)
    hist = m.train(d_tr, ntr, eos, epochs=6, batch=2048, lr=3e-3)
    print("loss first/last", hist[:4], hist[-4:])
    assert hist[-16:].mean() < 0.1 * hist[:16].mean()
    inp = workloads.mgfm_faces(nte, seed=0x5EED0005)
    d_in = P.DeviceArray.from_numpy(inp)
    d_out = P.DeviceArray((3, nte), np.float64)
    d_it = P.DeviceArray((nte,), np.int32)
    P.mgfm_batch_dev(d_in, d_out, d_it, nte, *eos, tol_mode=P.TOL_RELATIVE, rtol=1e-12,
                     max_iter=100)
    P.device_sync()
    out0, it0 = d_out.numpy(), d_it.numpy()
    m.mgfm_dev(d_in, d_out, d_it, nte, eos)
    P.device_sync()
    out1, it1 = d_out.numpy(), d_it.numpy()
    print(f"its/face plain {it0.mean():.3f} learned {it1.mean():.3f}")
    ok = (it0 < 100) & (it1 < 100)
    assert ok.mean() > 0.999
    rel = np.abs(out1[1][ok] - out0[1][ok]) / np.abs(out0[1][ok])
    print("p* max rel diff", rel.max())
    assert rel.max() <= 1e-10
    assert it1.mean() < it0.mean() - 2.0
    m.close()


@pytest.mark.gpu
def test_guess_kernel_matches_restatement(cuda):
    # the inference kernel (transposed weights, persistent tiles) against the
    # restatement's forward pass on the same standardised features
    import ctypes as C
    import pgdg_p2_amd as P
    from pgdg_p2_amd.mlp import Mlp, MlpNorm
    eos = workloads.MGFM_EOS
    n = 5000
    inp = workloads.mgfm_faces(n, seed=21)
    m = Mlp()
    Pm = R.init_params(9).astype(np.float32)
    nm = MlpNorm()
    mu = [15.0, 15.0, 3.0, -1.0, 0.0, 0.0, 0.0, 0.0, 4.4, 20.0, 1.4, 0.0]
    sd = [4.0, 4.0, 1.0, 2.0, 300.0, 300.0, 0.5, 400.0, 1.0, 1.0, 1.0, 1.0]
    nm.feat_mu[:], nm.feat_sd[:] = mu, sd
    nm.out_mu[:], nm.out_sd[:] = [-1.0, 0.0], [2.0, 100.0]
    m.set_params(Pm, nm)
    d_in = P.DeviceArray.from_numpy(inp)
    d_g = P.DeviceArray((n,), np.float64)
    m.guess_dev(d_in, n, eos, d_g)
    P.device_sync()
    g = d_g.numpy()
    f32 = np.float32
    raw = np.stack([np.log(np.maximum(inp[2].astype(f32), f32(1e-30))),
                    np.log(np.maximum(inp[3].astype(f32), f32(1e-30))),
                    np.log(np.maximum(inp[4].astype(f32), f32(1e-30))),
                    np.log(np.maximum(inp[5].astype(f32), f32(1e-30))),
                    inp[0].astype(f32), inp[1].astype(f32),
                    (inp[2] - inp[3]).astype(f32) * (1 / (inp[2] + inp[3]).astype(f32)),
                    (inp[0] - inp[1]).astype(f32),
                    np.full(n, eos[0], f32), np.full(n, np.log1p(eos[1]), f32),
                    np.full(n, eos[2], f32), np.full(n, np.log1p(eos[3]), f32)], axis=1)
    X = np.zeros((n, R.NPAD))
    X[:, :R.NIN] = (raw - np.array(mu, f32)) / np.array(sd, f32)
    y, _ = R.forward(Pm.astype(np.float64), X, emulate=True)
    pm = np.maximum(inp[2], inp[3])
    gr = pm * np.exp(y[:, 0] * 2.0 - 1.0)
    rel = np.abs(g - gr) / gr
    print("guess rel err max", rel.max(), "mean", rel.mean())
    assert np.median(rel) <= 1e-5 and rel.max() <= 0.05
    m.close()
