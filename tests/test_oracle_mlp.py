"""The numpy restatement of the guess network: its gradient is the exact
derivative of its loss (central differences in FP64), and the BF16 emulation
rounds like the hardware format."""
import numpy as np

from oracle import mlp_ref as R


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = R.init_params(1)
    X = np.zeros((32, R.NPAD))
    X[:, :R.NIN] = rng.normal(size=(32, R.NIN))
    T = rng.normal(size=(32, R.NOUT))
    L, g = R.loss_grad(P, X, T, emulate=False)
    live = np.flatnonzero(R.live_mask())
    pick = rng.choice(live, 200, replace=False)
    # include every bias block and the output layer
    pick = np.unique(np.concatenate([pick, [R.OB0, R.OB1 + 5, R.OB2 + 63, R.OB3, R.OB3 + 1,
                                            R.OW3, R.OW3 + 1, R.OW0 + 9 * R.H + 3]]))
    h = 1e-6
    for p in pick:
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        fd = (R.loss(Pp, X, T) - R.loss(Pm, X, T)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-7 + 1e-5 * abs(fd), (p, fd, g[p])
    assert np.all(g[~R.live_mask()] == 0)
    assert L > 0


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-40])
    r = R.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0  # halfway: to even
    assert r[1] == 1.0  # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: even is 1
    assert r[2] == 1.0
    assert r[3] == 1.0 + 2 ** -7  # above halfway
    assert r[4] == -2.5
    # every result has at most 8 significant bits
    m, _ = np.frexp(r[np.isfinite(r) & (r != 0)])
    assert np.all((m * 2 ** 8) == np.round(m * 2 ** 8))
