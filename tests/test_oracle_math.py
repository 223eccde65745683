"""The deterministic pow(x, 1./D) shared by the oracle and the HIP library,
pinned against glibc's pow (the reference's function) and an extended
precision evaluation."""
import math

import numpy as np

from oracle import pyoracle


def test_root3_close_to_glibc_pow_and_correctly_rounded():
    rng = np.random.default_rng(7)
    x = np.exp(rng.uniform(-40, 10, 400_000))
    y = pyoracle.root_dim(x, 3)
    # the C library's pow, as the reference calls it (math.pow -> libm pow)
    g = np.array([math.pow(v, 1.0 / 3.0) for v in x])
    ulp = np.abs(y - g) / np.spacing(g)
    assert ulp.max() <= 1.0
    assert (y == g).mean() > 0.995      # libm pow misrounds ~0.1% of inputs
    # extended precision reference of x^(1./3.) (the double 1/3)
    xl = x.astype(np.longdouble)
    e = np.longdouble(1.0 / 3.0)
    yl = np.power(xl, e).astype(np.float64)
    assert (y == yl).mean() > 0.999


def test_root2_is_sqrt():
    x = np.exp(np.random.default_rng(8).uniform(-30, 30, 10000))
    np.testing.assert_array_equal(pyoracle.root_dim(x, 2), np.sqrt(x))


def test_root3_cubes():
    # x^(1./3.) of a cube k^3 is k (1 - 3 eps ln k) with 1./3. = 1/3 - eps
    # (eps = 1.85e-17): at most a few ulps below k, never above
    k = np.arange(1, 2000, dtype=np.float64)
    y = pyoracle.root_dim(k ** 3, 3)
    kl = k.astype(np.longdouble)
    expect = kl * (1 - 3 * np.longdouble(1.850371707708594e-17) * np.log(kl))
    assert np.all(y <= k)
    err = np.abs(y.astype(np.longdouble) - expect) / np.spacing(y).astype(np.longdouble)
    assert np.all(err <= 0.5001)       # correctly rounded
    g = np.array([math.pow(v, 1.0 / 3.0) for v in k ** 3])
    assert (y == g).mean() > 0.99
    assert pyoracle.root_dim(np.array([0.0]), 3)[0] == 0.0
