"""Seeded synthetic workloads for the benchmark configurations.

MGFM batch (interface face pairs of stiffened-gas states): left medium
gamma=4.4, P_inf=6e8 (water-like, the UX case constants UX:mpm_ini.cpp:80-88),
rho log-uniform [500,1500]; right medium gamma=1.4, P_inf=0, rho log-uniform
[0.01,10]; pressures log-uniform [1e3,1e10] Pa; velocities uniform
[-500,500] m/s.
"""
from __future__ import annotations

import numpy as np

MGFM_SEED = 0x5EED0003
MGFM_EOS = (4.4, 6e8, 1.4, 0.0)


def _logu(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def mgfm_faces(n: int, seed: int = MGFM_SEED) -> np.ndarray:
    """[6, n] float64 = (ul, ur, pl, pr, rhol, rhor)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((6, n), dtype=np.float64)
    out[0] = rng.uniform(-500.0, 500.0, n)
    out[1] = rng.uniform(-500.0, 500.0, n)
    out[2] = _logu(rng, 1e3, 1e10, n)
    out[3] = _logu(rng, 1e3, 1e10, n)
    out[4] = _logu(rng, 500.0, 1500.0, n)
    out[5] = _logu(rng, 0.01, 10.0, n)
    return out
