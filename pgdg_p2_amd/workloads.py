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


# ---------------------------------------------------------------------------
# Isolated MLS P2G + per-cell least-squares solve: n^3 cells on [0,1]^3,
# Poisson(8) particles per cell clipped to [1, 27], uniform in the cell; six
# fields (rho, rho u, rho v, rho w, E, level set): the first five each a sum
# of 8 seeded random Fourier modes plus a unit jump across the sphere
# |x - (1/2,1/2,1/2)| = 0.3, the sixth the level set of that sphere (signed
# distance 0.3 - |x - (1/2,1/2,1/2)|, positive inside).
# ---------------------------------------------------------------------------
MLS_SEED = 0x5EED0001
MLS_FIELDS = 6


def mls_fields(x, nfields: int = MLS_FIELDS, seed: int = MLS_SEED):
    """The seeded fields at points x[3][m]: 8 Fourier modes each plus the unit
    jump inside |x - (1/2,1/2,1/2)| < 0.3; with six fields the last is the
    sphere's level set 0.3 - |x - (1/2,1/2,1/2)|.  Returns [nfields][m]."""
    x = np.asarray(x, dtype=np.float64)
    r = np.sqrt((x[0] - 0.5) ** 2 + (x[1] - 0.5) ** 2 + (x[2] - 0.5) ** 2)
    jump = (r < 0.3).astype(np.float64)
    val = np.empty((nfields, x.shape[1]))
    for f in range(nfields):
        if nfields == MLS_FIELDS and f == MLS_FIELDS - 1:
            val[f] = 0.3 - r  # level set of the interface
            continue
        fr = np.random.Generator(np.random.PCG64(seed + f))
        a = fr.uniform(-0.5, 0.5, 8)
        kv = fr.integers(-3, 4, (8, 3))
        ph = fr.uniform(0, 2 * np.pi, 8)
        acc = jump.copy()
        for m in range(8):
            arg = 2 * np.pi * (kv[m, 0] * x[0] + kv[m, 1] * x[1] + kv[m, 2] * x[2]) + ph[m]
            acc += a[m] * np.sin(arg)
        val[f] = acc
    return val


def mls_particles(n: int = 128, seed: int = MLS_SEED, nfields: int = MLS_FIELDS):
    """Cell-sorted particles.  Returns dict with
    off[nc+1] (int64), cell[np] (int32), u[3][np] (cell-local coordinates in
    [0,1)), x[3][np] (absolute), vol[np], val[nf][np], dx, n."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nc = n * n * n
    cnt = np.clip(rng.poisson(8.0, nc), 1, 27).astype(np.int64)
    off = np.zeros(nc + 1, dtype=np.int64)
    np.cumsum(cnt, out=off[1:])
    npart = int(off[-1])
    cell = np.repeat(np.arange(nc, dtype=np.int64), cnt)
    u = rng.random((3, npart))
    dx = 1.0 / n
    k = cell % n
    j = (cell // n) % n
    i = cell // (n * n)
    x = np.empty((3, npart))
    x[0] = (i + u[0]) * dx
    x[1] = (j + u[1]) * dx
    x[2] = (k + u[2]) * dx
    vol = (dx * dx * dx) / cnt[cell]
    val = mls_fields(x, nfields, seed)
    return {"n": n, "dx": dx, "off": off, "cell": cell.astype(np.int32), "u": u, "x": x,
            "vol": vol, "val": val}
