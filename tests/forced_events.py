"""Initial states that force the 3-D topology events of the step.

The seeded lattice states of `pgdg_p2_amd.cases` never split or merge a
particle in 3-D within the few steps a parity test can afford (every particle
starts at the lattice volume).  The builders here take the 3-D shock-bubble
channel and change a few particles so that the first steps reach every branch
of Split_particle_fast and Merge_particle_fast (S3:mpm.cpp:1036-1417,
1536-1706):

* split: particles above the split volume 1.6 p_sz^3 (S3:mpm.cpp:1068); three
  of them share one cell, so that a per-cell upper bound has to refuse some;
* merge: the children of a split are 1/8 of their parent's volume, below the
  merge volume (2/3) p_sz^3 (S3:mpm.cpp:1540) and about 0.7 p_sz apart, so
  the next step's merge passes pair them up; a few close same-phase pairs
  below the merge volume are added elsewhere;
* outlet: particles just beyond the outlet plane x = box_r (S3:mpm.cpp:1388);
* inlet: the channel's inlet copies (mark_in particles, S3:mpm.cpp:1091).
"""
from __future__ import annotations

import numpy as np

from pgdg_p2_amd import cases


def shock_bubble_forced(nx=16, ny=8, nz=8, scheme=None, seed=0x5EED00F0, big_frac=0.03,
                        small_pairs=24, outlet=12):
    case, st = cases.shock_bubble_3d(nx=nx, ny=ny, nz=nz, ppc_axis=2, jitter=0.25,
                                     scheme=scheme, cap_factor=4.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    n = st.n
    P = st.p
    ps3 = (case.p_sz * case.p_sz) * case.p_sz
    x = P["x"]
    dx = (case.box_r[0] - case.box_l[0]) / case.n[0]
    # interior particles: away from the inlet and outlet columns
    inner = np.nonzero((x[0, :n] > 2 * dx) & (x[0, :n] < case.box_r[0] - 2 * dx))[0]
    big = rng.choice(inner, size=max(1, int(big_frac * n)), replace=False)
    # three large particles in one interior cell (a per-cell bound refuses some)
    cell = (np.floor(x[0, :n] / dx).astype(int) * case.n[1]
            + np.floor(x[1, :n] / dx).astype(int)) * case.n[2] + np.floor(x[2, :n] / dx).astype(int)
    c0 = cell[inner[len(inner) // 2]]
    same = np.nonzero(cell == c0)[0][:3]
    big = np.unique(np.concatenate([big, same]))
    P["vol"][big] = ps3 * rng.uniform(2.0, 3.0, big.size)
    # close same-phase pairs below the merge volume: particle i and its
    # lattice neighbour along z, moved to half a spacing from i
    rest = np.setdiff1d(inner, np.concatenate([big, big + 1, big - 1]))
    cand = rest[(rest + 1 < n)]
    cand = cand[(P["phase"][cand] == P["phase"][cand + 1])
                & (np.abs(x[2, cand + 1] - x[2, cand]) < 1.6 * case.p_sz)
                & (x[2, cand + 1] > x[2, cand])]
    pick = rng.choice(cand, size=min(small_pairs, cand.size), replace=False)
    pick = pick[np.argsort(pick)]
    keep = [pick[0]]
    for i in pick[1:]:
        if i - keep[-1] > 2:
            keep.append(i)
    pick = np.array(keep)
    for i in pick:
        j = i + 1
        x[:, j] = x[:, i]
        x[2, j] = x[2, i] + 0.5 * case.p_sz
        P["vol"][i] = P["vol"][j] = 0.45 * ps3
    # outlet: particles of the last column moved past x = box_r
    last = np.nonzero(x[0, :n] > case.box_r[0] - case.p_sz)[0]
    out = rng.choice(last, size=min(outlet, last.size), replace=False)
    x[0, out] = case.box_r[0] + 0.05 * case.p_sz
    P["mass"][:n] = P["vol"][:n] * P["rho"][:n]
    return case, st
