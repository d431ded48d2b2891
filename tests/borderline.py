"""Site pairs whose cutoff decision depends on FMA contraction (shared by the CPU
oracle pin and the GPU parity test)."""
from fractions import Fraction

import numpy as np

from komd_amd import systems as S


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def borderline_pair(origin, rng, rc=2.5):
    """Two sites near `origin` at distance rc whose FP64 r^2 = (dx dx + dy dy) + dz dz
    (LJ12_6.cpp:24-28 order, unfused) lies on the other side of rc^2 than the same sum
    with the last two products fused into FMAs (evaluated exactly with fractions): a
    contracted distance test would decide the pair differently.  Returns the sites and
    whether the reference keeps the pair (r^2 <= rc^2)."""
    rc2 = rc * rc
    while True:
        a = origin + rng.uniform(0.0, 0.5, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        b = a + rc * u
        dx, dy, dz = a - b
        q = dx * dx
        q = q + dy * dy
        q = q + dz * dz
        if (q > rc2) != (_fma(dz, dz, _fma(dy, dy, dx * dx)) > rc2):
            return a, b, bool(q <= rc2)


def borderline_system(npairs=27, seed=5, span=10.0):
    """npairs such pairs, one per block of a 3 x 3 x 3 lattice of span-sized blocks
    (pairs of different blocks are > 4 apart).  Returns the system and the reference's
    keep decision of every pair (sites 2k, 2k + 1)."""
    rng = np.random.default_rng(seed)
    r, keep = [], []
    for k in range(npairs):
        o = np.array([k % 3, (k // 3) % 3, k // 9], float) * span + 3.0
        a, b, kp = borderline_pair(o, rng)
        r += [a, b]
        keep.append(kp)
    r = np.array(r)
    n = r.shape[0]
    L = 3 * span
    s = S.System(r=r, v=np.zeros((n, 3)), mass=np.ones(n), eps=np.ones(n), sig=np.ones(n),
                 mol_id=np.arange(n, dtype=np.uint64), low=np.zeros(3), high=np.full(3, L),
                 cell_size=S.cell_size_for(L, 2.5))
    return s, keep
