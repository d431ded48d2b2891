"""CPU checks of the numeric bounds behind the three-body kernel's candidate filter.

The rotation filter of k_atm (komd_amd/csrc/k_cell.hip) tests r_jk^2 = |y - x|^2 of two
list entries (x = r_j - r_i, y = r_k - r_i, |x|, |y| <= rc) with (x, y) components
rounded to FP16, their difference rounded to FP16 (v_pk_add_f16), the two squares as
exact FP16 products summed in FP32 (v_dot2_f32_f16) and z in FP32, against
rc^2 * 1.01 + band.  Every pair with r_jk <= rc must pass (the batch then decides the
cutoff exactly).  The rounding moves r_jk^2 by at most 2|z|u(|x|+|y|+|z|) <= 6u rc^2,
u = 2^-11.  These tests emulate that arithmetic in numpy and check the 1 % margin
holds, with the worst cases (all three vectors of length ~rc) sampled densely."""
import numpy as np

BAND = 1e-5  # band3 = 1e-5 rc^2 of the kernel


def filter_r2(x, y):
    """r_jk^2 as the FP16 rotation filter computes it (x, y: (..., 3) float32)."""
    xh = x[..., :2].astype(np.float16)
    yh = y[..., :2].astype(np.float16)
    d = (yh - xh).astype(np.float16)          # v_pk_add_f16 (y, -x): one FP16 rounding
    d32 = d.astype(np.float32)
    dz = (y[..., 2] - x[..., 2]).astype(np.float32)
    sq = d32[..., 0] * d32[..., 0]            # FP16 x FP16 is exact in FP32
    sq = (sq + d32[..., 1] * d32[..., 1]).astype(np.float32)
    return (sq + dz * dz).astype(np.float32)  # dz * dz in FP32, then the dot's FP32 add


def exact_r2(x, y):
    d = y.astype(np.float64) - x.astype(np.float64)
    return (d * d).sum(-1)


def random_unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def check(rc, n, seed, worst):
    rng = np.random.default_rng(seed)
    rc2 = np.float32(rc * rc)
    lim = np.float32(rc2 * np.float32(1.01) + np.float32(BAND) * rc2)
    if worst:
        # |x| = |y| = rc (up to an ulp), |z| from 0.9 rc to rc: the bound's worst case
        x = random_unit(rng, n) * rc
        ang = np.arccos(1 - 0.5 * rng.uniform(0.81, 1.0, n))  # |z|^2 = 2 rc^2 (1 - cos)
        axis = random_unit(rng, n)
        axis -= (axis * x).sum(1, keepdims=True) * x / rc ** 2
        axis /= np.linalg.norm(axis, axis=1, keepdims=True)
        y = x * np.cos(ang)[:, None] + rc * axis * np.sin(ang)[:, None]
    else:
        x = random_unit(rng, n) * rc * rng.uniform(0.0, 1.0, (n, 1)) ** (1 / 3)
        y = random_unit(rng, n) * rc * rng.uniform(0.0, 1.0, (n, 1)) ** (1 / 3)
    x = x.astype(np.float32)
    y = y.astype(np.float32)
    inside = exact_r2(x, y) <= float(rc2)
    z = filter_r2(x, y)
    missed = inside & ~(z <= lim)
    # the error of the filter's r^2 relative to rc^2 for the pairs within rc (|z| <= rc),
    # where the bound 6u applies
    err = np.abs(z.astype(np.float64) - exact_r2(x, y))[inside] / float(rc2)
    return int(inside.sum()), int(missed.sum()), float(err.max())


def test_fp16_filter_keeps_every_pair_within_rc():
    for rc in (2.5, 0.05, 1.0, 3.0, 12.0, 250.0, 9999.0):
        for worst in (True, False):
            n_in, missed, err = check(rc, 200_000, 17, worst)
            assert n_in > 1000
            assert missed == 0, (rc, worst, missed)
            assert err <= 6 * 2.0 ** -11, (rc, worst, err)  # 0.0029 rc^2 < the 1 % margin


def test_fp16_filter_bound_is_not_vacuous():
    """A margin far below the FP16 error bound would drop pairs: with the 1 % margin
    replaced by 1e-5 some worst-case pairs within rc fail the candidate test."""
    rng_rc = 2.5
    rng = np.random.default_rng(5)
    x = (random_unit(rng, 200_000) * rng_rc).astype(np.float32)
    y = (random_unit(rng, 200_000) * rng_rc).astype(np.float32)
    inside = exact_r2(x, y) <= rng_rc ** 2
    z = filter_r2(x, y)
    tight = np.float32(rng_rc ** 2 * (1 + 1e-5))
    assert (inside & ~(z <= tight)).sum() > 0
