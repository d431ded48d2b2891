"""Pins of the oracle's tabulated pair force (IBI: FunctionPL::evaluateAt, IBI_Math.h:30-72,
and IBI_Force, IBI_ForceFunctor.cpp:24-42).  CPU only.

Each check is chosen so that a plausible mistake fails it: a wrong segment or an
off-by-one in the node search fails the linear-function and hand cases; a swapped
default fails the out-of-range cases; a wrong sign or a missed image fails the brute
force, the LJ bound and the momentum sum."""
import numpy as np
import pytest

import oracle as O
from komd_amd import systems as S


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def lj_force(r):
    """|F| of LJ 12-6 (sigma = eps = 1) along r_ij: F(r) = 24 (2 r^-13 - r^-7)."""
    return 24.0 * (2.0 * r ** -13 - r ** -7)


def lj_energy(r):
    return 4.0 * (r ** -12 - r ** -6)


def test_evaluator_hand_cases():
    t = O.Table([1.0, 2.0, 3.0], [10.0, 20.0, 0.0], low=-1.0, high=5.0)
    got = t([0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.0 + 1e-12, 1.0 - 1e-12])
    np.testing.assert_array_equal(got, [-1.0, 10.0, 15.0, 20.0, 10.0, 0.0, 5.0, -1.0])


def test_evaluator_is_exact_for_linear_functions_on_uneven_nodes():
    rng = np.random.default_rng(1)
    x = np.cumsum(rng.uniform(0.01, 0.3, 40)) + 0.7
    y = 3.5 - 1.25 * x
    t = O.Table(x, y, low=7.0, high=-7.0)
    q = rng.uniform(x[0], x[-1], 2000)
    np.testing.assert_allclose(t(q), 3.5 - 1.25 * q, rtol=0, atol=1e-13)
    # every node reproduces its value
    np.testing.assert_allclose(t(x), y, rtol=0, atol=1e-14)


def test_lj_table_reproduces_lj_to_the_interpolation_bound():
    """Linear interpolation of a smooth F on nodes of spacing h errs by at most
    h^2/8 max|F''| on each segment (closed form)."""
    x = np.linspace(0.85, 2.5, 331)
    h = x[1] - x[0]
    t = O.Table(x, lj_force(x), low=1e6, high=0.0)
    q = np.linspace(0.85, 2.5, 20011)
    fpp = np.abs(24.0 * (2 * 13 * 14 * q ** -15 - 7 * 8 * q ** -9))
    err = np.abs(t(q) - lj_force(q))
    # bound per point with the largest |F''| of its segment
    seg = np.minimum(np.searchsorted(x, q, side="left"), x.size - 1)
    bound = np.array([h * h / 8 * fpp[(q >= x[max(k - 1, 0)]) & (q <= x[k])].max() for k in seg])
    assert (err <= bound * (1 + 1e-9) + 1e-12).all()
    assert err.max() > 0.1 * bound.max()  # the bound is tight, not vacuous


def test_table_walk_equals_brute_force_and_conserves_momentum():
    s = S.fcc(5, 0.8, 1.0, 7, jitter=0.15)
    r, b = O.update_container(s)
    x = np.linspace(0.8, 2.3, 61)   # ends inside the cutoff: the upper default acts
    t = O.Table(x, lj_force(x) + 0.3, low=50.0, high=-0.2)
    f, _, _ = O.table_c08(s, r, b, t)
    fb = O.table_brute(s, r, t)
    assert np.abs(f - fb).max() < 1e-12 * np.abs(fb).max()
    assert np.abs(f.sum(axis=0)).max() < 1e-10 * np.abs(f).max()
    fp, _, _ = O.table_c08(s, r, b, t, nthreads=8)
    assert np.abs(fp - f).max() < 1e-12 * np.abs(f).max()


def test_lj_sampled_table_matches_lj_forces():
    """A table sampled from LJ on fine nodes gives the LJ pair forces of the same pair set
    up to the interpolation error (argon sites: no molecule exclusion in play)."""
    s = S.fcc(6, 0.8, 1.0, 8, jitter=0.1)
    r, b = O.update_container(s)
    x = np.linspace(0.7, 2.5, 4001)
    f_t, _, _ = O.table_c08(s, r, b, O.Table(x, lj_force(x), low=1e6, high=0.0), nthreads=8)
    f_lj = O.lj_c08(s, r, b, nthreads=8)
    rel = np.abs(f_t - f_lj).max() / np.sqrt(np.mean(f_lj ** 2))
    assert rel < 1e-4


def test_energy_and_virial_against_minimum_image_sums():
    s = S.fcc(5, 0.8, 1.0, 9, jitter=0.15)
    r, b = O.update_container(s)
    x = np.linspace(0.8, 2.5, 171)
    ft, ut = O.Table(x, lj_force(x), 1e6, 0.0), O.Table(x, lj_energy(x), 0.0, 0.0)
    _, u, w = O.table_c08(s, r, b, ft, potential=ut)
    box = s.box
    d = r[:, None, :] - r[None, :, :]
    d -= box * np.round(d / box)
    iu = np.triu_indices(s.n, 1)
    dv = d[iu]
    rr = np.linalg.norm(dv, axis=1)
    m = rr * rr <= 2.5 ** 2
    dv, rr = dv[m], rr[m]
    assert abs(u - ut(rr).sum()) < 1e-10 * abs(u)
    fv = dv / rr[:, None] * ft(rr)[:, None]
    ref = np.array([(dv[:, a] * fv[:, c]).sum() for a, c in ((0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2))])
    np.testing.assert_allclose(w, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
