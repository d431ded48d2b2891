"""GPU parity of the tabulated pair force (KOMD_PAIR_TABLE, the IBI force functor:
IBI_ForceFunctor.cpp:24-42, FunctionPL::evaluateAt IBI_Math.h:30-72) through the C ABI
against the FP64 oracle.

Bars: FP32 pair arithmetic on relative coordinates, FP64 accumulation -> max and RMS
force error relative to the RMS of the oracle force, as for LJ (max < 2e-4, RMS < 5e-6);
energy and virial within 1e-5 relative; the range decisions at the table ends (where the
defaults make F jump) exact; the graphed run bit-equal to the per-functor loop."""
import numpy as np
import pytest

import komd_amd as K
from komd_amd import systems as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    import oracle
    oracle.build()
    return oracle


def rel_err(a, b):
    scale = np.sqrt(np.mean(b * b))
    d = np.linalg.norm(a - b, axis=1)
    return float(d.max() / scale), float(np.sqrt(np.mean(d * d)) / scale)


def lj_force(r):
    return 24.0 * (2.0 * r ** -13 - r ** -7)


def lj_energy(r):
    return 4.0 * (r ** -12 - r ** -6)


def tables(kind):
    """(force table, potential table or None) as oracle.Table objects."""
    import oracle as O
    if kind == "uniform":      # LJ sampled on 171 nodes up to the cutoff
        x = np.linspace(0.8, 2.5, 171)
        return O.Table(x, lj_force(x), 1e6, 0.0), O.Table(x, lj_energy(x), 0.0, 0.0)
    if kind == "uneven":       # uneven nodes ending inside the cutoff: the upper default acts
        rng = np.random.default_rng(3)
        x = np.concatenate([[0.85], 0.85 + np.cumsum(rng.uniform(0.005, 0.05, 60))])
        x = x[x < 2.2]
        y = lj_force(x) + 0.4 * np.sin(3 * x)
        u = lj_energy(x) + 0.1 * np.cos(2 * x)
        return O.Table(x, y, 40.0, -0.3), O.Table(x, u, 5.0, -0.05)
    if kind == "beyond":       # nodes past the cutoff (the cutoff decides), no potential
        x = np.linspace(0.7, 3.4, 55)
        return O.Table(x, lj_force(x) - 0.05, -1.0, 2.0), None
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["uniform", "uneven", "beyond"])
@pytest.mark.parametrize("cells,seed", [(5, 401), (10, 402), (20, 403)])
def test_table_forces_match_oracle(orc, kind, cells, seed):
    """cells 20: the 32,000-atom argon workload"""
    s = S.fcc(cells, 0.8, 0.8, seed, jitter=0.12)
    print(f"table {kind} n={s.n}", flush=True)
    ft, ut = tables(kind)
    r, b = orc.update_container(s)
    f_or, u_or, w_or = orc.table_c08(s, r, b, ft, potential=ut, nthreads=8)
    with K.Context(s) as c:
        c.set_pair_table(ft.x, ft.y, ft.low, ft.high, None if ut is None else ut.y,
                         0.0 if ut is None else ut.low, 0.0 if ut is None else ut.high)
        c.update_container()
        c.forces_pair(K.PAIR_TABLE)
        f = c.forces()
        c.select_pair(K.PAIR_TABLE)
        o = c.observables(0)
        assert np.array_equal(c.forces(), f)   # the energy pass leaves the forces as they were
    mx, rms = rel_err(f, f_or)
    print(f"table {kind} max {mx:.3e} rms {rms:.3e}")
    assert mx < 2e-4 and rms < 5e-6
    if ut is not None:
        assert abs(o["u_lj"] - u_or) < 1e-5 * abs(u_or)
    else:
        assert o["u_lj"] == 0.0
    np.testing.assert_allclose(o["virial_tensor"], w_or, rtol=0, atol=1e-5 * np.abs(w_or).max())


def _end_pairs(x_end, npairs=27, seed=11, span=10.0):
    """Isolated site pairs (one per block of a 3 x 3 x 3 lattice) at distance x_end up to
    FP64 rounding: which side of x_end each lies on is decided by the last bits of the
    FP64 distance, below the resolution of the FP32 list vectors."""
    rng = np.random.default_rng(seed)
    r = []
    for k in range(npairs):
        o = np.array([k % 3, (k // 3) % 3, k // 9], float) * span + 3.0
        a = o + rng.uniform(0.0, 0.5, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        r += [a, a + x_end * u]
    r = np.array(r)
    n = r.shape[0]
    L = 3 * span
    return S.System(r=r, v=np.zeros((n, 3)), mass=np.ones(n), eps=np.ones(n), sig=np.ones(n),
                    mol_id=np.arange(n, dtype=np.uint64), low=np.zeros(3), high=np.full(3, L),
                    cell_size=S.cell_size_for(L, 2.5))


@pytest.mark.parametrize("end", ["first", "last"])
def test_table_end_decisions_are_exact(orc, end):
    """The defaults outside [x0, x_last] make F jump (here 10 -> 50 below x0, 1 -> -3 above
    x_last): for pairs within an ulp of either end the GPU takes the FP64 decision of the
    reference's operation order, so every force equals the oracle's to FP32 rounding."""
    import oracle as O
    x = np.linspace(0.9, 2.2, 27)
    y = np.linspace(10.0, 1.0, 27)
    ft = O.Table(x, y, 50.0, -3.0)
    s = _end_pairs(x[0] if end == "first" else x[-1])
    r, b = orc.update_container(s)
    f_or, _, _ = orc.table_c08(s, r, b, ft)
    mag_or = np.linalg.norm(f_or[0::2], axis=1)
    side = mag_or > 2.0 if end == "last" else mag_or > 30.0   # the default's side of the jump
    print(f"{end}: {int(side.sum())} of {side.size} pairs outside the table", flush=True)
    assert 0 < side.sum() < side.size
    with K.Context(s) as c:
        c.set_pair_table(ft.x, ft.y, ft.low, ft.high)
        c.update_container()
        c.forces_pair(K.PAIR_TABLE)
        f = c.forces()
    mag = np.linalg.norm(f[0::2], axis=1)
    np.testing.assert_allclose(mag, mag_or, rtol=1e-5)


@pytest.mark.parametrize("limit", [0.0, 20.0])
def test_table_run_matches_functor_loop(limit):
    """komd_run with the table selected (graph path): without the clamp the table kernel
    applies the second half-kick, with it a separate kick follows the clamp.  Both equal the
    per-functor loop (integrate0, updateContainer, table force, Limit, integrate1) bit for bit."""
    s = S.fcc(6, 0.8, 0.8, 404, jitter=0.05)
    ft, ut = tables("uniform")
    with K.Context(s, limit_factor=limit) as a, K.Context(s, limit_factor=limit) as b:
        for c in (a, b):
            c.set_pair_table(ft.x, ft.y, ft.low, ft.high, ut.y, ut.low, ut.high)
            c.select_pair(K.PAIR_TABLE)
            c.compute_forces(0)
        a.run(12)
        for _ in range(12):
            b.integrate0()
            b.update_container()
            b.forces_pair(K.PAIR_TABLE)
            b.forces_limit()
            b.integrate1()
        ra, va, fa = a.get_state()
        rb, vb, fb = b.get_state()
    assert np.array_equal(fa, fb)
    assert np.array_equal(va, vb)
    assert np.array_equal(ra, rb)


def test_lj_table_run_follows_lj_run():
    """A table sampled finely from LJ selected in komd_run stays on the LJ trajectory over 50
    steps (interpolation error only), with the exact three-body force added to both."""
    s = S.fcc(8, 0.8, 0.8, 405, jitter=0.05)
    x = np.linspace(0.7, 2.5, 4001)
    with K.Context(s, nu=0.073) as a, K.Context(s, nu=0.073) as b:
        b.set_pair_table(x, lj_force(x), 1e6, 0.0, lj_energy(x), 0.0, 0.0)
        b.select_pair(K.PAIR_TABLE)
        for c in (a, b):
            c.compute_forces(K.ATM_EXACT)
            c.run(50, kind3b=K.ATM_EXACT)
        ra, va, _ = a.get_state()
        rb, vb, _ = b.get_state()
        oa, ob = a.observables(K.ATM_EXACT), b.observables(K.ATM_EXACT)
    assert np.abs(ra - rb).max() < 2e-4
    assert abs(oa["u_lj"] - ob["u_lj"]) < 1e-4 * abs(oa["u_lj"])


def test_table_argument_checks():
    s = S.fcc(5, 0.8, 0.8, 406)
    with K.Context(s) as c:
        with pytest.raises(K.KomdError, match="no pair table"):
            c.select_pair(K.PAIR_TABLE)
        with pytest.raises(K.KomdError, match="no pair table"):
            c.forces_pair(K.PAIR_TABLE)
        with pytest.raises(K.KomdError, match="increase strictly"):
            c.set_pair_table([1.0, 1.0, 2.0], [0.0, 1.0, 2.0])
        with pytest.raises(K.KomdError, match="non-finite"):
            c.set_pair_table([1.0, 2.0], [0.0, np.nan])
        with pytest.raises(K.KomdError, match="2 to 2"):
            c.set_pair_table([1.0], [0.0])


def test_atm2b_forces_and_run(orc):
    """KOMD_PAIR_ATM2B (potentials/ATM2B.cpp:9-56): the LJ force and energy times
    1 - 2 nu rho / (3 eps sig^6) (rho = sites per volume here), against the oracle's LJ
    times that factor; the graphed run with ATM2B selected equals the per-functor loop."""
    from oracle import surrogate as SO
    s = S.fcc(10, 0.8, 0.8, 407, jitter=0.1)
    nu = 0.073
    fac = SO.atm2b_scale(nu, s.n / float(np.prod(s.box)))
    r, b = orc.update_container(s)
    f_lj, u_lj, _ = orc.lj_c08(s, r, b, energy=True, nthreads=8)
    with K.Context(s, nu=nu) as c:
        c.update_container()
        c.forces_pair(K.PAIR_ATM2B)
        f = c.forces()
        c.select_pair(K.PAIR_ATM2B)
        o = c.observables(0)
    mx, rms = rel_err(f, fac * f_lj)
    print(f"ATM2B factor {fac:.6f}: max {mx:.3e} rms {rms:.3e}")
    assert abs(fac - (1 - 2 * nu * 0.8 / 3)) < 1e-12
    assert mx < 2e-4 and rms < 5e-6
    assert abs(o["u_lj"] - fac * u_lj) < 1e-5 * abs(u_lj)
    small = S.fcc(6, 0.8, 0.8, 408, jitter=0.05)
    with K.Context(small, nu=nu) as a, K.Context(small, nu=nu) as b2:
        for c in (a, b2):
            c.select_pair(K.PAIR_ATM2B)
            c.compute_forces(0)
        a.run(12)
        for _ in range(12):
            b2.integrate0()
            b2.update_container()
            b2.forces_pair(K.PAIR_ATM2B)
            b2.integrate1()
        ra, va, _ = a.get_state()
        rb, vb, _ = b2.get_state()
    assert np.array_equal(ra, rb) and np.array_equal(va, vb)
