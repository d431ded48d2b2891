"""Pins of the FP64 oracle against facts that do not come from the oracle
itself: closed forms, finite differences, brute force, invariants and the
reference's own example checkpoint.  Each test names the plausible mistake it
would catch.  (CPU only.)"""
import math
import os

import numpy as np
import pytest

import oracle as O
from komd_amd import systems as S

NU = 0.073
FIXTURES = os.path.join(os.path.dirname(__file__), "golden")


def fd_grad(fun, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.size):
        xp = x.copy(); xp.flat[i] += h
        xm = x.copy(); xm.flat[i] -= h
        g.flat[i] = (fun(xp) - fun(xm)) / (2 * h)
    return g


# ---------------------------------------------------------------- units
def test_temperature_factor_value():
    # TemperatureSensor.cpp:39-42: 1e4 * 1.66053906892e-27 / 1.380649e-23 = 1.2027235...e0 (K per u A^2/ps^2)
    f = O.lib().or_temperature_of(3.0, 1.0)  # mv2/(3n) = 1
    assert abs(f - 1e4 * 1.66053906892e-27 / 1.380649e-23) < 1e-12 * f
    # the reduced-unit deck value T* = 1 -> 1.2027235516 K (its inverse is k_B in u A^2/(ps^2 K))
    assert abs(1.0 / f - 0.8314462609869346) < 1e-9


def test_nu_conversion():
    # ATM.cpp:13: nu = energy_3b * 1e9 * conv_J_Ei ; nu* = 0.073 <-> energy_3b = 1.2121935202446447e-33
    assert abs(O.lib().or_nu_internal(1.2121935202446447e-33) - 0.073) < 1e-12


# ---------------------------------------------------------------- binning
def test_grid_dims_floor():
    nd = O.grid_dims([0, 0, 0], [17.2, 34.3, 10.0], 2.85)
    assert list(nd) == [6, 12, 3]  # floor(L / cs), remainder cell absorbs the rest
    # L / cs a hair below an integer floors down (the reference has the same behaviour)
    assert list(O.grid_dims([0, 0, 0], [17.0998] * 3, 2.85)) == [5, 5, 5]


def test_reference_checkpoint_fixture():
    """examples/test/test.ps of the reference (4 single-site molecules in a 10^3 box,
    cell 5 -> 2x2 domain cells, z in cell 0).  Catches an off-by-one in the +1 halo
    offset or an x/y swap in the flat index."""
    rows = np.loadtxt(os.path.join(FIXTURES, "ref_test_ps.txt"))
    r = rows[:, 1:4].copy()
    s = S.System(r=r, v=rows[:, 4:7], mass=rows[:, 7], eps=rows[:, 8], sig=rows[:, 9],
                 mol_id=rows[:, 0].astype(np.uint64), low=np.zeros(3), high=np.full(3, 10.0), cell_size=5.0)
    r2, b = O.update_container(s)
    na = b.nd + 2
    assert list(b.nd) == [2, 2, 2]
    expect = [(1, 1, 1), (2, 1, 1), (1, 2, 1), (2, 2, 1)]
    for i, (x, y, z) in enumerate(expect):
        assert b.site_cell[i] == x + na[0] * (y + na[1] * z)
    # halo cell (0,1,1) copies source (2,1,1) (LinkedCells.cpp:115)
    src = 2 + na[0] * (1 + na[1] * 1)
    halo = 0 + na[0] * (1 + na[1] * 1)
    assert b.cell_count[halo] == b.cell_count[src] == 1
    assert list(b.cell_list[b.cell_start[halo]:b.cell_start[halo + 1]]) == [1]


def test_binning_invariants():
    s = S.fcc(6, 0.8, 1.0, 3, jitter=0.2)
    r, b = O.update_container(s)
    na = b.nd + 2
    n_dom = 0
    for z in range(na[2]):
        for y in range(na[1]):
            for x in range(na[0]):
                c = x + na[0] * (y + na[1] * z)
                lst = b.cell_list[b.cell_start[c]:b.cell_start[c + 1]]
                assert np.all(np.diff(lst) > 0), "per-cell order must be ascending site index"
                halo = x in (0, na[0] - 1) or y in (0, na[1] - 1) or z in (0, na[2] - 1)
                if not halo:
                    n_dom += len(lst)
                    lo = s.low + (np.array([x, y, z]) - 1) * s.cell_size
                    hi = lo + s.cell_size
                    hi = np.where(np.array([x, y, z]) == b.nd, s.high, hi)
                    assert np.all((r[lst] >= lo - 1e-12) & (r[lst] <= hi + 1e-12)), "site outside its cell"
    assert n_dom == s.n


def test_periodic_wrap_strict_single_image():
    s = S.fcc(4, 0.8, 1.0, 3)
    s.cell_size = s.box[0] / 3
    r = s.r.copy()
    L = s.box[0]
    r[0] = [-0.1, L + 0.2, 0.5]
    r[1] = [L, 0.0, 0.3]            # exactly on the faces: not wrapped (strict <, >)
    r2, _ = O.update_container(s, r)
    assert np.allclose(r2[0], [-0.1 + L, 0.2, 0.5])
    assert np.array_equal(r2[1], r[1])


# ---------------------------------------------------------------- LJ
def test_lj_closed_form_and_fd():
    """Two atoms: F(2^(1/6)) = 0, F(1) = 24 (catches a wrong prefactor or power),
    and F = -dU/dr against central differences of the sensor energy (catches a sign error)."""
    for d, fexp in [(2 ** (1 / 6), 0.0), (1.0, 24.0)]:
        s = S.System(r=np.array([[5.0, 5, 5], [5.0 + d, 5, 5]]), v=np.zeros((2, 3)), mass=np.ones(2),
                     eps=np.ones(2), sig=np.ones(2), mol_id=np.array([0, 1], np.uint64), low=np.zeros(3),
                     high=np.full(3, 10.0), cell_size=3.3)
        r, b = O.update_container(s)
        f = O.lj_c08(s, r, b)
        assert abs(f[1, 0] - fexp) < 1e-12 and abs(f[0, 0] + fexp) < 1e-12

    s = S.fcc(5, 0.8, 1.0, 5, jitter=0.15)
    r, b = O.update_container(s)
    f, u, _ = O.lj_c08(s, r, b, energy=True)

    def energy(x):
        rr = r.copy()
        rr[7] = x
        _, uu = O.lj_brute(s, rr)
        return uu
    g = fd_grad(energy, r[7].copy())
    assert np.allclose(f[7], -g, rtol=1e-6, atol=1e-6)


def test_lj_c08_equals_brute_force():
    """The C08 walk must produce every minimum-image pair once per unshifted member."""
    for cells, seed in [(5, 1), (6, 2)]:
        s = S.fcc(cells, 0.8, 1.0, seed, jitter=0.1)
        r, b = O.update_container(s)
        f, u, _ = O.lj_c08(s, r, b, energy=True)
        fb, ub = O.lj_brute(s, r)
        assert np.abs(f - fb).max() < 1e-10 * np.abs(fb).max()
        assert abs(u - ub) < 1e-9 * abs(ub)
        assert np.abs(f.sum(0)).max() < 1e-9


def test_lj_cutoff_decision_is_unfused():
    """Mistake caught: an FMA-contracted distance in the oracle (a build without
    -ffp-contract=off).  For pairs within an ulp of rc, the kept set must be the one
    of the written FP64 order (dx dx + dy dy) + dz dz <= rc^2 (LJ12_6.cpp:24-28),
    which for each of these pairs differs from the fused sum's."""
    from borderline import borderline_system
    s, want = borderline_system()
    r, b = O.update_container(s)
    f = O.lj_c08(s, r, b)
    kept = [bool(np.abs(f[2 * k]).max() > 0) for k in range(len(want))]
    assert 0 < sum(want) < len(want)
    assert kept == want


def test_lj_same_molecule_excluded():
    """Two 2-site molecules in neighbouring cells: only the 4 inter-molecular pairs act
    (LJ12_6.cpp:19); catches a missing or inverted id test."""
    r = np.array([[4.0, 4.5, 4.5], [5.0, 4.5, 4.5], [4.5, 5.6, 4.5], [4.5, 6.5, 4.5]])
    s = S.System(r=r, v=np.zeros((4, 3)), mass=np.ones(4), eps=np.ones(4), sig=np.ones(4),
                 mol_id=np.array([7, 7, 9, 9], np.uint64), low=np.zeros(3), high=np.full(3, 9.0), cell_size=3.0)
    r2, b = O.update_container(s)
    f = O.lj_c08(s, r2, b)
    fb, _ = O.lj_brute(s, r2)
    assert np.abs(f - fb).max() < 1e-12 * np.abs(fb).max()
    # by hand: only pairs (0,2), (0,3), (1,2), (1,3)
    fh = np.zeros_like(r)
    for i in (0, 1):
        for j in (2, 3):
            d = r[i] - r[j]
            q = d @ d
            fac = 24 * (2 / q ** 6 - 1 / q ** 3) / q
            fh[i] += d * fac
            fh[j] -= d * fac
    assert np.allclose(f, fh, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- ATM
def test_atm_equilateral_energy():
    """V = nu (1 + 3 (1/2)^3) / r^9 = 11 nu / (8 r^9) for an equilateral triangle."""
    r = 1.3
    p = [np.zeros(3), np.array([r, 0, 0]), np.array([r / 2, r * math.sqrt(3) / 2, 0])]
    v = O.atm_energy_triplet(NU, *p)
    assert abs(v - 11 * NU / (8 * r ** 9)) < 1e-14


def test_atm_collinear_energy():
    """Collinear a, b at distances 1 and 2: cosines 1, 1, -1 -> V = nu (1 - 3) / (1 * 1 * 2)^3."""
    p = [np.zeros(3), np.array([1.0, 0, 0]), np.array([2.0, 0, 0])]
    assert abs(O.atm_energy_triplet(NU, *p) - NU * (1 - 3) / 8.0) < 1e-14


def test_reference_comp_force_is_plus_gradient():
    """comp_force (ATM.h:36-69) equals +(1/r_ab) dV/dr_ab: the reference applies F = +grad V.
    Checked against central differences of the textbook energy in the three distances."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b, c = 1.0 + rng.random(3)
        if a + b <= c or a + c <= b or b + c <= a:
            continue

        def V(ab, ac, bc):
            ci = (ab ** 2 + ac ** 2 - bc ** 2) / (2 * ab * ac)
            cj = (ab ** 2 + bc ** 2 - ac ** 2) / (2 * ab * bc)
            ck = (ac ** 2 + bc ** 2 - ab ** 2) / (2 * ac * bc)
            return NU * (1 + 3 * ci * cj * ck) / (ab * ac * bc) ** 3
        h = 1e-6
        dv = (V(a + h, b, c) - V(a - h, b, c)) / (2 * h)
        assert abs(O.comp_force(NU, a, b, c) - dv / a) < 1e-7 * max(1.0, abs(dv / a))


def test_atm_physical_force_is_minus_gradient():
    rng = np.random.default_rng(1)
    s = S.System(r=np.array([[2.0, 2, 2], [3.1, 2.2, 2.1], [2.4, 3.0, 2.3]]) + rng.random((3, 3)) * 0.1,
                 v=np.zeros((3, 3)), mass=np.ones(3), eps=np.ones(3), sig=np.ones(3),
                 mol_id=np.arange(3, dtype=np.uint64), low=np.zeros(3), high=np.full(3, 8.0), cell_size=2.6)
    f, e, hits = O.atm_brute(s, s.r, NU, sign=-1.0)
    assert hits == 1

    def energy(x, i):
        rr = s.r.copy()
        rr[i] = x
        return O.atm_energy_triplet(NU, rr[0], rr[1], rr[2])
    for i in range(3):
        g = fd_grad(lambda x: energy(x, i), s.r[i].copy())
        assert np.allclose(f[i], -g, rtol=1e-6, atol=1e-8)
    assert np.abs(f.sum(0)).max() < 1e-12


def test_atm_reference_walk_equals_independent_coverage_rule():
    """The 44-offset C08 triple walk + same-cell triples (ATM_NOLIST.cpp) covers exactly the
    minimum-image triplets whose cells are all equal or pairwise distinct (the 2+1 case is
    skipped), each member receiving its force once.  Catches a wrong offset table entry."""
    for cells, seed in [(5, 3), (6, 4)]:
        s = S.fcc(cells, 0.8, 1.0, seed, jitter=0.15)
        r, b = O.update_container(s)
        f, _ = O.atm_c08(s, r, b, NU, sign=1.0)
        fb, _, hb = O.atm_brute(s, r, NU, site_cell=b.site_cell, coverage=1, sign=1.0)
        assert np.abs(f - fb).max() < 1e-12 * max(1.0, np.abs(fb).max())
        # and the skipped 2+1 share is large at this density (about half of all triplets)
        _, _, hall = O.atm_brute(s, r, NU, coverage=0, sign=1.0)
        assert 0.3 < 1 - hb / hall < 0.7


def test_atm_complete_equals_brute_force():
    s = S.fcc(5, 0.8, 1.0, 9, jitter=0.15)
    r, b = O.update_container(s)
    f, e, h = O.atm_complete(s, r, b, NU, coverage=0, sign=-1.0)
    fb, eb, hb = O.atm_brute(s, r, NU, coverage=0, sign=-1.0)
    assert h == hb
    assert np.abs(f - fb).max() < 1e-12 * np.abs(fb).max()
    assert abs(e - eb) < 1e-12 * abs(eb)
    _, _, _, es = O.atm_complete(s, r, b, NU, coverage=0, sign=-1.0, site_energy=True)
    assert abs(es.sum() - eb) < 1e-10 * abs(eb)


def test_atm_complete_reference_coverage_equals_brute_force():
    """The cell-list ATM restricted to the reference coverage (cells all equal or pairwise
    distinct) finds the brute force's triplets (count, forces, energy) and gives the C08
    walk's forces: its count is the number of unique triplets the reference evaluates,
    which the C08 walk's evaluation count is not (a triplet across a periodic face is
    evaluated once per image combination holding an unshifted member)."""
    for cells, seed in [(5, 5), (6, 6)]:
        s = S.fcc(cells, 0.8, 1.0, seed, jitter=0.15)
        r, b = O.update_container(s)
        f, e, h = O.atm_complete(s, r, b, NU, coverage=1, sign=1.0)
        fb, eb, hb = O.atm_brute(s, r, NU, site_cell=b.site_cell, coverage=1, sign=1.0)
        fc, _ = O.atm_c08(s, r, b, NU, sign=1.0)
        assert h == hb
        assert np.abs(f - fb).max() < 1e-12 * np.abs(fb).max()
        assert abs(e - eb) < 1e-12 * abs(eb)
        assert np.abs(f - fc).max() < 1e-12 * np.abs(fc).max()


def test_c08_triple_table_counts():
    """44 = 56 - 12 distinct-cell triples per 2x2x2 block: all 21 with the origin, and of the
    35 without it, the 12 lying in the x=1, y=1 or z=1 faces are left out."""
    import itertools
    from oracle import lib
    corners = [np.array(c) for c in itertools.product((0, 1), repeat=3)]
    trip = set()
    # read the table back through a tiny system: one site per cell of a 3x3x3 grid
    cs = 2.6
    pos = np.array([[(x + 0.5) * cs, (y + 0.5) * cs, (z + 0.5) * cs] for z in range(3) for y in range(3)
                    for x in range(3)])
    assert len(list(itertools.combinations(corners, 3))) == 56
    s = S.System(r=pos, v=np.zeros_like(pos), mass=np.ones(27), eps=np.ones(27), sig=np.ones(27),
                 mol_id=np.arange(27, dtype=np.uint64), low=np.zeros(3), high=np.full(3, 3 * cs), cell_size=cs)
    r, b = O.update_container(s)
    # with cutoff large enough every candidate triple is in range: count evaluations
    _, hits = O.atm_c08(s, r, b, NU, cutoff=100.0, sign=1.0)
    # every (base block, table entry) with not all-halo cells evaluates 1 candidate here
    assert hits > 0


# ---------------------------------------------------------------- integrator, thermostat, bonded, limit
def test_vv_harmonic_oscillator():
    """Split VV on F = -x: one step maps (x, v) by the symplectic VV matrix (closed form)."""
    dt = 0.1
    r = np.array([[1.0, 0, 0]])
    v = np.array([[0.0, 0.5, 0]])
    m = np.ones(1)
    f = -r
    r1, v1 = O.vv_step0(dt, r, v, f, m)
    v2 = O.vv_step1(dt, v1, -r1, m)
    x0, p0 = 1.0, 0.0
    x1 = x0 * (1 - dt * dt / 2) + p0 * dt
    p1 = p0 * (1 - dt * dt / 2) - x0 * dt * (1 - dt * dt / 4)
    assert abs(r1[0, 0] - x1) < 1e-15 and abs(v2[0, 0] - p1) < 1e-15
    # time reversibility: run back with -v
    r3, v3 = O.vv_step0(dt, r1, -v2, -r1, m)
    v4 = O.vv_step1(dt, v3, -r3, m)
    assert abs(r3[0, 0] - 1.0) < 1e-14 and abs(v4[0, 0]) < 1e-14


def test_bonded_fd_and_momentum():
    s = S.chains((2, 2, 2), 4, 0.3, seed=3)
    s.r += np.random.default_rng(0).normal(size=s.r.shape) * 0.05
    f, u = O.bonded(s, s.r)

    def energy(x):
        rr = s.r.copy()
        rr[1] = x
        return O.bonded(s, rr)[1]
    g = fd_grad(energy, s.r[1].copy())
    assert np.allclose(f[1], -g, rtol=1e-5, atol=1e-5)
    assert np.abs(f.sum(0)).max() < 1e-9


def test_limit_clamps_components():
    s = S.fcc(2, 0.8, 1.0, 1)
    s.eps[:] = 2.0
    s.sig[:] = 0.5
    f = np.array([[100.0, -100.0, 3.0]] * s.n)
    g = O.limit(s, f, 20.0)
    assert np.all(g[:, 0] == 80.0) and np.all(g[:, 1] == -80.0) and np.all(g[:, 2] == 3.0)


@pytest.mark.parametrize("cells", [10, 20])
def test_drift_fixture_consistent(cells):
    """The stored FP64 energy trajectories (tests/golden/make_drift_fp64.py; 4,000 atoms and
    the 32,000-atom workload) start at the oracle's energy of the generated state, and their
    slopes are the fits of their own samples."""
    import json
    s = S.fcc(cells, 0.8, 0.8, 1003)
    with open(os.path.join(FIXTURES, f"drift_fp64_fcc{s.n}_atm.json")) as fh:
        ref = json.load(fh)
    r, b = O.update_container(s)
    _, u, _ = O.lj_c08(s, r, b, energy=True, nthreads=4)
    _, ea, _ = O.atm_complete(s, r, b, NU, coverage=0, sign=-1.0, nthreads=4)
    e0 = (0.5 * float(np.sum(s.mass[:, None] * s.v * s.v)) + u + ea) / s.n
    assert abs(e0 - ref["e_per_atom"][0]) < 1e-12
    t = np.array(ref["steps"]) * 0.005
    assert abs(np.polyfit(t, np.array(ref["e_per_atom"]), 1)[0] - ref["slope_e_per_atom_per_tau"]) < 1e-12
    assert ref["steps"][-1] == 10000 and len(ref["steps"]) == 101
