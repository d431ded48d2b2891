"""Pins of the FP64 surrogate oracle (CPU): the factorised angular invariants equal the
literal double sum, forces equal central differences of the energy, the training
gradient equals central differences of the loss, and the descriptors are invariant
under rotation and translation."""
import numpy as np
import pytest

from oracle import surrogate as SO
from komd_amd import systems as S


def small_system(cells=3, seed=3, jitter=0.08):
    s = S.fcc(cells, 0.8, 0.8, seed, jitter=jitter)
    return s.r.copy(), np.asarray(s.box, np.float64)


def random_params(seed=1):
    rng = np.random.default_rng(seed)
    r, box = small_system()
    i, j, d = SO.neighbour_pairs(r, box)
    D, _ = SO.descriptors(len(r), i, d)
    p = SO.unpack(SO.init_params(seed, shift=D.mean(0), scale=1.0 / (D.std(0) + 1e-9), e_mean=0.3, e_std=0.05))
    p["b1"] = rng.normal(size=SO.H1) * 0.3
    p["b2"] = rng.normal(size=SO.H2) * 0.3
    return SO.pack(p)


def test_angular_factorised_equals_literal():
    r, box = small_system(2, 5)
    i, j, d = SO.neighbour_pairs(r, box)
    D, _ = SO.descriptors(len(r), i, d)
    lit = SO.angular_literal(len(r), i, d)
    np.testing.assert_allclose(D[:, SO.NR:], lit, rtol=1e-11, atol=1e-14)


def test_descriptor_invariance():
    r, box = small_system(3, 6)
    n = len(r)
    i, j, d = SO.neighbour_pairs(r, box)
    D, _ = SO.descriptors(n, i, d)
    # rotate every displacement: invariants unchanged
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    Rx = np.array([[1.0, 0, 0], [0, np.cos(1.1), -np.sin(1.1)], [0, np.sin(1.1), np.cos(1.1)]])
    D2, _ = SO.descriptors(n, i, d @ (Rz @ Rx).T)
    np.testing.assert_allclose(D2, D, rtol=1e-10, atol=1e-13)
    # translate the configuration: same pairs, same descriptors
    i3, j3, d3 = SO.neighbour_pairs(r + np.array([0.37, -1.2, 2.9]), box)
    D3, _ = SO.descriptors(n, i3, d3)
    np.testing.assert_allclose(np.sort(D3, 0), np.sort(D, 0), rtol=1e-10, atol=1e-12)


def test_basis_derivatives():
    r = np.linspace(0.8, 2.49, 400)
    h = 1e-6
    f, df = SO.radial_basis(r)
    fp, _ = SO.radial_basis(r + h)
    fm, _ = SO.radial_basis(r - h)
    np.testing.assert_allclose(df, (fp - fm) / (2 * h), rtol=1e-5, atol=1e-7)
    g, dg = SO.angular_basis(r)
    gp, _ = SO.angular_basis(r + h)
    gm, _ = SO.angular_basis(r - h)
    np.testing.assert_allclose(dg, (gp - gm) / (2 * h), rtol=1e-5, atol=1e-7)
    # cutoff: basis and derivatives vanish at rc
    f0, df0 = SO.radial_basis(np.array([SO.RC]))
    assert np.abs(f0).max() < 1e-15 and np.abs(df0).max() < 1e-10


def test_forces_are_minus_gradient():
    r, box = small_system(3, 7)
    n = len(r)
    p = random_params(2)
    i, j, d = SO.neighbour_pairs(r, box)
    E, Ei, F = SO.energy_forces(p, n, i, j, d)
    assert abs(F.sum(0)).max() < 1e-10 * np.abs(F).max()   # translation invariance
    rng = np.random.default_rng(0)
    h = 1e-5
    for s in rng.choice(n, 6, replace=False):
        for ax in range(3):
            rp = r.copy(); rp[s, ax] += h
            rm = r.copy(); rm[s, ax] -= h
            Ep = SO.energy_forces(p, n, *SO.neighbour_pairs(rp, box))[0]
            Em = SO.energy_forces(p, n, *SO.neighbour_pairs(rm, box))[0]
            fd = -(Ep - Em) / (2 * h)
            assert abs(fd - F[s, ax]) < 1e-6 * max(1.0, abs(F[s, ax])), (s, ax, fd, F[s, ax])


def test_training_gradient_matches_finite_differences():
    r, box = small_system(3, 8)
    n = len(r)
    p = random_params(3)
    i, j, d = SO.neighbour_pairs(r, box)
    rng = np.random.default_rng(1)
    E_ref = rng.normal(0.3, 0.05, n)
    F_ref = rng.normal(0, 0.3, (n, 3))
    wE, wF = 2.0, 5.0
    L, g, _, _ = SO.loss_and_grad(p, n, i, j, d, E_ref, F_ref, wE, wF)
    off = {}
    o = 0
    for name, k in (("W1", SO.H1 * SO.NDESC), ("b1", SO.H1), ("W2", SO.H2 * SO.H1), ("b2", SO.H2), ("w3", SO.H2),
                    ("b3", 1)):
        off[name] = (o, k)
        o += k
    h = 1e-6
    for name, (o, k) in off.items():
        for idx in rng.choice(k, min(k, 5), replace=False):
            q = o + idx
            pp = p.copy(); pp[q] += h
            pm = p.copy(); pm[q] -= h
            Lp = SO.loss_and_grad(pp, n, i, j, d, E_ref, F_ref, wE, wF)[0]
            Lm = SO.loss_and_grad(pm, n, i, j, d, E_ref, F_ref, wE, wF)[0]
            fd = (Lp - Lm) / (2 * h)
            assert abs(fd - g[q]) < 1e-6 * max(1e-3, abs(g[q])) + 1e-9, (name, idx, fd, g[q])
    # force-only and energy-only parts are each exact
    for we, wf in ((1.0, 0.0), (0.0, 1.0)):
        L0, g0, _, _ = SO.loss_and_grad(p, n, i, j, d, E_ref, F_ref, we, wf)
        q = off["W1"][0] + 17
        pp = p.copy(); pp[q] += h
        pm = p.copy(); pm[q] -= h
        fd = (SO.loss_and_grad(pp, n, i, j, d, E_ref, F_ref, we, wf)[0] -
              SO.loss_and_grad(pm, n, i, j, d, E_ref, F_ref, we, wf)[0]) / (2 * h)
        assert abs(fd - g0[q]) < 1e-6 * max(1e-3, abs(g0[q])) + 1e-9


def test_atm2b_scale():
    # ATM2B.cpp:9-20 at rho = 0.8, nu = 0.073, eps = sigma = 1
    assert SO.atm2b_scale(0.073, 0.8) == pytest.approx(1 - 2 * 0.073 * 0.8 / 3, abs=1e-15)
    assert SO.NPARAM == 64 * 48 + 64 + 64 * 64 + 64 + 64 + 1 + 48 + 48


def test_atm_triplet_terms_against_closed_forms():
    # equilateral triangle of side s: V = 11 nu / (8 s^9); gradient by central differences
    s_, nu = 1.1, 0.073
    r = np.array([[1.0, 1.0, 1.0], [1.0 + s_, 1.0, 1.0], [1.0 + s_ / 2, 1.0 + s_ * np.sqrt(3) / 2, 1.0]])
    box = np.array([20.0, 20.0, 20.0])
    trip = np.array([[0, 1, 2]])
    V, gi, gj, gk = SO.atm_triplet_terms(r, box, trip, nu)
    assert V[0] == pytest.approx(11 * nu / (8 * s_ ** 9), rel=1e-12)
    rng = np.random.default_rng(2)
    r = r + rng.normal(size=r.shape) * 0.1
    V, gi, gj, gk = SO.atm_triplet_terms(r, box, trip, nu)
    h = 1e-6
    for m, g in enumerate((gi, gj, gk)):
        for ax in range(3):
            rp = r.copy(); rp[m, ax] += h
            rm = r.copy(); rm[m, ax] -= h
            fd = (SO.atm_triplet_terms(rp, box, trip, nu)[0] - SO.atm_triplet_terms(rm, box, trip, nu)[0]) / (2 * h)
            assert fd[0] == pytest.approx(g[0, ax], rel=1e-6, abs=1e-10)


def test_hybrid_forces_are_minus_gradient():
    r, box = small_system(3, 9)
    n = len(r)
    p = random_params(5)
    args = (0.073, 0, box[0] * 0.5, 0.8, 1.2)   # nu, axis, center, half width, blend
    E, F = SO.hybrid_energy_forces(p, r, box, *args)
    lam, _ = SO.lam_of(r[:, 0], box[0], box[0] * 0.5, 0.8, 1.2)
    assert (lam > 0).any() and (lam < 1).any() and ((lam > 0) & (lam < 1)).any()
    rng = np.random.default_rng(4)
    h = 1e-5
    # sites in the blend region carry the weight-gradient terms
    blend_sites = np.nonzero((lam > 0.05) & (lam < 0.95))[0]
    for s in list(rng.choice(blend_sites, min(3, len(blend_sites)), replace=False)) + list(rng.choice(n, 2)):
        for ax in range(3):
            rp = r.copy(); rp[s, ax] += h
            rm = r.copy(); rm[s, ax] -= h
            fd = -(SO.hybrid_energy_forces(p, rp, box, *args)[0] - SO.hybrid_energy_forces(p, rm, box, *args)[0]) / (2 * h)
            assert abs(fd - F[s, ax]) < 2e-6 * max(1.0, abs(F[s, ax])), (s, ax, fd, F[s, ax])


@pytest.mark.parametrize("case", ["fcc", "slab", "chains"])
def test_cell_list_neighbours_equal_brute_force(case):
    """The O(N) cell-list search returns the brute force's pairs, FP64 displacements and
    (i, j) order, for all sites and for a subset of centres (a missed neighbour cell, a
    wrong periodic image or a double-counted cell would change them)."""
    from komd_amd import systems as S
    s = {"fcc": lambda: S.fcc(6, 0.8, 1.0, 21, jitter=0.2),
         "slab": lambda: S.simple_cubic(24, 8, 8, 0.8, 1.0, 22),
         "chains": lambda: S.chains((5, 5, 5), 4, 0.7, seed=23)}[case]()
    box = np.asarray(s.box, np.float64)
    r = np.mod(s.r - s.low, s.box) + s.low + np.random.default_rng(0).normal(size=s.r.shape) * 0.05
    a = SO.neighbour_pairs(r, box)
    b = SO.neighbour_pairs_cells(r, box)
    assert all(np.array_equal(x, y) for x, y in zip(a, b)) and len(a[0]) > 10 * s.n
    cen = np.random.default_rng(1).choice(s.n, 57, replace=False)
    keep = np.isin(a[0], cen)
    c = SO.neighbour_pairs_cells(r, box, centers=cen)
    assert all(np.array_equal(x[keep], y) for x, y in zip(a, c))


def test_subset_evaluation_equals_full():
    """Descriptors, site energies and forces of a site subset (the network on the subset's
    neighbourhood only) equal the whole-system evaluation, bit for bit."""
    from komd_amd import systems as S
    s = S.fcc(8, 0.8, 0.8, 24, jitter=0.1)
    box = np.asarray(s.box, np.float64)
    r = np.mod(s.r - s.low, s.box) + s.low
    p = SO.init_params(5, e_mean=0.2, e_std=0.05)
    i, j, d = SO.neighbour_pairs(r, box)
    D, _ = SO.descriptors(s.n, i, d)
    _, E, F = SO.energy_forces(p, s.n, i, j, d)
    sub = np.unique(np.random.default_rng(2).choice(s.n, 120, replace=False))
    Ds, Es, Fs = SO.energy_forces_subset(p, r, box, sub)
    assert np.array_equal(Ds, D[sub]) and np.array_equal(Es, E[sub]) and np.array_equal(Fs, F[sub])
