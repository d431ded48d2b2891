"""Pins of the oracle's sensor restatements (virial tensor, radial distribution)
against facts that do not come from the oracle's own formulas.  (CPU only.)

- Virial: for an isolated cluster, W_ab = sum_m r_m,a F_m,b = -dU/de under the
  homogeneous strain r -> (I + e E_ab) r, where U comes from the brute-force energy
  functions.  A dropped ATM pair term, a wrong sign, a transposed component or a
  missing zz term fails it.
- RDF: the coordination shells of a perfect FCC lattice (12, 6, 24, 12, 24, 8, 48
  neighbours at a sqrt(n/2)), and a numpy histogram of minimum-image distances.
"""
import math

import numpy as np
import pytest

import oracle as O
from komd_amd import systems as S

NU = 0.073
TH0 = math.radians(109.5)
# order of the symmetric components (PressureSensor.h:18)
SYM = [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]


def _cluster(cells=3, seed=5, box=40.0, jitter=0.08):
    """Jittered FCC block of 4 cells^3 atoms in the middle of a large periodic box, so
    no pair or triplet reaches a periodic image."""
    s = S.fcc(cells, 0.8, 1.0, seed, jitter=jitter)
    s.r = s.r - s.r.mean(0) + box / 2
    s.low, s.high = np.zeros(3), np.full(3, box)
    s.cell_size = S.cell_size_for(box, 2.5)
    return s


def _chain_cluster(box=40.0):
    """8 four-site chains (bonds, angles) packed near the box centre."""
    c = S.chains((2, 2, 2), sites=4, site_density=0.7, seed=11)
    c.r = c.r - c.r.mean(0) + box / 2
    c.low, c.high = np.zeros(3), np.full(3, box)
    c.cell_size = S.cell_size_for(box, 2.5)
    return c


def _energy(s, r, nu, bonded):
    _, u = O.lj_brute(s, r)
    if nu:
        _, e, _ = O.atm_brute(s, r, nu, sign=-1.0)
        u += e
    if bonded:
        _, ub = O.bonded(s, r)
        u += ub
    return u


def _strain_virial(s, nu, bonded, h=1e-6):
    r0 = s.r - s.r.mean(0)
    centre = s.r.mean(0)
    w = np.zeros(6)
    for k, (a, b) in enumerate(SYM):
        def u_of(e):
            r = r0.copy()
            r[:, a] += e * r0[:, b]
            return _energy(s, r + centre, nu, bonded)
        w[k] = -(u_of(h) - u_of(-h)) / (2 * h)
    return w


def _no_pair_near_cutoff(s, rc=2.5, gap=1e-5):
    d = s.r[:, None, :] - s.r[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    return np.all(np.abs(r[np.triu_indices(s.n, 1)] - rc) > gap)


@pytest.mark.parametrize("nu", [0.0, NU])
def test_virial_is_minus_strain_derivative_lj_atm(nu):
    s = _cluster()
    assert _no_pair_near_cutoff(s), "choose another seed: a pair sits on the cutoff"
    w = O.virial(s, s.r, nu=nu, bonded=False)
    ref = _strain_virial(s, nu, bonded=False)
    scale = np.abs(ref).max()
    assert np.abs(w - ref).max() < 1e-6 * scale, (w, ref)
    if nu:
        # the three-body part alone is non-zero in every component, zz included
        w_atm = w - O.virial(s, s.r, nu=0.0, bonded=False)
        assert np.all(np.abs(w_atm[[0, 2, 5]]) > 1e-3 * np.abs(w_atm).max())


def test_virial_is_minus_strain_derivative_bonded():
    s = _chain_cluster()
    assert _no_pair_near_cutoff(s)
    w = O.virial(s, s.r, nu=0.0, bonded=True)
    ref = _strain_virial(s, 0.0, bonded=True)
    assert np.abs(w - ref).max() < 1e-6 * np.abs(ref).max(), (w, ref)


def test_virial_trace_equals_pair_virial_periodic():
    """Periodic box: the trace equals the scalar sum r_ij . F_ij of the C08 pair walk."""
    s = S.fcc(5, 0.8, 1.0, 3, jitter=0.1)
    r, b = O.update_container(s)
    _, _, wscal = O.lj_c08(s, r, b, energy=True)
    w = O.virial(s, r)
    assert abs(w[0] + w[2] + w[5] - wscal) < 1e-9 * abs(wscal)


def test_rdf_fcc_coordination_shells():
    """Perfect FCC, a = (4/rho)^(1/3), L = 4a: shells at a sqrt(n/2) hold 12, 6, 24, 12, 24,
    8, 48 neighbours (n = 1..7), all below L/2 so minimum image does not fold them."""
    s = S.fcc(4, 0.8, 1.0, 1)
    a = (4.0 / 0.8) ** (1.0 / 3.0)
    n = s.n
    dr = 0.01
    max_r = 1.95 * a  # between shell 7 (1.87 a) and shell 8 (2 a)
    nb = int(max_r / dr)
    com = O.molecule_coms(s, s.r)
    cnt = O.rdf(com, s.box, max_r, dr, nb)
    shells = {1: 12, 2: 6, 3: 24, 4: 12, 5: 24, 6: 8, 7: 48}
    expect = np.zeros(nb, np.uint64)
    for k, z in shells.items():
        x = a * math.sqrt(k / 2.0) / dr
        assert abs(x - round(x)) > 1e-6  # no shell on a bin edge
        expect[int(x)] += n * z
    assert np.array_equal(cnt, expect)


def test_rdf_matches_numpy_histogram():
    rng = np.random.default_rng(7)
    box = np.array([8.0, 9.0, 7.5])
    com = rng.uniform(0, 1, (300, 3)) * box
    max_r, dr = 4.2, 0.1
    nb = int(max_r / dr)
    cnt = O.rdf(com, box, max_r, dr, nb)
    d = com[:, None, :] - com[None, :, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d ** 2).sum(-1))[~np.eye(300, dtype=bool)]
    r = r[r <= max_r]
    ref = np.bincount(np.minimum((r / dr).astype(np.int64), nb - 1), minlength=nb)
    assert np.array_equal(cnt, ref.astype(np.uint64))
    assert cnt.sum() % 2 == 0  # ordered pairs come in twos


def test_molecule_coms_of_chains():
    c = S.chains((2, 2, 2), sites=4, seed=3)
    com = O.molecule_coms(c, c.r)
    ref = c.r.reshape(-1, 4, 3).mean(1)
    assert np.allclose(com, ref, rtol=0, atol=1e-12)
