"""GPU sensors through the C ABI against the oracle: stress tensor (kinetic and
virial parts of komd_measure_observables) and the radial distribution (komd_rdf).

Bars:
  * RDF counts: bit-exact (integer histogram of FP64 distances computed with the
    reference's operation order);
  * kinetic tensor: FP64 sums, 1e-12 relative;
  * virial tensor of LJ, ATM and bonded terms: FP32 pair/triplet arithmetic with FP64
    accumulation, 2e-5 of the largest component;
  * surrogate virial: against sum_i r_i (x) F_i of the FP64 network on an isolated
    cluster, 2e-2 of the largest component (the BF16 network's force error).
"""
import ctypes

import numpy as np
import pytest

import komd_amd as K
from komd_amd import systems as S

pytestmark = pytest.mark.gpu

NU = 0.073


@pytest.fixture(scope="module")
def orc():
    import oracle
    oracle.build()
    return oracle


def _tensor_close(a, b, tol):
    scale = np.abs(b).max()
    err = np.abs(a - b).max() / scale
    assert err < tol, (err, a, b)
    return err


@pytest.mark.parametrize("kind", [K.ATM_EXACT, K.ATM_REF])
def test_virial_tensor_lj_atm(orc, kind):
    s = S.fcc(6, 0.8, 0.8, 21, jitter=0.12)
    print(f"virial lj+atm kind={kind} n={s.n}", flush=True)
    r, b = orc.update_container(s)
    ref = orc.virial(s, r, nu=NU, site_cell=b.site_cell, coverage=1 if kind == K.ATM_REF else 0)
    with K.Context(s, nu=NU) as c:
        o = c.observables(kind)
        o2 = c.observables(0)
    _tensor_close(o["virial_tensor"], ref, 2e-5)
    _tensor_close(o2["virial_tensor"], orc.virial(s, r), 2e-5)
    assert abs(o["virial"] - (ref[0] + ref[2] + ref[5])) < 6e-5 * np.abs(ref).max()


def test_virial_tensor_bonded_chains(orc):
    s = S.chains((6, 6, 6), 4, 0.7, seed=23)
    s.r = s.r + np.random.default_rng(23).normal(size=s.r.shape) * 0.05
    print(f"virial chains n={s.n}", flush=True)
    ref = orc.virial(s, s.r)
    with K.Context(s, bin_mode=K.BIN_SITE, topology={"bonds": s.bonds, "angles": s.angles}) as c:
        o = c.observables(0)
    _tensor_close(o["virial_tensor"], ref, 2e-5)


def test_kinetic_tensor(orc):
    s = S.fcc(5, 0.8, 1.3, 24)
    s.mass = 1.0 + 0.5 * np.random.default_rng(24).uniform(size=s.n)
    with K.Context(s) as c:
        o = c.observables(0)
    ref = np.einsum("i,ia,ib->ab", s.mass, s.v, s.v)
    ref6 = np.array([ref[0, 0], ref[0, 1], ref[1, 1], ref[0, 2], ref[1, 2], ref[2, 2]])
    _tensor_close(o["kinetic_tensor"], ref6, 1e-12)
    assert abs(0.5 * (ref6[0] + ref6[2] + ref6[5]) - o["kinetic"]) < 1e-12 * o["kinetic"]


def test_virial_surrogate_cluster(orc):
    from komd_amd.train import DEFAULT_WEIGHTS
    from oracle import surrogate as SO
    s = S.fcc(3, 0.8, 1.0, 25, jitter=0.08)
    box = 40.0
    s.r = s.r - s.r.mean(0) + box / 2
    s.low, s.high = np.zeros(3), np.full(3, box)
    s.cell_size = S.cell_size_for(box, 2.5)
    p = np.load(DEFAULT_WEIGHTS)
    print(f"virial surrogate cluster n={s.n}", flush=True)
    with K.Context(s, nu=NU) as c:
        c.surrogate_set_params(p)
        w_all = c.observables(K.SURROGATE)["virial_tensor"]
        w_lj = c.observables(0)["virial_tensor"]
    i, j, d = SO.neighbour_pairs(s.r, np.asarray(s.box, np.float64))
    _, _, f = SO.energy_forces(p.astype(np.float64), s.n, i, j, d)
    x = s.r - s.r.mean(0)
    w = np.einsum("ia,ib->ab", x, f)
    w = 0.5 * (w + w.T)
    ref6 = np.array([w[0, 0], w[0, 1], w[1, 1], w[0, 2], w[1, 2], w[2, 2]])
    _tensor_close(w_all - w_lj, ref6, 2e-2)


@pytest.mark.parametrize("case", ["fcc4000", "chains", "wide", "uneven", "fcc32000"])
def test_rdf_bit_exact(orc, case):
    if case == "fcc4000":
        s, max_r, dr = S.fcc(10, 0.8, 1.0, 31, jitter=0.1), 5.0, 0.1
    elif case == "chains":
        s, max_r, dr = S.chains((8, 8, 8), 4, 0.7, seed=32), 6.0, 0.05
    elif case == "wide":  # max_r beyond L/2: minimum image folds the far shells
        s, max_r, dr = S.fcc(10, 0.8, 1.0, 33, jitter=0.1), 15.0, 0.25
    elif case == "uneven":  # max_r / dr not an integer: the last bin absorbs [49 dr, max_r]
        s, max_r, dr = S.fcc(10, 0.8, 1.0, 34, jitter=0.1), 5.07, 0.1
    else:
        s, max_r, dr = S.fcc(20, 0.8, 1.0, 35, jitter=0.1), 5.0, 0.1
    nb = K.rdf_bins(max_r, dr)
    print(f"rdf {case} n={s.n} max_r={max_r} dr={dr} bins={nb}", flush=True)
    with K.Context(s) as c:
        cnt = c.rdf(max_r, dr)
        r, _, _ = c.get_state()
    ref = orc.rdf(orc.molecule_coms(s, r), s.box, max_r, dr, nb, nthreads=16)
    assert np.array_equal(cnt, ref)
    m = len(s.mol_start()) - 1
    assert int(cnt.sum()) <= m * (m - 1) and int(cnt.sum()) % 2 == 0


def _raw_rdf(c, max_r, dr, nbins):
    out = np.zeros(max(nbins, 1), np.uint64)
    c._check(c.lib.komd_rdf(c._ctx, float(max_r), float(dr), int(nbins), out.ctypes.data_as(ctypes.c_void_p)))
    return out


def test_rdf_argument_checks():
    """Host-side checks: every call below is rejected before any kernel runs."""
    s = S.fcc(5, 0.8, 1.0, 36)
    with K.Context(s) as c:
        with pytest.raises(K.KomdError):
            _raw_rdf(c, 5.0, 0.1, 49)  # nbins must be (uint64)(max_r / dr) = 50
        with pytest.raises(K.KomdError):
            _raw_rdf(c, 5000.0, 0.1, 50000)  # above the histogram capacity
        with pytest.raises(K.KomdError):
            _raw_rdf(c, 5.0, -0.1, 50)
