"""Context state across calls, through the C ABI: thermostat targets of captured steps,
site-count changes on a live context, force restore after a failed observable pass,
errors raised inside a graphed run, and the in-box site count of the temperature rule."""
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


def test_thermostat_target_follows_each_run():
    """komd_run captures one step as a HIP graph and replays it; the rescale target t0
    is read from device memory, so a later run with another t0 replays the same graph
    with the new target.  Graph path (timing off) and eager path (timing on) must agree
    through a sequence of targets, and each run ends at its own target."""
    s = S.fcc(6, 0.8, 1.0, 301, jitter=0.05)
    print(f"thermostat targets n={s.n}", flush=True)
    with K.Context(s, nu=NU) as a, K.Context(s, nu=NU) as b:
        a.compute_forces(K.ATM_EXACT)
        b.compute_forces(K.ATM_EXACT)
        for t0 in (1.0, 0.7, 0.9, 0.7):
            a.run(10, kind3b=K.ATM_EXACT, t0=t0)      # graph path
            b.timing(True)
            b.run(10, kind3b=K.ATM_EXACT, t0=t0)      # eager path, kernels launched per step
            b.timing(False)
            ra, va, _ = a.get_state()
            rb, vb, _ = b.get_state()
            # a replay with a stale target would rescale toward the previous t0 at 9 of the
            # 10 steps: an O(1e-2) velocity difference, not rounding
            assert np.abs(ra - rb).max() < 1e-8 and np.abs(va - vb).max() < 1e-6, t0
            assert abs(a.temperature() - t0) < 1e-9 * t0
            assert abs(b.temperature() - t0) < 1e-9 * t0


def test_set_sites_with_more_sites_keeps_surrogate_buffers_sized():
    """komd_set_sites with a larger count on a context that already ran surrogate and
    training passes: every per-site buffer follows the new count, the old training set
    is dropped, and the results equal those of a fresh context."""
    from komd_amd.train import DEFAULT_WEIGHTS
    p = np.load(DEFAULT_WEIGHTS)
    big = S.fcc(8, 0.8, 0.8, 312, jitter=0.05)
    small = big.copy()  # the same box (the context's grid), fewer sites
    keep = np.sort(np.random.default_rng(311).choice(big.n, 900, replace=False))
    for name in ("r", "v", "mass", "eps", "sig", "mol_id"):
        setattr(small, name, getattr(big, name)[keep].copy())
    print(f"resize {small.n} -> {big.n}", flush=True)
    with K.Context(small, nu=NU) as c:
        c.surrogate_set_params(p)
        c.atm_site_energies()
        c.train_add_snapshot()
        c.surrogate_judge()
        assert c.train_size() == 1
        c.set_sites(big)
        assert c.train_size() == 0
        e_big = c.atm_site_energies()
        c.train_add_snapshot()
        c.train_add_snapshot()
        assert c.train_size() == 2
        loss, g = c.train_gradient(1, 0.01, 1.0)
        judge = c.surrogate_judge()
        c.update_container()
        c.forces_3b(K.SURROGATE)
        f = c.forces()
    with K.Context(big, nu=NU) as d:
        d.surrogate_set_params(p)
        e_ref = d.atm_site_energies()
        d.train_add_snapshot()
        d.train_add_snapshot()
        loss_ref, g_ref = d.train_gradient(1, 0.01, 1.0)
        judge_ref = d.surrogate_judge()
        d.update_container()
        d.forces_3b(K.SURROGATE)
        f_ref = d.forces()
    np.testing.assert_allclose(e_big, e_ref, rtol=1e-12, atol=1e-15)
    assert abs(loss - loss_ref) <= 1e-6 * abs(loss_ref)
    np.testing.assert_allclose(g, g_ref, rtol=1e-4, atol=1e-6 * np.abs(g_ref).max())
    assert abs(judge[0] - judge_ref[0]) <= 1e-9
    assert np.array_equal(f, f_ref)


def test_observables_restore_forces_after_a_failed_pass():
    """komd_measure_observables rebuilds the forces in energy mode; when a pass fails (a
    surrogate kind with no parameters set) the forces are still restored, so the next
    komd_run starts from them."""
    s = S.fcc(5, 0.8, 0.8, 321, jitter=0.05)
    with K.Context(s, nu=NU) as c:
        c.compute_forces(K.ATM_EXACT)
        f0 = c.forces()
        assert np.abs(f0).max() > 0
        with pytest.raises(K.KomdError, match="surrogate parameters not set"):
            c.observables(K.SURROGATE)
        assert np.array_equal(c.forces(), f0)
        o = c.observables(K.ATM_EXACT)
        assert np.array_equal(c.forces(), f0)
        assert o["fmax"] == np.abs(f0).max()


def test_graphed_run_reports_list_overflow():
    """A neighbour-list overflow inside a graphed komd_run comes back as KOMD_ERR_OVERFLOW
    (the flags are read every 32 steps), and the context runs normally afterwards from
    a valid state."""
    dense = S.fcc(6, 1.7, 0.8, 45, jitter=0.02)   # 111-128 neighbours, over the 96-entry list
    print(f"run overflow n={dense.n}", flush=True)
    with K.Context(dense, nu=NU, dt=0.001) as c:
        with pytest.raises(K.KomdError) as ei:
            c.run(100, kind3b=K.ATM_EXACT)
        assert ei.value.status in (4, 5)    # overflow (or non-finite after capped forces)
        ok = dense.copy()  # the context's box, at density 0.79
        keep = np.sort(np.random.default_rng(46).choice(dense.n, 400, replace=False))
        for name in ("r", "v", "mass", "eps", "sig", "mol_id"):
            setattr(ok, name, getattr(dense, name)[keep].copy())
        c.set_sites(ok)
        c.compute_forces(K.ATM_EXACT)
        c.run(100, kind3b=K.ATM_EXACT)
        r, v, f = c.get_state()
        assert np.isfinite(r).all() and np.isfinite(v).all() and np.isfinite(f).all()


def test_temperature_counts_only_sites_inside_the_box(orc):
    """TemperatureSensor::measure divides by the sites inside [low, high]
    (TemperatureSensor.cpp:32-36, Geometry.h:14-17).  Chains binned by their centre of mass
    keep sites outside the box after the wrap when the centre lies just inside a face."""
    s = S.chains((5, 5, 5), 4, 0.7, seed=331)
    # shift three chains so that their centres sit just inside the low x face
    for m in (0, 7, 19):
        idx = np.arange(4 * m, 4 * m + 4)
        com = s.r[idx].mean(axis=0)
        s.r[idx] += np.array([0.05 - com[0], 0.0, 0.0])
    s.v = s.v * 1.7
    r_or, _ = orc.update_container(s)
    mv2, cnt = orc.kinetic_sum(s, r_or, s.v)
    assert cnt < s.n, "the case needs sites outside the box"
    with K.Context(s) as c:
        c.update_container()
        r, _, _ = c.get_state()
        assert np.array_equal(r, r_or)
        T = c.temperature()
        o = c.observables(0)
    assert o["n_in_box"] == int(cnt)
    assert abs(T - mv2 / (3 * cnt)) <= 1e-12 * T
