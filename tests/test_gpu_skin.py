"""Verlet-list (skinned) runs of the exact three-body argon path, through the C ABI.

The reference reuses its pair list for pair_steps steps without a skin
(PairList.h:22-24, LinkedCells.cpp:130-131), so its stale steps miss pairs; the
skinned run keeps every site within cutoff + skin at its last rebuild, tests the
cutoff at the current positions every step and rebuilds before any site has moved
half the skin.  Bars:
  * forces after steps on a list built several steps earlier equal the FP64
    oracle's LJ (C08 walk) + ATM at the same positions: max 2e-4, RMS 5e-6 of the
    RMS force (the bars of the one-step force tests);
  * the unique in-range pairs and triplet visits summed over a run equal those of
    a run that rebuilds the container every step from the same state, and the
    trajectories agree to 1e-9 (only FP32/FP64 summation orders differ);
  * a list over its capacity is an error, runs the lists cannot serve ignore the
    skin (bit-identical to a context without one).
"""
import numpy as np
import pytest

import komd_amd as K
from komd_amd import systems as S

pytestmark = pytest.mark.gpu

NU = 0.073
SKIN = 0.3


def rel_err(a, b):
    scale = np.sqrt(np.mean(b * b))
    d = np.linalg.norm(a - b, axis=1)
    return float(d.max() / scale), float(np.sqrt(np.mean(d * d)) / scale)


@pytest.fixture(scope="module")
def orc():
    import oracle
    oracle.build()
    return oracle


def argon(cells, seed, skin=SKIN, rho=0.8, T=0.8):
    """FCC argon whose cells have an edge of at least cutoff + skin."""
    return S.fcc(cells, rho, T, seed, rc=2.5 + skin)


def oracle_forces(orc, s, r, nthreads=8):
    t = s.copy()
    t.r = r
    rr, b = orc.update_container(t)
    f_lj = orc.lj_c08(t, rr, b, nthreads=nthreads)
    f_atm = orc.atm_complete(t, rr, b, NU, coverage=0, sign=-1.0, nthreads=nthreads)[0]
    return f_lj + f_atm


@pytest.mark.parametrize("cells,seed,steps", [(10, 61, 6), (10, 62, 60), (20, 63, 40), (40, 1004, 40)])
def test_skin_forces_match_oracle(orc, cells, seed, steps):
    s = argon(cells, seed)
    print(f"skin forces n={s.n} steps={steps}", flush=True)
    with K.Context(s, nu=NU) as c:
        c.set_skin(SKIN)
        c.compute_forces(K.ATM_EXACT)
        c.run(steps, kind3b=K.ATM_EXACT)
        reb = c.skin_rebuilds()
        r, _, f = c.get_state()
    assert np.all((r >= s.low) & (r <= s.high)), "the run must end with wrapped positions"
    f_or = oracle_forces(orc, s, r, nthreads=16 if s.n > 100000 else 8)
    mx, rms = rel_err(f, f_or)
    print(f"  rebuilds {reb}; forces vs oracle max {mx:.3e} rms {rms:.3e}")
    assert 1 <= reb < steps
    if steps == 6:
        assert reb == 1, "the case needs its last step on the list of the first"
    assert mx < 2e-4 and rms < 5e-6


def fp64_run(orc, s, steps, dt=0.005):
    """The FP64 trajectory of the oracle: vv_step0 -> container update -> LJ (C08 walk) +
    ATM over every triplet -> vv_step1 (the order of Simulation.cpp:49-74)."""
    def forces(r):
        rb, b = orc.update_container(s, r)
        f = orc.lj_c08(s, rb, b, nthreads=8)
        return rb, f + orc.atm_complete(s, rb, b, NU, coverage=0, sign=-1.0, nthreads=8)[0]

    r, f = forces(s.r)
    v = s.v.copy()
    for _ in range(steps):
        r, v = orc.vv_step0(dt, r, v, f, s.mass)
        r, f = forces(r)
        v = orc.vv_step1(dt, v, f, s.mass)
    return r, v


def min_image(d, box):
    return d - box * np.round(d / box)


@pytest.mark.parametrize("skin,steps", [(SKIN, 80), (0.02, 30)])
def test_skin_run_follows_plain_run(orc, skin, steps):
    """skin 0.02: a rebuild nearly every step (the rebuild path itself); 0.3: rare.
    Both GPU runs round positions to FP32 relative to a cell centre (the per-step run to
    the current cell's, the skinned run to the build-time cell's), so their trajectories
    part at the FP32 level; the skinned run must track the FP64 trajectory as closely as
    the per-step run does, with the same interaction set every step."""
    s = argon(10, 64, skin=skin)
    print(f"skin vs plain n={s.n} skin={skin} steps={steps}", flush=True)
    with K.Context(s, nu=NU) as a, K.Context(s, nu=NU) as b:
        b.set_skin(skin)
        out = []
        for c in (a, b):
            c.compute_forces(K.ATM_EXACT)
            c.counters(reset=True)
            c.run(steps, kind3b=K.ATM_EXACT)
            out.append((c.get_state(), c.counters()))
        reb = b.skin_rebuilds()
    (ra, va, _), ca = out[0]
    (rb, vb, _), cb = out[1]
    r64, v64 = fp64_run(orc, s, steps)
    box = s.box
    ea = float(np.abs(min_image(ra - r64, box)).max())
    eb = float(np.abs(min_image(rb - r64, box)).max())
    eva, evb = float(np.abs(va - v64).max()), float(np.abs(vb - v64).max())
    print(f"  rebuilds {reb}; vs FP64: per-step |dr| {ea:.2e} |dv| {eva:.2e}, skinned |dr| {eb:.2e} "
          f"|dv| {evb:.2e}; counts {ca} {cb}")
    assert ca == cb, "the same in-range pairs and triplets every step"
    assert eb < 3.0 * ea + 1e-9 and evb < 3.0 * eva + 1e-8
    assert eb < 1e-5 and evb < 1e-4
    if skin < 0.1:
        assert reb > steps // 2
    else:
        assert reb < steps // 4


def test_skin_many_rebuild_segments_in_one_run(orc):
    """A long run at T* 1.5 crosses many rebuilds; its end state still has exact forces."""
    s = argon(10, 65, T=1.5)
    with K.Context(s, nu=NU) as c:
        c.set_skin(SKIN)
        c.compute_forces(K.ATM_EXACT)
        c.run(300, kind3b=K.ATM_EXACT)
        reb = c.skin_rebuilds()
        r, _, f = c.get_state()
    mx, rms = rel_err(f, oracle_forces(orc, s, r))
    print(f"skin 300 steps T1.5: rebuilds {reb}; forces max {mx:.3e} rms {rms:.3e}")
    assert 10 <= reb < 150
    assert mx < 2e-4 and rms < 5e-6


@pytest.mark.parametrize("skin,kind", [(SKIN, K.ATM_EXACT), (0.0, K.ATM_EXACT), (SKIN, 0)])
def test_graph_replay_matches_eager_steps(skin, kind):
    """komd_run replays graphs of eight steps (from step parity 0 on Verlet lists) and of one
    step; with kernel timing on it queues the same kernels step by step.  A run of 5 steps
    and then one of 27 (which starts at the odd parity) must give bit-identical states."""
    s = argon(10, 76)
    out = []
    for eager in (False, True):
        with K.Context(s, nu=NU) as c:
            c.set_skin(skin)
            c.timing(eager)
            c.compute_forces(kind)
            c.run(5, kind3b=kind)
            c.run(27, kind3b=kind)
            out.append(c.get_state())
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


def test_skin_list_overflow_is_an_error():
    """Density 1.1 with cutoff + skin = 3.1: about 136 sites within the list radius, over the
    list capacity of 128."""
    s = S.fcc(8, 1.1, 0.8, 66, rc=3.1)
    with K.Context(s, nu=NU) as c:
        c.set_skin(0.6)
        c.compute_forces(K.ATM_EXACT)
        with pytest.raises(K.KomdError) as ei:
            c.run(5, kind3b=K.ATM_EXACT)
    assert ei.value.status == 4 and "Verlet-list capacity" in str(ei.value)


def test_skin_arguments():
    s = argon(10, 67)
    with K.Context(s, nu=NU) as c:
        for bad in (-0.1, float("nan"), float("inf")):
            with pytest.raises(K.KomdError) as ei:
                c.set_skin(bad)
            assert ei.value.status == 1
        # the cells of this system are 2.8 + a little: a skin of 0.5 needs 3.0
        with pytest.raises(K.KomdError) as ei:
            c.set_skin(0.5)
        assert ei.value.status == 1 and "cell edge" in str(ei.value)
        c.set_skin(0.3)
        c.set_skin(0.0)


def test_skin_is_ignored_where_it_does_not_apply():
    """Thermostatted runs and the reference-mode ATM keep their per-step container update:
    bit-identical to a context without a skin."""
    s = argon(10, 68)
    with K.Context(s, nu=NU) as a, K.Context(s, nu=NU) as b:
        b.set_skin(SKIN)
        res = []
        for c in (a, b):
            c.compute_forces(K.ATM_EXACT)
            c.run(20, kind3b=K.ATM_EXACT, t0=0.9)
            c.run(20, kind3b=0, t0=0.9)
            c.run(10, kind3b=K.ATM_REF)
            res.append(c.get_state())
        assert b.skin_rebuilds() == 0
    for x, y in zip(res[0], res[1]):
        assert np.array_equal(x, y)


def test_skin_is_ignored_for_chains_it_cannot_serve():
    """Multi-site molecules binned by their centre of mass, or whose molecule id reappears
    after another molecule (the same-molecule exclusion then spans several runs of sites),
    keep the per-step container update (bit-identical)."""
    t, topo = warmed_chains((6, 6, 6), 71)
    u = t.copy()
    u.mol_id = t.mol_id.copy()
    u.mol_id[-4:] = t.mol_id[0]  # the last chain takes the first chain's id
    for sysm, mode in ((t, K.BIN_COM), (u, K.BIN_SITE)):
        out = []
        for skin in (0.0, SKIN):
            with K.Context(sysm, topology=topo, bin_mode=mode) as c:
                c.set_skin(skin)
                c.compute_forces(0, bonded=True)
                c.run(20, kind3b=0, bonded=True)
                out.append(c.get_state())
                assert c.skin_rebuilds() == 0
        for x, y in zip(out[0], out[1]):
            assert np.array_equal(x, y)


def test_set_sites_drops_a_topology_it_no_longer_fits():
    """The bonded terms index sites and are grouped by molecule: new sites with the same count
    and molecules keep them, any other count drops them (no bonded force, no bonded energy)."""
    t, topo = warmed_chains((6, 6, 6), 74)
    with K.Context(t, topology=topo, bin_mode=K.BIN_SITE) as c:
        c.update_container()
        c.forces_bonded()
        f_b = c.forces()
        c.set_sites(t)                      # same sites and molecules: terms kept
        c.update_container()
        c.forces_bonded()
        assert np.array_equal(c.forces(), f_b) and np.abs(f_b).max() > 0
        assert c.observables(0)["u_bond"] > 0
        s2 = S.chains((4, 4, 4), 4, 0.7, seed=75, rc=2.5 + SKIN)
        c.set_sites(s2)                     # another count: terms dropped
        c.update_container()
        c.forces_bonded()
        assert not c.forces().any()
        o = c.observables(0)
        assert o["u_bond"] == 0 and o["u_angle"] == 0


def trained_params():
    from komd_amd.train import DEFAULT_WEIGHTS
    return np.load(DEFAULT_WEIGHTS)


@pytest.mark.parametrize("kind,steps,cells", [(K.SURROGATE, 60, 10), (K.HYBRID, 40, 10), (K.SURROGATE, 40, 40)])
def test_skin_learned_run_matches_fresh_forces_and_plain_run(kind, steps, cells):
    """The learned and hybrid kinds on Verlet lists: each step's rows come from the lists at
    the current positions (k_skin_refresh), so the forces at the end of the run equal a fresh
    force pass (container update, lists, LJ, three-body) at the same positions to the FP32
    rounding of the list vectors, and the run follows the per-step run to the FP32 level of
    the relative coordinates (the surrogate sums its descriptors in the list order)."""
    s = argon(cells, 70)
    p = trained_params()
    center = float(s.low[0] + 0.5 * s.box[0])
    print(f"skin learned kind={kind} n={s.n} steps={steps}", flush=True)
    with K.Context(s, nu=NU) as a, K.Context(s, nu=NU) as b:
        b.set_skin(SKIN)
        out = []
        for c in (a, b):
            c.surrogate_set_params(p)
            if kind == K.HYBRID:
                c.set_hybrid(0, center, 2.0, 2.0)
            c.compute_forces(kind)
            c.run(steps, kind3b=kind)
            out.append(c.get_state())
        reb = b.skin_rebuilds()
        rb, vb, fb = out[1]
        b.compute_forces(kind)               # a fresh pass at the run's final positions
        f_fresh = b.forces()
    ra, va, fa = out[0]
    mx, rms = rel_err(fb, f_fresh)
    dr = float(np.abs(min_image(ra - rb, s.box)).max())
    dv = float(np.abs(va - vb).max())
    print(f"  rebuilds {reb}; forces vs fresh max {mx:.2e} rms {rms:.2e}; vs per-step run |dr| {dr:.2e} |dv| {dv:.2e}")
    assert 1 <= reb < steps // 2
    if s.n > 100000:
        # The descriptors enter the MLP in BF16 (8-bit mantissa): an FP32-level difference of
        # a descriptor between the two frames of reference flips its BF16 rounding at a few of
        # the 256,000 sites, a step of 2^-8 in one input (256k: max 1.4e-4, RMS 1.6e-6; the
        # model's own error against the exact ATM is 3e-2)
        assert mx < 1e-3 and rms < 1e-5
        assert dr < 1e-4 and dv < 1e-3
    else:
        assert mx < 1e-4 and rms < 1e-5
        assert dr < 1e-5 and dv < 1e-4


# ---- runs without a three-body term: LJ (and the bonded terms) from the Verlet lists

@pytest.mark.parametrize("cells", [10, 30])
def test_skin_lj_only_run(orc, cells):
    """Single-site LJ without a three-body term: the forces after the run equal the oracle's
    LJ at the final positions, and every step has the pair set of a per-step run.  4,000
    atoms walk their lists with 16 lanes per site, 108,000 with 8."""
    s = argon(cells, 72)
    print(f"skin LJ-only n={s.n}", flush=True)
    with K.Context(s) as a, K.Context(s) as b:
        b.set_skin(SKIN)
        out = []
        for c in (a, b):
            c.compute_forces(0)
            c.counters(reset=True)
            c.run(80, kind3b=0)
            out.append((c.get_state(), c.counters()))
        reb = b.skin_rebuilds()
    (ra, va, _), ca = out[0]
    (rb, vb, fb), cb = out[1]
    t = s.copy()
    t.r = rb
    rr, bb = orc.update_container(t)
    mx, rms = rel_err(fb, orc.lj_c08(t, rr, bb, nthreads=8))
    dr = float(np.abs(min_image(ra - rb, s.box)).max())
    print(f"  rebuilds {reb}; forces vs oracle max {mx:.2e} rms {rms:.2e}; vs per-step |dr| {dr:.2e}; "
          f"counts {ca} {cb}")
    assert ca == cb
    assert 1 <= reb < 25
    assert mx < 2e-4 and rms < 5e-6
    assert dr < 1e-5


def warmed_chains(nmol_xyz, seed, sites=4):
    """Chains on a site grid of cells >= cutoff + skin after the capped warm-up of the
    benchmark (force cap raised from 20 to 2e5 in five 300-step stages at dt 0.001 with
    rescaling to T* 1.0), sites binned on their own positions."""
    s = S.chains(nmol_xyz, sites, 0.7, seed=seed, rc=2.5 + SKIN)
    topo = {"bonds": s.bonds, "angles": s.angles}
    for lim in (20.0, 200.0, 2000.0, 2e4, 2e5):
        with K.Context(s, cutoff=2.5, dt=0.001, limit_factor=lim, topology=topo, bin_mode=K.BIN_SITE) as c:
            c.compute_forces(0, bonded=True)
            c.run(300, kind3b=0, bonded=True, t0=1.0)
            s.r, s.v, _ = c.get_state()
    return s, topo


def chain_oracle_forces(orc, s, r, nthreads=8, bonded=True):
    """FP64 LJ between molecules over the site grid (the chains' ids exclude same-molecule
    pairs, LJ12_6.cpp:19) plus the harmonic bonds and angles; returns (wrapped r, f)."""
    single = s.copy()
    single.mol_id = np.arange(s.n, dtype=np.uint64)
    rr, b = orc.update_container(single, r)
    f = orc.lj_c08(s, rr, b, nthreads=nthreads)
    return rr, (f + orc.bonded(s, rr)[0]) if bonded else f


def chain_runs(s, topo, steps, bonded=True):
    """The same NVE run per step and on Verlet lists: (state, counters) of each, rebuilds."""
    with K.Context(s, topology=topo, bin_mode=K.BIN_SITE) as a, \
            K.Context(s, topology=topo, bin_mode=K.BIN_SITE) as b:
        b.set_skin(SKIN)
        out = []
        for c in (a, b):
            c.compute_forces(0, bonded=bonded)
            c.counters(reset=True)
            c.run(steps, kind3b=0, bonded=bonded)
            out.append((c.get_state(), c.counters()))
        return out, b.skin_rebuilds()


@pytest.mark.parametrize("bonded,steps,sites", [(True, 60, 4), (False, 30, 4), (True, 30, 6), (False, 20, 6)])
def test_skin_chains_follow_fp64_run(orc, bonded, steps, sites):
    """512 chains of four or six sites, with their bonds and angles or without (LJ between
    molecules only): the same pair set every step as the per-step run (the lists leave
    same-molecule pairs out; the bonded kernels count the in-range ones, six-site molecules
    past their four-site register block), forces at the end equal to the oracle's, and both
    GPU runs track the FP64 trajectory of the oracle (reference step order) equally well."""
    s, topo = warmed_chains((8, 8, 8), 73, sites)
    print(f"skin chains n={s.n} sites={sites} bonded={bonded} steps={steps}", flush=True)
    out, reb = chain_runs(s, topo, steps, bonded)
    (ra, va, _), ca = out[0]
    (rb, vb, fb), cb = out[1]
    _, f_or = chain_oracle_forces(orc, s, rb, bonded=bonded)
    mx, rms = rel_err(fb, f_or)
    dt = 0.005
    r, f = chain_oracle_forces(orc, s, s.r, bonded=bonded)
    v = s.v.copy()
    for _ in range(steps):
        r, v = orc.vv_step0(dt, r, v, f, s.mass)
        r, f = chain_oracle_forces(orc, s, r, bonded=bonded)
        v = orc.vv_step1(dt, v, f, s.mass)
    ea = float(np.abs(min_image(ra - r, s.box)).max())
    eb = float(np.abs(min_image(rb - r, s.box)).max())
    eva, evb = float(np.abs(va - v).max()), float(np.abs(vb - v).max())
    print(f"  rebuilds {reb}; forces vs oracle max {mx:.2e} rms {rms:.2e}; vs FP64: per-step |dr| {ea:.2e} "
          f"|dv| {eva:.2e}, skinned |dr| {eb:.2e} |dv| {evb:.2e}; counts {ca} {cb}")
    assert ca == cb, "the same in-range pairs every step"
    assert 1 <= reb < steps // 3
    assert mx < 2e-4 and rms < 1e-5
    assert eb < 3.0 * ea + 1e-9 and evb < 3.0 * eva + 1e-8
    assert eb < 1e-5 and evb < 1e-4


def test_skin_chains_262144_full_size(orc):
    """The chain workload at its full size (65,536 chains) on Verlet lists after the
    benchmark's warm-up, 40 steps: forces at the end equal to the oracle's, and the run
    tracks the FP64 trajectory as closely as the per-step run does.  The two GPU runs part at
    the FP32 level and the chains amplify that (|dr| ~2e-5 after 100 steps), so a pair that
    close to the cutoff may be counted on either side: the counts agree to 1e-8 here, exactly
    in the 2,048-site test."""
    s, topo = warmed_chains((64, 32, 32), 1002)
    steps = 40
    print(f"skin chains full size n={s.n} steps={steps}", flush=True)
    out, reb = chain_runs(s, topo, steps)
    (ra, va, _), ca = out[0]
    (rb, vb, fb), cb = out[1]
    _, f_or = chain_oracle_forces(orc, s, rb, nthreads=16)
    mx, rms = rel_err(fb, f_or)
    dt = 0.005
    r, f = chain_oracle_forces(orc, s, s.r, nthreads=16)
    v = s.v.copy()
    for _ in range(steps):
        r, v = orc.vv_step0(dt, r, v, f, s.mass)
        r, f = chain_oracle_forces(orc, s, r, nthreads=16)
        v = orc.vv_step1(dt, v, f, s.mass)
    ea = float(np.abs(min_image(ra - r, s.box)).max())
    eb = float(np.abs(min_image(rb - r, s.box)).max())
    eva, evb = float(np.abs(va - v).max()), float(np.abs(vb - v).max())
    print(f"  rebuilds {reb}; forces vs oracle max {mx:.2e} rms {rms:.2e}; vs FP64: per-step |dr| {ea:.2e} "
          f"|dv| {eva:.2e}, skinned |dr| {eb:.2e} |dv| {evb:.2e}; counts {ca} {cb}")
    assert abs(ca[0] - cb[0]) <= 1e-8 * ca[0] and ca[1] == cb[1] == 0
    assert 1 <= reb < steps // 3
    assert mx < 2e-4 and rms < 1e-5
    assert eb < 3.0 * ea + 1e-9 and evb < 3.0 * eva + 1e-8
    assert eb < 1e-4 and evb < 1e-3
