"""GPU parity of the learned three-body surrogate against the FP64 oracle
(oracle/surrogate.py), through the C ABI.

Bars (stated per test):
  * descriptors: FP32 sums of FP32 basis values -> relative 2e-4 of each column's scale;
  * energies / forces / training gradient: BF16 MFMA operands with FP32 accumulation
    -> RMS error a few 1e-3 of the quantity's spread, tolerances written in each test.
A wrong MFMA operand layout or a wrong backward term shows up as O(1) errors.
"""
import numpy as np
import pytest

import komd_amd as K
from komd_amd import systems as S
from oracle import surrogate as SO

pytestmark = pytest.mark.gpu

NU = 0.073


def system(cells=5, seed=3, jitter=0.1):
    return S.fcc(cells, 0.8, 0.8, seed, jitter=jitter)


def oracle_inputs(s):
    r = np.mod(s.r - s.low, s.box) + s.low
    return SO.neighbour_pairs(r, np.asarray(s.box, np.float64))


def params_for(s, seed=1):
    i, j, d = oracle_inputs(s)
    D, _ = SO.descriptors(s.n, i, d)
    rng = np.random.default_rng(seed)
    p = SO.unpack(SO.init_params(seed, shift=D.mean(0), scale=1.0 / (D.std(0) + 1e-9), e_mean=0.25, e_std=0.05))
    p["b1"] = rng.normal(size=SO.H1) * 0.3
    p["b2"] = rng.normal(size=SO.H2) * 0.3
    return SO.pack(p).astype(np.float32)


def test_descriptors_match_oracle():
    s = system(5, 11)
    print(f"descriptors n={s.n}", flush=True)
    with K.Context(s) as c:
        Dg = c.surrogate_descriptors()
    i, j, d = oracle_inputs(s)
    Do, _ = SO.descriptors(s.n, i, d)
    scale = np.abs(Do).max(0) + 1e-12
    err = np.abs(Dg - Do) / scale
    print("max col-relative error", err.max())
    assert err.max() < 2e-4


def test_energies_match_oracle():
    s = system(5, 12)
    p = params_for(s, 2)
    with K.Context(s) as c:
        c.surrogate_set_params(p)
        Eg = c.surrogate_energies()
        assert np.array_equal(c.surrogate_get_params(), p)
    i, j, d = oracle_inputs(s)
    _, Eo, _ = SO.energy_forces(p.astype(np.float64), s.n, i, j, d)
    spread = Eo.std()
    rms = np.sqrt(np.mean((Eg - Eo) ** 2)) / spread
    mx = np.abs(Eg - Eo).max() / spread
    print(f"E rms {rms:.2e} max {mx:.2e} (of the spread {spread:.3e})")
    # BF16 operands: relative rounding 2^-9 per operand, FP32 accumulation
    assert rms < 2e-2 and mx < 1e-1


@pytest.mark.parametrize("cells,seed", [(5, 13), (8, 14)])
def test_forces_match_oracle(cells, seed):
    s = system(cells, seed)
    p = params_for(s, 3)
    print(f"surrogate forces n={s.n}", flush=True)
    with K.Context(s) as c:
        c.surrogate_set_params(p)
        c.update_container()
        c.forces_3b(K.SURROGATE)
        Fg = c.forces()
        obs = c.observables(K.SURROGATE)
    i, j, d = oracle_inputs(s)
    Eo, Ei, Fo = SO.energy_forces(p.astype(np.float64), s.n, i, j, d)
    scale = np.sqrt(np.mean(Fo ** 2))
    dd = np.linalg.norm(Fg - Fo, axis=1)
    rms = np.sqrt(np.mean(dd ** 2)) / scale
    print(f"F rms {rms:.2e} max {dd.max() / scale:.2e}; sum F {np.abs(Fg.sum(0)).max():.2e}")
    assert rms < 2e-2 and dd.max() / scale < 2e-1
    assert np.abs(Fg.sum(0)).max() < 1e-3 * scale * np.sqrt(s.n)
    assert abs(obs["u_3b"] - Eo) < 1e-2 * abs(Eo)


def test_forces_bitwise_repeatable():
    """The per-cell force rows are summed in fixed point, so two evaluations of one
    state (each in a fresh context, forces starting from zero) give the same bits
    whatever order the pair adds ran in."""
    s = system(8, 15)
    p = params_for(s, 4)
    F = []
    for _ in range(2):
        with K.Context(s) as c:
            c.surrogate_set_params(p)
            c.update_container()
            c.forces_3b(K.SURROGATE)
            F.append(c.forces())
    print(f"max |F| {np.abs(F[0]).max():.3e}", flush=True)
    assert np.array_equal(F[0], F[1])


def test_step_forces_lj_in_descriptor_pass():
    """The surrogate NVE step of a uniform single-site system sums LJ in the descriptor
    pass (its pair pass only builds lists).  Its forces equal the separate pair and
    surrogate passes (both pinned to their oracles elsewhere) to rounding: the LJ terms
    are summed in FP64 in another order, and with these parameters some per-pair
    surrogate terms exceed the fixed-point range, so blocks sum in float in any order.
    A missing or doubled LJ term would differ by the LJ force itself."""
    s = system(8, 17)
    p = params_for(s, 7)
    with K.Context(s) as c, K.Context(s) as d:
        c.surrogate_set_params(p)
        c.compute_forces(K.SURROGATE)
        f = c.forces()
        d.surrogate_set_params(p)
        d.update_container()
        d.forces_pair()
        d.forces_3b(K.SURROGATE)
        f_sep = d.forces()
    scale = np.abs(f_sep).max()
    print(f"max |F| {scale:.3e}, max diff {np.abs(f - f_sep).max():.3e}", flush=True)
    assert scale > 0.1
    assert np.abs(f - f_sep).max() <= 1e-6 * scale


def test_forces_large_scale_float_path():
    """Final-layer weights scaled by 1e4 scale every force by 1e4, beyond the fixed-point
    accumulator's per-add range: the blocks fall back to float adds and the result still
    matches the oracle to the bar of the unscaled case."""
    s = system(5, 16)
    p = SO.unpack(params_for(s, 5).astype(np.float64))
    p["w3"] = p["w3"] * 1e4
    p = SO.pack(p).astype(np.float32)
    with K.Context(s) as c:
        c.surrogate_set_params(p)
        c.update_container()
        c.forces_3b(K.SURROGATE)
        Fg = c.forces()
    i, j, d = oracle_inputs(s)
    _, _, Fo = SO.energy_forces(p.astype(np.float64), s.n, i, j, d)
    scale = np.sqrt(np.mean(Fo ** 2))
    dd = np.linalg.norm(Fg - Fo, axis=1)
    rms = np.sqrt(np.mean(dd ** 2)) / scale
    print(f"F rms scale {scale:.3e} rel rms {rms:.2e} max {dd.max() / scale:.2e}")
    assert scale > 100.0  # site forces far beyond the per-add limit of 16 of the fixed-point sums
    assert rms < 2e-2 and dd.max() / scale < 2e-1


def test_training_gradient_matches_oracle():
    s = system(5, 15)
    p = params_for(s, 4)
    wE, wF = 40.0, 10.0
    with K.Context(s, nu=NU) as c:
        c.surrogate_set_params(p)
        c.train_add_snapshot()
        assert c.train_size() == 1
        # the stored targets: exact ATM per-site energies and forces
        E_ref = c.atm_site_energies()
        c.update_container()
        c.forces_3b(K.ATM_EXACT)
        F_ref = c.forces()
        loss_g, grad_g = c.train_gradient(0, wE, wF)
        r_after, _, _ = c.get_state()
    i, j, d = oracle_inputs(s)
    loss_o, grad_o, _, _ = SO.loss_and_grad(p.astype(np.float64), s.n, i, j, d, E_ref, F_ref, wE, wF)
    print(f"loss gpu {loss_g:.6e} oracle {loss_o:.6e}")
    assert abs(loss_g - loss_o) < 2e-2 * loss_o
    sl = slice(0, SO.NPARAM - 2 * SO.NDESC)
    gg, go = grad_g[sl].astype(np.float64), grad_o[sl]
    rel = np.linalg.norm(gg - go) / np.linalg.norm(go)
    cos = gg @ go / (np.linalg.norm(gg) * np.linalg.norm(go))
    print(f"gradient relative error {rel:.3e}, cosine {cos:.6f}")
    assert rel < 5e-2 and cos > 0.998
    # each block of the gradient separately (a missing term in one block shows here)
    o = 0
    for name, k in (("W1", 64 * 48), ("b1", 64), ("W2", 64 * 64), ("b2", 64), ("w3", 64), ("b3", 1)):
        bg, bo = gg[o:o + k], go[o:o + k]
        rb = np.linalg.norm(bg - bo) / (np.linalg.norm(bo) + 1e-30)
        print(f"  {name}: relative error {rb:.3e}")
        assert rb < 1e-1, name
        o += k
    assert not np.array_equal(r_after, None)


def test_training_reduces_loss():
    s = system(5, 16)
    with K.Context(s, nu=NU, dt=0.005) as c:
        for k in range(6):
            c.run(10, kind3b=K.ATM_EXACT, t0=0.8)
            c.train_add_snapshot()
        c.train_standardize()
        e_mean, e_std, f_rms = c.train_stats()
        p = SO.unpack(c.surrogate_get_params().astype(np.float64))
        p0 = SO.unpack(SO.init_params(5, shift=p["shift"], scale=p["scale"], e_mean=e_mean, e_std=e_std))
        c.surrogate_set_params(SO.pack(p0).astype(np.float32))
        wE, wF = 0.1 / e_std ** 2, 1.0 / f_rms ** 2
        losses = [c.train_epoch(2e-3, wE, wF, seed=k) for k in range(25)]
        print("losses", [f"{x:.3e}" for x in losses[::4]])
        assert losses[-1] < 0.5 * losses[0]
        rel, mx = c.surrogate_judge()
        print(f"judge after training: rel rms {rel:.3f} max {mx:.3f}")
        assert rel < 1.0


def test_judge_and_fallback():
    s = system(5, 17)
    p = params_for(s, 6)
    with K.Context(s, nu=NU, dt=0.005) as c:
        c.surrogate_set_params(p)
        rel, mx = c.surrogate_judge()
        assert rel > 0 and mx >= rel
        c.compute_forces(K.SURROGATE)
        c.set_judge(2, 1e-9)          # any error triggers the fallback
        c.run(6, kind3b=K.SURROGATE)
        rep = c.judge_report()
        assert rep["checks"] == 1 and rep["fell_back"]
        # after the fallback the forces are the exact ATM ones
        r, v, f = c.get_state()
    with K.Context(s, nu=NU, dt=0.005) as e:
        e.set_state(r, v)
        e.compute_forces(K.ATM_EXACT)
        fe = e.forces()
    # same configuration, same kernels; only the FP32 summation order of the pair pass may differ
    assert np.abs(f - fe).max() < 1e-5 * np.abs(fe).max()


def test_trained_surrogate_gain():
    """What the trained network gains over the methods without it: three-body force error
    against the exact ATM on a thermostatted argon state (rho 0.8, T* 0.8, 4,000 atoms) for
    the committed weights, and for comparators that cost one pair pass: the mean-field
    ATM2B pair force (ATM2B.cpp:9-56, KOMD_PAIR_ATM2B), an IBI table fitted to the exact
    path's g(r) (the IBI plugin's iteration, IBI.cpp:26-171, on KOMD_PAIR_TABLE) and LJ
    alone.  The comparators' three-body estimate is their pair force minus LJ's."""
    from komd_amd import ibi
    from komd_amd.train import DEFAULT_WEIGHTS
    s = S.fcc(10, 0.8, 0.8, 2024)
    vol = float(np.prod(s.box))
    with K.Context(s, nu=NU, dt=0.005) as c:
        c.surrogate_set_params(np.load(DEFAULT_WEIGHTS))
        c.compute_forces(K.ATM_EXACT)
        c.run(400, kind3b=K.ATM_EXACT, t0=0.8)
        g_target = ibi.measure_g(c, s.n, vol, 2.5, 100, 50, 10, K.ATM_EXACT, 0.8)
        r, v, _ = c.get_state()
        rel, mx = c.surrogate_judge()
        c.update_container()
        c.forces_3b(K.ATM_EXACT)
        f_atm = c.forces()
        c.update_container()
        c.forces_pair()
        f_lj = c.forces()
        c.update_container()
        c.forces_pair(K.PAIR_ATM2B)
        f_2b = c.forces()
    s2 = s.copy()
    s2.r, s2.v = r, v
    with K.Context(s2, dt=0.005) as d:
        _, _, _, hist = ibi.run(d, g_target, s.n, vol, 2.5, 0.8, iterations=8)
        d.set_state(r=r, v=v)
        d.update_container()
        d.forces_pair(K.PAIR_TABLE)
        f_ibi = d.forces()

    def rms(a):
        return float(np.sqrt(np.mean(np.sum(a ** 2, axis=1))))
    scale = rms(f_atm)
    err_2b = rms(f_2b - f_lj - f_atm) / scale
    err_ibi = rms(f_ibi - f_lj - f_atm) / scale
    print(f"three-body force error: surrogate {rel:.4f} (max {mx:.3f}), ATM2B {err_2b:.4f}, IBI {err_ibi:.4f} "
          f"(g RMS error {hist[0]:.4f} -> {hist[-1]:.4f}), LJ only 1.0")
    assert rel < 0.06
    assert rel < 0.25 * err_2b and rel < 0.25 * err_ibi
    assert hist[-1] < hist[0]  # the iteration moves g(r) toward the target


@pytest.mark.parametrize("half,blend", [(1.5, 2.0), (0.8, 1.0)])
def test_hybrid_matches_oracle(half, blend):
    """Exact ATM weighted by the mean lambda of each triplet plus surrogate site energies
    weighted by 1 - lambda, forces including the weight gradients (config-5 blend)."""
    s = system(5, 18, jitter=0.1)
    p = params_for(s, 7)
    box = np.asarray(s.box, np.float64)
    center = s.low[0] + 0.5 * box[0]
    with K.Context(s, nu=NU) as c:
        c.surrogate_set_params(p)
        c.set_hybrid(0, center, half, blend)
        c.update_container()
        c.forces_3b(K.HYBRID)
        Fg = c.forces()
        obs = c.observables(K.HYBRID)
        r_w, _, _ = c.get_state()
    E_o, F_o = SO.hybrid_energy_forces(p.astype(np.float64), r_w, box, NU, 0, center, half, blend)
    scale = np.sqrt(np.mean(np.sum(F_o ** 2, axis=1)))
    dd = np.linalg.norm(Fg - F_o, axis=1)
    rms = np.sqrt(np.mean(dd ** 2)) / scale
    print(f"hybrid F rms {rms:.2e} max {dd.max() / scale:.2e}; E gpu {obs['u_3b']:.6f} oracle {E_o:.6f}")
    # ATM part FP32 (1e-5), surrogate part BF16 operands
    assert rms < 2e-2 and dd.max() / scale < 2e-1
    assert abs(obs["u_3b"] - E_o) < 1e-2 * abs(E_o)


def test_hybrid_limits():
    """lambda = 1 everywhere reproduces the exact ATM; lambda = 0 everywhere the surrogate."""
    s = system(5, 19)
    p = params_for(s, 8)
    with K.Context(s, nu=NU) as c:
        c.surrogate_set_params(p)
        c.update_container()
        c.forces_3b(K.ATM_EXACT)
        f_atm = c.forces()
        c.update_container()
        c.forces_3b(K.SURROGATE)
        f_sur = c.forces()
        c.set_hybrid(1, 0.0, 1e3, 1.0)          # every site inside the exact region
        c.update_container()
        c.forces_3b(K.HYBRID)
        f1 = c.forces()
        c.set_hybrid(1, s.low[1] + 0.123, 0.0, 1e-4)  # no site within 1e-4 of the plane
        c.update_container()
        c.forces_3b(K.HYBRID)
        f0 = c.forces()
    assert np.abs(f1 - f_atm).max() <= 1e-6 * np.abs(f_atm).max()
    assert np.abs(f0 - f_sur).max() <= 1e-5 * np.abs(f_sur).max()
