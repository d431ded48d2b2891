"""The IBI driver's update rules (komd_amd/ibi.py) on the CPU, against hand cases of the
reference plugin's arithmetic (plugins/IBI/IBI.cpp:110-171, RDF_Profiler.cpp:36-54,
IBI_Math.cpp:22-35)."""
import numpy as np

from komd_amd import ibi


def test_nodes_and_normalisation_of_an_ideal_gas():
    """Uniform random points: g(r) -> 1 in every bin (the shell volumes and N (N - 1) / V)."""
    rng = np.random.default_rng(0)
    L, n, rc, bins = 10.0, 1500, 2.5, 25
    r = rng.uniform(0, L, (n, 3))
    d = r[:, None, :] - r[None, :, :]
    d -= L * np.round(d / L)
    dist = np.linalg.norm(d, axis=-1)[~np.eye(n, dtype=bool)]
    counts = np.bincount(np.minimum((dist[dist <= rc] / (rc / bins)).astype(int), bins - 1), minlength=bins)
    g = ibi.normalise(counts.astype(np.uint64), 1, n, L ** 3, rc)
    assert np.allclose(ibi.nodes(rc, bins), (np.arange(bins) + 0.5) * 0.1)
    assert abs(g[5:].mean() - 1.0) < 0.02


def test_initial_potential_continues_linearly_below_the_first_finite_value():
    g = np.array([0.0, 0.0, 0.5, 1.0, 1.2])
    v = ibi.initial_potential(g, 2.0)
    v2, v3 = -2.0 * np.log(0.5), 0.0
    dy = v3 - v2
    np.testing.assert_allclose(v, [v2 - 2 * dy, v2 - dy, v2, v3, -2.0 * np.log(1.2)])


def test_correction_skips_undefined_logs_and_continues():
    g_t = np.array([0.0, 0.4, 1.0, 1.1])
    g_i = np.array([0.0, 0.5, 0.8, 0.0])
    v = np.array([9.0, 5.0, 1.0, 0.5])
    out = ibi.add_correction(v, g_i, g_t, 1.5, 0.2)
    u1, u2 = 0.2 * 1.5 * np.log(0.5 / 0.4), 0.2 * 1.5 * np.log(0.8 / 1.0)
    new1, new2 = 5.0 + u1, 1.0 + u2
    # bin 3: log 0 -> not finite -> unchanged; bin 0 continued from bins 1, 2
    np.testing.assert_allclose(out, [new1 - (new2 - new1), new1, new2, 0.5])


def test_force_is_minus_the_finite_difference_derivative():
    """Central differences inside and one-sided ones at the ends are exact for a quadratic
    inside and for the secant slopes at the ends."""
    x = np.linspace(0.5, 2.5, 41)
    v = 3.0 * x * x - 2.0 * x
    f = ibi.force_of(x, v)
    np.testing.assert_allclose(f[1:-1], -(6.0 * x[1:-1] - 2.0), rtol=1e-12)
    np.testing.assert_allclose(f[0], -(v[1] - v[0]) / (x[1] - x[0]))
    np.testing.assert_allclose(f[-1], -(v[-1] - v[-2]) / (x[-1] - x[-2]))
