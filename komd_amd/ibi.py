"""Iterative Boltzmann inversion (IBI) on the GPU: a tabulated pair force whose radial
distribution function matches a target, the reference's IBI plugin (plugins/IBI/IBI.cpp,
RDF_Profiler.cpp) restated as a driver of this library's tabulated pair force
(KOMD_PAIR_TABLE) and RDF sample (komd_rdf).

  nodes    r_k = (k + 1/2) dr, dr = rc / bins                (RDF_Profiler.cpp:8-16)
  g(r)     ordered-pair counts per sample, divided by the samples and by
           N (N - 1) V_shell / V                                (RDF_Profiler.cpp:36-54)
  V_0      -kT ln g_target, continued linearly below the first finite value
                                                                (IBI.cpp:110-133)
  update   V += alpha kT ln(g_i / g_target) where finite, the same continuation
                                                                (IBI.cpp:135-165)
  force    F = -dV/dr by FunctionPL::derivative (IBI_Math.cpp:22-35), 1e6 below the first
           node and 0 above the last                           (IBI.cpp:166-171)

kT: the reference multiplies ln g by the deck temperature in K (IBI.cpp:19, 113, 138);
here kT is the thermal energy in the caller's unit (T* in reduced units).  Each iteration
equilibrates and then samples g(r) with the rescale thermostat at the target temperature,
as the reference's always-on thermostat does (main.cpp:93).
"""
from __future__ import annotations

import numpy as np

from . import PAIR_LJ12_6, PAIR_TABLE

LOW_FORCE, HIGH_FORCE = 1e6, 0.0   # IBI.cpp:167-168


def nodes(rc: float, bins: int) -> np.ndarray:
    dr = rc / bins
    return (np.arange(bins) + 0.5) * dr


def normalise(counts: np.ndarray, samples: int, n_mol: int, volume: float, rc: float) -> np.ndarray:
    """RDF_Profiler::getRDF: counts per sample over N (N - 1) V_shell / V."""
    bins = counts.shape[0]
    dr = rc / bins
    rmin = np.arange(bins) * dr
    rmax = (np.arange(bins) + 1) * dr
    shell = 4.0 / 3.0 * np.pi * (rmax ** 3 - rmin ** 3)
    den = n_mol * (n_mol - 1.0) * shell / volume
    return counts.astype(np.float64) / samples / den


def measure_g(ctx, n_mol: int, volume: float, rc: float, bins: int, samples: int, every: int, kind3b: int,
              t0: float) -> np.ndarray:
    """g(r) over `samples` RDF samples taken every `every` steps of komd_run."""
    acc = np.zeros(bins, np.uint64)
    for _ in range(samples):
        ctx.run(every, kind3b=kind3b, t0=t0)
        acc += ctx.rdf(rc, rc / bins)
    return normalise(acc, samples, n_mol, volume, rc)


def _continue_below(v: np.ndarray, first: int) -> np.ndarray:
    """Linear continuation of v below index `first` with the slope of its first two values."""
    if first < 0 or first + 1 >= v.size:
        return v
    dy = v[first + 1] - v[first]
    y = v[first]
    for k in range(first, -1, -1):
        v[k] = y
        y -= dy
    return v


def initial_potential(g_target: np.ndarray, kT: float) -> np.ndarray:
    """IBI::initializePotentialValues (IBI.cpp:110-133)."""
    with np.errstate(divide="ignore"):
        v = -kT * np.log(g_target)
    finite = np.nonzero(np.isfinite(v))[0]
    return _continue_below(v, int(finite[0]) if finite.size else -1)


def add_correction(v: np.ndarray, g_now: np.ndarray, g_target: np.ndarray, kT: float, alpha: float) -> np.ndarray:
    """IBI::addPotentialCorrection (IBI.cpp:135-165)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        upd = alpha * kT * np.log(g_now / g_target)
    ok = np.isfinite(upd)
    v = v.copy()
    v[ok] += upd[ok]
    nz = np.nonzero(g_target != 0.0)[0]
    return _continue_below(v, int(nz[0]) if nz.size else -1)


def force_of(x: np.ndarray, v: np.ndarray) -> np.ndarray:
    """-dV/dr: FunctionPL::derivative (IBI_Math.cpp:22-35) times -1 (IBI.cpp:166-171)."""
    d = np.empty_like(v)
    d[0] = (v[1] - v[0]) / (x[1] - x[0])
    d[-1] = (v[-1] - v[-2]) / (x[-1] - x[-2])
    d[1:-1] = (v[2:] - v[:-2]) / (x[2:] - x[:-2])
    return -d


def run(ctx, g_target: np.ndarray, n_mol: int, volume: float, rc: float, kT: float, iterations: int = 8,
        equil: int = 200, samples: int = 50, every: int = 10, alpha: float = 0.2, log=None):
    """IBI iterations on a context whose state is the starting configuration.  The context
    is left with the final table selected.  Returns (x, F, V, per-iteration RMS g error)."""
    bins = g_target.shape[0]
    x = nodes(rc, bins)
    v = initial_potential(g_target, kT)
    history = []
    for it in range(iterations):
        ctx.set_pair_table(x, force_of(x, v), LOW_FORCE, HIGH_FORCE, v, 1e6, 0.0)
        ctx.select_pair(PAIR_TABLE)
        ctx.compute_forces(0)
        ctx.run(equil, kind3b=0, t0=kT)
        g = measure_g(ctx, n_mol, volume, rc, bins, samples, every, 0, kT)
        err = float(np.sqrt(np.mean((g - g_target) ** 2)))
        history.append(err)
        if log:
            log(f"IBI iteration {it}: RMS g error {err:.4f}")
        v = add_correction(v, g, g_target, kT, alpha)
    f = force_of(x, v)
    ctx.set_pair_table(x, f, LOW_FORCE, HIGH_FORCE, v, 1e6, 0.0)
    ctx.select_pair(PAIR_TABLE)
    return x, f, v, history


def restore_lj(ctx):
    ctx.select_pair(PAIR_LJ12_6)
