"""Seeded synthetic phase-space generators (host side, numpy).

The reference builds its start states with ``PhasespaceGenerator::generate``
(util/PhasespaceGenerator.cpp:27-82: simple-cubic lattice, libstdc++
``default_random_engine``, no momentum removal) or reads ``.ps`` checkpoints
(IO/CheckpointIO.cpp:38-76).  This module replaces the generator with
counter-based, platform-independent ones (SplitMix64 + Box-Muller) so that the
same seed gives the same state on every machine.  Arrays follow the
reference's site model (molecule/Site.h, container/SOA.h): interleaved
positions r[N,3], velocities v[N,3], per-site mass/epsilon/sigma and the
molecule id of every site.

Reduced units throughout (sigma = epsilon = m = 1 unless given otherwise).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive SplitMix64 outputs (Steele, Lea, Flood 2014) as uint64."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = (np.uint64(seed & MASK64) + k * np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """Doubles in (0, 1) from the top 53 bits of SplitMix64."""
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller on SplitMix64 uniforms."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1, u2 = u[:m], u[m:]
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = rad * np.cos(2 * np.pi * u2)
    out[1::2] = rad * np.sin(2 * np.pi * u2)
    return out[:n]


@dataclass
class System:
    """A phase-space state plus the box and cell grid it lives in."""

    r: np.ndarray                  # [N,3] float64
    v: np.ndarray                  # [N,3] float64
    mass: np.ndarray               # [N]
    eps: np.ndarray                # [N]
    sig: np.ndarray                # [N]
    mol_id: np.ndarray             # [N] uint64, consecutive equal ids = one molecule
    low: np.ndarray                # [3]
    high: np.ndarray               # [3]
    cell_size: float
    bonds: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.uint32))
    angles: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.uint32))
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.r.shape[0])

    @property
    def box(self) -> np.ndarray:
        return self.high - self.low

    def mol_start(self) -> np.ndarray:
        """CSR start offsets of molecules (runs of equal consecutive ids, CheckpointIO.cpp:63-72)."""
        ids = self.mol_id
        brk = np.nonzero(ids[1:] != ids[:-1])[0] + 1
        return np.concatenate([[0], brk, [len(ids)]]).astype(np.int64)

    def copy(self) -> "System":
        return System(self.r.copy(), self.v.copy(), self.mass.copy(), self.eps.copy(), self.sig.copy(),
                      self.mol_id.copy(), self.low.copy(), self.high.copy(), self.cell_size,
                      self.bonds.copy(), self.angles.copy(), self.name)


def cell_size_for(length: float, rc: float) -> float:
    """Cell edge L / floor(L / rc), nudged up by ulps until floor(L / cs)
    reproduces that count, so the reference's grid (LinkedCells.cpp:19) has
    no remainder cell and every cell is >= rc."""
    n = int(math.floor(length / rc))
    cs = length / n
    while math.floor(length / cs) < n:
        cs = math.nextafter(cs, 0.0)
    while cs < rc:
        n -= 1
        cs = length / n
    return cs


def maxwell_boltzmann(n: int, temperature: float, mass: np.ndarray, seed: int) -> np.ndarray:
    """MB velocities, zero net momentum, rescaled to exactly `temperature`
    with 3N degrees of freedom (TemperatureSensor.cpp:41 counts 3 per site)."""
    g = normal(seed, 3 * n).reshape(n, 3)
    v = g * np.sqrt(temperature / mass)[:, None]
    p = (v * mass[:, None]).sum(0) / mass.sum()
    v -= p[None, :]
    t_now = float((mass[:, None] * v * v).sum() / (3.0 * n))
    if t_now > 0:
        v *= math.sqrt(temperature / t_now)
    return v


def fcc(cells: int, density: float, temperature: float, seed: int, rc: float = 2.5,
        jitter: float = 0.0) -> System:
    """FCC argon: cells^3 unit cells of 4 atoms at number density `density`.
    Lattice shifted by a/4 so no site starts on a cell face.  `jitter` adds a
    seeded uniform displacement of that amplitude (in units of a)."""
    a = (4.0 / density) ** (1.0 / 3.0)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], dtype=np.float64)
    g = np.arange(cells, dtype=np.float64)
    gz, gy, gx = np.meshgrid(g, g, g, indexing="ij")
    origins = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    r = ((origins[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25) * a
    n = r.shape[0]
    if jitter > 0:
        r = r + (uniform01(seed ^ 0x5DEECE66D, 3 * n).reshape(n, 3) - 0.5) * (2 * jitter * a)
    length = cells * a
    r = np.mod(r, length)
    mass = np.ones(n)
    v = maxwell_boltzmann(n, temperature, mass, seed)
    return System(r=r, v=v, mass=mass, eps=np.ones(n), sig=np.ones(n), mol_id=np.arange(n, dtype=np.uint64),
                  low=np.zeros(3), high=np.full(3, length), cell_size=cell_size_for(length, rc),
                  name=f"fcc{cells}^3x4 rho={density} T*={temperature}")


def simple_cubic(nx: int, ny: int, nz: int, density: float, temperature: float, seed: int,
                 rc: float = 2.5) -> System:
    """Simple-cubic lattice nx*ny*nz at `density` (spacing (1/density)^(1/3)),
    box low = 0; cubic cells are chosen per dimension."""
    a = (1.0 / density) ** (1.0 / 3.0)
    gz, gy, gx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = (np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1).astype(np.float64) + 0.25) * a
    n = r.shape[0]
    mass = np.ones(n)
    v = maxwell_boltzmann(n, temperature, mass, seed)
    high = np.array([nx * a, ny * a, nz * a])
    cs = min(cell_size_for(h, rc) for h in high)
    cs = max(cs, rc)
    return System(r=r, v=v, mass=mass, eps=np.ones(n), sig=np.ones(n), mol_id=np.arange(n, dtype=np.uint64),
                  low=np.zeros(3), high=high, cell_size=cs, name=f"sc{nx}x{ny}x{nz} rho={density}")


def _random_rotations(n: int, seed: int) -> np.ndarray:
    """Uniform random rotation matrices from normalised Gaussian quaternions."""
    q = normal(seed, 4 * n).reshape(n, 4)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], axis=1)


def chains(nmol_xyz=(64, 32, 32), sites: int = 4, site_density: float = 0.7, bond: float = 1.0,
           theta0_deg: float = 109.5, temperature: float = 1.0, seed: int = 1002, rc: float = 2.5) -> System:
    """Linear `sites`-site chains (bond length `bond`, planar zig-zag with bend
    angle theta0) with centres on an nx*ny*nz lattice in a cubic box of the
    given site density, each chain randomly oriented.  Bonds (i, i+1) and
    angles (i, i+1, i+2) inside every chain."""
    nx, ny, nz = nmol_xyz
    nmol = nx * ny * nz
    n = nmol * sites
    length = (n / site_density) ** (1.0 / 3.0)
    th = math.radians(theta0_deg)
    # zig-zag template in the xy plane
    tmpl = np.zeros((sites, 3))
    half = (math.pi - th) / 2.0
    for s in range(1, sites):
        sgn = 1.0 if s % 2 else -1.0
        tmpl[s] = tmpl[s - 1] + bond * np.array([math.cos(half), sgn * math.sin(half), 0.0])
    tmpl -= tmpl.mean(0)
    rot = _random_rotations(nmol, seed)
    gz, gy, gx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    centers = (np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1) + 0.5) * np.array(
        [length / nx, length / ny, length / nz])
    # sites stay around their chain centre (the container wraps whole molecules by COM)
    r = (centers[:, None, :] + np.einsum("mij,sj->msi", rot, tmpl)).reshape(-1, 3)
    mass = np.ones(n)
    v = maxwell_boltzmann(n, temperature, mass, seed + 7)
    mol = np.repeat(np.arange(nmol, dtype=np.uint64), sites)
    base = (np.arange(nmol) * sites).astype(np.uint32)
    bonds = np.concatenate([np.stack([base + s, base + s + 1], 1) for s in range(sites - 1)]).astype(np.uint32)
    angles = np.concatenate([np.stack([base + s, base + s + 1, base + s + 2], 1)
                             for s in range(sites - 2)]).astype(np.uint32)
    return System(r=r, v=v, mass=mass, eps=np.ones(n), sig=np.ones(n), mol_id=mol, low=np.zeros(3),
                  high=np.full(3, length), cell_size=cell_size_for(length, rc), bonds=bonds, angles=angles,
                  name=f"chains{nmol}x{sites} rho_site={site_density}")


# The workloads the benchmark and the parity tests use (reduced units).
def config_lj_4000(seed: int = 1001) -> System:
    """4,000 LJ argon atoms, FCC 10^3 x 4, rho 0.8, T* 1.0."""
    return fcc(10, 0.8, 1.0, seed)


def config_chains_262144(seed: int = 1002, list_radius: float = 2.5) -> System:
    """65,536 linear 4-site chains (262,144 sites), site density 0.7; cells of edge at least
    list_radius (cutoff 2.5 plus the Verlet-list skin, if any)."""
    return chains((64, 32, 32), 4, 0.7, seed=seed, rc=list_radius)


def config_atm_32000(seed: int = 1003, list_radius: float = 2.5) -> System:
    """32,000 argon atoms, FCC 20^3 x 4, rho 0.8, T* 0.8 (LJ + ATM); cells of edge at
    least list_radius (cutoff 2.5 plus the Verlet-list skin, if any)."""
    return fcc(20, 0.8, 0.8, seed, rc=list_radius)


def config_atm_256000(seed: int = 1004, list_radius: float = 2.5) -> System:
    """256,000 argon atoms, FCC 40^3 x 4, rho 0.8, T* 0.8."""
    return fcc(40, 0.8, 0.8, seed, rc=list_radius)


def config_slab_1048576(seed: int = 1005, list_radius: float = 2.5) -> System:
    """1,048,576 argon atoms, simple cubic 256 x 64 x 64 at rho 0.8 in a 4:1:1 box."""
    return simple_cubic(256, 64, 64, 0.8, 0.8, seed, rc=list_radius)
