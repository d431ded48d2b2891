"""FP64 CPU oracle for KoMD's force hot path -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this package; the product (komd_amd) never does.  ``liboracle.so`` is a
plain-C restatement of the reference algorithm (see komd_oracle.c for the
file:line each function follows); this module wraps it with ctypes and adds
convenience drivers that run whole reference-order passes on a
``komd_amd.systems.System``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        i64, f64, i32 = C.c_int64, C.c_double, C.c_int
        sig = {
            "or_temperature_of": (f64, [f64, f64]),
            "or_nu_internal": (f64, [f64]),
            "or_conv_J_Ei": (f64, []),
            "or_grid_dims": (None, [P, P, f64, P]),
            "or_cell_halo": (None, [P, P, f64, P, P]),
            "or_update_com": (None, [i64, P, P, P, P]),
            "or_periodic": (None, [i64, P, P, P, P]),
            "or_bin": (i64, [i64, P, P, P, P, f64, P, P, P, P, P]),
            "or_lj_c08": (None, [i64, P, P, P, P, f64, P, P, P, P, P, P, P, P, i32]),
            "or_lj_brute": (None, [i64, P, P, P, P, f64, P, P, P]),
            "or_atm_comp_force": (f64, [f64, f64, f64, f64]),
            "or_atm_energy_triplet": (f64, [f64, P, P, P]),
            "or_atm_c08": (i64, [i64, P, f64, f64, P, P, P, P, P, f64, P, i32]),
            "or_atm_brute": (i64, [i64, P, f64, f64, P, P, i32, f64, P, P]),
            "or_atm_complete": (i64, [i64, P, f64, f64, P, P, P, P, P, P, i32, f64, P, P, P, i32]),
            "or_limit": (None, [i64, f64, P, P, P]),
            "or_vv_step0": (None, [i64, f64, P, P, P, P]),
            "or_vv_step1": (None, [i64, f64, P, P, P]),
            "or_kinetic_sum": (None, [i64, P, P, P, P, P, P, P]),
            "or_rescale": (None, [i64, f64, P]),
            "or_bonded": (f64, [i64, P, f64, f64, i64, P, f64, f64, P, P, P]),
            "or_run": (i32, [i64, P, P, P, P, P, P, P, P, P, f64, f64, f64, f64, i32, f64, f64, f64, f64, i64, i32]),
            "or_num_threads": (i32, []),
            "or_virial": (None, [i64, P, P, P, P, f64, P, f64, P, i32, i64, P, f64, f64, i64, P, f64, f64, P]),
            "or_rdf": (None, [i64, P, P, f64, f64, i64, P, i32]),
            "or_table_eval": (f64, [f64, i64, P, P, f64, f64]),
            "or_table_c08": (None, [i64, P, f64, P, P, P, P, P, i64, P, P, f64, f64, P, f64, f64, P, P, P, i32]),
            "or_table_brute": (None, [i64, P, f64, P, i64, P, P, f64, f64, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Binning:
    """Result of the reference's container update on one snapshot."""

    def __init__(self, nd, site_cell, cell_count, cell_start, cell_list):
        self.nd = nd
        self.site_cell = site_cell
        self.cell_count = cell_count
        self.cell_start = cell_start
        self.cell_list = cell_list

    @property
    def na(self):
        return self.nd + 2


def grid_dims(low, high, cell_size) -> np.ndarray:
    nd = np.zeros(3, np.int64)
    lib().or_grid_dims(_p(_c(low, np.float64)), _p(_c(high, np.float64)), float(cell_size), _p(nd))
    return nd


def update_container(sys, r=None):
    """COM -> periodic wrap -> COM -> writeIndices (LinkedCells.cpp:65-81).
    Returns (wrapped r, Binning)."""
    L = lib()
    r = _c(sys.r if r is None else r, np.float64).copy()
    n = r.shape[0]
    ms = sys.mol_start()
    nmol = len(ms) - 1
    m = _c(sys.mass, np.float64)
    low, high = _c(sys.low, np.float64), _c(sys.high, np.float64)
    com = np.zeros_like(r)
    L.or_update_com(nmol, _p(ms), _p(r), _p(m), _p(com))
    L.or_periodic(n, _p(com), _p(r), _p(low), _p(high))
    L.or_update_com(nmol, _p(ms), _p(r), _p(m), _p(com))
    nd = grid_dims(low, high, sys.cell_size)
    na = nd + 2
    ncell = int(np.prod(na))
    site_cell = np.zeros(n, np.int64)
    cnt = np.zeros(ncell, np.int64)
    start = np.zeros(ncell + 1, np.int64)
    tot = L.or_bin(nmol, _p(ms), _p(com), _p(low), _p(high), float(sys.cell_size), _p(nd), _p(site_cell), _p(cnt),
                   _p(start), None)
    lst = np.zeros(tot, np.int64)
    L.or_bin(nmol, _p(ms), _p(com), _p(low), _p(high), float(sys.cell_size), _p(nd), _p(site_cell), _p(cnt),
             _p(start), _p(lst))
    return r, Binning(nd, site_cell, cnt, start, lst)


def lj_c08(sys, r, b: Binning, cutoff=2.5, nthreads=1, energy=False):
    """LJ forces (and the sensor's total potential) over the reference pair set."""
    f = np.zeros_like(r)
    u = np.zeros(1)
    w = np.zeros(1)
    lib().or_lj_c08(r.shape[0], _p(r), _p(_c(sys.mol_id, np.uint64)), _p(_c(sys.sig, np.float64)),
                    _p(_c(sys.eps, np.float64)), float(cutoff), _p(_c(sys.low, np.float64)),
                    _p(_c(sys.high, np.float64)), _p(b.nd), _p(b.cell_start), _p(b.cell_list), _p(f),
                    _p(u) if energy else None, _p(w) if energy else None, int(nthreads))
    return (f, float(u[0]), float(w[0])) if energy else f


def lj_brute(sys, r, cutoff=2.5):
    f = np.zeros_like(r)
    u = np.zeros(1)
    lib().or_lj_brute(r.shape[0], _p(r), _p(_c(sys.mol_id, np.uint64)), _p(_c(sys.sig, np.float64)),
                      _p(_c(sys.eps, np.float64)), float(cutoff), _p(_c(sys.box, np.float64)), _p(f), _p(u))
    return f, float(u[0])


def atm_c08(sys, r, b: Binning, nu, cutoff=2.5, sign=1.0, nthreads=1):
    f = np.zeros_like(r)
    hits = lib().or_atm_c08(r.shape[0], _p(r), float(nu), float(cutoff), _p(_c(sys.low, np.float64)),
                            _p(_c(sys.high, np.float64)), _p(b.nd), _p(b.cell_start), _p(b.cell_list), float(sign),
                            _p(f), int(nthreads))
    return f, int(hits)


def atm_brute(sys, r, nu, cutoff=2.5, site_cell=None, coverage=0, sign=-1.0):
    f = np.zeros_like(r)
    e = np.zeros(1)
    sc = _c(site_cell if site_cell is not None else np.zeros(r.shape[0]), np.int64)
    hits = lib().or_atm_brute(r.shape[0], _p(r), float(nu), float(cutoff), _p(_c(sys.box, np.float64)), _p(sc),
                              int(coverage), float(sign), _p(f), _p(e))
    return f, float(e[0]), int(hits)


def atm_complete(sys, r, b: Binning, nu, cutoff=2.5, coverage=0, sign=-1.0, nthreads=8, site_energy=False):
    f = np.zeros_like(r)
    e = np.zeros(1)
    es = np.zeros(r.shape[0]) if site_energy else None
    hits = lib().or_atm_complete(r.shape[0], _p(r), float(nu), float(cutoff), _p(_c(sys.low, np.float64)),
                                 _p(_c(sys.high, np.float64)), _p(b.nd), _p(b.site_cell), _p(b.cell_start),
                                 _p(b.cell_list), int(coverage), float(sign), _p(f), _p(e),
                                 _p(es) if site_energy else None, int(nthreads))
    if site_energy:
        return f, float(e[0]), int(hits), es
    return f, float(e[0]), int(hits)


def comp_force(nu, rab, rac, rbc):
    return lib().or_atm_comp_force(float(nu), float(rab), float(rac), float(rbc))


def atm_energy_triplet(nu, ri, rj, rk):
    return lib().or_atm_energy_triplet(float(nu), _p(_c(ri, np.float64)), _p(_c(rj, np.float64)),
                                       _p(_c(rk, np.float64)))


def bonded(sys, r, kb=500.0, r0=1.0, ka=50.0, theta0=np.radians(109.5)):
    f = np.zeros_like(r)
    bonds = _c(sys.bonds, np.uint32)
    angles = _c(sys.angles, np.uint32)
    u = lib().or_bonded(bonds.shape[0], _p(bonds), float(kb), float(r0), angles.shape[0], _p(angles), float(ka),
                        float(theta0), _p(_c(r, np.float64)), _p(_c(sys.box, np.float64)), _p(f))
    return f, float(u)


def virial(sys, r, cutoff=2.5, nu=0.0, site_cell=None, coverage=0, kb=500.0, r0=1.0, ka=50.0,
           theta0=np.radians(109.5), bonded=True):
    """Virial tensor (XX XY YY XZ YZ ZZ) of LJ + physical ATM (nu != 0) + harmonic
    bonds/angles (when the system has them and bonded=True), minimum image."""
    r = _c(r, np.float64)
    n = r.shape[0]
    bonds = _c(sys.bonds if bonded else np.zeros((0, 2)), np.uint32)
    angles = _c(sys.angles if bonded else np.zeros((0, 3)), np.uint32)
    sc = _c(site_cell if site_cell is not None else np.zeros(n), np.int64)
    w = np.zeros(6)
    lib().or_virial(n, _p(r), _p(_c(sys.mol_id, np.uint64)), _p(_c(sys.sig, np.float64)),
                    _p(_c(sys.eps, np.float64)), float(cutoff), _p(_c(sys.box, np.float64)), float(nu), _p(sc),
                    int(coverage), bonds.shape[0], _p(bonds), float(kb), float(r0), angles.shape[0], _p(angles),
                    float(ka), float(theta0), _p(w))
    return w


def molecule_coms(sys, r):
    """Centre of mass of every molecule (MoleculeContainer.cpp:55-91), shape (M, 3)."""
    r = _c(r, np.float64)
    ms = sys.mol_start()
    com = np.zeros_like(r)
    lib().or_update_com(len(ms) - 1, _p(ms), _p(r), _p(_c(sys.mass, np.float64)), _p(com))
    return com[ms[:-1]].copy()


def rdf(com, box, max_r, delta_r, nbins, nthreads=8):
    """RDFSensor::measure counts over ordered molecule pairs (RDFSensor.cpp:59-72)."""
    com = _c(com, np.float64)
    out = np.zeros(int(nbins), np.uint64)
    lib().or_rdf(com.shape[0], _p(com), _p(_c(box, np.float64)), float(max_r), float(delta_r), int(nbins), _p(out),
                 int(nthreads))
    return out


class Table:
    """A piecewise-linear function as the reference's FunctionPL holds it (IBI_Math.h:13-41):
    nodes x (ascending), values y, and the values returned below the first and above the
    last node."""

    def __init__(self, x, y, low=0.0, high=0.0):
        self.x = _c(x, np.float64)
        self.y = _c(y, np.float64)
        self.low = float(low)
        self.high = float(high)

    def __call__(self, r):
        L = lib()
        r = np.atleast_1d(np.asarray(r, np.float64))
        out = np.array([L.or_table_eval(float(v), self.x.size, _p(self.x), _p(self.y), self.low, self.high)
                        for v in r.ravel()])
        return out.reshape(r.shape)


def table_c08(sys, r, b: Binning, force: Table, cutoff=2.5, potential: Table | None = None, nthreads=1):
    """Tabulated pair forces (IBI_ForceFunctor.cpp:24-42) over the reference pair set; with a
    potential table on the same nodes also the total energy and the virial tensor
    (XX XY YY XZ YZ ZZ), each physical pair once.  Returns (f, u, w)."""
    r = _c(r, np.float64)
    f = np.zeros_like(r)
    u = np.zeros(1)
    w = np.zeros(6)
    pu = _c(potential.y, np.float64) if potential is not None else None
    lib().or_table_c08(r.shape[0], _p(r), float(cutoff), _p(_c(sys.low, np.float64)), _p(_c(sys.high, np.float64)),
                       _p(b.nd), _p(b.cell_start), _p(b.cell_list), force.x.size, _p(force.x), _p(force.y), force.low,
                       force.high, _p(pu) if pu is not None else None,
                       potential.low if potential is not None else 0.0,
                       potential.high if potential is not None else 0.0, _p(f), _p(u), _p(w), int(nthreads))
    return f, float(u[0]), w


def table_brute(sys, r, force: Table, cutoff=2.5):
    r = _c(r, np.float64)
    f = np.zeros_like(r)
    lib().or_table_brute(r.shape[0], _p(r), float(cutoff), _p(_c(sys.box, np.float64)), force.x.size, _p(force.x),
                         _p(force.y), force.low, force.high, _p(f))
    return f


def limit(sys, f, limit_factor):
    f = _c(f, np.float64).copy()
    lib().or_limit(f.shape[0], float(limit_factor), _p(_c(sys.eps, np.float64)), _p(_c(sys.sig, np.float64)), _p(f))
    return f


def vv_step0(dt, r, v, f, m):
    r, v = _c(r, np.float64).copy(), _c(v, np.float64).copy()
    lib().or_vv_step0(r.shape[0], float(dt), _p(r), _p(v), _p(_c(f, np.float64)), _p(_c(m, np.float64)))
    return r, v


def vv_step1(dt, v, f, m):
    v = _c(v, np.float64).copy()
    lib().or_vv_step1(v.shape[0], float(dt), _p(v), _p(_c(f, np.float64)), _p(_c(m, np.float64)))
    return v


def kinetic_sum(sys, r, v):
    a, b = np.zeros(1), np.zeros(1)
    lib().or_kinetic_sum(r.shape[0], _p(_c(r, np.float64)), _p(_c(v, np.float64)), _p(_c(sys.mass, np.float64)),
                         _p(_c(sys.low, np.float64)), _p(_c(sys.high, np.float64)), _p(a), _p(b))
    return float(a[0]), float(b[0])


def run(sys, nsteps, cutoff=2.5, dt=0.005, nu=0.0, atm=0, atm_sign=1.0, limit_factor=0.0, T0=0.0,
        temp_factor=1.0, nthreads=1, f=None, r=None):
    """Reference-order MD loop (Simulation.cpp:49-74) on single-site molecules."""
    r = _c(sys.r if r is None else r, np.float64).copy()
    v = _c(sys.v, np.float64).copy()
    f = np.zeros_like(r) if f is None else _c(f, np.float64).copy()
    lib().or_run(r.shape[0], _p(r), _p(v), _p(f), _p(_c(sys.mass, np.float64)), _p(_c(sys.mol_id, np.uint64)),
                 _p(_c(sys.sig, np.float64)), _p(_c(sys.eps, np.float64)), _p(_c(sys.low, np.float64)),
                 _p(_c(sys.high, np.float64)), float(sys.cell_size), float(cutoff), float(dt), float(nu), int(atm),
                 float(atm_sign), float(limit_factor), float(T0), float(temp_factor), int(nsteps), int(nthreads))
    return r, v, f
