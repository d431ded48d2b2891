"""komd_amd -- MI355X-native KoMD force hot path.

The compute path is libkomd_hip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/komd.h).  This package is its Python host layer: a ctypes
binding (``Context``) used by the tests, the benchmark and the surrogate
training driver, plus the seeded phase-space generators (``systems``).

There is no fallback: if the library is missing or no GPU is present, calls
fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkomd_hip.so")

KOMD_OK = 0
STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_STATE", 4: "ERR_OVERFLOW", 5: "ERR_NONFINITE",
          6: "ERR_NOMEM"}
BIN_COM, BIN_SITE = 0, 1
ATM_EXACT, ATM_REF, SURROGATE, HYBRID = 1, 2, 3, 4
PAIR_LJ12_6, PAIR_TABLE, PAIR_ATM2B = 1, 2, 3
SUR_NDESC, SUR_H1, SUR_H2 = 48, 64, 64
SUR_NPARAM = 64 * 48 + 64 + 64 * 64 + 64 + 64 + 1 + 48 + 48


class KomdError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Box(C.Structure):
    _fields_ = [("low", C.c_double * 3), ("high", C.c_double * 3), ("cell_size", C.c_double)]


class Params(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("delta_t", C.c_double), ("limit_factor", C.c_double),
                ("nu", C.c_double), ("temp_factor", C.c_double), ("bin_mode", C.c_int32),
                ("max_neighbors", C.c_int32), ("reserved", C.c_int32 * 4)]


class Observables(C.Structure):
    _fields_ = [("kinetic", C.c_double), ("temperature", C.c_double), ("u_lj", C.c_double), ("u_3b", C.c_double),
                ("u_bond", C.c_double), ("u_angle", C.c_double), ("virial", C.c_double), ("fmax", C.c_double),
                ("n_in_box", C.c_int64), ("pairs", C.c_int64), ("triplets", C.c_int64),
                ("kinetic_tensor", C.c_double * 6), ("virial_tensor", C.c_double * 6)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = np.array(v[:], np.float64) if k.endswith("_tensor") else v
        return out


# Every entry point of include/komd.h with its ctypes signature.
_P = C.c_void_p
_SIGNATURES = {
    "komd_create": (C.c_int, [_P, _P, _P]),
    "komd_destroy": (None, [_P]),
    "komd_last_error": (C.c_char_p, [_P]),
    "komd_device_name": (C.c_int, [C.c_char_p, C.c_int32]),
    "komd_grid_dims": (C.c_int, [_P, _P]),
    "komd_set_sites": (C.c_int, [_P, C.c_int64, _P, _P, _P, _P, _P, _P]),
    "komd_set_state": (C.c_int, [_P, _P, _P, _P]),
    "komd_get_sites": (C.c_int, [_P, _P, _P, _P]),
    "komd_set_topology": (C.c_int, [_P, C.c_int64, _P, C.c_double, C.c_double, C.c_int64, _P, C.c_double,
                                    C.c_double]),
    "komd_update_container": (C.c_int, [_P]),
    "komd_get_cells": (C.c_int, [_P, _P, _P, _P, _P]),
    "komd_forces_pair": (C.c_int, [_P, C.c_int32]),
    "komd_set_pair_table": (C.c_int, [_P, C.c_int64, _P, _P, C.c_double, C.c_double, _P, C.c_double, C.c_double]),
    "komd_select_pair": (C.c_int, [_P, C.c_int32]),
    "komd_forces_bonded": (C.c_int, [_P]),
    "komd_forces_limit": (C.c_int, [_P]),
    "komd_forces_3b": (C.c_int, [_P, C.c_int32]),
    "komd_integrate0": (C.c_int, [_P]),
    "komd_integrate1": (C.c_int, [_P]),
    "komd_measure_temperature": (C.c_int, [_P, _P]),
    "komd_rescale": (C.c_int, [_P, C.c_double]),
    "komd_measure_observables": (C.c_int, [_P, C.c_int32, _P]),
    "komd_rdf": (C.c_int, [_P, C.c_double, C.c_double, C.c_int64, _P]),
    "komd_rdf_noalloc": (C.c_int, [_P, C.c_double, C.c_double, C.c_int64, _P]),
    "komd_run": (C.c_int, [_P, C.c_int64, C.c_int32, C.c_int32, C.c_double]),
    "komd_compute_forces": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "komd_timing_enable": (C.c_int, [_P, C.c_int32]),
    "komd_timing_get": (C.c_int, [_P, C.c_char_p, _P, _P]),
    "komd_counters": (C.c_int, [_P, _P, _P, C.c_int32]),
    "komd_surrogate_set_params": (C.c_int, [_P, _P, C.c_int64]),
    "komd_surrogate_get_params": (C.c_int, [_P, _P, C.c_int64]),
    "komd_surrogate_descriptors": (C.c_int, [_P, _P]),
    "komd_surrogate_energies": (C.c_int, [_P, _P]),
    "komd_atm_site_energies": (C.c_int, [_P, _P]),
    "komd_train_add_snapshot": (C.c_int, [_P]),
    "komd_train_clear": (C.c_int, [_P]),
    "komd_train_size": (C.c_int, [_P, _P]),
    "komd_train_standardize": (C.c_int, [_P]),
    "komd_train_stats": (C.c_int, [_P, _P]),
    "komd_train_gradient": (C.c_int, [_P, C.c_int64, C.c_float, C.c_float, _P, _P]),
    "komd_train_epoch": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.c_uint64, _P]),
    "komd_surrogate_judge": (C.c_int, [_P, _P]),
    "komd_set_judge": (C.c_int, [_P, C.c_int64, C.c_double]),
    "komd_judge_report": (C.c_int, [_P, _P, _P, _P, _P]),
    "komd_set_hybrid": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double, C.c_double]),
    "komd_set_skin": (C.c_int, [_P, C.c_double]),
    "komd_skin_rebuilds": (C.c_int, [_P, _P]),
}

_LIB = None


def load_library(path: str = LIB_PATH):
    """Loads libkomd_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m komd_amd.build` "
                           f"(the HIP library is the only compute path)")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def declared_symbols(header: str | None = None) -> list[str]:
    """Function names declared in include/komd.h (for the ABI export check)."""
    import re
    header = header or os.path.join(os.path.dirname(HERE), "include", "komd.h")
    text = open(header).read()
    return sorted(set(re.findall(r"^(?:komd_status|void|const char\*)\s+(komd_\w+)\s*\(", text, re.M)))


def device_name() -> str:
    buf = C.create_string_buffer(256)
    st = load_library().komd_device_name(buf, 256)
    if st != KOMD_OK:
        raise KomdError(st, "no HIP device")
    return buf.value.decode()


def rdf_bins(max_r: float, delta_r: float) -> int:
    """Bin count of RDFSensor (RDFSensor.cpp:19): the double quotient truncated to uint64."""
    q = float(max_r) / float(delta_r)
    if not (q >= 1.0):
        raise ValueError("rdf: max_r / delta_r must be at least 1")
    return int(q)


def rdf_normalise(counts: np.ndarray, samples: int, n_mol: int, volume: float, delta_r: float):
    """RDFSensor::write (RDFSensor.cpp:37-57): bin centres r and g(r) from accumulated
    ordered-pair counts, g = (n_bin V) / (V_shell n (n - 1)), n_bin = counts / samples."""
    nb = counts.shape[0]
    h = delta_r / 2.0
    r = np.add.accumulate(np.r_[h, np.full(nb - 1, float(delta_r))]) if nb > 1 else np.array([h])
    v_bin = 4.0 / 3.0 * np.pi * ((r + h) ** 3 - (r - h) ** 3)
    n_bin = counts.astype(np.float64) / samples
    g = (n_bin * volume) / (v_bin * (float(n_mol) * (n_mol - 1)))
    return r, g


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One device-resident simulation state (wraps a komd_ctx).

    Mirrors the reference's container + force functor + integrator objects:
    update_container() ~ LinkedCells::updateContainer, forces_pair() ~ LJ12_6,
    forces_3b() ~ ATM_NOLIST/ATM, integrate0/1() ~ Integrator, rescale() ~
    VelocityScaling.  `system` is a komd_amd.systems.System.
    """

    def __init__(self, system, cutoff: float = 2.5, dt: float = 0.005, nu: float = 0.0,
                 limit_factor: float = 0.0, temp_factor: float = 1.0, bin_mode: int = BIN_COM,
                 max_neighbors: int = 0, topology: dict | None = None):
        self.lib = load_library()
        self.sys = system
        box = Box()
        for k in range(3):
            box.low[k] = float(system.low[k])
            box.high[k] = float(system.high[k])
        box.cell_size = float(system.cell_size)
        prm = Params()
        prm.cutoff, prm.delta_t, prm.limit_factor, prm.nu = cutoff, dt, limit_factor, nu
        prm.temp_factor, prm.bin_mode, prm.max_neighbors = temp_factor, bin_mode, max_neighbors
        self._ctx = C.c_void_p()
        st = self.lib.komd_create(C.byref(box), C.byref(prm), C.byref(self._ctx))
        if st != KOMD_OK:
            msg = self.lib.komd_last_error(self._ctx).decode() if self._ctx else "komd_create"
            if self._ctx:
                self.lib.komd_destroy(self._ctx)
            self._ctx = None
            raise KomdError(st, msg)
        self.n = system.n
        self.set_sites(system)
        if topology is not None:
            self.set_topology(**topology)

    # -- plumbing
    def _check(self, st: int):
        if st != KOMD_OK:
            raise KomdError(st, self.lib.komd_last_error(self._ctx).decode())

    def close(self):
        if getattr(self, "_ctx", None):
            self.lib.komd_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state
    def set_sites(self, s):
        r = np.ascontiguousarray(s.r, np.float64)
        v = np.ascontiguousarray(s.v, np.float64)
        m = np.ascontiguousarray(s.mass, np.float64)
        e = np.ascontiguousarray(s.eps, np.float64)
        g = np.ascontiguousarray(s.sig, np.float64)
        i = np.ascontiguousarray(s.mol_id, np.uint64)
        self.n = r.shape[0]
        self._check(self.lib.komd_set_sites(self._ctx, self.n, _p(r), _p(v), _p(m), _p(e), _p(g), _p(i)))

    def set_state(self, r=None, v=None, f=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (r, v, f)]
        self._check(self.lib.komd_set_state(self._ctx, *[None if a is None else _p(a) for a in arrs]))

    def get_state(self):
        r = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        f = np.empty((self.n, 3))
        self._check(self.lib.komd_get_sites(self._ctx, _p(r), _p(v), _p(f)))
        return r, v, f

    def forces(self) -> np.ndarray:
        f = np.empty((self.n, 3))
        self._check(self.lib.komd_get_sites(self._ctx, None, None, _p(f)))
        return f

    def set_topology(self, bonds, angles, kb=500.0, r0=1.0, ka=50.0, theta0=np.radians(109.5)):
        b = np.ascontiguousarray(bonds, np.uint32)
        a = np.ascontiguousarray(angles, np.uint32)
        self._check(self.lib.komd_set_topology(self._ctx, b.shape[0], _p(b), kb, r0, a.shape[0], _p(a), ka,
                                               float(theta0)))

    def grid_dims(self):
        d = np.zeros(3, np.int64)
        self._check(self.lib.komd_grid_dims(self._ctx, _p(d)))
        return d

    # -- container and forces
    def update_container(self):
        self._check(self.lib.komd_update_container(self._ctx))

    def get_cells(self):
        nd = self.grid_dims()
        nc = int(np.prod(nd))
        site_cell = np.empty(self.n, np.int64)
        count = np.empty(nc, np.uint32)
        start = np.empty(nc + 1, np.uint32)
        srt = np.empty(self.n, np.uint32)
        self._check(self.lib.komd_get_cells(self._ctx, _p(site_cell), _p(count), _p(start), _p(srt)))
        return site_cell, count, start, srt

    def forces_pair(self, kind: int = PAIR_LJ12_6):
        self._check(self.lib.komd_forces_pair(self._ctx, kind))

    def set_pair_table(self, x, force, force_low: float = 0.0, force_high: float = 0.0, potential=None,
                       potential_low: float = 0.0, potential_high: float = 0.0):
        """Piecewise-linear pair force F(r) on nodes x (FunctionPL of the IBI functor) with
        its values below the first and above the last node; optionally the potential on the
        same nodes for energies."""
        x = np.ascontiguousarray(x, np.float64)
        f = np.ascontiguousarray(force, np.float64)
        if f.shape != x.shape:
            raise ValueError("force values must match the nodes")
        u = None if potential is None else np.ascontiguousarray(potential, np.float64)
        if u is not None and u.shape != x.shape:
            raise ValueError("potential values must match the nodes")
        self._check(self.lib.komd_set_pair_table(self._ctx, x.size, _p(x), _p(f), float(force_low),
                                                 float(force_high), None if u is None else _p(u),
                                                 float(potential_low), float(potential_high)))

    def select_pair(self, kind: int):
        """The pair force of run / compute_forces / observables: PAIR_LJ12_6 or PAIR_TABLE."""
        self._check(self.lib.komd_select_pair(self._ctx, kind))

    def forces_bonded(self):
        self._check(self.lib.komd_forces_bonded(self._ctx))

    def forces_limit(self):
        self._check(self.lib.komd_forces_limit(self._ctx))

    def forces_3b(self, kind=ATM_EXACT):
        self._check(self.lib.komd_forces_3b(self._ctx, kind))

    def compute_forces(self, kind3b=0, bonded=False):
        self._check(self.lib.komd_compute_forces(self._ctx, kind3b, 1 if bonded else 0))

    # -- integration
    def integrate0(self):
        self._check(self.lib.komd_integrate0(self._ctx))

    def integrate1(self):
        self._check(self.lib.komd_integrate1(self._ctx))

    def temperature(self) -> float:
        t = C.c_double()
        self._check(self.lib.komd_measure_temperature(self._ctx, C.byref(t)))
        return t.value

    def rescale(self, t0: float):
        self._check(self.lib.komd_rescale(self._ctx, t0))

    def run(self, nsteps: int, kind3b: int = 0, bonded: bool = False, t0: float = 0.0):
        self._check(self.lib.komd_run(self._ctx, nsteps, kind3b, 1 if bonded else 0, t0))

    def set_skin(self, skin: float):
        """Verlet lists of width cutoff + skin for NVE komd_run with LJ 12-6 (exact ATM,
        surrogate, hybrid, or no three-body term incl. site-binned bonded chains); 0: off.
        Needs cell_size >= cutoff + skin."""
        self._check(self.lib.komd_set_skin(self._ctx, float(skin)))

    def skin_rebuilds(self) -> int:
        """List rebuilds of skinned runs since set_skin."""
        out = C.c_int64(0)
        self._check(self.lib.komd_skin_rebuilds(self._ctx, C.byref(out)))
        return int(out.value)

    def observables(self, kind3b: int = 0) -> dict:
        o = Observables()
        self._check(self.lib.komd_measure_observables(self._ctx, kind3b, C.byref(o)))
        return o.as_dict()

    def rdf(self, max_r: float = 5.0, delta_r: float = 0.1) -> np.ndarray:
        """One RDFSensor::measure sample (RDFSensor.cpp:23-35): ordered-pair counts per bin
        of the molecules' minimum-image COM distances, (uint64)(max_r / delta_r) bins."""
        nbins = rdf_bins(max_r, delta_r)
        out = np.zeros(nbins, np.uint64)
        self._check(self.lib.komd_rdf(self._ctx, float(max_r), float(delta_r), nbins, _p(out)))
        return out

    def rdf_noalloc(self, max_r: float = 5.0, delta_r: float = 0.1) -> np.ndarray:
        """The same sample as rdf() from the GPU's static scratch (komd_rdf_noalloc: no
        device allocation; at most 262,144 molecules)."""
        nbins = rdf_bins(max_r, delta_r)
        out = np.zeros(nbins, np.uint64)
        self._check(self.lib.komd_rdf_noalloc(self._ctx, float(max_r), float(delta_r), nbins, _p(out)))
        return out

    # -- measurement
    def timing(self, enable: bool):
        self._check(self.lib.komd_timing_enable(self._ctx, 1 if enable else 0))

    def timer(self, name: str):
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.lib.komd_timing_get(self._ctx, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def counters(self, reset: bool = False):
        p = C.c_int64()
        t = C.c_int64()
        self._check(self.lib.komd_counters(self._ctx, C.byref(p), C.byref(t), 1 if reset else 0))
        return p.value, t.value

    # -- surrogate
    def surrogate_set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        self._check(self.lib.komd_surrogate_set_params(self._ctx, _p(p), p.size))

    def surrogate_get_params(self) -> np.ndarray:
        p = np.empty(SUR_NPARAM, np.float32)
        self._check(self.lib.komd_surrogate_get_params(self._ctx, _p(p), p.size))
        return p

    def surrogate_descriptors(self) -> np.ndarray:
        d = np.empty((self.n, SUR_NDESC), np.float32)
        self._check(self.lib.komd_surrogate_descriptors(self._ctx, _p(d)))
        return d

    def surrogate_energies(self) -> np.ndarray:
        e = np.empty(self.n, np.float32)
        self._check(self.lib.komd_surrogate_energies(self._ctx, _p(e)))
        return e

    def atm_site_energies(self) -> np.ndarray:
        e = np.empty(self.n, np.float64)
        self._check(self.lib.komd_atm_site_energies(self._ctx, _p(e)))
        return e

    def train_add_snapshot(self):
        self._check(self.lib.komd_train_add_snapshot(self._ctx))

    def train_clear(self):
        self._check(self.lib.komd_train_clear(self._ctx))

    def train_size(self) -> int:
        n = C.c_int64()
        self._check(self.lib.komd_train_size(self._ctx, C.byref(n)))
        return n.value

    def train_standardize(self):
        self._check(self.lib.komd_train_standardize(self._ctx))

    def train_stats(self):
        """(mean E_i, std E_i, RMS force component) of the stored targets."""
        out = np.zeros(3)
        self._check(self.lib.komd_train_stats(self._ctx, _p(out)))
        return float(out[0]), float(out[1]), float(out[2])

    def train_gradient(self, snapshot: int, w_energy: float, w_force: float):
        g = np.zeros(SUR_NPARAM, np.float32)
        loss = C.c_double(0)
        self._check(self.lib.komd_train_gradient(self._ctx, snapshot, w_energy, w_force, _p(g), C.byref(loss)))
        return loss.value, g

    def train_epoch(self, lr: float, w_energy: float, w_force: float, seed: int) -> float:
        loss = C.c_double(0)
        self._check(self.lib.komd_train_epoch(self._ctx, lr, w_energy, w_force, seed, C.byref(loss)))
        return loss.value

    def surrogate_judge(self):
        out = np.zeros(2)
        self._check(self.lib.komd_surrogate_judge(self._ctx, _p(out)))
        return float(out[0]), float(out[1])

    def set_judge(self, every: int, max_rel_rms: float):
        self._check(self.lib.komd_set_judge(self._ctx, every, max_rel_rms))

    def judge_report(self):
        n, last, worst, fell = C.c_int64(0), C.c_double(0), C.c_double(0), C.c_int32(0)
        self._check(self.lib.komd_judge_report(self._ctx, C.byref(n), C.byref(last), C.byref(worst), C.byref(fell)))
        return {"checks": n.value, "last_rel_rms": last.value, "worst_rel_rms": worst.value,
                "fell_back": bool(fell.value)}

    def set_hybrid(self, axis: int, center: float, half_width: float, blend: float):
        self._check(self.lib.komd_set_hybrid(self._ctx, axis, center, half_width, blend))
