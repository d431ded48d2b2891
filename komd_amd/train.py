"""Training driver of the learned three-body surrogate (runs on the GPU through the C ABI).

Data: snapshots of an argon LJ + exact-ATM run (physical forces, every triplet within
the cutoff) with the rescale thermostat at temperatures drawn from a range; the stored
targets are the exact per-site ATM energy shares V/3 and the exact ATM forces, both
produced by the library's own ATM kernel.  Training: one Adam step per snapshot with
the library's BF16 MFMA forward/backward kernels, loss
  L = wE/N sum (E_i - E*_i)^2 + wF/(3N) sum |F_i - F*_i|^2,
wE = lam_E / std(E*)^2, wF = 1 / rms(F*)^2, cosine learning-rate decay.

  python -m komd_amd.train --atoms 32000 --snapshots 2000 --out weights.npy
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import komd_amd as K  # noqa: E402
from komd_amd import systems as S  # noqa: E402

NDESC, H1, H2 = 48, 64, 64
NPARAM = H1 * NDESC + H1 + H2 * H1 + H2 + H2 + 1 + NDESC + NDESC
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "atm_surrogate.npy")


def init_params(shift, scale, e_mean, e_std, seed=0):
    """Glorot-uniform hidden layers, zero biases, output layer at the target spread."""
    rng = np.random.default_rng(seed)
    l1 = math.sqrt(6.0 / (NDESC + H1))
    l2 = math.sqrt(6.0 / (H1 + H2))
    parts = [rng.uniform(-l1, l1, H1 * NDESC), np.zeros(H1), rng.uniform(-l2, l2, H2 * H1), np.zeros(H2),
             rng.uniform(-1, 1, H2) * e_std / math.sqrt(H2), [e_mean], shift, scale]
    p = np.concatenate([np.asarray(x, np.float64).ravel() for x in parts]).astype(np.float32)
    assert p.size == NPARAM
    return p


def collect(ctx, nsnap, every, t_range, seed, equil, log=print):
    rng = np.random.default_rng(seed)
    ctx.compute_forces(K.ATM_EXACT)
    if equil:
        ctx.run(equil, kind3b=K.ATM_EXACT, t0=t_range[1])
    t0 = time.time()
    for k in range(nsnap):
        ctx.run(every, kind3b=K.ATM_EXACT, t0=float(rng.uniform(*t_range)))
        ctx.train_add_snapshot()
        if (k + 1) % max(1, nsnap // 10) == 0:
            log(f"snapshots {k + 1}/{nsnap}  {time.time() - t0:.1f}s")


def fit(ctx, epochs, lr0, lr1, lam_e=0.1, seed=0, params=None, log=print):
    ctx.train_standardize()
    e_mean, e_std, f_rms = ctx.train_stats()
    cur = ctx.surrogate_get_params()
    if params is None:
        params = init_params(cur[NPARAM - 2 * NDESC:NPARAM - NDESC], cur[NPARAM - NDESC:], e_mean, e_std, seed)
    ctx.surrogate_set_params(params)
    wE, wF = lam_e / e_std ** 2, 1.0 / f_rms ** 2
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        lr = lr1 + 0.5 * (lr0 - lr1) * (1 + math.cos(math.pi * ep / max(1, epochs - 1)))
        loss = ctx.train_epoch(lr, wE, wF, seed=seed * 100003 + ep)
        hist.append(loss)
        log(f"epoch {ep:3d} lr {lr:.2e} loss {loss:.4e}  {time.time() - t0:.1f}s")
    return ctx.surrogate_get_params(), {"e_mean": e_mean, "e_std": e_std, "f_rms": f_rms, "wE": wE, "wF": wF,
                                        "loss": hist}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--atoms", type=int, default=32000, choices=[4000, 32000, 256000])
    ap.add_argument("--snapshots", type=int, default=2000)
    ap.add_argument("--every", type=int, default=20, help="MD steps between snapshots")
    ap.add_argument("--equil", type=int, default=1000, help="melting steps at the top temperature")
    ap.add_argument("--tmin", type=float, default=0.7)
    ap.add_argument("--tmax", type=float, default=1.0)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr0", type=float, default=2e-3)
    ap.add_argument("--lr1", type=float, default=5e-5)
    ap.add_argument("--lam-e", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=1004)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--log", default=None)
    a = ap.parse_args(argv)
    cells = {4000: 10, 32000: 20, 256000: 40}[a.atoms]
    sysm = S.fcc(cells, 0.8, a.tmax, a.seed)
    logf = open(a.log, "w") if a.log else None

    def log(msg):
        print(msg, flush=True)
        if logf:
            logf.write(msg + "\n")
            logf.flush()

    log(f"device {K.device_name()}; {sysm.n} atoms, {a.snapshots} snapshots every {a.every} steps, "
        f"T* in [{a.tmin}, {a.tmax}]")
    with K.Context(sysm, cutoff=2.5, dt=0.005, nu=0.073) as ctx:
        collect(ctx, a.snapshots, a.every, (a.tmin, a.tmax), a.seed, a.equil, log)
        params, info = fit(ctx, a.epochs, a.lr0, a.lr1, a.lam_e, a.seed, log=log)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.save(a.out, params)
        info.update({"atoms": sysm.n, "snapshots": a.snapshots, "every": a.every, "t_range": [a.tmin, a.tmax],
                     "epochs": a.epochs, "lam_e": a.lam_e, "seed": a.seed})
        with open(os.path.splitext(a.out)[0] + ".json", "w") as fh:
            json.dump(info, fh, indent=1)
        log(f"saved {a.out}")


if __name__ == "__main__":
    main()
