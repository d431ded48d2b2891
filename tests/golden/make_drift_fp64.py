"""FP64 energy trajectory of the exact three-body argon system from the oracle (CPU).

System: FCC argon (default 4,000 atoms, FCC 10^3 x 4; --cells 20 gives the 32,000-atom
workload), rho 0.8, T* 0.8, seed 1003, LJ 12-6 rc 2.5 +
Axilrod-Teller-Muto nu 0.073 over every triplet within rc (physical sign, -grad V),
velocity Verlet dt 0.005, NVE, 10,000 steps, all in FP64 in the reference's operation
order (oracle: vv_step0 -> update_container -> lj_c08 + atm_complete -> vv_step1).
Writes E_total / N every 100 steps and the slope of a linear fit against time.

  python tests/golden/make_drift_fp64.py [--cells 10] [--steps 10000] [--threads 8] [--out FILE]
                                         [--state FILE]

--state keeps the trajectory state (r, v, f and the energies so far) every 1,000 steps
and resumes from it when the file exists, so an interrupted hour-long run continues.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from komd_amd import systems as S  # noqa: E402

NU, DT, EVERY = 0.073, 0.005, 100


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--cells", type=int, default=10)
    ap.add_argument("--out", default=None)
    ap.add_argument("--state", default=None)
    a = ap.parse_args()
    O.build()
    s = S.fcc(a.cells, 0.8, 0.8, 1003)
    if a.out is None:
        a.out = os.path.join(HERE, f"drift_fp64_fcc{s.n}_atm.json")
    n = s.n

    def forces(r):
        rb, b = O.update_container(s, r)
        f, u, _ = O.lj_c08(s, rb, b, energy=True, nthreads=a.threads)
        fa, ea, _ = O.atm_complete(s, rb, b, NU, coverage=0, sign=-1.0, nthreads=a.threads)
        return rb, f + fa, u + ea

    r, v = s.r.copy(), s.v.copy()
    r, f, U = forces(r)
    steps, e = [0], [(0.5 * float(np.sum(s.mass[:, None] * v * v)) + U) / n]
    first = 1
    if a.state and os.path.exists(a.state):
        z = np.load(a.state)
        r, v, f = z["r"], z["v"], z["f"]
        steps, e = [int(x) for x in z["steps"]], [float(x) for x in z["e"]]
        first = steps[-1] + 1
        print(f"resumed at step {steps[-1]}", flush=True)
    t0 = time.time()
    for k in range(first, a.steps + 1):
        r, v = O.vv_step0(DT, r, v, f, s.mass)
        r, f, U = forces(r)
        v = O.vv_step1(DT, v, f, s.mass)
        if k % EVERY == 0:
            steps.append(k)
            e.append((0.5 * float(np.sum(s.mass[:, None] * v * v)) + U) / n)
            if k % 1000 == 0:
                print(f"step {k}  E/N {e[-1]:.10f}  {time.time() - t0:.0f}s", flush=True)
                if a.state:
                    np.savez(a.state + ".tmp.npz", r=r, v=v, f=f, steps=np.array(steps), e=np.array(e))
                    os.replace(a.state + ".tmp.npz", a.state)
    t = np.array(steps) * DT
    slope = float(np.polyfit(t, np.array(e), 1)[0])
    out = {"system": f"{n:,} argon atoms, FCC {a.cells}^3 x 4, rho 0.8, T* 0.8 (seed 1003), LJ rc 2.5 + ATM nu 0.073 "
                     "(every triplet within rc, -grad V), dt 0.005, NVE",
           "precision": "FP64 throughout (oracle)", "every": EVERY, "steps": steps, "e_per_atom": e,
           "slope_e_per_atom_per_tau": slope, "max_dev_per_atom": float(np.max(np.abs(np.array(e) - e[0])))}
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"slope {slope:.4e}  written {a.out}")


if __name__ == "__main__":
    main()
