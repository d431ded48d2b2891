#!/usr/bin/env python3
"""Benchmark of the KoMD force hot path on one MI355X.

Default workload (the exact three-body argon configuration): 32,000 argon
atoms, FCC 20^3 x 4 at reduced density 0.8, T* = 0.8, LJ 12-6 (rc 2.5) +
Axilrod-Teller-Muto (nu 0.073) over every triplet with all three distances
below 2.5, velocity Verlet dt = 0.005, NVE.  A step is one pass of the hot
path: kick-drift and wrap, counting-sort binning, the cell-tile pass building
the full neighbour lists, per-site ATM with the LJ force of each list, kick.

Prints one JSON line:
  value        atom-timesteps/s of K graph-replayed steps (inputs resident in HBM)
  roofline     the dominant kernel (per-site ATM + LJ): algorithmic FP32 FLOPs
               (147 per unique in-range triplet and 26 per unique LJ pair, counted
               by the kernels and pinned to the oracle's counts in the tests) over
               its average launch time,
               measured with HIP events on the library's stream over a second
               pass of the same K steps
  cpu_baseline the FP64 OpenMP restatement of the reference algorithm (C08 cell
               pairs, ATM_NOLIST triplet walk, per-interaction atomics), timed on
               this host on a bounded sample of the same workload
  force_error  max / RMS |F_gpu - F_fp64| / RMS(F_fp64) on an equilibrated
               snapshot, total force and three-body force separately (the FP64
               forces come from the same CPU leg as cpu_baseline)
  energy_drift NVE drift of the total energy per atom over --drift-steps steps
  energy_drift_vs_fp64, energy_drift_vs_fp64_32000
               the GPU path's drift beside the FP64 run of the same 4,000-atom
               and 32,000-atom LJ + ATM systems over 10,000 steps (stored fixtures)
  learned      the learned three-body path (descriptor-MLP surrogate, weights in
               komd_amd/weights/) at 32,000 and 256,000 atoms: throughput beside the
               exact path on the same state, its three-body force error against the
               exact ATM (and that of the mean-field ATM2B scaling and of LJ alone), the
               surrogate kernels' times and its NVE energy drift
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-timesteps/s at fixed force error vs FP64: exact ATM vs learned, 1 GPU"
NU = 0.073
FLOP_PAIR = 26      # per unique in-range LJ pair
FLOP_TRIPLET = 147  # per unique in-range ATM triplet
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK = 8.0e12  # B/s
# algorithmic HBM bytes per site of the binning / integration kernels (FP64 state, u32 indices):
#   bin_key     Integrator::Step0 + container update: read r v f 72 + mass 8, write r v 48,
#               f = 0 24, cell key and arrival slot 8
#   bin_scatter read key and slot 8, write the CSR slot 4
#   bin_sort    read key 4 + its CSR entry 4 + r 24, write sorted, inverse map 8 and FP32 position 16
#   kick        (thermostatted steps only) Integrator::Step1: read v f mass 56, write v 24
KERNEL_BYTES_PER_SITE = {"bin_key": 160, "bin_scatter": 12, "bin_sort": 56, "kick": 80}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "atm_32000_traffic.json")
DRIFT_FIXTURES = {n: os.path.join(ROOT, "tests", "golden", f"drift_fp64_fcc{n}_atm.json") for n in (4000, 32000)}
JUDGE_EVERY, JUDGE_MAX = 100, 0.06  # exact-ATM check of the surrogate: period (steps) and RMS bound
WEIGHTS = os.path.join(ROOT, "komd_amd", "weights", "atm_surrogate.npy")
RHO = 0.8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--atoms", type=int, default=32000, choices=[32000, 256000],
                    help="32000: FCC 20^3 x 4 (default); 256000: FCC 40^3 x 4")
    ap.add_argument("--skin", type=float, default=0.3,
                    help="Verlet-list skin of the timed exact-ATM run (cells of edge >= 2.5 + skin; 0: the "
                         "container and lists rebuilt every step)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 force-error check")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--equil-steps", type=int, default=400, help="steps before the force-error snapshot")
    ap.add_argument("--drift-steps", type=int, default=10000, help="NVE steps for the energy drift (0: skip)")
    ap.add_argument("--learned-atoms", default="32000,256000",
                    help="comma list of system sizes for the learned leg ('' to skip)")
    ap.add_argument("--no-hybrid", action="store_true", help="skip the 1,048,576-atom hybrid slab leg")
    ap.add_argument("--no-other", action="store_true", help="skip the LJ-4000 and chains-262144 legs")
    ap.add_argument("--no-ibi", action="store_true", help="skip the IBI table fit of the learned leg")
    ap.add_argument("--hybrid-steps", type=int, default=2000)
    ap.add_argument("--drift-seeds", type=int, default=5,
                    help="seeds of the exact / surrogate drift comparison at 32,000 atoms (0: skip)")
    ap.add_argument("--structure-steps", type=int, default=1000,
                    help="NVE steps of the g(r) / pressure comparison, surrogate against exact (0: skip)")
    ap.add_argument("--only-two-body", action="store_true",
                    help="only run the LJ-4000 and chains-262144 legs and print their JSON (profiling)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (counter calibration) and exit")
    return ap.parse_args()


def rel_err(a, b):
    scale = float(np.sqrt(np.mean(b * b)))
    d = np.linalg.norm(a - b, axis=1)
    return {"max": float(d.max() / scale), "rms": float(np.sqrt(np.mean(d * d)) / scale)}


def hbm_rates(kern_us, n):
    """Achieved GB/s of the binning / integration kernels from their algorithmic bytes."""
    return {k: round(KERNEL_BYTES_PER_SITE[k] * n / (us * 1e-6) / 1e9, 1)
            for k, us in kern_us.items() if k in KERNEL_BYTES_PER_SITE and us > 0}


def drift_of(ctx, K, n, kind, steps):
    """Slope of E_total / N against time over `steps` NVE steps (11 samples)."""
    chunk = max(1, steps // 10)
    ts, es = [], []
    for k in range(11):
        o = ctx.observables(kind)
        ts.append(k * chunk * 0.005)
        es.append((o["kinetic"] + o["u_lj"] + o["u_3b"]) / n)
        if k < 10:
            ctx.run(chunk, kind3b=kind)
    return {"steps": 10 * chunk, "slope_e_per_atom_per_tau": float(np.polyfit(ts, es, 1)[0]),
            "max_dev_per_atom": float(np.max(np.abs(np.array(es) - es[0])))}


def structure_of(ctx, K, kind, n, vol, steps, every, max_r=4.0, dr=0.05):
    """g(r) (RDFSensor counts normalised as RDFSensor::write) and the scalar pressure
    (kinetic + virial tensor trace / 3V) sampled every `every` NVE steps."""
    nb = K.rdf_bins(max_r, dr)
    acc = np.zeros(nb)
    ps = []
    ctx.compute_forces(kind)
    for _ in range(steps // every):
        ctx.run(every, kind3b=kind)
        acc += ctx.rdf(max_r, dr)
        o = ctx.observables(kind)
        t = o["kinetic_tensor"] + o["virial_tensor"]
        ps.append((t[0] + t[2] + t[5]) / (3.0 * vol))
    r, g = K.rdf_normalise(acc, steps // every, n, vol, dr)
    return r, g, np.array(ps)


def learned_leg(n_atoms, args):
    """Exact ATM and the surrogate on the same equilibrated state of an argon system."""
    import komd_amd as K
    from komd_amd import systems as S
    cfg = S.config_atm_32000 if n_atoms == 32000 else S.config_atm_256000
    sysm = cfg(list_radius=2.5 + args.skin)
    n = sysm.n
    out = {"n_atoms": n, "skin": args.skin}
    with K.Context(sysm, cutoff=2.5, dt=0.005, nu=NU) as ctx:
        ctx.surrogate_set_params(np.load(WEIGHTS))
        ctx.compute_forces(K.ATM_EXACT)
        ctx.run(args.equil_steps, kind3b=K.ATM_EXACT)
        r0, v0, _ = ctx.get_state()
        # three-body force error of the surrogate and of the comparators without it against the
        # exact ATM: ATM2B (KOMD_PAIR_ATM2B), an IBI table fitted to the exact path's g(r)
        # (KOMD_PAIR_TABLE, 32,000 atoms) and LJ alone; a comparator's three-body estimate is
        # its pair force minus LJ's
        rel, mx = ctx.surrogate_judge()
        ctx.update_container()
        ctx.forces_3b(K.ATM_EXACT)
        f_atm = ctx.forces()
        ctx.update_container()
        ctx.forces_pair()
        f_lj = ctx.forces()
        ctx.update_container()
        ctx.forces_pair(K.PAIR_ATM2B)
        f_2b = ctx.forces()
        scale = float(np.sqrt(np.mean(np.sum(f_atm ** 2, axis=1))))

        def err3(f_pair):
            return float(np.sqrt(np.mean(np.sum((f_pair - f_lj - f_atm) ** 2, axis=1)))) / scale
        out["three_body_force_error"] = {
            "surrogate": {"rms": rel, "max": mx}, "atm2b_mean_field": {"rms": err3(f_2b)}, "lj_only": {"rms": 1.0},
            "reference": "exact ATM (FP32 arithmetic, FP64 accumulation; 1.6e-6 RMS from the FP64 oracle) "
                         f"after {args.equil_steps} NVE steps from the lattice"}
        ibi_table = None
        if n_atoms == 32000 and not args.no_ibi:
            from komd_amd import ibi
            vol = float(np.prod(sysm.box))
            kT = ctx.observables(K.ATM_EXACT)["temperature"]
            ctx.compute_forces(K.ATM_EXACT)
            g_t = ibi.measure_g(ctx, n, vol, 2.5, 100, 50, 10, K.ATM_EXACT, 0.0)   # NVE samples of the exact path
            ctx.set_state(r0, v0)
            t_ibi = time.perf_counter()
            x_t, f_t, u_t, hist = ibi.run(ctx, g_t, n, vol, 2.5, kT, iterations=8)
            t_ibi = time.perf_counter() - t_ibi
            ctx.set_state(r0, v0)
            ctx.update_container()
            ctx.forces_pair(K.PAIR_TABLE)
            out["three_body_force_error"]["ibi_table"] = {
                "rms": err3(ctx.forces()), "g_rms_error_first_last": [hist[0], hist[-1]], "iterations": len(hist),
                "fit_s": round(t_ibi, 2), "kT": kT,
                "note": "IBI (plugins/IBI/IBI.cpp) on the GPU table force: 100 nodes to rc, target g(r) from "
                        "50 NVE samples of the exact path, 8 iterations of 200 + 500 thermostatted steps"}
            ibi_table = (x_t, f_t, u_t)
            ctx.select_pair(K.PAIR_LJ12_6)
        # the comparators cost one pair pass: their own throughput on the same state
        comp = {}
        for name, pair in (("lj_only", K.PAIR_LJ12_6), ("atm2b", K.PAIR_ATM2B), ("ibi_table", K.PAIR_TABLE)):
            if pair == K.PAIR_TABLE and ibi_table is None:
                continue
            if pair == K.PAIR_TABLE:
                ctx.set_pair_table(ibi_table[0], ibi_table[1], 1e6, 0.0, ibi_table[2], 1e6, 0.0)
            ctx.select_pair(pair)
            ctx.set_state(r0, v0)
            ctx.compute_forces(0)
            ctx.run(args.warmup, kind3b=0)
            t0 = time.perf_counter()
            ctx.run(args.steps, kind3b=0)
            dt = time.perf_counter() - t0
            ctx.timing(True)
            ctx.run(min(args.steps, 200), kind3b=0)
            kern = {}
            for k in ("bin_key", "bin_scan", "bin_scatter", "bin_sort", "cell_lj", "cell_nl", "table", "limit"):
                ms, cnt = ctx.timer(k)
                if cnt:
                    kern[k] = round(ms / cnt * 1e3, 2)
            ctx.timing(False)
            comp[name] = {"value": round(n * args.steps / dt, 1), "ms_per_step": round(dt / args.steps * 1e3, 4),
                          "kernels_us": kern}
        ctx.select_pair(K.PAIR_LJ12_6)
        out["comparator_throughput"] = comp
        legs = {}
        if args.skin > 0:  # both timed legs on Verlet lists rebuilt on demand
            ctx.set_skin(args.skin)
        for kind, name in ((K.ATM_EXACT, "exact"), (K.SURROGATE, "surrogate")):
            ctx.set_state(r0, v0)
            ctx.compute_forces(kind)
            ctx.run(args.warmup, kind3b=kind)
            if kind == K.SURROGATE:  # the exact ATM judges the surrogate every 100 steps (fallback above 6 %)
                ctx.set_judge(JUDGE_EVERY, JUDGE_MAX)
            t0 = time.perf_counter()
            ctx.run(args.steps, kind3b=kind)
            dt = time.perf_counter() - t0
            if kind == K.SURROGATE:
                judge = ctx.judge_report()
                ctx.set_judge(0, JUDGE_MAX)
            ctx.timing(True)
            ctx.run(min(args.steps, 200), kind3b=kind)
            kern = {}
            for k in ("bin_key", "bin_scan", "bin_scatter", "bin_sort", "cell_nl", "cell_lj_nl", "skin_refresh", "atm",
                      "sur_desc", "sur_mlp", "sur_force", "sur_apply"):
                ms, cnt = ctx.timer(k)
                if cnt:
                    kern[k] = round(ms / cnt * 1e3, 2)
            ctx.timing(False)
            legs[name] = {"value": round(n * args.steps / dt, 1), "ms_per_step": round(dt / args.steps * 1e3, 4),
                          "kernels_us": kern}
            if args.skin <= 0:  # with Verlet lists the binning kernels return at once on most steps
                legs[name]["binning_GBps"] = hbm_rates(kern, n)
        out["exact"], out["surrogate"] = legs["exact"], legs["surrogate"]
        out["surrogate"]["judge"] = dict(judge, every=JUDGE_EVERY, max_rel_rms=JUDGE_MAX,
                                         note="exact ATM on the current state every 100 steps inside the "
                                              "timed run; the surrogate's steps fall back to the exact ATM "
                                              "if its three-body RMS error exceeds the bound")
        out["speedup_surrogate_vs_exact"] = round(legs["surrogate"]["value"] / legs["exact"]["value"], 3)
        out["three_body_part_us"] = {
            "exact": legs["exact"]["kernels_us"].get("atm"),
            "surrogate": round(sum(legs["surrogate"]["kernels_us"].get(k, 0.0)
                                   for k in ("sur_desc", "sur_mlp", "sur_force", "sur_apply")), 2),
            "note": "the exact path's k_atm also sums LJ over the lists; the surrogate path sums LJ in its "
                    "descriptor pass (or, on Verlet lists, in skin_refresh): force_part_us compares like with like"}
        # pair pass + three-body kernels: the whole force part of a step but binning
        out["force_part_us"] = {
            name: round(sum(legs[name]["kernels_us"].get(k, 0.0)
                            for k in ("cell_nl", "cell_lj_nl", "skin_refresh", "atm", "sur_desc", "sur_mlp",
                                      "sur_force", "sur_apply")), 2) for name in ("exact", "surrogate")}
        if args.drift_steps > 0 and n_atoms == 32000:
            ctx.set_state(r0, v0)
            ctx.compute_forces(K.SURROGATE)
            out["energy_drift_surrogate"] = drift_of(ctx, K, n, K.SURROGATE, args.drift_steps)
        if n_atoms == 32000 and args.structure_steps > 0:
            # structure and pressure of the surrogate's trajectory against the exact one's,
            # both from the same equilibrated state
            vol = float(np.prod(sysm.box))
            st = {}
            for kind, name in ((K.ATM_EXACT, "exact"), (K.SURROGATE, "surrogate")):
                ctx.set_state(r0, v0)
                st[name] = structure_of(ctx, K, kind, n, vol, args.structure_steps, 10)
            r_, g_e, p_e = st["exact"]
            _, g_s, p_s = st["surrogate"]
            k1 = int(np.argmax(g_e))
            ctx.timing(True)
            ctx.rdf(4.0, 0.05)
            rdf_ms, rdf_n = ctx.timer("rdf")
            ctx.timing(False)
            out["structure"] = {
                "samples": args.structure_steps // 10, "steps": args.structure_steps,
                "rdf": "max_r 4.0, dr 0.05, molecule COMs, ordered pairs (RDFSensor.cpp:59-72)",
                "g_rms_difference": float(np.sqrt(np.mean((g_s - g_e) ** 2))),
                "g_max_difference": float(np.abs(g_s - g_e).max()),
                "first_peak": {"r": float(r_[k1]), "g_exact": float(g_e[k1]), "g_surrogate": float(g_s[k1])},
                "pressure_reduced": {"exact": [float(p_e.mean()), float(p_e.std())],
                                     "surrogate": [float(p_s.mean()), float(p_s.std())]},
                "rdf_sample_us": round(rdf_ms / max(rdf_n, 1) * 1e3, 1)}
    return out


def hybrid_leg(args):
    """Config 5: 1,048,576 argon atoms (simple cubic 256 x 64 x 64, rho 0.8, 4:1:1 box); exact
    ATM in the central 20 % of x, cos^2 blend over 2 sigma, surrogate elsewhere.  After a
    200-step thermostatted melt, hybrid_steps NVE steps in segments of hybrid_steps / 10;
    after each segment: three-body force error of the hybrid against the exact ATM
    everywhere, per region, and the temperature of each region; the total (blended,
    conservative) energy gives the drift.  The same run with the exact ATM everywhere is
    the reference.  Throughput counts the timed segments only."""
    import komd_amd as K
    from komd_amd import systems as S
    sysm = S.config_slab_1048576(list_radius=2.5 + args.skin)
    n = sysm.n
    Lx = float(sysm.box[0])
    center, half, blend = float(sysm.low[0]) + 0.5 * Lx, 0.1 * Lx, 2.0
    out = {"n_atoms": n, "box": [float(x) for x in sysm.box], "exact_slab": [center - half, center + half],
           "blend_width": blend, "steps": args.hybrid_steps}
    nseg = 10
    seg = max(1, args.hybrid_steps // nseg)

    def region_masks(r):
        d = r[:, 0] - center
        d -= Lx * np.rint(d / Lx)
        ad = np.abs(d)
        return {"exact": ad <= half, "blend": (ad > half) & (ad < half + blend), "surrogate": ad >= half + blend}

    with K.Context(sysm, cutoff=2.5, dt=0.005, nu=NU) as ctx:
        ctx.surrogate_set_params(np.load(WEIGHTS))
        ctx.set_hybrid(0, center, half, blend)
        ctx.compute_forces(K.HYBRID)
        ctx.run(200, kind3b=K.HYBRID, t0=0.8)   # melt the lattice
        r0, v0, _ = ctx.get_state()
        if args.skin > 0:  # the NVE segments on Verlet lists rebuilt on demand
            ctx.set_skin(args.skin)
        out["skin"] = args.skin
        legs = {}
        for kind, name in ((K.HYBRID, "hybrid"), (K.ATM_EXACT, "exact_everywhere")):
            ctx.set_state(r0, v0)
            ctx.compute_forces(kind)
            ctx.run(5, kind3b=kind)
            ctx.compute_forces(kind)
            track = {"step": [], "energy_per_atom": [], "temperature": {k: [] for k in ("exact", "blend", "surrogate")},
                     "three_body_rms_error": {k: [] for k in ("exact", "blend", "surrogate")}}
            timed = 0.0
            for k in range(nseg + 1):
                o = ctx.observables(kind)
                track["step"].append(k * seg)
                track["energy_per_atom"].append((o["kinetic"] + o["u_lj"] + o["u_3b"]) / n)
                r, v, f_run = ctx.get_state()
                m = region_masks(r)
                for reg, mk in m.items():
                    track["temperature"][reg].append(float(np.sum(v[mk] ** 2) / (3.0 * mk.sum())))
                if kind == K.HYBRID:
                    # per-region three-body error of the state reached (forces restored after)
                    ctx.update_container()
                    ctx.forces_3b(K.ATM_EXACT)
                    f_atm = ctx.forces()
                    ctx.update_container()
                    ctx.forces_3b(K.HYBRID)
                    f_hyb = ctx.forces()
                    ctx.set_state(f=f_run)
                    scale = float(np.sqrt(np.mean(np.sum(f_atm ** 2, axis=1))))
                    for reg, mk in m.items():
                        track["three_body_rms_error"][reg].append(
                            float(np.sqrt(np.mean(np.sum((f_hyb[mk] - f_atm[mk]) ** 2, axis=1))) / scale))
                if k < nseg:
                    t0 = time.perf_counter()
                    ctx.run(seg, kind3b=kind)
                    timed += time.perf_counter() - t0
            t = np.array(track["step"]) * 0.005
            drift = {"total_slope_e_per_atom_per_tau": float(np.polyfit(t, track["energy_per_atom"], 1)[0]),
                     "temperature_slope_per_tau": {reg: float(np.polyfit(t, track["temperature"][reg], 1)[0])
                                                   for reg in track["temperature"]}}
            ctx.timing(True)
            ctx.run(20, kind3b=kind)
            kern = {}
            for kk in ("cell_lj_nl", "cell_nl", "skin_refresh", "atm", "sur_desc", "sur_mlp", "sur_force", "sur_apply"):
                ms, cnt = ctx.timer(kk)
                if cnt:
                    kern[kk] = round(ms / cnt * 1e3, 2)
            ctx.timing(False)
            leg = {"value": round(n * seg * nseg / timed, 1), "ms_per_step": round(timed / (seg * nseg) * 1e3, 4),
                   "kernels_us": kern, "drift": drift, "along_run": track}
            if kind != K.HYBRID:
                del leg["along_run"]["three_body_rms_error"]
            legs[name] = leg
        out.update(legs)
        m = region_masks(r0)
        out["regions"] = {reg: {"sites_at_start": int(mk.sum()),
                                "three_body_rms_error_mean": float(np.mean(legs["hybrid"]["along_run"]
                                                                           ["three_body_rms_error"][reg])),
                                "three_body_rms_error_max": float(np.max(legs["hybrid"]["along_run"]
                                                                         ["three_body_rms_error"][reg])),
                                "temperature_slope_per_tau": legs["hybrid"]["drift"]["temperature_slope_per_tau"][reg]}
                          for reg, mk in m.items()}
        k = legs["hybrid"]["kernels_us"]
        ns_exact = out["regions"]["exact"]["sites_at_start"] + out["regions"]["blend"]["sites_at_start"]
        ns_sur = out["regions"]["surrogate"]["sites_at_start"] + out["regions"]["blend"]["sites_at_start"]
        # region throughput: the sites a region's three-body kernels serve over those kernels' time
        out["region_three_body_throughput"] = {
            "exact_atm_kernel": {"sites": ns_exact, "us": k.get("atm"),
                                 "site_steps_per_s": round(ns_exact / (k["atm"] * 1e-6), 1) if k.get("atm") else None},
            "surrogate_kernels": {"sites": ns_sur,
                                  "us": round(sum(k.get(x, 0.0) for x in ("sur_desc", "sur_mlp", "sur_force",
                                                                          "sur_apply")), 2)}}
        su = out["region_three_body_throughput"]["surrogate_kernels"]
        su["site_steps_per_s"] = round(ns_sur / (su["us"] * 1e-6), 1) if su["us"] else None
        out["speedup_hybrid_vs_exact"] = round(legs["hybrid"]["value"] / legs["exact_everywhere"]["value"], 3)
    return out


def drift_seeds(args):
    """NVE energy drift of the exact and the surrogate path over several seeds of the
    32,000-atom argon system (each equilibrated for equil_steps with the exact ATM, then
    drift_steps of each path from the same state): mean and spread of the slopes."""
    import komd_amd as K
    from komd_amd import systems as S
    res = {"exact": [], "surrogate": []}
    for seed in range(1003, 1003 + args.drift_seeds):
        sysm = S.fcc(20, 0.8, 0.8, seed)
        with K.Context(sysm, cutoff=2.5, dt=0.005, nu=NU) as ctx:
            ctx.surrogate_set_params(np.load(WEIGHTS))
            ctx.compute_forces(K.ATM_EXACT)
            ctx.run(args.equil_steps, kind3b=K.ATM_EXACT)
            r0, v0, _ = ctx.get_state()
            for kind, name in ((K.ATM_EXACT, "exact"), (K.SURROGATE, "surrogate")):
                ctx.set_state(r0, v0)
                ctx.compute_forces(kind)
                res[name].append(drift_of(ctx, K, sysm.n, kind, args.drift_steps)["slope_e_per_atom_per_tau"])
    out = {"seeds": list(range(1003, 1003 + args.drift_seeds)), "steps": args.drift_steps}
    for name, sl in res.items():
        out[name] = {"slopes": sl, "mean": float(np.mean(sl)), "std": float(np.std(sl, ddof=1)) if len(sl) > 1 else 0.0}
    out["difference_in_spreads"] = ((out["surrogate"]["mean"] - out["exact"]["mean"])
                                    / max(out["exact"]["std"], 1e-300))
    return out


def other_legs(args):
    """The two-body configurations: plain LJ argon (4,000 atoms, 1,000 steps) and the
    multi-site chains (65,536 x 4 sites, harmonic bonds and angles, LJ between molecules)."""
    import komd_amd as K
    from komd_amd import systems as S
    out = {}
    s1 = S.config_lj_4000()
    # FCC 10^3 x 4 at rho 0.8: the 6^3 cells of 2.85 hold cutoff 2.5 and a skin up to 0.35

    def lj_leg(skin):
        with K.Context(s1, cutoff=2.5, dt=0.005) as ctx:
            if skin > 0:
                ctx.set_skin(skin)
            ctx.compute_forces(0)
            o0 = ctx.observables(0)
            r0 = ctx.skin_rebuilds()
            t0 = time.perf_counter()
            ctx.run(1000, kind3b=0)
            dt = time.perf_counter() - t0
            reb = ctx.skin_rebuilds() - r0
            o1 = ctx.observables(0)
        e0 = (o0["kinetic"] + o0["u_lj"]) / s1.n
        e1 = (o1["kinetic"] + o1["u_lj"]) / s1.n
        return {"value": round(s1.n * 1000 / dt, 1), "ms_per_step": round(dt, 4), "energy_per_atom_start": e0,
                "energy_per_atom_end": e1, "rebuilds": int(reb)}

    lj_skin = args.skin > 0 and np.floor(s1.box[0] / s1.cell_size) == np.floor(s1.box[0] / (2.5 + args.skin))
    main = lj_leg(args.skin if lj_skin else 0.0)
    out["lj_4000"] = {"workload": "4,000 LJ argon atoms, FCC 10^3 x 4, rho 0.8, T* 1.0, rc 2.5, dt 0.005, NVE",
                      "steps": 1000, **main}
    if lj_skin:
        # the timed run includes its graph captures (1,000 steps from the lattice, as before)
        out["lj_4000"]["neighbour_lists"] = {"kind": "Verlet lists within cutoff + skin, rebuilt on demand",
                                             "skin": args.skin}
        per = lj_leg(0.0)
        out["lj_4000"]["per_step_rebuild"] = {"value": per["value"], "ms_per_step": per["ms_per_step"]}
    # cells of edge >= cutoff + skin for the Verlet-list run (the per-step run below uses its
    # own grid of cells >= cutoff)
    s2 = S.config_chains_262144(list_radius=2.5 + max(args.skin, 0.0))
    topo = {"bonds": s2.bonds, "angles": s2.angles}
    # capped-force warm-up from the lattice (the clamp would also cap the stiff bonds later)
    # sites binned on their own positions: a chain reaches ~1.2 sigma from its centre, so
    # centre-of-mass cells of edge ~rc would miss pairs between neighbouring chains.
    # Warm-up from the lattice with the force cap raised in stages (a single low cap lets
    # pairs close to ~0.45 sigma, which the uncapped force then blows apart).
    for lim in (20.0, 200.0, 2000.0, 2e4, 2e5):
        with K.Context(s2, cutoff=2.5, dt=0.001, limit_factor=lim, topology=topo, bin_mode=K.BIN_SITE) as ctx:
            ctx.compute_forces(0, bonded=True)
            ctx.run(300, kind3b=0, bonded=True, t0=1.0)
            s2.r, s2.v, _ = ctx.get_state()
    chain_kernels = ("bin_key", "bin_scan", "bin_scatter", "bin_sort", "cell_lj", "cell_nl", "bonded", "skin_lj",
                     "kick")

    def chain_leg(sysc, skin):
        """Rescaled warm-up steps, then the timed NVE steps and 100 steps under kernel timers."""
        with K.Context(sysc, cutoff=2.5, dt=0.005, topology=topo, bin_mode=K.BIN_SITE) as ctx:
            ctx.compute_forces(0, bonded=True)
            ctx.run(args.warmup, kind3b=0, bonded=True, t0=1.0)
            if skin > 0:
                ctx.set_skin(skin)
                ctx.run(args.warmup, kind3b=0, bonded=True)  # the Verlet-list graphs, captured
            r0 = ctx.skin_rebuilds()
            t0 = time.perf_counter()
            ctx.run(args.steps, kind3b=0, bonded=True)
            dt = time.perf_counter() - t0
            reb = ctx.skin_rebuilds() - r0
            ctx.timing(True)
            ctx.run(100, kind3b=0, bonded=True)
            kern = {}
            for k in chain_kernels:
                ms, cnt = ctx.timer(k)
                if cnt:
                    kern[k] = round(ms / cnt * 1e3, 2)
            ctx.timing(False)
            obs = ctx.observables(0)
        return dt, reb, kern, obs

    dt, reb, kern, obs = chain_leg(s2, args.skin)
    out["chains_262144"] = {
        "workload": "65,536 linear 4-site chains (262,144 sites), site density 0.7, harmonic bonds k 500 r0 1, "
                    "angles k 50 theta0 109.5 deg, LJ rc 2.5 between molecules, sites binned on their own "
                    "positions, dt 0.005, NVE after a warm-up with the force cap raised from 20 to 2e5 "
                    "in five 300-step stages at dt 0.001 with rescaling to T* 1.0",
        "steps": args.steps, "value": round(s2.n * args.steps / dt, 1),
        "ms_per_step": round(dt / args.steps * 1e3, 4), "kernels_us": kern,
        "u_bond_per_site": obs["u_bond"] / s2.n, "u_angle_per_site": obs["u_angle"] / s2.n,
        "temperature": obs["temperature"],
        "equipartition": "3 bonds and 2 angles per 4 sites: kT/2 each gives 0.375 T and 0.25 T per site"}
    if args.skin > 0:
        out["chains_262144"]["neighbour_lists"] = {
            "kind": "Verlet lists of sites of other molecules within cutoff + skin, rebuilt on demand",
            "skin": args.skin, "grid_cells_per_dim": int(np.floor(s2.box[0] / s2.cell_size)),
            "rebuilds_in_timed_steps": int(reb)}
        # the same state on a grid of cells >= cutoff, container update every step
        s3 = s2.copy()
        s3.cell_size = S.cell_size_for(float(s2.box[0]), 2.5)
        dtp, _, kernp, _ = chain_leg(s3, 0.0)
        out["chains_262144"]["per_step_rebuild"] = {
            "value": round(s3.n * args.steps / dtp, 1), "ms_per_step": round(dtp / args.steps * 1e3, 4),
            "grid_cells_per_dim": int(np.floor(s3.box[0] / s3.cell_size)), "kernels_us": kernp}
    return out


def cpu_reference_leg(sysm, snapshot, args):
    """The CPU side, the only place the bench runs the FP64 oracle (checker and
    baseline, never the measured path):
      cpu_baseline  the oracle's reference-order loop (C08 cell pairs, ATM_NOLIST triplet
                    walk, FP64 atomics) timed on a bounded sample of the workload;
      force_error   FP64 forces of the GPU snapshot's positions against the GPU's."""
    cpu = force_error = None
    if args.no_cpu and snapshot is None:
        return cpu, force_error
    import oracle
    threads = os.cpu_count() or 1
    n = sysm.n
    if not args.no_cpu:
        oracle.run(sysm, 1, dt=0.005, nu=NU, atm=1, atm_sign=-1.0, nthreads=threads)  # warm (first binning)
        c0 = time.perf_counter()
        oracle.run(sysm, args.cpu_steps, dt=0.005, nu=NU, atm=1, atm_sign=-1.0, nthreads=threads)
        c1 = time.perf_counter()
        cpu = {"value": round(n * args.cpu_steps / (c1 - c0), 1), "unit": "atom-timesteps/s", "cores": threads,
               "kind": "port",
               "sample": f"{args.cpu_steps} steps of the {n}-atom LJ+ATM workload after 1 warm step: C08 cell "
                         f"pairs (LinkedCells.cpp:164-224) + ATM_NOLIST triplet walk (ATM_NOLIST.cpp:20-159) with "
                         f"FP64 atomics, {threads} OpenMP threads, {c1 - c0:.2f} s wall"}
    if snapshot is not None:
        r_snap, v_snap, f_tot, f_atm = snapshot
        snap = sysm.copy()
        snap.r, snap.v = r_snap, v_snap
        r, b = oracle.update_container(snap)
        f_lj = oracle.lj_c08(snap, r, b, nthreads=threads)
        f_3b, _, _ = oracle.atm_complete(snap, r, b, NU, coverage=0, sign=-1.0, nthreads=threads)
        force_error = {"total": rel_err(f_tot, f_lj + f_3b), "three_body": rel_err(f_atm, f_3b),
                       "snapshot": f"after {args.equil_steps} NVE steps from the lattice",
                       "reference": "FP64 oracle (C restatement) on the same positions"}
    return cpu, force_error


def drift_against_fp64(K, n_atoms):
    """NVE energy trajectory of the GPU path against the FP64 run of the same system
    (n_atoms FCC argon, rho 0.8, T* 0.8, seed 1003, LJ + complete ATM, 10,000 steps), whose
    E/N every 100 steps is stored in tests/golden/drift_fp64_fcc<n>_atm.json (written by
    tests/golden/make_drift_fp64.py on the CPU: 4,000 atoms and the 32,000-atom workload)."""
    from komd_amd import systems as S
    path = DRIFT_FIXTURES[n_atoms]
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        ref = json.load(fh)
    s = S.fcc(int(round((n_atoms / 4) ** (1 / 3))), 0.8, 0.8, 1003)
    every, steps = ref["every"], ref["steps"]
    e = []
    with K.Context(s, cutoff=2.5, dt=0.005, nu=NU) as ctx:
        ctx.compute_forces(K.ATM_EXACT)
        for k, st in enumerate(steps):
            if k:
                ctx.run(st - steps[k - 1], kind3b=K.ATM_EXACT)
            o = ctx.observables(K.ATM_EXACT)
            e.append((o["kinetic"] + o["u_lj"] + o["u_3b"]) / s.n)
    t = np.array(steps) * 0.005
    e = np.array(e)
    eref = np.array(ref["e_per_atom"])
    return {"system": ref["system"], "steps": steps[-1],
            "gpu": {"slope_e_per_atom_per_tau": float(np.polyfit(t, e, 1)[0]),
                    "max_dev_per_atom": float(np.max(np.abs(e - e[0])))},
            "fp64": {"slope_e_per_atom_per_tau": ref["slope_e_per_atom_per_tau"],
                     "max_dev_per_atom": ref["max_dev_per_atom"]},
            "e_per_atom_start": {"gpu": float(e[0]), "fp64": float(eref[0])},
            "max_abs_diff_first_1000_steps": float(np.max(np.abs(e[:11] - eref[:11])))}


def copy_calibration(mib: int):
    """A plain device copy of known size: rocprofv3 FETCH_SIZE / WRITE_SIZE read on
    it give the factor that turns those counters into HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "launches": 5}), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the force path is single-GPU here (spatial decomposition "
              "over several GPUs is designed but not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_mib:
        copy_calibration(args.copy_mib)
        return
    if args.only_two_body:
        print(json.dumps(other_legs(args)), flush=True)
        return

    import komd_amd as K
    from komd_amd import systems as S

    cfg = S.config_atm_32000 if args.atoms == 32000 else S.config_atm_256000
    sysm = cfg(list_radius=2.5 + args.skin)
    n = sysm.n
    dev = K.device_name()
    ctx = K.Context(sysm, cutoff=2.5, dt=0.005, nu=NU)
    ctx.compute_forces(K.ATM_EXACT)

    # ---- GPU forces on an equilibrated snapshot; the FP64 reference forces of the same
    # positions are computed in the CPU reference leg at the end
    snapshot = None
    if not args.no_check:
        ctx.run(args.equil_steps, kind3b=K.ATM_EXACT)
        r_snap, v_snap, _ = ctx.get_state()
        ctx.compute_forces(K.ATM_EXACT)
        f_tot = ctx.forces()
        ctx.update_container()
        ctx.forces_3b(K.ATM_EXACT)
        snapshot = (r_snap, v_snap, f_tot, ctx.forces())
        ctx.compute_forces(K.ATM_EXACT)

    # ---- throughput: K graph-replayed steps, inputs resident in HBM.  With a skin the run
    # keeps Verlet lists of radius 2.5 + skin and the device rebuilds the container and the
    # lists before any atom has moved skin/2: every step has the interaction set of a
    # rebuild at that step (the GPU tests check the counts and the forces on stale lists)
    if args.skin > 0:
        ctx.set_skin(args.skin)
    ctx.run(args.warmup, kind3b=K.ATM_EXACT)
    ctx.counters(reset=True)
    reb0 = ctx.skin_rebuilds()
    t0 = time.perf_counter()
    ctx.run(args.steps, kind3b=K.ATM_EXACT)   # synchronises at the end
    t1 = time.perf_counter()
    rebuilds = ctx.skin_rebuilds() - reb0
    ctx.counters(reset=True)
    dt = t1 - t0
    value = n * args.steps / dt

    def kernel_times(c, natoms):
        """Per-kernel device time with HIP events over a second pass of K steps."""
        c.timing(True)
        c.run(args.steps, kind3b=K.ATM_EXACT)
        cnt_pt = c.counters(reset=True)
        out = {}
        for name in ("bin_key", "bin_scan", "bin_scatter", "bin_sort", "cell_nl", "cell_lj_nl", "atm", "kick"):
            ms, cnt = c.timer(name)
            if not cnt:  # fused: the cell-count scan into the scatter (grids up to 4096 cells), the
                continue  # NVE second half-kick into the three-body kernel
            out[name] = {"us_per_launch": round(ms / max(cnt, 1) * 1e3, 2), "launches": cnt}
        raw = {name: c.timer(name) for name in ("atm", "cell_nl", "cell_lj_nl")}
        c.timing(False)
        return out, cnt_pt, raw

    kernels, (pairs, trips), raw = kernel_times(ctx, n)
    if args.skin > 0:
        for name in ("bin_key", "bin_scan", "bin_scatter", "bin_sort", "cell_nl"):
            if name in kernels:
                kernels[name]["note"] = ("average over every step: the Verlet-list step's kick-drift and position "
                                         "refresh (bin_key) or, on steps that keep their lists, an immediate return"
                                         if name == "bin_key" else
                                         "average over every step, most of them an immediate return (lists kept)")

    # ---- the same workload with the container and lists rebuilt every step (the reference's
    # pair_steps 1), on the reference-sized grid (cells of edge >= 2.5)
    per_step = None
    if args.skin > 0:
        s0 = cfg()
        with K.Context(s0, cutoff=2.5, dt=0.005, nu=NU) as c0:
            c0.compute_forces(K.ATM_EXACT)
            c0.run(args.warmup, kind3b=K.ATM_EXACT)
            ta = time.perf_counter()
            c0.run(args.steps, kind3b=K.ATM_EXACT)
            tb = time.perf_counter()
            k0, _, _ = kernel_times(c0, n)
        per_step = {"value": round(n * args.steps / (tb - ta), 1), "ms_per_step": round((tb - ta) / args.steps * 1e3, 4),
                    "grid_cells_per_dim": float(s0.box[0] / s0.cell_size), "kernels": k0}
        hbm_src = k0
    else:
        hbm_src = kernels
    for name, kk in hbm_src.items():
        if name in KERNEL_BYTES_PER_SITE:
            us = kk["us_per_launch"]
            b = KERNEL_BYTES_PER_SITE[name] * n
            kk.update({"algorithmic_bytes": b, "achieved_GBps": round(b / (us * 1e-6) / 1e9, 1),
                       "frac_of_hbm_peak": round(b / (us * 1e-6) / HBM_PEAK, 4)})
    atm_ms, atm_n = raw["atm"]
    # uniform argon without the clamp: the pair pass only builds the lists ("cell_nl") and
    # k_atm also sums LJ over them
    lj_in_atm = raw["cell_nl"][1] > 0
    cell_ms, cell_n = raw["cell_nl" if lj_in_atm else "cell_lj_nl"]
    flop_atm = FLOP_TRIPLET * trips + (FLOP_PAIR * pairs if lj_in_atm else 0)
    achieved = flop_atm / (atm_ms * 1e-3) / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and args.atoms == 32000:
        with open(TRAFFIC_FILE) as fh:
            tj = json.load(fh)
        traffic = tj.get("hbm_bytes_per_launch_skin" if args.skin > 0 else "hbm_bytes_per_launch")
    roofline = {
        "kernel": "k_atm (per-site Axilrod-Teller-Muto" + (" + LJ 12-6 over the lists" if lj_in_atm else "")
                  + ", FP32 VALU)", "bound": "mfma",
        "bound_note": "compute-bound (FLOP/s roofline, the 'mfma' class of the contract): the kernel issues FP32 "
                      "vector instructions, not matrix ones, and is limited by VALU issue (SQ_INSTS_VALU x 4 "
                      "cycles over the 1,024 SIMDs equals its time); 157.3 TFLOP/s is MI355X's FP32 figure",
        "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
        "avg_launch_us": round(atm_ms / max(atm_n, 1) * 1e3, 2),
        "algorithmic_flop_per_launch": int(round(flop_atm / max(atm_n, 1))),
        "flop_count": f"{FLOP_TRIPLET} per unique in-range triplet"
                      + (f" + {FLOP_PAIR} per unique in-range LJ pair" if lj_in_atm else ""),
        "unique_triplets_per_step": int(trips // max(atm_n, 1)),
        "peak_note": "MI355X FP32 vector peak 157.3 TFLOP/s counts packed FMA; the kernels are built "
                     "without packed FP32, whose own ceiling is 78.6 TFLOP/s",
        "timing": "HIP events on the library stream, second pass of the same steps",
        "second_kernel": {"kernel": ("k_cell (Verlet lists within 2.5 + skin: 27-cell distance tests, list "
                                     "compaction; an immediate return on steps that keep their lists)")
                                    if args.skin > 0 else
                                    ("k_cell (neighbour lists: 27-cell distance tests, list compaction)"
                                     if lj_in_atm else "k_cell (LJ 12-6 + neighbour lists)"),
                          "achieved_tflops": None if lj_in_atm
                          else round(FLOP_PAIR * pairs / (cell_ms * 1e-3) / 1e12, 3),
                          "avg_launch_us": round(cell_ms / max(cell_n, 1) * 1e3, 2),
                          "unique_pairs_per_step": int(pairs // max(cell_n, 1))},
    }

    # ---- NVE energy drift over drift_steps (E per atom, linear fit against time)
    drift = None
    if args.drift_steps > 0:
        drift = drift_of(ctx, K, n, K.ATM_EXACT, args.drift_steps)
        drift["note"] = ("LJ and ATM truncated at rc without shift, as the reference's energy "
                         "(LJ12_6_Sensor.cpp:76): pairs and triplets crossing rc contribute energy steps")

    # ---- learned three-body path beside the exact one
    learned = {}
    if args.learned_atoms and os.path.exists(WEIGHTS):
        for na in [int(x) for x in args.learned_atoms.split(",") if x]:
            learned[str(na)] = learned_leg(na, args)

    hybrid = None
    if not args.no_hybrid and os.path.exists(WEIGHTS):
        hybrid = hybrid_leg(args)
    other = None if args.no_other else other_legs(args)
    seeds = drift_seeds(args) if args.drift_seeds > 0 and args.drift_steps > 0 and os.path.exists(WEIGHTS) else None

    cpu, force_error = cpu_reference_leg(sysm, snapshot, args)
    drift_fp64 = None if args.drift_steps <= 0 else drift_against_fp64(K, 4000)
    drift_fp64_32k = None if args.drift_steps <= 0 else drift_against_fp64(K, 32000)

    cells = "20^3" if n == 32000 else "40^3"
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "atom-timesteps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "dtype_note": "f32 pair and triplet arithmetic; f64 positions, velocities, force accumulation and reductions",
        "data": "synthetic: seeded FCC argon lattice, Maxwell-Boltzmann velocities",
        "config": {"workload": f"exact 3-body argon, {n:,} atoms (FCC {cells} x 4), rho 0.8, T* 0.8, LJ 12-6 "
                               f"rc 2.5 + ATM nu 0.073 over every triplet within rc, dt 0.005, NVE, 1 GPU",
                   "n_atoms": n, "device": dev, "path": "exact ATM (physical sign and complete triplet set)",
                   "neighbour_lists": ({"kind": "Verlet lists rebuilt on demand (the reference's pair-list reuse, "
                                                "PairList.h:22-24, with a skin so no pair is missed)",
                                        "skin": args.skin, "list_radius": 2.5 + args.skin,
                                        "grid_cells_per_dim": float(sysm.box[0] / sysm.cell_size),
                                        "rebuilds_in_timed_steps": rebuilds}
                                       if args.skin > 0 else
                                       {"kind": "container update and lists every step (pair_steps 1)"})},
        "per_step_rebuild": per_step,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "force_error": force_error,
        "energy_drift_vs_fp64": drift_fp64,
        "energy_drift_vs_fp64_32000": drift_fp64_32k,
        "energy_drift": drift,
        "kernels": kernels,
        "learned": learned or None,
        "hybrid_slab": hybrid,
        "energy_drift_seeds": seeds,
        "two_body_configs": other,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
