"""NP-MD throughput of the per-step path on one MI355X, FP64 plain mode and FP32 mode.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128] [--precision both|fp32|fp64]

Workload: BaTiO3 bulk 128^3 supercell (2,097,152 cells), the bulk
phase-transition parameter set, Nose-Poincare MD ('md') at 300 K, dt = 2 fs,
acoustic displacement optimised each step.  Initial u is uniform in
[-0.05, 0.05] A per component from a fixed seed, p = 0.  Q_Nose grows with N
(1.0 at 64^3 as the reference's 64^3 md benchmark, scaled by N).  State and
tables are resident in HBM before the timed region; K steps run back to back
on the library's stream, bracketed by device synchronisation.

Two modes are timed on the same workload: FP64 plain (everything FP64, the
reference's arithmetic) and FP32 (spectra, FFTs, the k-space stage and the
stored state u, p in FP32; local forces, drift and kicks computed in FP64
registers; thermostat, energy sums and homogeneous strain in FP64).  ``value`` is the FP64 plain-mode rate; the FP32 rate and the ratio
are in ``modes``.  Each mode's force and energy error against the FP64 oracle
(the C restatement of the reference) after the prologue and after the first
two steps is reported per run in ``modes.<mode>.vs_oracle``: the oracle runs
the same prologue and two steps from the same doubles in the CPU-baseline leg.

``roofline`` is the dominant kernel's algorithmic HBM bytes per launch over
its average launch time, taken with HIP events on the library's stream in a
separate pass of max(K, 100) steps (direct launches) between the warm-up and
the timed steps; ``roofline.copy_GBs`` is a 1 GiB device copy timed in the same run (the
attainable HBM rate on this box).  ``cpu_baseline`` times the CPU FP64 oracle (OpenMP over
the host threads) on a bounded sample of the same workload.  ``ensemble`` is
the batched 48-replica sweep ({12,16,24,32}^3, T and E_z drawn per replica,
graph-replayed steps) in each mode: replica-steps/s and the equilibrated
replicas per hour at 5000 steps per replica.  ``small_16cubed`` is the 16^3
NP-MD case (FP64, 1000 steps in one call) beside the reference's published
rate for that size.
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

HBM_PEAK_GBS = 8000.0
ENERGY_TERMS = ("short_range", "long_range", "dipole_E_field", "unharmonic", "homo_strain",
                "homo_coupling", "inho_strain", "inho_coupling", "inho_modulation")


def kernel_bytes(name: str, N: int, Kh: int, nf: int, ws: int, materialize_frac: float) -> float:
    """Algorithmic HBM bytes of one launch (compulsory reads + writes).

    A packed half spectrum holds N/2 complex points (kx = 0 and kx = Lx/2
    share column 0); ws: bytes of one word (8 FP64, 4 FP32) of the spectra
    and of the state fields, which follow the mode; a complex point is 2 ws;
    the L(k) table has Kh = (Lx/2+1) Ly Lz points of 6 words."""
    cw = 2 * ws
    M = N // 2
    if name == "xfwd_drift_r2c":      # read u, pk, code; write u, nf spectra
        return 3 * ws * N * 3 + N + nf * cw * M
    if name in ("y_forward", "y_inverse"):
        return 2 * nf * cw * M
    if name == "z_dipole_dipole":     # 3 spectra in/out + 6-word L(k) table
        return 2 * 3 * cw * M + 6 * ws * Kh
    if name == "z_strain":            # 6 spectra in/out
        return 2 * (nf - 3) * cw * M
    if name == "z_pair":              # both z groups in one launch
        return 2 * 3 * cw * M + 6 * ws * Kh + 2 * (nf - 3) * cw * M
    if name == "xinv_c2r_local_kick":  # read nf spectra, u, pk, code; write pk (+p, F when materialised)
        return nf * cw * M + 3 * ws * N * 3 + N + materialize_frac * 2 * 3 * ws * N
    return 0.0


def step_bytes(N: int, Kh: int, ws: int) -> float:
    """Compulsory bytes of a 3-sweep fused md step: 18 N w_s + 78 Kh w_f, the
    state word w_s and the spectral word w_f both ws (FP64 or FP32 mode)."""
    return 18 * N * ws + 78 * Kh * ws


def cpu_baseline(p, u0, steps: int):
    """The oracle (FP64, OpenMP over the host threads) on the same workload:
    prologue untimed, `steps` NP steps timed.  Returns the baseline record and
    the oracle's state after the prologue and after two steps."""
    import oracle
    oracle.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.Oracle(p)
    o.field("dipoR")[...] = u0
    e = o.md_prologue()
    ref = {"e0": e.as_dict(), "F0": o.field("dVddi").copy()}
    t0 = time.perf_counter()
    for i in range(steps):
        o.nose_poincare(e)
        if i == 1:
            ref.update(e2=e.as_dict(), u2=o.field("dipoR").copy(), p2=o.field("dipoP").copy(),
                       F2=o.field("dVddi").copy())
    dt = time.perf_counter() - t0
    o.close()
    rec = {"value": steps / dt, "unit": "steps/s", "cores": threads, "kind": "port",
           "sample": f"{steps} NP steps of the same {p.Lx}^3 workload with the C FP64 oracle "
                     f"(OpenMP, {threads} threads), after its untimed table set-up and prologue"}
    return rec, (ref if steps >= 2 else None)


def errors_vs(first, ref) -> dict:
    """Force and energy errors of a device run against a reference run
    (normwise max|dF| / max|F|; energies against the sum of |terms|)."""
    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    def erel(a, b):
        scale = sum(abs(b[t]) for t in ENERGY_TERMS)
        return float(max(abs(a[t] - b[t]) for t in ENERGY_TERMS) / scale)

    return {"force_rel_err_prologue": rel(first["F0"], ref["F0"]),
            "energy_rel_err_prologue": erel(first["e0"], ref["e0"]),
            "force_rel_err_step2": rel(first["F2"], ref["F2"]),
            "energy_rel_err_step2": erel(first["e2"], ref["e2"]),
            "u_rel_err_step2": rel(first["u2"], ref["u2"]),
            "p_rel_err_step2": rel(first["p2"], ref["p2"])}


def copy_rate(nbytes: int = 1 << 30, reps: int = 10):
    """GB/s (read + write bytes over the time per copy) of the library's plain
    16-byte-word copy kernel between two device buffers (fhip_copy_rate)."""
    import ctypes as C
    from feram_amd import _lib
    L = _lib.load()
    gbs, us = C.c_double(), C.c_double()
    rc = L.fhip_copy_rate(int(nbytes), int(reps), C.byref(gbs), C.byref(us))
    if rc != 0:
        raise RuntimeError(f"fhip_copy_rate: status {rc}")
    return gbs.value, us.value


def measure(p, u0, precision: int, steps: int, warmup: int, sync_extra):
    from feram_amd.system import System
    g = System(p, precision=precision)
    g.put("dipoR", u0)
    e0 = g.md_prologue()
    F0 = g.get("dVddi")
    # the first two steps, compared with the oracle's (the CPU-baseline leg)
    e2 = g.nose_poincare(2)
    first = {"e0": e0, "F0": F0, "e2": e2, "u2": g.get("dipoR"), "p2": g.get("dipoP"), "F2": g.get("dVddi")}

    def sync():
        g.synchronize()
        sync_extra()

    if warmup > 0:
        g.nose_poincare(warmup)
    # per-kernel times: max(K, 100) steps with HIP events around every launch
    # (direct launches), between the warm-up and the timed steps
    prof_steps = max(steps, 100)
    g.profile(True)
    g.nose_poincare(prof_steps)
    sync()
    stats = g.kernel_stats()
    g.profile(False)
    sync()
    t0 = time.perf_counter()
    e = g.nose_poincare(steps)
    sync()
    dt = time.perf_counter() - t0
    if not all(np.isfinite(v) for v in e.values()):
        raise RuntimeError(f"non-finite energies after the timed steps: {e}")
    L = p.Lx
    N, Kh = p.N, (L // 2 + 1) * L * L
    ws = 4 if precision == 1 else 8
    per = {}
    for name, (ms, cnt) in stats.items():
        avg_ms = ms / cnt
        b = kernel_bytes(name, N, Kh, 9, ws, 1.0 / prof_steps)
        per[name] = {"avg_us": avg_ms * 1e3, "launches": cnt,
                     "GBs": (b / (avg_ms * 1e-3) / 1e9) if b else None, "bytes": b}
    dom = max((k for k in per if per[k]["bytes"]), key=lambda k: per[k]["avg_us"] * per[k]["launches"])
    sb = step_bytes(N, Kh, ws)
    res = {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps,
           "kernel_sum_ms_per_step": sum(ms for ms, cnt in stats.values()) / steps,
           "dominant": dom, "kernels": per,
           "step_roofline": {"algorithmic_bytes_per_step": sb, "achieved_GBs": sb * steps / dt / 1e9,
                             "frac": sb * steps / dt / 1e9 / HBM_PEAK_GBS},
           "device_bytes": g.device_bytes, "total_energy_eV_per_uc": e["total"] / N}
    g.close()
    return res, first


def sweep_replicas(n: int, precision: int, seed: int = 20241203):
    """The batched-ensemble workload (feram_amd.sweeps.replica_ensemble): n
    BaTiO3 bulk replicas with edges drawn from {12, 16, 24, 32} (seeded,
    uneven mix), T ~ U[150, 450] K, static E_z ~ U[0, 100] kV/cm, NP-MD,
    u0 ~ U[-0.05, 0.05] A."""
    from feram_amd.sweeps import replica_ensemble
    from feram_amd.system import System
    params, u0 = replica_ensemble(n, seed)
    systems = []
    for p, u in zip(params, u0):
        g = System(p, precision=precision)
        g.put("dipoR", u)
        g.md_prologue()
        systems.append(g)
    edges = np.array([p.Lx for p in params])
    temps = np.array([p.kelvin for p in params])
    fields = np.array([p.external_E_field[2] for p in params])
    return systems, edges, temps, fields


# The reference's published NP-MD rate at 16^3: 40000 + 20000 steps
# (18example-benchmark/md16x16x16.feram:32-33) in 23.22 s of simulation time
# on one Xeon X5650 chip (18example-benchmark/md16x16x16.X5650.r1463.dat:4-5).
REF_16_STEPS_PER_S = 60000 / 23.22


def small_case(steps: int, warmup: int) -> dict:
    """The 16^3 BaTiO3 NP-MD case (FP64, 300 K, u0 ~ U[-0.05, 0.05] A seeded)
    on the device: one context, `steps` steps in one call after `warmup`."""
    from feram_amd.params import batio3_zvr
    from feram_amd.system import System
    p = batio3_zvr(16, method="md", kelvin=300.0)
    rng = np.random.default_rng(2024)
    g = System(p)
    g.put("dipoR", rng.uniform(-0.05, 0.05, size=(16, 16, 16, 3)))
    g.md_prologue()
    g.nose_poincare(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    e = g.nose_poincare(steps)
    g.synchronize()
    dt = time.perf_counter() - t0
    g.close()
    if not all(np.isfinite(v) for v in e.values()):
        raise RuntimeError("non-finite energies after the timed 16^3 steps")
    return {"L": 16, "steps": steps, "steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps,
            "reference_published_steps_per_s": REF_16_STEPS_PER_S,
            "reference_published": "Xeon X5650, one chip (18example-benchmark/md16x16x16.X5650.r1463.dat)"}


def ensemble_rate(n: int, precision: int, steps: int, warmup: int) -> dict:
    """Replica-steps/s of the n-replica batched sweep with graph-replayed
    steps, and the equilibrated replicas per hour that gives at 5000 steps
    per replica (the ensemble workload's own length)."""
    from feram_amd.system import Batch
    systems, edges, temps, fields = sweep_replicas(n, precision)
    cells = int(sum(int(L) ** 3 for L in edges))
    with Batch(systems) as bt:
        bt.step(warmup)
        t0 = time.perf_counter()
        e = bt.step(steps)
        dt = time.perf_counter() - t0
    for g in systems:
        g.close()
    if not all(np.isfinite(v) for x in e for v in x.values()):
        raise RuntimeError("non-finite energies after the timed ensemble steps")
    rate = n * steps / dt
    return {"replicas": n, "cells": cells, "steps": steps, "replica_steps_per_s": rate,
            "ms_per_step": 1e3 * dt / steps, "cell_updates_per_s": cells * steps / dt,
            "replicas_per_hour_at_5000_steps": 3600.0 * rate / 5000.0}


def bench_replicas(args, sync_extra):
    from feram_amd.system import Batch
    n = args.replicas
    modes = {}
    for name, prec in (("fp64", 0), ("fp32", 1)):
        if args.precision not in ("both", name):
            continue
        systems, edges, temps, fields = sweep_replicas(n, prec)
        cells = int(sum(int(L) ** 3 for L in edges))
        res = {"replicas": n, "cells": cells,
               "edges": {str(L): int(np.sum(edges == L)) for L in (12, 16, 24, 32)}}
        with Batch(systems) as bt:
            res["size_classes"] = bt.classes
            for graph in (True, False):
                bt.use_graph(graph)
                bt.step(args.warmup)
                sync_extra()
                print(f"[bench] {n} replicas {name} graph={graph}: timed {args.steps}", file=sys.stderr, flush=True)
                t0 = time.perf_counter()
                e = bt.step(args.steps)
                sync_extra()
                dt = time.perf_counter() - t0
                if not all(np.isfinite(v) for x in e for v in x.values()):
                    raise RuntimeError("non-finite energies after the timed steps")
                key = "graph" if graph else "direct"
                res[key] = {"replica_steps_per_s": n * args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
                            "cell_updates_per_s": cells * args.steps / dt}
                print(f"[bench] {key}: {res[key]['replica_steps_per_s']:.0f} replica-steps/s",
                      file=sys.stderr, flush=True)
        # the same replicas one context at a time (the per-context call per replica)
        ks = max(1, args.steps // 4)
        for g in systems:
            g.nose_poincare(2)
        sync_extra()
        t0 = time.perf_counter()
        for g in systems:
            g.nose_poincare(ks)
        sync_extra()
        dt = time.perf_counter() - t0
        res["sequential"] = {"replica_steps_per_s": n * ks / dt, "steps_per_replica": ks}
        res["batch_speedup_vs_sequential"] = res["graph"]["replica_steps_per_s"] / res["sequential"]["replica_steps_per_s"]
        res["replicas_per_hour_at_5000_steps"] = 3600.0 * res["graph"]["replica_steps_per_s"] / 5000.0
        for g in systems:
            g.close()
        modes[name] = res
    head_name = "fp32" if "fp32" in modes else "fp64"
    head = modes[head_name]
    out = {
        "metric": f"replica NP-MD steps/s, {n}-replica batched sweep ({head_name}; both modes alongside)",
        "value": head["graph"]["replica_steps_per_s"],
        "unit": "replica-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["graph"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if head_name == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"{n} BaTiO3 bulk replicas, edges from {{12,16,24,32}} seeded, "
                               "T ~ U[150,450] K, E_z ~ U[0,100] kV/cm, NP-MD dt=2 fs, batched "
                               "(one launch per stage and size class, graph-replayed steps)",
                   "replicas": n},
        "modes": modes,
    }
    print(json.dumps(out))


def bench_warmstart(args, sync_extra):
    """Learned warm start on a heating scan: train the network on the heating
    branch that chained plain-MD scans produce (16^3, several fields and
    pressures, all chains in one batch), then run the scan from the
    reference's default start and from the warm start in one batch."""
    from feram_amd.warmstart import WarmStart, chain_samples, heating_scan
    prec = 1 if args.precision == "fp32" else 0
    train_temps = np.arange(150, 501, 10)
    print(f"[bench] warm start: chained training scans {args.train_size}^3, {len(train_temps)} T",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    cond, targets = chain_samples(train_temps, L=args.train_size, n_thermalize=args.train_therm,
                                  n_average=args.train_avg, precision=prec)
    sync_extra()
    t_data = time.perf_counter() - t0
    ws = WarmStart(seed=2)
    t0 = time.perf_counter()
    hist = ws.fit(cond, targets, iters=args.train_iters, lr=2e-3)
    t_fit = time.perf_counter() - t0
    temps = np.linspace(250, 450, args.temps)
    print(f"[bench] heating scan {args.temps} T at {args.size}^3, {args.scan_steps} steps", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    res = heating_scan(ws, temps, L=args.size, n_steps=args.scan_steps, record=10, precision=prec, seed=7,
                       Q_Nose=args.scan_q)
    t_scan = time.perf_counter() - t0
    res.pop("series")
    print(f"[bench] gain {res['gain']:.2f}", file=sys.stderr, flush=True)
    # the plain MD judges: the reference's chained protocol at the same size
    from feram_amd.params import batio3_zvr
    from feram_amd.sweeps import temperature_chain
    base = batio3_zvr(args.size, method="md", kelvin=float(temps[0]))
    base.Q_Nose = args.scan_q
    print(f"[bench] chained plain-MD scan {args.size}^3, 2 seeds, {args.chain_therm}+{args.train_avg} steps",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    chain = temperature_chain(temps, seeds=(1, 2), base=base, n_thermalize=args.chain_therm,
                              n_average=args.train_avg, precision=prec)
    t_chain = time.perf_counter() - t0
    uw, ew = np.array(res["warm"]["u"]), np.array(res["warm"]["u_err"])
    uc = chain["u_sorted"]                                 # [T][seed][3]
    cm, csd = uc.mean(axis=1), uc.std(axis=1, ddof=1)
    agree = (np.abs(uw - cm) <= 2 * np.sqrt(ew ** 2 + csd ** 2) + 2e-3).all(axis=1)
    res["chain"] = {"u": cm.tolist(), "u_sd": csd.tolist(), "phases": chain["phases"], "wall_s": t_chain,
                    "n_thermalize": args.chain_therm, "n_average": args.train_avg}
    res["warm_P_agrees_with_chain"] = agree.tolist()
    print(f"[bench] warm <u> agrees with the chained scan at {int(agree.sum())}/{len(agree)} temperatures",
          file=sys.stderr, flush=True)
    out = {
        "metric": f"thermalisation-step gain of learned warm starts, {args.temps}-temperature "
                  f"{args.size}^3 heating scan (sum of equilibration steps, default start / warm start)",
        "value": res["gain"], "unit": "x", "n_gpus": 1, "steps": args.scan_steps, "warmup": 0,
        "ms_per_step": 1e3 * t_scan / args.scan_steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if prec else "f64", "data": "synthetic",
        "config": {"workload": f"BaTiO3 bulk {args.size}^3 NP-MD heating scan 250-450 K, E=0, -5 GPa, "
                               f"Q_Nose {args.scan_q}; network trained on {len(cond)} equilibrated states of "
                               f"chained {args.train_size}^3 scans ({args.train_therm}+{args.train_avg} steps "
                               "per temperature)",
                   "train_samples": int(len(cond)), "train_size": args.train_size, "scan_size": args.size},
        "training": {"data_s": t_data, "fit_s": t_fit, "iters": args.train_iters,
                     "loss_first": float(hist[0]), "loss_last": float(hist[-1])},
        "scan_s": t_scan,
        "P_agreement": f"{int(agree.sum())}/{len(agree)} temperatures: warm-start <u> within 2 sigma of the "
                       "chained plain-MD scan",
        "scan": res,
    }
    print(json.dumps(out))


def bench_heating16(args, sync_extra):
    """The reference's own 16^3 BaTiO3 heating scan (its example protocol:
    170-350 K in 5 K steps, 40000 + 10000 NP-MD steps per temperature, each
    started from the previous final state) for four seeds in one replica
    batch, compared with the .avg file the reference ships."""
    from feram_amd.sweeps import phase_of, temperature_chain
    with open(os.path.join(ROOT, "tests", "golden", "bto16_phase_transition.json")) as f:
        ref = {r[0]: r for r in json.load(f)["heating"]}
    temps = np.arange(170, 351, 5)
    seeds = tuple(range(1, args.seeds + 1))
    prec = 1 if args.precision == "fp32" else 0
    t0 = time.perf_counter()

    def progress(i, T, u):
        print(f"[bench] {T:.0f} K  t={time.perf_counter() - t0:.0f}s", file=sys.stderr, flush=True)

    res = temperature_chain(temps, seeds=seeds, n_thermalize=args.scan_steps, n_average=args.avg_steps,
                            precision=prec, progress=progress)
    sync_extra()
    dt = time.perf_counter() - t0
    rows = []
    for i, T in enumerate(res["temps"]):
        r = ref[T]
        rows.append({"T": T, "u_sorted": res["u_sorted_mean"][i].tolist(), "u_err": res["u_sorted_err"][i].tolist(),
                     "phases": res["phases"][i], "ref_u_sorted": np.sort(np.abs(r[7:10])).tolist(),
                     "ref_phase": phase_of(np.sort(np.abs(r[7:10])))})
    agree = sum(1 for r in rows if all(ph == r["ref_phase"] for ph in r["phases"]))
    nrep = len(temps) * len(seeds)
    steps = args.scan_steps + args.avg_steps
    out = {
        "metric": "equilibrated replicas/hr, 16^3 BaTiO3 heating scan with the reference's example protocol",
        "value": 3600.0 * nrep / dt, "unit": "replicas/hr", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": 1e3 * dt / (len(temps) * steps), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if prec else "f64", "data": "synthetic",
        "config": {"workload": f"BaTiO3 bulk 16^3, -5 GPa, NP-MD, 170-350 K / 5 K, {args.scan_steps}+"
                               f"{args.avg_steps} steps per temperature, chained, {len(seeds)} seeds batched",
                   "replicas": nrep, "replica_steps_per_s": nrep * steps / dt},
        "phase_agreement": f"{agree}/{len(rows)} temperatures with every seed in the reference's phase",
        "scan": rows, "wall_s": dt,
    }
    print(json.dumps(out))


def bench_heating48(args, sync_extra):
    """The 48^3 acoustic heating scan: BaTiO3 bulk 48x48x48, FP64 plain mode,
    velocity scaling with acoustic MD (acoustic_mass_amu 46.44 as the
    reference's 26example-BaTiO3-acoustic-MD), 200 -> 400 K in 10 K steps,
    2000 thermalising + 2000 averaging steps per temperature, each started
    from the previous final state, seeded random initial u (init_dipo_avg
    0.1, dev 0.02); one chain per seed (default one)."""
    from feram_amd.params import batio3_zvr
    from feram_amd.sweeps import temperature_chain
    temps = np.arange(200, 401, 10)
    seeds = tuple(range(1, args.seeds + 1))
    prec = 1 if args.precision == "fp32" else 0
    base = batio3_zvr(48, method="vs", kelvin=float(temps[0]), acoustic_mass_amu=46.44)
    t0 = time.perf_counter()

    def progress(i, T, u):
        print(f"[bench] {T:.0f} K  t={time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    res = temperature_chain(temps, seeds=seeds, base=base, n_thermalize=args.scan_steps,
                            n_average=args.avg_steps, precision=prec, progress=progress)
    sync_extra()
    dt = time.perf_counter() - t0
    steps = len(temps) * (args.scan_steps + args.avg_steps)
    rows = [{"T": T, "u_sorted": res["u_sorted_mean"][i].tolist(), "phases": res["phases"][i],
             "strain": res["strain"][i].mean(axis=0).tolist()} for i, T in enumerate(res["temps"])]
    out = {
        "metric": "MD steps/s of the 48^3 acoustic heating scan (velocity scaling, acoustic MD)",
        "value": len(seeds) * steps / dt, "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": 1e3 * dt / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec else "f64", "data": "synthetic",
        "config": {"workload": f"BaTiO3 bulk 48^3, -5 GPa, vs with acoustic MD, 200-400 K / 10 K, "
                               f"{args.scan_steps}+{args.avg_steps} steps per temperature, chained, "
                               f"{len(seeds)} seed(s)",
                   "cells": 48 ** 3, "cell_updates_per_s": len(seeds) * steps * 48 ** 3 / dt},
        "scan": rows, "wall_s": dt,
    }
    print(json.dumps(out))


def bench_epit_hl(args, sync_extra):
    """The reference's epitaxial-film hysteresis example
    (11example-BaTiO3-new-hysteresis-loop/hysteresis16x16x128d0.feram):
    BaTiO3 epit 16x16x128, gap_id 0, epi_strain -1 %, 150 K, velocity
    scaling, E_z swept from -0.03 to +0.03 V/A over n_average steps after
    n_thermalize, a .hl row every n_hl_freq steps.  Its 00TIMING.txt lists
    the wall time of the whole run (20000 + 3000000 steps) on the reference's
    machines; the fastest is feram-0.18.05 on an SR11000 node (16 cores),
    7638 s."""
    from feram_amd.params import batio3_zvr
    from feram_amd.sweeps import hysteresis_loop
    from feram_amd.system import System
    p = batio3_zvr((16, 16, 128), method="vs", kelvin=150.0)
    p.bulk_or_film, p.gap_id, p.epi_strain = "epit", 0, -0.01
    p.dt, p.Q_Nose = 0.001, 0.05
    p.external_E_field = [0.0, 0.0, -0.03]
    p.E_field_always_zero = False
    prec = 1 if args.precision == "fp32" else 0
    g = System(p, precision=prec)
    rng = np.random.default_rng(20111113)
    g.put("dipoR", np.array([0.0, 0.0, -0.10]) + 0.02 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
    g.md_prologue()
    n_therm, n_avg, freq = args.hl_therm, args.hl_avg, args.hl_freq
    t0 = time.perf_counter()

    def progress(step, rows):
        if (step - n_therm) % (50 * freq) == 0:
            print(f"[bench] hl step {step}  t={time.perf_counter() - t0:.0f}s", file=sys.stderr, flush=True)

    rows = np.array(hysteresis_loop([g], n_therm, n_avg, freq, p.external_E_field, i_start=1,
                                    progress=progress)[0])
    sync_extra()
    dt = time.perf_counter() - t0
    nsteps = n_therm + n_avg
    # P_z in uC/cm^2 from <u_z> (hysteresis16x16x128d0.gp: 1.6e3 Z* / a0^3)
    c = 1.6e3 * p.Z_star[0] / p.a0 ** 3
    E, Pz = rows[:, 4], rows[:, 13] * c
    flips = [float(E[i]) for i in range(1, len(Pz)) if np.sign(Pz[i]) != np.sign(Pz[i - 1])]
    out = {
        "metric": "MD steps/s, the reference's 16x16x128 epit hysteresis-loop example",
        "value": nsteps / dt, "unit": "steps/s", "n_gpus": 1, "steps": nsteps, "warmup": 0,
        "ms_per_step": 1e3 * dt / nsteps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if prec else "f64", "data": "synthetic",
        "config": {"workload": f"BaTiO3 epit 16x16x128 gap_id 0, epi_strain -0.01, 150 K, vs, "
                               f"E_z -0.03 -> +0.03 V/A, {n_therm}+{n_avg} steps, .hl every {freq}"},
        "wall_s": dt,
        "reference_wall_s_full_run": {"feram-0.18.05 SR11000 16 cores": 7638,
                                      "feram-0.18.00 Xeon X5690 12 cores": 9582},
        "full_run_s_at_this_rate": 3020000 / (nsteps / dt),
        "loop": {"Pz_uC_cm2_start": float(Pz[0]), "Pz_uC_cm2_end": float(Pz[-1]),
                 "E_at_sign_changes_V_A": flips, "rows": len(rows)},
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--precision", choices=("both", "fp32", "fp64"), default="both")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 48-replica ensemble and 16^3 figures of the default workload")
    ap.add_argument("--q-nose", type=float, default=0.0, help="<= 0: 1.0 * N / 64^3")
    ap.add_argument("--workload", choices=("md128", "replicas", "warmstart", "heating16", "heating48", "epit_hl"),
                    default="md128",
                    help="md128: the single-supercell NP-MD step (default); replicas: the "
                         "batched replica sweep; warmstart: the learned warm start on a heating scan; "
                         "heating16: the reference's 16^3 heating-scan example against its .avg output; "
                         "heating48: the 48^3 acoustic (velocity-scaling) heating scan, 200-400 K; "
                         "epit_hl: the reference's 16x16x128 epit hysteresis-loop example")
    ap.add_argument("--hl-therm", type=int, default=20000)
    ap.add_argument("--hl-avg", type=int, default=3000000)
    ap.add_argument("--hl-freq", type=int, default=10000)
    ap.add_argument("--seeds", type=int, default=4)
    ap.add_argument("--avg-steps", type=int, default=10000)
    ap.add_argument("--replicas", type=int, default=48)
    ap.add_argument("--temps", type=int, default=20)
    ap.add_argument("--scan-steps", type=int, default=4000)
    ap.add_argument("--train-size", type=int, default=16)
    ap.add_argument("--train-therm", type=int, default=6000)
    ap.add_argument("--train-avg", type=int, default=3000)
    ap.add_argument("--train-iters", type=int, default=4000)
    ap.add_argument("--chain-therm", type=int, default=10000)
    ap.add_argument("--scan-q", type=float, default=1.0,
                    help="Q_Nose of the warm-start scan (the reference's 64^3 md benchmark uses 1.0)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's plain device copy of MIB MiB (5 timed launches after 2 "
                         "of warm-up) to calibrate the FETCH_SIZE / WRITE_SIZE counters, then exit")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        nbytes = args.copy_calibration * (1 << 20)
        gbs, us = copy_rate(nbytes, 5)
        print(json.dumps({"copy_bytes_per_launch_read": nbytes, "copy_bytes_per_launch_write": nbytes,
                          "kernel": "k_copy16", "GBs": gbs, "us": us}))
        return
    if args.gpus != 1:
        print(f"bench.py: the step is measured on one GPU; --gpus {args.gpus} is not run "
              "(replica ensembles would shard across GPUs with no exchange, one supercell as z-slabs "
              "with one all-to-all per transform direction)", file=sys.stderr)
        sys.exit(2)

    def sync_extra():
        pass

    if args.workload == "replicas":
        bench_replicas(args, sync_extra)
        return
    if args.workload == "epit_hl":
        bench_epit_hl(args, sync_extra)
        return
    if args.workload == "heating48":
        if args.scan_steps == 4000:
            args.scan_steps = 2000
        if args.avg_steps == 10000:
            args.avg_steps = 2000
        if args.seeds == 4:
            args.seeds = 1
        bench_heating48(args, sync_extra)
        return
    if args.workload == "heating16":
        if args.scan_steps == 4000:
            args.scan_steps = 40000
        bench_heating16(args, sync_extra)
        return
    if args.workload == "warmstart":
        if args.size == 128:
            args.size = 64
        if args.scan_steps == 4000:
            args.scan_steps = 10000
        bench_warmstart(args, sync_extra)
        return

    from feram_amd.params import batio3_zvr

    L = args.size
    p = batio3_zvr(L, method="md", kelvin=300.0)
    # eq. 38's root needs CB dt/Q < 1 with CB ~ 3N kT dt/2, so Q_Nose grows with N:
    # the reference's 64^3 md benchmark uses Q_Nose = 1.0
    # (18example-benchmark/md64x64x64.feram:6); scale it with N.
    p.Q_Nose = args.q_nose if args.q_nose > 0 else 1.0 * p.N / 64 ** 3
    rng = np.random.default_rng(20241203)
    u0 = rng.uniform(-0.05, 0.05, size=(L, L, L, 3))

    modes = {}
    firsts = {}
    for name, prec in (("fp64", 0), ("fp32", 1)):
        if args.precision not in ("both", name):
            continue
        print(f"[bench] {L}^3 {name}: warmup {args.warmup}, timed {args.steps}", file=sys.stderr, flush=True)
        res, first = measure(p, u0, prec, args.steps, args.warmup, sync_extra)
        print(f"[bench] {name}: {res['steps_per_s']:.1f} steps/s", file=sys.stderr, flush=True)
        modes[name] = res
        firsts[name] = first
    if "fp64" in modes and "fp32" in modes:
        modes["fp32"]["vs_fp64_mode"] = errors_vs(firsts["fp32"], firsts["fp64"])
        modes["fp32"]["speedup_vs_fp64"] = modes["fp32"]["steps_per_s"] / modes["fp64"]["steps_per_s"]

    head_name = "fp64" if "fp64" in modes else "fp32"
    head = modes[head_name]
    dom = head["dominant"]
    kd = head["kernels"][dom]
    achieved = kd["GBs"]
    traffic = None
    tr_path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tr_path):
        with open(tr_path) as f:
            traffic = json.load(f).get(f"{L}:{head_name}:{dom}")
    copy_gbs, _ = copy_rate()

    cpu = None
    if not args.no_cpu:
        print(f"[bench] cpu baseline ({args.cpu_steps} steps) and oracle comparison", file=sys.stderr, flush=True)
        cpu, ref = cpu_baseline(p, u0, args.cpu_steps)
        if ref is not None:
            for name in modes:
                modes[name]["vs_oracle"] = errors_vs(firsts[name], ref)
                print(f"[bench] {name} vs oracle: force {modes[name]['vs_oracle']['force_rel_err_step2']:.2e}, "
                      f"energy {modes[name]['vs_oracle']['energy_rel_err_step2']:.2e}", file=sys.stderr, flush=True)

    # the batched-ensemble metric (equilibrated replicas per hour) beside the
    # single-supercell one: 48 replicas of {12,16,24,32}^3, graph-replayed
    ensemble = None
    if not args.no_ensemble:
        ensemble = {}
        for name, prec in (("fp64", 0), ("fp32", 1)):
            if args.precision not in ("both", name):
                continue
            print(f"[bench] 48-replica ensemble {name}", file=sys.stderr, flush=True)
            ensemble[name] = ensemble_rate(48, prec, 100, 20)
            print(f"[bench] ensemble {name}: {ensemble[name]['replica_steps_per_s']:.0f} replica-steps/s",
                  file=sys.stderr, flush=True)

    small = None
    if not args.no_ensemble:
        print("[bench] 16^3 NP-MD", file=sys.stderr, flush=True)
        small = small_case(1000, 100)
        print(f"[bench] 16^3: {small['steps_per_s']:.0f} steps/s", file=sys.stderr, flush=True)

    N = p.N
    out = {
        "metric": f"NP-MD steps/s at {L}^3 ({head_name.upper()} plain mode; the other mode alongside)",
        "value": head["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head_name == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": f"BaTiO3 bulk {L}x{L}x{L} NP-MD ('md'), 300 K, dt=2 fs, optimised "
                               "acoustic strain, u0 ~ U[-0.05,0.05] A seeded",
                   "L": L, "cells": N, "method": "md", "precision": head_name, "Q_Nose": p.Q_Nose,
                   "cell_updates_per_s": head["steps_per_s"] * N},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": traffic, "bytes_per_launch": kd["bytes"], "avg_launch_us": kd["avg_us"],
                     "copy_GBs": copy_gbs},
        "step_roofline": head["step_roofline"],
        "modes": modes,
        "ensemble": ensemble,
        "small_16cubed": small,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
