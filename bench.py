#!/usr/bin/env python3
"""Throughput of the batched shower transport on one MI355X.

Workload (one step): 100,000 primaries drawn uniformly from {e-, e+, gamma},
energies log-uniform in [0.1, 100] GeV along +z with a Gaussian angular spread
of 1 mrad per projected angle, in lead, tracking cutoff 0.010 GeV, Philox
seed 1234, shipped 100 GeV tables.  A step transports every primary and all
its descendants to the end (PETITE's generate_shower for each primary).

Prints one JSON line: particle-steps/s (value), showers/s, the roofline of
the dominant kernel (from HIP events on the run's stream) and a CPU baseline
(the float64 oracle, one core, on a bounded sample of the same primaries).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_PRIMARIES = 100_000
SEED = 1234
MATERIAL = "lead"
CUTOFF = 0.010
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic bytes per unit of work: a particle-step reads and
# writes a 64-byte live record; an emitted particle writes a 96-byte record;
# a primary is read once (32 bytes)
B_STEP, B_OUT, B_PRIM = 128, 96, 32


def make_primaries(n=N_PRIMARIES, seed=SEED):
    from petite_amd.batch import PRIMARY_DTYPE
    from petite_amd.constants import m_electron
    rng = np.random.default_rng(seed)
    pid = rng.choice(np.array([11, -11, 22], dtype=np.int32), size=n)
    E = np.exp(rng.uniform(np.log(0.1), np.log(100.0), n))
    tx = rng.normal(0.0, 1e-3, n)
    ty = rng.normal(0.0, 1e-3, n)
    m = np.where(pid == 22, 0.0, m_electron)
    p = np.sqrt(E * E - m * m)
    nrm = np.sqrt(tx * tx + ty * ty + 1.0)
    prim = np.zeros(n, dtype=PRIMARY_DTYPE)
    prim["E"] = E
    prim["px"], prim["py"], prim["pz"] = p * tx / nrm, p * ty / nrm, p / nrm
    prim["mass"] = m
    prim["weight"] = 1.0
    prim["pid"] = pid
    prim["shower"] = np.arange(n, dtype=np.int32)
    prim["sid"] = np.arange(n, dtype=np.uint64)
    return prim


def cpu_baseline(prim, maxf, budget_s=20.0):
    """Oracle (float64, scalar Python, per-candidate draws) on the first primaries."""
    from oracle.shower import OracleShower
    from oracle.tables import SMTables
    T = SMTables(MATERIAL)
    osh = OracleShower(T, CUTOFF, seed=SEED, maxf=maxf, vectorised_draw=False)
    t0 = time.perf_counter()
    n = 0
    steps0 = osh.stats.steps
    while n < len(prim) and time.perf_counter() - t0 < budget_s:
        r = prim[n]
        osh.generate_shower(int(r["pid"]), [r["E"], r["px"], r["py"], r["pz"]], sid=int(r["sid"]), shower=n,
                            mass=float(r["mass"]))
        n += 1
    dt = time.perf_counter() - t0
    steps = osh.stats.steps - steps0
    return {"value": steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "showers_per_s": n / dt,
            "sample": f"first {n} of the {len(prim)} primaries, {dt:.1f} s on one core "
                      f"(oracle float64 restatement, per-candidate accept/reject)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--primaries", type=int, default=N_PRIMARIES)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): showers are independent, so a multi-GPU "
              "run is a sharding of primaries with no exchange; see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch
    from petite_amd import Shower

    prim = make_primaries(args.primaries)
    sh = Shower("100GeV", MATERIAL, CUTOFF, seed=SEED)
    ctx = sh.device_context()
    ctx.set_timing(True)
    print(f"bench: {len(prim)} primaries, lead, cutoff {CUTOFF}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        b = sh.generate_showers(prim, seed=SEED)
        print(f"warmup {w}: {b.n} particles, {b.stats['n_iterations']} sweeps", file=sys.stderr, flush=True)
        del b
    torch.cuda.synchronize()
    totals = {"steps": 0, "particles": 0, "draws": 0, "trials": 0}
    kern = {"transport": [0.0, 0], "draw": [0.0, 0], "emit": [0.0, 0]}
    t0 = time.perf_counter()
    for k in range(args.steps):
        b = sh.generate_showers(prim, seed=SEED)
        st = b.stats
        totals["steps"] += st["n_steps"]
        totals["particles"] += st["n_particles"]
        totals["draws"] += st["n_draws"]
        totals["trials"] += st["n_trials"]
        for key in kern:
            kern[key][0] += st[f"ms_{key}"]
            kern[key][1] += st[f"launches_{key}"]
        print(f"step {k}: {b.n} particles, {st['n_steps']} particle-steps", file=sys.stderr, flush=True)
        del b
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    K = args.steps
    steps_per_s = totals["steps"] / dt
    showers_per_s = K * len(prim) / dt

    # dominant kernel and its roofline
    dom = max(kern, key=lambda k: kern[k][0])
    t_dom = kern[dom][0] / 1e3
    if dom == "transport":
        alg = B_STEP * totals["steps"]
        ach = alg / t_dom / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": None, "kernel": "k_transport", "per_unit": "128 B per particle-step"}
    elif dom == "emit":
        alg = B_OUT * totals["particles"]
        ach = alg / t_dom / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": None, "kernel": "k_emit", "per_unit": "96 B per emitted particle"}
    else:
        # accept/reject candidates: vector-ALU bound; flops counted per candidate
        flops = FLOPS_PER_CANDIDATE * totals["trials"]
        ach = flops / t_dom / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / FP32_PEAK_TFLOPS, "traffic": None, "kernel": "k_draw",
                "per_unit": f"{FLOPS_PER_CANDIDATE} flop per candidate (vector ALU; no MFMA used)"}
    roof["avg_launch_ms"] = kern[dom][0] / max(kern[dom][1], 1)
    kernels = {k: {"ms_total": v[0] / K, "launches": v[1] // K} for k, v in kern.items()}

    cpu = None
    if not args.no_cpu_baseline:
        maxf = {(P, ie): float(v) for P, arr in ctx.maxf.items() for ie, v in enumerate(arr)}
        cpu = cpu_baseline(prim, maxf, args.cpu_budget)

    out = {
        "metric": "particle-steps/s (batched SM showers, 1 MI355X)",
        "value": steps_per_s,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": dt / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(prim)} primaries e-/e+/gamma, E log-uniform [0.1,100] GeV, 1 mrad spread, "
                               f"lead, cutoff {CUTOFF} GeV, Philox seed {SEED}, 100 GeV tables",
                   "primaries": len(prim), "material": MATERIAL, "cutoff_GeV": CUTOFF, "parallelism": "single GPU"},
        "showers_per_s": showers_per_s,
        "particles_per_step": totals["particles"] / K,
        "particle_steps_per_step": totals["steps"] / K,
        "trials_per_draw": totals["trials"] / max(totals["draws"], 1),
        "kernels": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu_showers"] = showers_per_s / cpu["showers_per_s"]
    print(json.dumps(out))


# float operations of one accept/reject candidate, averaged over the mix of
# processes (map transform + integrand; Philox integer work excluded)
FLOPS_PER_CANDIDATE = 120

if __name__ == "__main__":
    main()
