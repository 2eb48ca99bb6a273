#!/usr/bin/env python3
"""Throughput of the batched shower transport on one MI355X.

Workload (one step): 100,000 primaries drawn uniformly from {e-, e+, gamma},
energies log-uniform in [0.1, 100] GeV along +z with a Gaussian angular spread
of 1 mrad per projected angle, in lead, tracking cutoff 0.010 GeV, Philox
seed 1234, shipped 100 GeV tables.  A step transports every primary and all
its descendants to the end (PETITE's generate_shower for each primary).

Prints one JSON line: particle-steps/s (value), showers/s, the roofline of
the dominant kernel (from HIP events on the run's stream) and a CPU baseline
(the float64 oracle on a stratified sample of the same primaries: one core,
and all cores of a process pool), per-process draw statistics with the
max_F violations, and the reference's own CPU case beside the GPU under
"c1_graphite_10GeV" (1,000 showers of 10 GeV e- in graphite; the oracle per
shower against the published 2.59 s).
The dark-yield scan follows under "dark_yield_scan" (a fixed record of 50,000
e- showers in lead, yield and d/dm_V at 32 masses: gradient evaluations/s and
the gradient/forward cost ratio) with its CPU baseline; --dark-forward adds
"dark_forward" (20,000 e- at 10 GeV in graphite, m_V = 1 MeV: vectors/s).
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
# algorithmic bytes of one particle-step: it reads and writes a 64-byte live
# particle record
B_STEP = 128


def measured_valu(kernel):
    """The instruction roofline of `kernel` from the committed rocprofv3 SQ
    counter pass of this workload (profiles/valu.json, made by the command it
    records): VALU issue as a fraction of the SIMDs' issue peak, SIMD VALU-busy
    fraction and resident waves per SIMD."""
    path = os.path.join(ROOT, "profiles", "valu.json")
    try:
        with open(path) as f:
            t = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None
    return dict(t, source="profiles/valu.json")


def measured_traffic(kernel, alg_bytes_per_launch):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement of this workload (profiles/traffic.json, made by the
    command it records), with the algorithmic bytes per launch beside it."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None
    return {"bytes_per_launch": t["bytes_per_launch"], "fetch_bytes": t["fetch_bytes"],
            "write_bytes": t["write_bytes"], "over_algorithmic": t["bytes_per_launch"] / alg_bytes_per_launch,
            "source": "profiles/traffic.json"}


def make_primaries(n=N_PRIMARIES, seed=SEED):
    from petite_amd.workloads import batch_primaries
    return batch_primaries(n, seed)


def tuning_report(ctx):
    """Where the run's launch sizes came from, and the autotuner's table entry
    for this device (its best time against the built-in defaults)."""
    from petite_amd import autotune
    out = {"source": ctx.tuning_source, "point": ctx.get_tuning()}
    try:
        import json as _json
        with open(autotune.TABLE) as f:
            ent = _json.load(f).get(autotune.device_key())
        if ent:
            out.update({"table_ms": ent["ms"], "table_default_ms": ent["default_ms"], "table_workload": ent["workload"]})
    except (OSError, ValueError):
        pass
    return out


def _oracle_shower(args):
    """One oracle shower in a worker process: (particle-steps, seconds)."""
    material, cutoff, maxf, pid, p4, sid, mass = args
    from oracle.shower import OracleShower
    from oracle.tables import SMTables
    osh = OracleShower(SMTables(material), cutoff, seed=SEED, maxf=maxf, vectorised_draw=False)
    t0 = time.perf_counter()
    osh.generate_shower(pid, p4, sid=sid, shower=0, mass=mass)
    return osh.stats.steps, time.perf_counter() - t0


def _oracle_pool(jobs, workers):
    """Run oracle showers on `workers` processes (spawned: they start a fresh
    interpreter with one BLAS thread each, never a fork of this GPU process).
    Returns (per-job results, wall seconds)."""
    import multiprocessing as mp
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in env:
        os.environ[k] = "1"
    try:
        t0 = time.perf_counter()
        with mp.get_context("spawn").Pool(workers) as pool:
            res = pool.map(_oracle_shower, jobs, chunksize=1)
        wall = time.perf_counter() - t0
    finally:
        for k, v in env.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return res, wall


def _workers():
    n = os.cpu_count() or 1
    return max(1, min(16, n))  # the GPU box gives one GPU's share of 16 cores


def cpu_baseline(prim, maxf, steps_per_shower, strata=20):
    """Oracle (float64, scalar Python, per-candidate draws like draw_sample) on
    a stratified sample of the workload: the first primary of the list in each
    of `strata` equal log-energy strata of [0.1, 100] GeV.  One core: summed
    particle-steps over summed per-shower CPU time; all cores: the same
    showers on a process pool, over its wall time.  Showers/s: the one-core
    rate over the workload's mean particle-steps per shower (the GPU's showers
    are the oracle's, particle for particle)."""
    edges = np.geomspace(0.1, 100.0, strata + 1)
    picks = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        idx = np.nonzero((prim["E"] >= lo) & (prim["E"] < hi))[0]
        if len(idx):
            picks.append(int(idx[0]))
    picks.sort(key=lambda i: -float(prim["E"][i]))  # the longest first
    jobs = [(MATERIAL, CUTOFF, maxf, int(prim["pid"][i]), [float(prim["E"][i]), float(prim["px"][i]),
             float(prim["py"][i]), float(prim["pz"][i])], int(prim["sid"][i]), float(prim["mass"][i])) for i in picks]
    workers = _workers()
    res, wall = _oracle_pool(jobs, workers)
    steps = sum(r[0] for r in res)
    cpu_s = sum(r[1] for r in res)
    one = steps / cpu_s
    return {"value": one, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "showers_per_s": one / steps_per_shower,
            "all_cores": {"cores": workers, "value": steps / wall, "unit": "particle-steps/s",
                          "showers_per_s": steps / wall / steps_per_shower, "wall_s": wall},
            "sample": f"{len(picks)} primaries, the first of the workload list in each of {strata} equal log-energy "
                      f"strata of [0.1, 100] GeV ({steps} particle-steps, {cpu_s:.1f} s of CPU); showers/s = "
                      f"particle-steps/s over the workload's {steps_per_shower:.0f} particle-steps per shower; "
                      f"oracle float64 restatement, per-candidate accept/reject"}


PROC_NAMES = ("Brem", "Ann", "PairProd", "Comp", "Moller", "Bhabha", "MuonE", "MuonBrem")


def bench_c5(n=10_000, reps=2):
    """The 400 GeV workload: n primaries at 400 GeV along +z, half muons and
    half electrons, lead, cutoff 0.100 GeV, the 400 GeV table set with its
    GPU-trained Brem/PairProd/MuonBrem maps (best of `reps` runs)."""
    import torch

    from petite_amd import Shower
    from petite_amd.batch import PRIMARY_DTYPE
    from petite_amd.constants import mass_dict
    prim = np.zeros(n, dtype=PRIMARY_DTYPE)
    pid = np.where(np.arange(n) < n // 2, 13, 11).astype(np.int32)
    m = np.array([mass_dict[int(q)] for q in pid])
    prim["E"], prim["pz"], prim["mass"], prim["weight"], prim["pid"] = 400.0, np.sqrt(400.0 ** 2 - m * m), m, 1.0, pid
    prim["shower"] = np.arange(n)
    prim["sid"] = np.arange(n)
    sh = Shower("400GeV", "lead", 0.100, seed=SEED)
    b = sh.generate_showers(prim, seed=SEED)
    best = 1e9
    for _ in range(reps):
        del b
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b = sh.generate_showers(prim, seed=SEED)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    st = b.stats
    return {"workload": f"{n} primaries at 400 GeV along +z (half mu-, half e-), lead, cutoff 0.100 GeV, 400 GeV tables "
                        f"with GPU-trained maps",
            "showers_per_s": n / best, "particle_steps_per_s": st["n_steps"] / best, "ms": best * 1e3,
            "particles": st["n_particles"], "maps": dict(sh._tables.map_source),
            "sweeps": st["n_iterations"], "tail_launches": st["launches_tail"],
            "draws_by_process": {name: {"draws": d, "trials_per_draw": t / d, "clipped_fraction": v / d}
                                 for name, d, t, v in zip(PROC_NAMES, st["n_draws_proc"], st["n_trials_proc"],
                                                          st["n_viol_proc"]) if d > 0}}


def bench_c1(cpu=True, n=1000, reps=3):
    """The reference's own CPU case: 10 GeV e- along +z in graphite, cutoff
    0.010 GeV, one shower per seed 0..n-1 (Philox stream id = seed).  GPU:
    all n showers in one batched run (best of `reps`); CPU: the oracle on
    one shower per worker, against the published 2.59 s per shower (the
    reference under cProfile on the author's machine)."""
    import torch

    from petite_amd import Shower
    from petite_amd.batch import PRIMARY_DTYPE
    me = 0.51099895e-3
    prim = np.zeros(n, dtype=PRIMARY_DTYPE)
    prim["E"], prim["pz"], prim["mass"], prim["weight"], prim["pid"] = 10.0, math.sqrt(100.0 - me * me), me, 1.0, 11
    prim["shower"] = np.arange(n)
    prim["sid"] = np.arange(n)
    sh = Shower("100GeV", "graphite", CUTOFF, seed=SEED)
    b = sh.generate_showers(prim, seed=SEED)
    best = 1e9
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b = sh.generate_showers(prim, seed=SEED)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    st = b.stats
    out = {"workload": f"{n} showers of 10 GeV e- along +z, graphite, cutoff {CUTOFF} GeV, 100 GeV tables, "
                       f"stream ids 0..{n - 1}",
           "showers_per_s": n / best, "particle_steps_per_s": st["n_steps"] / best,
           "particles_per_shower": st["n_particles"] / n, "ms": best * 1e3}
    if cpu:
        ctx = sh.device_context()
        maxf = {(P, ie): float(v) for P, arr in ctx.maxf.items() for ie, v in enumerate(arr)}
        workers = _workers()
        jobs = [("graphite", CUTOFF, maxf, 11, [10.0, 0.0, 0.0, math.sqrt(100.0 - me * me)], i, me)
                for i in range(workers)]
        res, wall = _oracle_pool(jobs, workers)
        s_per = sum(r[1] for r in res) / len(res)
        out["cpu_baseline"] = {"s_per_shower": s_per, "showers_per_s": 1.0 / s_per, "cores": 1, "kind": "port",
                               "all_cores": {"cores": workers, "showers_per_s": len(res) / wall},
                               "published_reference_s_per_shower": 2.59,
                               "sample": f"stream ids 0..{workers - 1}, one shower per worker process, oracle "
                                         f"float64 restatement with per-candidate accept/reject"}
        out["gpu_over_cpu_showers"] = out["showers_per_s"] / out["cpu_baseline"]["showers_per_s"]
    return out


def copy_calibration():
    """Three copies of a 256 MiB device buffer: a kernel of known traffic
    (256 MiB read + 256 MiB written per copy) against which the rocprofv3
    FETCH_SIZE / WRITE_SIZE readings of the shower kernels are scaled."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {n * 4} bytes x3", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--primaries", type=int, default=N_PRIMARIES)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-dark", action="store_true", help="skip the dark-yield workload")
    ap.add_argument("--showers-only", action="store_true",
                    help="only the value workload (no C1, detector, map-training, meson or dark legs): for profiling")
    ap.add_argument("--dark-forward", action="store_true",
                    help="also run the dark-vector workload (20,000 e- at 10 GeV, m_V = 1 MeV)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB device buffer three times (the reference load for FETCH_SIZE and "
                         "WRITE_SIZE counter scaling) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.steps < 1:
        sys.exit("bench.py: --steps must be at least 1")
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): showers are independent, so a multi-GPU "
              "run is a sharding of primaries with no exchange between GPUs", file=sys.stderr)
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
    totals = {"steps": 0, "particles": 0, "draws": 0, "trials": 0, "tail_steps": 0}
    # stage times from HIP events on the run's stream: the transport kernel,
    # the draws with their emissions (per-process streams, fork to join), and
    # the depth-first walk of the last particles
    kern = {"transport": [0.0, 0], "draw_emit": [0.0, 0], "tail": [0.0, 0]}
    t0 = time.perf_counter()
    for k in range(args.steps):
        b = sh.generate_showers(prim, seed=SEED)
        st = b.stats
        totals["steps"] += st["n_steps"]
        totals["particles"] += st["n_particles"]
        totals["draws"] += st["n_draws"]
        totals["trials"] += st["n_trials"]
        kern["transport"][0] += st["ms_transport"]
        kern["transport"][1] += st["launches_transport"]
        kern["draw_emit"][0] += st["ms_draw"] + st["ms_emit"]
        kern["draw_emit"][1] += st["launches_draw"]
        kern["tail"][0] += st["ms_tail"]
        kern["tail"][1] += st["launches_tail"]
        totals["tail_steps"] += st["n_steps_tail"]
        print(f"step {k}: {b.n} particles, {st['n_steps']} particle-steps", file=sys.stderr, flush=True)
        if k + 1 < args.steps:
            del b
    torch.cuda.synchronize()
    last_stats = b.stats
    dt = time.perf_counter() - t0
    K = args.steps
    steps_per_s = totals["steps"] / dt
    showers_per_s = K * len(prim) / dt

    # roofline of the dominant kernel, k_transport (the largest single kernel in
    # the rocprofv3 summary; the draw kernels of the six processes run side by side)
    t_dom = kern["transport"][0] / 1e3
    alg = B_STEP * (totals["steps"] - totals["tail_steps"])  # the steps k_transport took
    ach = alg / t_dom / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "kernel": "k_transport", "per_unit": "128 B per particle-step"}
    roof["avg_launch_ms"] = kern["transport"][0] / max(kern["transport"][1], 1)
    roof["traffic"] = measured_traffic(roof["kernel"], alg / max(kern["transport"][1], 1))
    # the bound that holds: k_transport is VALU-issue bound (its HBM traffic is
    # a fifth of the algorithmic bytes: a particle stays in registers)
    roof["valu"] = measured_valu(roof["kernel"])
    kernels = {k: {"ms_total": v[0] / K, "launches": v[1] // K} for k, v in kern.items()}

    cpu = None
    if not args.no_cpu_baseline:
        maxf = {(P, ie): float(v) for P, arr in ctx.maxf.items() for ie, v in enumerate(arr)}
        cpu = cpu_baseline(prim, maxf, totals["steps"] / K / len(prim))
        print(f"cpu baseline: {cpu['value']:.0f} particle-steps/s on one core", file=sys.stderr, flush=True)

    c1_out = c5_out = det_out = train_out = meson_out = None
    if not args.showers_only:
        c1_out = bench_c1(cpu=not args.no_cpu_baseline)
        print(f"c1: {c1_out['showers_per_s']:.0f} showers/s", file=sys.stderr, flush=True)
        c5_out = bench_c5()
        print(f"c5: {c5_out['showers_per_s']:.0f} showers/s", file=sys.stderr, flush=True)
        det_out = bench_detector(b, cpu=not args.no_cpu_baseline)
        print(f"detector: {det_out['records']} records, {det_out['ms']:.3f} ms", file=sys.stderr, flush=True)
    del b
    if not args.showers_only:
        train_out = bench_map_training(cpu=not args.no_cpu_baseline)
        print(f"map training: {train_out['samples']} samples, {train_out['s_total']:.2f} s", file=sys.stderr,
              flush=True)
        meson_out = bench_meson_beam(cpu=not args.no_cpu_baseline)
        print(f"meson beam: {meson_out['records']} records, {meson_out['ms']:.3f} ms", file=sys.stderr, flush=True)

    dark_out = yield_out = None
    if args.showers_only:
        args.no_dark, args.dark_forward = True, False
    if args.dark_forward:
        del sh, ctx
        torch.cuda.empty_cache()
        dark_out, ds, rec, dark = bench_dark_forward()
        print(f"dark: {dark_out['vectors']} vectors from {dark_out['sm_records']} records", file=sys.stderr, flush=True)
        if not args.no_cpu_baseline:
            dark_out["cpu_baseline"] = cpu_dark_baseline(ds, rec)
            dark_out["gpu_over_cpu_vectors"] = dark_out["vectors_per_s"] / dark_out["cpu_baseline"]["value"]
        del ds, rec, dark
    if not args.no_dark:
        torch.cuda.empty_cache()
        yield_out, ds, rec, model = bench_yield()
        print(f"yield: {yield_out['sm_records']} records, grad/forward {yield_out['grad_over_forward']:.2f}",
              file=sys.stderr, flush=True)
        if not args.no_cpu_baseline:
            yield_out["cpu_baseline"] = cpu_yield_baseline(rec, model)
            yield_out["gpu_over_cpu_grad_evals"] = yield_out["grad_evals_per_s"] / yield_out["cpu_baseline"]["value"]
        del ds, rec, model

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
        "tail_particle_steps_per_step": totals["tail_steps"] / K,
        "trials_per_draw": totals["trials"] / max(totals["draws"], 1),
        "draws_by_process": {name: {"draws": d, "trials_per_draw": t / d, "maxf_violations": v,
                                    "clipped_fraction": v / d}
                             for name, d, t, v in zip(PROC_NAMES, last_stats["n_draws_proc"],
                                                      last_stats["n_trials_proc"], last_stats["n_viol_proc"])
                             if d > 0},
        "maxf_violations": last_stats["n_maxf_viol"],
        "tuning": tuning_report(ctx),
        "kernels": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu_showers"] = showers_per_s / cpu["showers_per_s"]
    for key, val in (("c1_graphite_10GeV", c1_out), ("c5_400GeV", c5_out), ("detector_cut", det_out),
                     ("map_training", train_out),
                     ("meson_beam", meson_out)):
        if val is not None:
            out[key] = val
    if dark_out:
        out["dark_forward"] = dark_out
    if yield_out:
        out["dark_yield_scan"] = yield_out
    print(json.dumps(out))


def bench_detector(batch, reps=10, cpu=True):
    """detector_cut (shower.py:825-865) over the whole SM record of the last
    step: three planes, an annulus and an energy window; HIP events around
    `reps` launches on the current stream.  40 algorithmic bytes per record
    (p0 and r0 read, the 64-bit pass mask written)."""
    import torch

    from petite_amd.detector import cut_masks
    planes, radius, inner, ecut = [1.0, 2.0, 5.0], 0.5, 0.01, (0.1, 1000.0)
    p, r = batch.p0, batch.r0
    n = batch.n
    cut_masks(p, r, planes, radius, ecut, inner)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        mask, wsum = cut_masks(p, r, planes, radius, ecut, inner)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    ach = 40.0 * n / (ms / 1e3) / 1e9
    out = {"workload": f"the {n}-record SM record of the value workload; planes z = {planes} m, radius {radius} m, "
                       f"inner {inner} m, energy window {ecut} GeV",
           "records": n, "ms": ms, "records_per_s": n / (ms / 1e3),
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": "k_detector_cut",
                        "per_unit": "40 B per record"},
           "passing_weight": wsum.cpu().tolist()}
    if cpu:
        from oracle import detector as od
        m = min(n, 4_000_000)
        pp = p[:m].double().cpu().numpy()
        rr = r[:m].double().cpu().numpy()
        t0 = time.perf_counter()
        od.pass_matrix(pp, rr, planes, radius, inner, ecut)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": m / dt, "unit": "records/s", "cores": 1, "kind": "port",
                               "sample": f"first {m} records, {dt:.2f} s (oracle numpy float64)"}
    return out


def bench_meson_beam(n=4_000_000, reps=5, cpu=True):
    """TwoBody_BSMDecay of a meson beam (dark_shower.py:633-640, 836-844): a
    synthetic 3:1 pi0/eta beam of (E, px, py, pz) rows, E log-uniform in
    [0.2, 120] GeV within 20 mrad, as a device record (ParticleBatch.from_arrays),
    m_V = 30 MeV; host-timed generate_dark_showers calls (launches, the
    count readback and the yield sums included), and k_meson_decay's HIP-event
    time: 100 algorithmic bytes per meson record (pid, r0, sid, pf, rf read;
    p, r, parent, meta written), 4 per other record (pid)."""
    import torch

    from petite_amd.batch import ParticleBatch
    from petite_amd.constants import mass_dict
    from petite_amd.dark_shower import DarkShower
    rng = np.random.default_rng(SEED)
    pid = np.where(rng.uniform(size=n) < 0.75, 111, 221).astype(np.int32)
    m = np.where(pid == 111, mass_dict[111], mass_dict[221])
    E = np.exp(rng.uniform(np.log(0.2), np.log(120.0), n))
    E = np.maximum(E, 1.001 * m)
    th, ph = np.abs(rng.normal(0, 0.02, n)), rng.uniform(0, 2 * np.pi, n)
    p = np.sqrt(E * E - m * m)
    beam = np.stack([E, p * np.sin(th) * np.cos(ph), p * np.sin(th) * np.sin(ph), p * np.cos(th)], axis=1)
    ds = DarkShower("100GeV", "graphite", CUTOFF, 0.03, seed=SEED, active_processes=["TwoBody_BSMDecay"])
    rec = ParticleBatch.from_arrays(beam, pid)
    dark = ds.generate_dark_showers(rec)
    ds.device_context().set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_kernel = 0.0
    for _ in range(reps):
        dark = ds.generate_dark_showers(rec)
        ms_kernel += dark.stats["ms_meson"]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    ms_kernel /= reps
    ds.device_context().set_timing(False)
    nv = int(dark.n)
    ach = (100.0 * nv + 4.0 * (n - nv)) / (ms_kernel / 1e3) / 1e9
    out = {"workload": f"{n} mesons (3:1 pi0/eta), E log-uniform [0.2, 120] GeV, m_V = 30 MeV, TwoBody_BSMDecay only",
           "records": n, "vectors": nv, "ms": ms, "records_per_s": n / (ms / 1e3),
           "kernel_ms": ms_kernel,
           "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                        "traffic": None, "kernel": "k_meson_decay",
                        "per_unit": "100 B per meson record, 4 B per other record"},
           "yield": float(dark.yields["TwoBody_BSMDecay"])}
    if cpu:
        from oracle import dark as od
        k = 20_000
        a = rec.to_numpy()
        t0 = time.perf_counter()
        for i in range(k):
            od.produce_meson({"pid": int(a["pid"][i]), "pf": a["pf"][i].astype(np.float64), "rf": a["rf"][i, :3],
                              "weight": float(a["r0"][i, 3]), "sid": np.uint64(i)}, ds.get_mV(), SEED)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k / dt, "unit": "records/s", "cores": 1, "kind": "port",
                               "sample": f"first {k} records, {dt:.2f} s (oracle produce_meson, float64 scalar)"}
    del dark, rec
    return out


def bench_map_training(n_iter=10, n_per=1 << 16, cpu=True):
    """VEGAS map training (vegas_train): the 400 GeV Brem map, missing from the
    shipped tables, put on the 150 cross-section energies of the 400 GeV set
    (lead) from the nearest 100 GeV maps and trained; the accept/reject
    efficiency (mean J f / max J f) above 100 GeV before and after."""
    import torch

    from petite_amd import Shower
    sh = Shower("400GeV", "lead", 0.100, seed=SEED)
    E0 = sh.device_context().tables_grid("Brem")[0]
    # the stand-in's nodes on the 400 GeV energies (nearest copy, no adaptation)
    sh.train_maps(["Brem"], energies="sigma", n_iter=0)
    ctx = sh.device_context()
    E = ctx.tables_grid("Brem")[0]
    hi = E > 100.0
    eff_copy = ctx.sigma_est["Brem"] / ctx.maxf["Brem"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sh.train_maps(["Brem"], n_iter=n_iter, n_per_energy=n_per)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    eff = ctx.sigma_est["Brem"] / ctx.maxf["Brem"]
    samples = len(E) * n_iter * n_per
    out = {"workload": f"400 GeV Brem map (lead): {len(E)} energies to {E[-1]:.0f} GeV (shipped stand-in: "
                       f"{len(E0)} to {E0[-1]:.0f} GeV), {n_iter} adaptations of {n_per} samples per energy",
           "samples": samples, "s_total": dt, "samples_per_s": samples / dt,
           "efficiency_above_100GeV_stand_in": float(np.median(eff_copy[hi])),
           "efficiency_above_100GeV_trained": float(np.median(eff[hi]))}
    if cpu:
        from oracle import vegas as ov
        from oracle.tables import SMTables
        T = SMTables("lead")
        grid = T.grids["Brem"]
        ie = len(grid.E) - 1
        m = 1 << 19
        t0 = time.perf_counter()
        ov.train_accumulate(grid, "Brem", ie, T.event_info("Brem", grid.E[ie]), 7, 0, m)
        t1 = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": m / t1, "unit": "samples/s", "cores": 1, "kind": "port",
                               "sample": f"{m} samples of the accumulation at one energy, {t1:.3f} s (oracle numpy "
                                         "float64; the adaptation itself is host code in both)"}
    return out


# dark workloads
DARK_N, DARK_E, DARK_MAT, DARK_MV = 20_000, 10.0, "graphite", 0.001
YIELD_N, YIELD_MAT = 50_000, "lead"
YIELD_MASSES = np.geomspace(1e-3, 0.1, 32)
# algorithmic bytes: the dark selection reads a 64-byte SM record and the
# emission writes a 48-byte vector record; the yield pass reads pid, E0 and
# the weight (12 bytes) of every record once per batch of 16 masses
B_DARK_SM, B_DARK_V, B_YIELD = 64, 48, 12


def _electrons(n, energies, seed):
    from petite_amd.batch import PRIMARY_DTYPE
    from petite_amd.constants import m_electron
    E = np.asarray(energies, dtype=np.float64)
    prim = np.zeros(n, dtype=PRIMARY_DTYPE)
    prim["E"] = E
    prim["pz"] = np.sqrt(E * E - m_electron ** 2)
    prim["mass"] = m_electron
    prim["weight"] = 1.0
    prim["pid"] = 11
    prim["shower"] = np.arange(n, dtype=np.int32)
    prim["sid"] = np.arange(n, dtype=np.uint64) + (np.uint64(seed) << np.uint64(40))
    return prim


def bench_dark_forward(n=DARK_N, reps=2):
    """20,000 e- at 10 GeV in graphite, m_V = 1 MeV (trained tables of 3 MeV,
    mode='approx'), DarkBrem/DarkAnn/DarkComp: SM record, then the vectors."""
    import torch
    from petite_amd.dark_shower import DarkShower
    ds = DarkShower("100GeV", DARK_MAT, CUTOFF, DARK_MV, mode="approx", seed=SEED,
                    active_processes=["DarkBrem", "DarkAnn", "DarkComp"])
    ctx = ds.device_context()
    ctx.set_timing(True)
    prim = _electrons(n, np.full(n, DARK_E), 3)
    rec = ds.generate_showers(prim)
    print(f"dark: SM record of {rec.n} particles", file=sys.stderr, flush=True)
    dark = ds.generate_dark_showers(rec)  # warm-up
    print(f"dark: warm-up reweight, {dark.n} entries", file=sys.stderr, flush=True)
    del dark
    torch.cuda.synchronize()
    t_sm = t_dk = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        rec = ds.generate_showers(prim)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        dark = ds.generate_dark_showers(rec)
        torch.cuda.synchronize()
        t_sm += t1 - t0
        t_dk += time.perf_counter() - t1
        print(f"dark: rep {t_dk:.3f} s", file=sys.stderr, flush=True)
    st = dark.stats
    n_sm, n_v = rec.n, st["n_vectors"]
    k_ms = st["ms_select"] + st["ms_draw"] + st["ms_emit"]
    alg = B_DARK_SM * n_sm + B_DARK_V * st["n_entries"]
    out = {"workload": f"{n} e- at {DARK_E} GeV, {DARK_MAT}, cutoff {CUTOFF} GeV, m_V={DARK_MV} GeV (approx on the "
                       f"3 MeV tables), DarkBrem/DarkAnn/DarkComp, seed {SEED}",
           "sm_records": n_sm, "entries": st["n_entries"], "vectors": n_v,
           "yield_per_eps2": sum(st["yield"]), "ms_sm_record": t_sm / reps * 1e3, "ms_dark": t_dk / reps * 1e3,
           "vectors_per_s": n_v / (t_dk / reps), "sm_records_reweighted_per_s": n_sm / (t_dk / reps),
           "showers_per_s_sm_plus_dark": n / ((t_sm + t_dk) / reps),
           "kernels_ms": {"select": st["ms_select"], "draw": st["ms_draw"], "emit": st["ms_emit"]},
           "trials_per_draw": st["n_trials"] / max(st["n_draws"], 1),
           "roofline_select_emit": {"bound": "hbm", "achieved": alg / ((st["ms_select"] + st["ms_emit"]) / 1e3) / 1e9,
                                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "per_unit": "64 B per SM record + 48 B per entry"},
           "kernel_ms_total": k_ms}
    return out, ds, rec, dark


def yield_finite_differences(model, rec, masses, h=1e-5):
    """dY/dm_V by finite differences of step h m_V at every mass: central
    between trained masses; at a trained mass (where the interpolation has a
    kink) second-order one-sided on the side the model differentiates -- the
    segment that starts there, or for the heaviest trained mass the one that
    ends there."""
    masses = np.asarray(masses, dtype=np.float64)
    knots = np.asarray(model.masses)
    out = np.empty_like(masses)
    for k, m in enumerate(masses):
        at = np.any(np.abs(m / knots - 1) < 1e-12)
        if not at:
            Yp, _ = model.evaluate(rec, [m * (1 + h)], grad=False)
            Ym, _ = model.evaluate(rec, [m * (1 - h)], grad=False)
            out[k] = (Yp[0] - Ym[0]) / (2 * h * m)
            continue
        sgn = -1.0 if m >= knots[-1] * (1 - 1e-12) else 1.0
        Y = [model.evaluate(rec, [m * (1 + sgn * i * h)], grad=False)[0][0] for i in range(3)]
        out[k] = sgn * (-3 * Y[0] + 4 * Y[1] - Y[2]) / (2 * h * m)
    return out


def bench_yield(reps=5):
    """Fixed record of 50,000 e- showers (1-20 GeV log-uniform, lead), yield
    and d/dm_V at 32 m_V in [1, 100] MeV; d/d(eps^2) is the yield itself."""
    import torch
    from petite_amd.dark_shower import DarkShower
    ds = DarkShower("100GeV", YIELD_MAT, CUTOFF, 0.01, seed=SEED)
    rng = np.random.default_rng(SEED + 1)
    prim = _electrons(YIELD_N, np.exp(rng.uniform(0.0, math.log(20.0), YIELD_N)), 4)
    rec = ds.generate_showers(prim)
    print(f"yield: record of {rec.n} particles", file=sys.stderr, flush=True)
    ds.dark_yield(rec, torch.tensor(YIELD_MASSES[:1]), 1.0)
    model = ds._yield_model
    model.evaluate(rec, YIELD_MASSES)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        Yf, _ = model.evaluate(rec, YIELD_MASSES, grad=False)
    t_f = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        Y, dY = model.evaluate(rec, YIELD_MASSES)
    t_g = (time.perf_counter() - t0) / reps
    fd = yield_finite_differences(model, rec, YIELD_MASSES)
    fd_err = float(np.max(np.abs(fd - dY) / np.maximum(np.abs(dY), 1e-300)))
    inner = YIELD_MASSES
    passes = math.ceil(len(YIELD_MASSES) / 16)
    alg = B_YIELD * rec.n * passes
    out = {"workload": f"{YIELD_N} e- showers, E log-uniform [1,20] GeV, {YIELD_MAT}, cutoff {CUTOFF} GeV; "
                       f"32 m_V in [1,100] MeV; eps^2 enters linearly",
           "sm_records": rec.n, "trained_masses": model.masses,
           "ms_forward": t_f * 1e3, "ms_forward_and_grad": t_g * 1e3, "grad_over_forward": t_g / t_f,
           "scans_per_s": 1.0 / t_g, "grad_evals_per_s": len(YIELD_MASSES) / t_g,
           "fd_max_rel_err": fd_err, "fd_points": len(inner),
           "roofline": {"bound": "hbm", "achieved": alg / t_g / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "per_unit": "12 B per record per 16-mass pass (host-timed, includes launch and readback)"},
           "yield_per_eps2_at_10MeV": float(np.interp(0.01, YIELD_MASSES, Y))}
    return out, ds, rec, model


def cpu_dark_baseline(ds, rec, budget_s=8.0):
    """Oracle produce_bsm_particle over the first records of the dark workload, one core."""
    from oracle.dark import OracleDark, OracleDarkSampler
    from oracle.tables import SMTables
    from petite_amd.constants import mass_dict
    ctx = ds.device_context()
    od = OracleDark(SMTables(DARK_MAT), DARK_MV, mode="approx", with_integrals=False)
    s = OracleDarkSampler(od, ds._dark.W, ds._dark.D, {P: ctx.maxf[P] for P in ("DarkBrem", "DarkAnn", "DarkComp")},
                          max_trials=ctx.desc.max_trials)
    a = {k: rec.tensor(k)[:20000].cpu().numpy() for k in ("pid", "p0", "pf", "rf", "r0", "sid")}
    t0 = time.perf_counter()
    n = nv = 0
    while n < len(a["pid"]) and time.perf_counter() - t0 < budget_s:
        pid = int(a["pid"][n])
        r = {"pid": pid, "mass": mass_dict[pid], "p0": a["p0"][n].astype(np.float64),
             "pf": a["pf"][n].astype(np.float64), "rf": a["rf"][n][:3].astype(np.float64),
             "weight": float(a["r0"][n][3]), "sid": np.uint64(a["sid"][n].view(np.uint64))}
        for P in ("DarkBrem", "DarkAnn", "DarkComp"):
            wg = s.weight(pid, r["p0"][0], P)
            if wg > 0 and s.produce(r, P, wg, SEED) is not None:
                nv += 1
        n += 1
    dt = time.perf_counter() - t0
    return {"value": nv / dt, "unit": "dark vectors/s", "cores": 1, "kind": "port",
            "sm_records_per_s": n / dt,
            "sample": f"first {n} SM records of the dark workload, {dt:.1f} s on one core (oracle float64)"}


def cpu_yield_baseline(rec, model, budget_records=200_000):
    """Oracle float64 yield + gradient over a bounded prefix of the record, one core."""
    from oracle.dark import OracleDark, OracleDarkSampler, yield_and_grad
    from oracle.tables import SMTables
    T = SMTables(YIELD_MAT)
    samplers = [OracleDarkSampler(OracleDark(T, m, with_integrals=False), model.dark[m].W, model.dark[m].D, None)
                for m in model.masses]
    n = min(rec.n, budget_records)
    pid = rec.tensor("pid")[:n].cpu().numpy()
    E0 = rec.tensor("p0")[:n, 0].double().cpu().numpy()
    w = rec.tensor("r0")[:n, 3].double().cpu().numpy()
    procs = ["DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"]
    t0 = time.perf_counter()
    yield_and_grad(samplers, model.masses, procs, pid, E0, w, YIELD_MASSES)
    dt = time.perf_counter() - t0
    scans = (n / rec.n) / dt
    return {"value": scans * len(YIELD_MASSES), "unit": "grad evals/s (full record)", "cores": 1, "kind": "port",
            "sample": f"first {n} of {rec.n} records, 32 masses, {dt:.2f} s (numpy float64), scaled to the record"}

if __name__ == "__main__":
    main()
