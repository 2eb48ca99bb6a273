"""Launch and queue sizes of the batched shower run, tuned on the live device.

The reference's driver is a sequential Python FIFO (src/PETITE/shower.py:
634-706) and has nothing to tune.  Here petite_run_showers has launch and
queue sizes (PetiteTuning: refill batch, hand-on substeps, waves per CU of
each kernel, the draw group minimum, the sweep group and bounds) that change
only its speed: every particle draws from its own Philox stream, so any
setting gives the same records.  ``autotune`` searches them by coordinate
descent on a caller's workload (each candidate timed over whole runs, the
best of `repeats`), and ``save`` writes the best point to a per-device
table that DeviceContext applies when it is created on a device of the same
kind (``load``).

Table (JSON, petite_amd/data/tuning.json by default): {device key: {"point":
{field: value}, "ms": best run time, "default_ms": time at the built-in
defaults, "workload": description}}.  The device key is the GPU's
architecture name and compute-unit count ("gfx950:256").
"""
import json
import os
import time

import numpy as np

from . import _lib

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning.json")

# candidate values per field; draw_waves entries are tuned per SM process code
SPACE = {
    "batch_lanes": [36, 40, 44, 48, 52],
    "yield_steps": [24, 32, 40, 48, 64],
    "transport_waves": [12, 16, 20, 24, 28],
    "emit_waves": [16, 24, 32, 40],
    "draw_waves.0": [8, 12, 16, 20],
    "draw_waves.2": [4, 6, 8, 12],
    "draw_waves.1": [1, 2, 4, 8],
    "draw_waves.3": [1, 2, 4],
    "draw_waves.4": [1, 2, 4],
    "draw_waves.5": [1, 2, 4],
    "min_group_lanes": [2, 4, 8],
    "group_requests": [1, 2, 4],
    "sweep_group": [8, 16, 32],
    "small_sweep": [1, 4096, 16384, 65536, 262144],
    "small_yield_steps": [24, 32, 40, 160, 1 << 30],
    "graph": [1, 2],
    "tail_sweep": [1, 1024, 4096, 16384, 65536],
    "tail_waves": [4, 8, 16],
    "tail_budget": [8, 16, 64, 1 << 30],
}
FIELDS = {f for f, _ in _lib.Tuning._fields_ if f != "pad"}


def device_key():
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(p, "gcnArchName", "") or p.name
    return f"{arch.split(':')[0]}:{p.multi_processor_count}"


def _get(point, name):
    if "." in name:
        f, i = name.split(".")
        return point[f][int(i)]
    return point[name]


def _set(point, name, value):
    q = {k: (list(v) if isinstance(v, list) else v) for k, v in point.items()}
    if "." in name:
        f, i = name.split(".")
        q[f][int(i)] = value
    else:
        q[name] = value
    return q


def validate(point):
    """The point as a PetiteTuning dict (unknown fields raise)."""
    bad = set(point) - FIELDS
    if bad:
        raise ValueError(f"unknown tuning fields {sorted(bad)}")
    return _lib.Tuning.from_dict(point).as_dict()


def load(path=TABLE, key=None):
    """The tuned point for this device (or `key`), or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    ent = table.get(key if key is not None else device_key())
    return None if ent is None else validate(ent["point"])


def save(point, ms, default_ms, workload, path=TABLE, key=None):
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    table[key if key is not None else device_key()] = {"point": validate(point), "ms": float(ms),
                                                       "default_ms": float(default_ms), "workload": str(workload)}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")


def time_point(run, ctx, point, repeats):
    import torch
    ctx.set_tuning(point)
    best = np.inf
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def autotune(shower, primaries, seed=1234, space=None, repeats=2, rounds=2, log=print):
    """Coordinate descent over `space` (default SPACE) from the built-in
    defaults on the workload shower.generate_showers(primaries, seed).
    Returns (best point, best ms, default ms, history of (point, ms))."""
    space = SPACE if space is None else space
    ctx = shower.device_context()
    run = lambda: shower.generate_showers(primaries, seed=seed)
    run()  # warm up (tables, allocations)
    ctx.set_tuning(None)
    point = ctx.get_tuning()
    default_ms = time_point(run, ctx, point, repeats)
    best_ms = default_ms
    history = [(point, default_ms)]
    log(f"default {default_ms:.2f} ms")
    for r in range(rounds):
        moved = False
        for name, values in space.items():
            for v in values:
                if v == _get(point, name):
                    continue
                cand = _set(point, name, v)
                ms = time_point(run, ctx, cand, repeats)
                history.append((cand, ms))
                if ms < best_ms * 0.995:
                    log(f"round {r}: {name} {_get(point, name)} -> {v}: {best_ms:.2f} -> {ms:.2f} ms")
                    point, best_ms, moved = cand, ms, True
        if not moved:
            break
    ctx.set_tuning(point)
    return point, best_ms, default_ms, history


def main():
    """Tune on the batched workload (petite_amd.workloads.batch_primaries,
    lead, cutoff 10 MeV, 100 GeV tables) and write the device's entry."""
    import argparse
    from .shower import Shower
    from .workloads import batch_primaries
    ap = argparse.ArgumentParser()
    ap.add_argument("--primaries", type=int, default=100_000)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args()
    prim = batch_primaries(a.primaries)
    sh = Shower("100GeV", "lead", 0.010, seed=1234)
    point, ms, dms, hist = autotune(sh, prim, repeats=a.repeats, rounds=a.rounds,
                                    log=lambda m: print(m, flush=True))
    save(point, ms, dms, f"{a.primaries} primaries e-/e+/gamma, E log-uniform [0.1, 100] GeV, lead, cutoff 0.01 GeV",
         path=a.out)
    print(json.dumps({"point": point, "ms": ms, "default_ms": dms, "evaluations": len(hist)}), flush=True)


if __name__ == "__main__":
    main()
