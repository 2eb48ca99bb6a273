"""Launch and queue sizes change speed, never results.

Every particle draws from its own Philox stream, so a run under any
PetiteTuning must give the same records, particle for particle (matched by
stream id), bit for bit.  The small settings below also drive the paths the
default sizes rarely reach on small inputs: sweeps limited by their bound
with hand-ons pending (the clamp of new particles to bound - resumed), a
resume buffer that fills up (the lane keeps stepping), single-lane refill
batches, one-wave launches and single-lane draw groups.  The defaults walk
these 48 showers depth-first after the first host check (k_tail); TINY and
WIDE never do, so they hold the walk to the sweeps' records; DEEP walks from
the second sweep with a one-entry stack, so most children go through the
spill lists and several launches, hand-ons of the first sweep included;
BUDGET leaves every subtree past 3 particles per claim to later launches.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle.shower import OracleShower
from oracle.tables import SMTables
from petite_amd import Particle, Shower
from petite_amd.batch import PRIMARY_DTYPE

ME = 0.51099895e-3
MMU = 0.1056583755

# TINY never takes the small-sweep launches (per-process draw streams
# throughout, each 3-sweep group replayed from a captured HIP graph); the defaults on these 48 primaries and WIDE take them after the
# first host check (every process's draws in one k_draw_small launch; WIDE
# hands particles on after 2 substeps there)
TINY = {"batch_lanes": 1, "yield_steps": 3, "transport_waves": 1, "emit_waves": 1, "draw_waves": [1] * 8,
        "min_group_lanes": 1, "sweep_group": 3, "sweep_bound": 64, "resume_cap": 8, "small_sweep": 1, "graph": 1,
        "tail_sweep": 1}
WIDE = {"batch_lanes": 64, "yield_steps": 1 << 30, "transport_waves": 8, "emit_waves": 4,
        "draw_waves": [4, 1, 3, 1, 1, 1, 1, 1], "min_group_lanes": 64, "sweep_group": 1, "small_sweep": 1 << 30,
        "small_yield_steps": 2, "tail_sweep": 1}
DEEP = {"yield_steps": 2, "sweep_group": 1, "small_sweep": 1, "tail_sweep": 1 << 30, "tail_waves": 1, "tail_depth": 1}
# the walk from the first host check with a budget of 3 particles per claim:
# every larger subtree continues in later launches
BUDGET = {"sweep_group": 2, "tail_sweep": 1 << 30, "tail_budget": 3}


def _primaries(n, seed):
    rng = np.random.default_rng(seed)
    prim = np.zeros(n, dtype=PRIMARY_DTYPE)
    pids = rng.choice([11, -11, 22, 13], size=n, p=[0.3, 0.3, 0.3, 0.1])
    E = np.exp(rng.uniform(np.log(0.3), np.log(20.0), n))
    m = np.where(np.abs(pids) == 11, ME, np.where(np.abs(pids) == 13, MMU, 0.0))
    E = np.maximum(E, m + 0.3)
    prim["E"] = E
    prim["pz"] = np.sqrt(E * E - m * m)
    prim["mass"] = m
    prim["weight"] = 1.0
    prim["pid"] = pids
    prim["shower"] = np.arange(n)
    prim["sid"] = np.arange(n)
    return prim


def _by_sid(batch):
    a = batch.to_numpy()
    order = np.argsort(a["sid"].astype(np.uint64), kind="stable")
    return {k: a[k][order] for k in ("sid", "pid", "p0", "r0", "pf", "rf", "meta", "shower")}


@pytest.fixture(scope="module")
def lead():
    return Shower("100GeV", "lead", 0.010, seed=5)


@pytest.fixture(scope="module")
def reference_run(lead):
    prim = _primaries(48, 3)
    ctx = lead.device_context()
    ctx.set_tuning(None)
    batch = lead.generate_showers(prim, seed=21)
    return prim, batch


@pytest.mark.parametrize("name,tuning", [("tiny", TINY), ("wide", WIDE), ("deep", DEEP), ("budget", BUDGET)])
def test_records_do_not_depend_on_tuning(lead, reference_run, name, tuning):
    prim, ref = reference_run
    ctx = lead.device_context()
    got = ctx.set_tuning(tuning)
    for k, v in tuning.items():
        assert got[k] == v, (k, got[k], v)
    try:
        batch = lead.generate_showers(prim, seed=21)
    finally:
        ctx.set_tuning(None)
    assert batch.n == ref.n
    a, b = _by_sid(ref), _by_sid(batch)
    for k in a:
        assert np.array_equal(a[k], b[k]), (name, k)
    ea, eb = ref.energy.cpu().numpy(), batch.energy.cpu().numpy()
    assert np.allclose(ea, eb, rtol=1e-12, atol=1e-12)
    if name == "tiny":
        # the small bound splits the run into many more sweeps
        assert batch.stats["n_iterations"] > ref.stats["n_iterations"]
    if name in ("tiny", "wide"):
        assert batch.stats["launches_tail"] == 0
    else:
        # one-entry stacks or a 3-particle budget: the spilled children take
        # further launches
        assert batch.stats["launches_tail"] > 1
    assert ref.stats["launches_tail"] >= 1
    assert batch.stats["n_steps"] == ref.stats["n_steps"]
    assert batch.stats["n_trials"] == ref.stats["n_trials"]


def test_tiny_tuning_matches_oracle(lead):
    """The oracle's particle-matched comparison under the smallest sizes."""
    sh = Shower("100GeV", "graphite", 0.010, seed=11)
    ctx = sh.device_context()
    ctx.set_tuning(TINY)
    energies = [2.0, 3.0, 1.0, 5.0]
    pids = [11, 22, -11, 11]
    prims = []
    for E, p in zip(energies, pids):
        m = ME if abs(p) == 11 else 0.0
        prims.append(Particle([E, 0.0, 0.0, math.sqrt(E * E - m * m)], [0, 0, 0], {"PID": p}))
    try:
        batch = sh.generate_showers(prims, seed=11)
    finally:
        ctx.set_tuning(None)
    maxf = {(P, ie): v for P, arr in ctx.maxf.items() for ie, v in enumerate(arr)}
    osh = OracleShower(SMTables("graphite"), 0.010, seed=11, maxf=maxf)
    a = batch.to_numpy()
    matched = total = 0
    for s, (E, p) in enumerate(zip(energies, pids)):
        m = ME if abs(p) == 11 else 0.0
        ref = osh.generate_shower(p, [E, 0, 0, math.sqrt(E * E - m * m)], sid=s, shower=s)
        sel = a["shower"] == s
        gsid = {int(v): i for i, v in zip(np.nonzero(sel)[0], a["sid"][sel])}
        for q in ref:
            i = gsid.get(int(q.sid))
            if i is not None and q.pid == a["pid"][i] and np.allclose(a["p0"][i], q.p0, rtol=1e-4, atol=1e-6):
                matched += 1
        total += len(ref)
    assert matched >= 0.9 * total, (matched, total)


def test_tuning_validation(lead):
    from petite_amd._lib import PetiteError
    ctx = lead.device_context()
    for bad in ({"batch_lanes": 65}, {"min_group_lanes": 3}, {"min_group_lanes": 128}, {"sweep_bound": 32}, {"sweep_group": 65},
                {"sweep_bound": 128, "resume_cap": 256}, {"transport_waves": 65}, {"tail_depth": 129},
                {"tail_waves": 65}):
        with pytest.raises(PetiteError):
            ctx.set_tuning(bad)
    d = ctx.set_tuning(None)
    assert d["batch_lanes"] == 44 and d["yield_steps"] == 40 and d["sweep_bound"] == 1 << 23
    assert d["resume_cap"] == 1 << 20 and d["draw_waves"][0] == 16
    assert d["tail_sweep"] == 16384 and d["tail_waves"] == 8 and d["tail_depth"] == 128 and d["tail_budget"] == 16
    assert ctx.set_tuning({"tail_budget": 1 << 30})["tail_budget"] == 1 << 30  # no limit
    ctx.set_tuning(None)
