"""The full-size shower workload (100,000 primaries e-/e+/gamma, E
log-uniform in [0.1, 100] GeV within 1 mrad, lead, cutoff 10 MeV: about 2.2e8
records) checked through properties that do not depend on the size, on the
device: finite and on-shell records; parent links, generations, shower ids
and child species consistent with each vertex's process (shower.py:87-96,
486-507); kept children above the cutoff (shower.py:703-705); energy
conservation at every vertex whose two children were kept (the atomic
electron's mass added for Compton, Moller, Bhabha, annihilation and mu e);
every shower's float64 energy budget closed; unique stream ids; and the
record's step, particle and dropped-child counts equal to the run
statistics."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from petite_amd import Shower
from petite_amd.batch import PRIMARY_DTYPE
from petite_amd.constants import mass_dict, process_code

N, CUT, ME = 100_000, 0.010, mass_dict[11]


def _primaries():
    rng = np.random.default_rng(1234)
    prim = np.zeros(N, dtype=PRIMARY_DTYPE)
    pid = rng.choice(np.array([11, -11, 22], dtype=np.int32), N)
    E = np.exp(rng.uniform(np.log(0.1), np.log(100.0), N))
    th, ph = rng.uniform(0, 1e-3, N), rng.uniform(0, 2 * np.pi, N)
    m = np.where(pid == 22, 0.0, ME)
    p = np.sqrt(E * E - m * m)
    prim["E"], prim["px"], prim["py"], prim["pz"] = E, p * np.sin(th) * np.cos(ph), p * np.sin(th) * np.sin(ph), \
        p * np.cos(th)
    prim["mass"], prim["weight"], prim["pid"] = m, 1.0, pid
    prim["shower"] = np.arange(N, dtype=np.int32)
    prim["sid"] = np.arange(N, dtype=np.uint64)
    return prim


@pytest.fixture(scope="module")
def run():
    prim = _primaries()
    sh = Shower("100GeV", "lead", CUT, seed=1234)
    b = sh.generate_showers(prim, seed=1234)
    return b, prim


def test_records_finite_and_on_shell(run):
    b, _ = run
    assert b.n > 1e8
    for k in ("p0", "pf", "r0", "rf"):
        assert bool(torch.isfinite(b.tensor(k)).all()), k
    pid = b.pid
    m = torch.zeros_like(b.p0[:, 0], dtype=torch.float64)
    m[pid.abs() == 11] = ME
    m[pid.abs() == 13] = mass_dict[13]
    for k in ("p0", "pf"):
        p = b.tensor(k).double()
        m2 = p[:, 0] ** 2 - (p[:, 1:] ** 2).sum(1)
        assert bool(((m2 - m * m).abs() <= 4e-6 * p[:, 0] ** 2 + 1e-12).all()), k
        assert bool((p[:, 0] >= m * (1 - 1e-6)).all()), k


def test_links_generations_and_species(run):
    b, _ = run
    n = b.n
    par = b.parent.long()
    idx = torch.arange(n, device=par.device)
    child = par >= 0
    assert int((~child).sum()) == N
    pc, ic = par[child], idx[child]
    assert bool((pc < ic).all())  # a parent is stored before its children
    gen, proc, shw = b.generation, b.process, b.shower
    assert bool((gen[ic] == gen[pc] + 1).all())
    assert bool((shw[ic] == shw[pc]).all())
    assert bool((proc[~child] == 0).all()) and bool((proc[child] >= 1).all())
    # kept children are above the cutoff: decided in float64, stored in float32
    # elementwise on the whole column: a row gather of the (n, 4) store past 2^31
    # bytes comes back wrong from torch on this stack
    low = child & (b.p0[:, 0].double() <= CUT * (1 - 1e-6))
    assert not bool(low.any()), int(low.sum())
    # child species per vertex (process_PIDS, 0 = the parent's PID)
    pid, ppid, k = b.pid[ic], b.pid[pc], b.child_index[ic]
    code = proc[ic] - 1
    want = {"PairProd": (-11, 11), "Brem": (0, 22), "Comp": (11, 22), "Ann": (22, 22), "Moller": (0, 11),
            "Bhabha": (0, 11), "MuonE": (0, 11), "MuonBrem": (0, 22)}
    for P, (a, c) in want.items():
        sel = code == process_code[P]
        for kk, w in ((0, a), (1, c)):
            s = sel & (k == kk)
            expect = ppid[s] if w == 0 else torch.full_like(pid[s], w)
            assert bool((pid[s] == expect).all()), (P, kk)


def test_vertex_energy_conservation(run):
    b, _ = run
    par = b.parent.long()
    k = b.child_index
    # the two children of a vertex are stored next to each other (k = 0, then 1)
    first = torch.nonzero((par[:-1] >= 0) & (par[:-1] == par[1:]) & (k[:-1] == 0) & (k[1:] == 1)).squeeze(1)
    assert first.numel() > 1e7
    pa = par[first]
    E0 = b.p0[:, 0].contiguous().double()  # 1-D columns: no row gathers of the (n, 4) store
    Ein = b.pf[:, 0].contiguous().double()[pa]
    code = b.process[first].long() - 1
    atomic = torch.zeros_like(Ein)
    for P in ("Comp", "Ann", "Moller", "Bhabha", "MuonE"):
        atomic[code == process_code[P]] = ME
    Eout = E0[first] + E0[first + 1]
    rel = (Eout - Ein - atomic).abs() / Ein
    assert float(rel.max()) < 2e-6, float(rel.max())


def test_energy_budget_and_counts(run):
    b, prim = run
    e = b.energy.cpu().numpy()
    bal = prim["E"] + e[:, 3] * ME - e[:, 0] - e[:, 1] - e[:, 2]
    assert np.all(np.abs(bal) < 1e-4 * np.maximum(prim["E"], 1.0)), np.abs(bal).max()
    st = b.stats
    assert st["n_particles"] == b.n and st["n_no_sample"] == 0
    # every vertex makes two children, kept or dropped
    assert 2 * st["n_draws"] == (b.n - N) + st["n_dropped"]
    assert int(b.rf[:, 3].double().sum()) == st["n_steps"]
    s = torch.sort(b.sid).values
    assert bool((s[1:] != s[:-1]).all())  # stream ids are unique
