"""The reference's lower-level public methods on the drop-in classes, each run
on the GPU and checked against the oracle on the same Philox stream:

* Shower.sample_scattering (shower.py:467-507): children of a moving,
  displaced particle for every SM process, four-vectors to 1e-6 of the
  parent energy, ids as the reference sets them;
* Shower.propagate_particle (509-601), as generate_shower calls it: final
  momentum and position of photons, e+-, mu+- to float32 rounding;
* DarkShower.draw_dark_sample (dark_shower.py:649-704): the accepted point and
  candidate count;
* DarkShower.draw_pe_sample (706-719): equal to the oracle's sampler on the
  same stream, and distributed as the 1s momentum density (KS);
* DarkShower.produce_bsm_particle (721-804): the vector and its weight.
"""
import math

import numpy as np
import pytest
from scipy import stats
from scipy.integrate import quad

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import dark as odark
from oracle import philox
from oracle import vegas as ovegas
from oracle.constants import mass_dict as omass
from oracle.shower import OParticle, OracleShower
from oracle.tables import SMTables
from petite_amd import Particle, Shower
from petite_amd.dark_shower import DarkShower

SEED = 31


def _moving(pid, E, rng):
    m = omass[pid]
    p = math.sqrt(E * E - m * m)
    th, ph = rng.uniform(0, 0.3), rng.uniform(0, 2 * math.pi)
    d = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    q = Particle([E, 0.0, 0.0, p], [0.0, 0.0, 0.0], {"PID": pid, "mass": m, "ID": 5, "generation_number": 3,
                                                    "weight": 0.75})
    q.set_pf(np.concatenate([[E], p * d]))
    q.set_rf(rng.uniform(-0.1, 0.1, 3))
    return q


@pytest.fixture(scope="module")
def graphite():
    sh = Shower("100GeV", "graphite", 0.010, seed=SEED)
    ctx = sh.device_context()
    maxf = {(P, ie): v for P, arr in ctx.maxf.items() for ie, v in enumerate(arr)}
    return sh, OracleShower(SMTables("graphite"), 0.010, seed=SEED, maxf=maxf)


CASES = [(11, "Brem"), (11, "Moller"), (-11, "Brem"), (-11, "Ann"), (-11, "Bhabha"), (22, "PairProd"), (22, "Comp"),
         (13, "MuonE"), (13, "MuonBrem")]


@pytest.mark.parametrize("pid,process", CASES)
def test_sample_scattering_matches_oracle(graphite, pid, process):
    sh, osh = graphite
    rng = np.random.default_rng(abs(pid) * 100 + len(process))
    for j in range(6):
        E = float(np.exp(rng.uniform(np.log(0.5), np.log(50.0))))
        q = _moving(pid, E, rng)
        sid = 1000 + j
        kids = sh.sample_scattering(q, process, sid=sid)
        op = OParticle(pid, omass[pid], q.get_p0(), q.get_r0(), np.uint64(sid))
        op.pf, op.rf = np.asarray(q.get_pf(), dtype=np.float64), np.asarray(q.get_rf(), dtype=np.float64)
        ref = osh.sample_scattering(op, process)
        assert len(kids) == 2
        for k in range(2):
            ids = kids[k].get_ids()
            assert ids["PID"] == ref[k][0]
            assert np.max(np.abs(np.asarray(kids[k].get_p0()) - ref[k][1])) <= 1e-6 * E + 1e-9, (process, j, k)
            assert ids["ID"] == 2 * 5 + k and ids["parent_ID"] == 5 and ids["generation_number"] == 4
            assert ids["generation_process"] == process and ids["weight"] == 0.75 and ids["parent_PID"] == pid
            assert ids["mass"] == omass[ids["PID"]]
            assert np.allclose(kids[k].get_r0(), q.get_rf())


def test_sample_scattering_below_threshold_is_none(graphite):
    sh, _ = graphite
    q = Particle([0.005, 0.0, 0.0, math.sqrt(0.005 ** 2 - omass[11] ** 2)], [0, 0, 0], {"PID": 11})
    assert sh.sample_scattering(q, "Brem") is None


@pytest.mark.parametrize("pid", [22, 11, -11, 13])
def test_propagate_particle_matches_oracle(graphite, pid):
    sh, osh = graphite
    rng = np.random.default_rng(7 + abs(pid))
    for j in range(8):
        E = float(np.exp(rng.uniform(np.log(0.3), np.log(30.0))))
        m = omass[pid]
        q = Particle([E, 0.0, 0.0, math.sqrt(E * E - m * m)], [0.0, 0.0, 0.0], {"PID": pid, "mass": m})
        sid = 5000 + j
        if pid == 22:
            out = sh.propagate_particle(q, sid=sid)
        else:
            out = sh.propagate_particle(q, Losses=sh._tables.dEdx, MS=True, sid=sid)
        assert out.get_ended()
        op = OParticle(pid, m, [E, 0.0, 0.0, math.sqrt(E * E - m * m)], [0.0, 0.0, 0.0], np.uint64(sid))
        osh.propagate(op)
        assert np.max(np.abs(np.asarray(out.get_pf()) - op.pf)) <= 2e-6 * E, (pid, j, out.get_pf(), op.pf)
        L = float(np.linalg.norm(op.rf)) + 1e-6
        assert np.max(np.abs(np.asarray(out.get_rf()) - op.rf)) <= 2e-6 * L, (pid, j, out.get_rf(), op.rf)
    with pytest.raises(NotImplementedError):
        q = Particle([1.0, 0.0, 0.0, 1.0], [0, 0, 0], {"PID": 22})
        sh.propagate_particle(q, MS=True)


@pytest.fixture(scope="module")
def dark():
    ds = DarkShower("100GeV", "graphite", 0.010, 0.03, seed=SEED)
    ctx = ds.device_context()
    od = odark.OracleDark(SMTables("graphite"), 0.03)
    return ds, ctx, od


@pytest.mark.parametrize("process,E", [("DarkBrem", 2.0), ("DarkBrem", 30.0), ("DarkComp", 5.0), ("DarkAnn", 3.0),
                                       ("DarkMuonBrem", 20.0)])
def test_draw_dark_sample_matches_oracle(dark, process, E):
    ds, ctx, od = dark
    grid = od._grids[process]
    for j in range(5):
        sid = 7000 + j
        x = ds.draw_dark_sample(E, process=process, VB=True, sid=sid)
        ie = grid.lu_key(E)
        T = od.T
        ev = T.event_info(process, E, od.mV)
        ev["Zeff"] = od.Zeff
        xo, to, _ = ovegas.draw(grid, process, E, ev, ctx.maxf[process][ie], SEED, np.uint64(sid))
        assert int(x[-1]) == to
        assert np.allclose(x[:-1], xo, rtol=2e-6, atol=2e-6)


def test_draw_pe_sample_matches_oracle_and_density(dark):
    ds, _, od = dark
    vals = np.array([ds.draw_pe_sample(sid=s) for s in range(40)])
    for s in range(40):
        assert abs(vals[s] - odark.draw_pe(SEED, np.uint64(s), ds.Zeff)) <= 1e-15
    from petite_amd.device import draw_pe
    many = draw_pe(np.arange(20000, dtype=np.uint64), SEED, ds.Zeff).cpu().numpy()
    assert np.all(many > 0)
    lam = odark.alpha_em * ds.Zeff * odark.m_electron
    dens = lambda p: odark.electron_wave_function(p, ds.Zeff)
    tot = quad(dens, 0, 1e-3, points=[lam], limit=200)[0]
    cdf = lambda p: np.array([quad(dens, 0, v, points=[lam] if v > lam else None, limit=200)[0] for v in np.atleast_1d(p)]) / tot
    assert stats.kstest(many[:3000], cdf).pvalue > 1e-3


@pytest.mark.parametrize("pid,process", [(11, "DarkBrem"), (-11, "DarkBrem"), (-11, "DarkAnn"), (22, "DarkComp"),
                                         (13, "DarkMuonBrem")])
def test_produce_bsm_particle_matches_oracle(dark, pid, process):
    ds, ctx, od = dark
    W = ds._dark.W
    D = ds._dark.D
    sampler = odark.OracleDarkSampler(od, W, D, {P: ctx.maxf[P] for P in odark.DARK_CODE},
                                      max_trials=ctx.desc.max_trials)
    rng = np.random.default_rng(3)
    done = 0
    for j in range(12):
        E = float(np.exp(rng.uniform(np.log(1.0), np.log(40.0))))
        q = _moving(pid, E, rng)
        q.set_p0(q.get_pf())
        wg = ds.GetBSMWeights(q, process)
        if not wg > 0:
            continue
        sid = 9000 + j
        v = ds.produce_bsm_particle(q, process, sid=sid)
        rec = {"pid": pid, "mass": omass[pid], "p0": np.asarray(q.get_p0(), dtype=np.float32).astype(np.float64),
               "pf": np.asarray(q.get_pf(), dtype=np.float32).astype(np.float64),
               "rf": np.asarray(q.get_rf(), dtype=np.float32).astype(np.float64),
               "weight": float(np.float32(0.75)), "sid": np.uint64(sid)}
        ref = sampler.produce(rec, process, sampler.weight(pid, rec["p0"][0], process), SEED)
        assert (v is None) == (ref is None), (process, j)
        if v is None:
            continue
        assert np.max(np.abs(np.asarray(v.get_p0()) - ref["p"])) <= 5e-5 * ref["p"][0] + 1e-7, (process, j)
        assert abs(v.get_ids()["weight"] - ref["weight"]) <= 1e-5 * abs(ref["weight"])
        assert v.get_ids()["PID"] == 4900022
        done += 1
    assert done >= 3
