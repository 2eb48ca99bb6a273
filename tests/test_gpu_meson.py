"""GPU TwoBody_BSMDecay of meson records (k_meson_decay) against the oracle
(produce_meson; dark_shower.py:633-640, 836-844, particle.py:187-254): one
vector per pi0 / eta / eta' heavier than m_V, none for the others, from the
same store values and dark stream, to float32 precision."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import dark as od
from petite_amd import Particle
from petite_amd.constants import mass_dict
from petite_amd.dark_shower import DarkShower

SEED = 31


def beam(n, rng):
    out = []
    for i in range(n):
        pid = (111, 221, 331, 11)[i % 4]
        m = mass_dict[pid]
        E = float(np.exp(rng.uniform(np.log(1.0), np.log(100.0))))
        th = abs(rng.normal(0, 0.05))
        ph = rng.uniform(0, 2 * np.pi)
        p = math.sqrt(E * E - m * m)
        out.append(Particle([E, p * math.sin(th) * math.cos(ph), p * math.sin(th) * math.sin(ph), p * math.cos(th)],
                            rng.normal(0, 0.01, 3), {"PID": pid, "weight": float(rng.uniform(0.5, 2.0)), "mass": m}))
    return out


@pytest.mark.parametrize("mV", [0.05, 0.3])
def test_meson_vectors_match_oracle(mV):
    ds = DarkShower("100GeV", "graphite", 0.010, mV, seed=SEED, active_processes=["TwoBody_BSMDecay"])
    mV = ds.get_mV()  # mode "exact" takes the closest lesser trained mass (dark_shower.py:167-170)
    parts = beam(400, np.random.default_rng(4))
    dark = ds.generate_dark_showers(parts)
    store = ds._store_from_particles(parts)
    pf = store["pf"].cpu().numpy().astype(np.float64)
    rf = store["rf"].cpu().numpy().astype(np.float64)
    w = store["r0"].cpu().numpy()[:, 3].astype(np.float64)
    par = dark.parent.cpu().numpy()
    proc = dark.process.cpu().numpy()
    pv = dark.p.cpu().numpy().astype(np.float64)
    rv = dark.tensor("r").cpu().numpy().astype(np.float64)
    assert np.all(proc == 12) and np.all(dark.valid.cpu().numpy())
    got = {int(par[e]): e for e in range(dark.n)}
    n_exp = 0
    for i, q in enumerate(parts):
        pid = q.get_ids()["PID"]
        if pid == 11:
            assert i not in got
            continue
        v = od.produce_meson({"pid": pid, "pf": pf[i], "rf": rf[i, :3], "weight": w[i], "sid": np.uint64(i)}, mV, SEED)
        if v is None:
            assert i not in got and mV >= od.MESON_MASS[pid]
            continue
        n_exp += 1
        e = got[i]
        np.testing.assert_allclose(pv[e], v["p"], rtol=2e-6, atol=2e-6 * v["p"][0])
        np.testing.assert_allclose(rv[e, :3], v["r"], rtol=1e-7, atol=1e-9)
        np.testing.assert_allclose(rv[e, 3], v["weight"], rtol=1e-6)
        np.testing.assert_allclose(rv[e, 3], ds.GetBSMWeights(q, "TwoBody_BSMDecay") * q.get_ids()["weight"], rtol=1e-6)
    assert dark.n == n_exp > 50
    np.testing.assert_allclose(dark.yields["TwoBody_BSMDecay"], rv[:, 3].sum(), rtol=1e-6)


def test_drop_in_list_ids_and_other_processes():
    ds = DarkShower("100GeV", "graphite", 0.010, 0.05, seed=SEED)  # every dark process, TwoBody_BSMDecay included
    parts = beam(40, np.random.default_rng(5))
    for i, q in enumerate(parts):
        q.get_ids()["ID"] = 10 + i
    sm, vecs = ds.generate_dark_shower(ExDir=parts)
    meson = [v for v in vecs if v.get_ids()["generation_process"] == "TwoBody_BSMDecay"]
    assert len(meson) == sum(1 for q in parts if q.get_ids()["PID"] != 11)
    for v in meson:
        d = v.get_ids()
        assert d["ID"] == 2 * d["parent_ID"] + 1 and d["parent_PID"] in (111, 221, 331)
    # the electrons' dark-brem entries are still there
    assert any(v.get_ids()["generation_process"] == "DarkBrem" for v in vecs)
