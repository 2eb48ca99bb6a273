"""Edge cases through the product API: an empty primary list and an empty
record; primaries below the cutoff (their showers are the primary alone,
shower.py:630-632); primaries at and above the tables' last energy, where
n sigma is 0 and the mean free path is get_mfp's 1e12 m (shower.py:386-387,
kept): a charged lepton's first substep then takes all its energy in dE/dx
and it stops far away without interacting, a photon flies off; a dark reweight of records that no active process applies to; and mesons all
lighter than m_V (no TwoBody_BSMDecay vector, dark_shower.py:636-637)."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from petite_amd import Particle, Shower
from petite_amd.batch import PRIMARY_DTYPE, ParticleBatch
from petite_amd.constants import mass_dict
from petite_amd.dark_shower import DarkShower

ME = mass_dict[11]


def _prim(pids, energies):
    n = len(pids)
    prim = np.zeros(n, dtype=PRIMARY_DTYPE)
    for i, (p, E) in enumerate(zip(pids, energies)):
        m = mass_dict[p]
        prim[i]["E"], prim[i]["pz"], prim[i]["mass"] = E, math.sqrt(E * E - m * m), m
        prim[i]["weight"], prim[i]["pid"], prim[i]["shower"], prim[i]["sid"] = 1.0, p, i, i
    return prim


def test_empty_inputs():
    sh = Shower("100GeV", "graphite", 0.010, seed=1)
    b = sh.generate_showers([])
    assert b.n == 0 and b.to_particles() == [] and b.energy.shape == (0, 4)
    b = sh.generate_showers(np.zeros(0, dtype=PRIMARY_DTYPE))
    assert b.n == 0
    ds = DarkShower("100GeV", "graphite", 0.010, 0.03, seed=1)
    d = ds.generate_dark_showers(ParticleBatch.from_arrays(np.zeros((0, 4)), 111))
    assert d.n == 0 and all(v == 0.0 for v in d.yields.values())
    d = ds.generate_dark_showers(b)
    assert d.n == 0


def test_primaries_below_the_cutoff_end_at_once():
    sh = Shower("100GeV", "lead", 0.100, seed=2)
    b = sh.generate_showers(_prim([11, -11, 22], [0.05, 0.08, 0.099]))
    a = b.to_numpy()
    assert b.n == 3 and np.all(a["parent"] == -1) and np.all(a["rf"][:, 3] == 0)
    np.testing.assert_array_equal(a["pf"], a["p0"])
    # the list API returns the primary alone, ended (shower.py:630-632)
    q = Particle([0.05, 0.0, 0.0, math.sqrt(0.05 ** 2 - ME ** 2)], [0, 0, 0], {"PID": 11})
    out = sh.generate_shower(q)
    assert len(out) == 1 and out[0].get_ended()


@pytest.mark.parametrize("E", [100.0, 150.0])
def test_at_and_above_the_last_table_energy(E):
    sh = Shower("100GeV", "lead", 0.100, seed=3)
    prim = _prim([11, -11, 22, 13], [E, E, E, E])
    b = sh.generate_showers(prim)
    e = b.energy.cpu().numpy()
    bal = prim["E"] + e[:, 3] * ME - e[:, 0] - e[:, 1] - e[:, 2]
    assert np.all(np.abs(bal) < 1e-4 * E), bal
    a = b.to_numpy()
    assert np.all(np.isfinite(a["pf"])) and np.all(np.isfinite(a["rf"]))
    if E > 100.0:
        # no table above 100 GeV: nothing interacts (the photon flies off, the
        # leptons stop in their first substep)
        g = a["pid"] == 22
        ph = np.nonzero(g & (a["parent"] == -1))[0][0]
        assert np.sum(a["parent"] == ph) == 0
        assert b.n == 4 and np.all(a["parent"] == -1)
        lep = ~g
        np.testing.assert_allclose(a["pf"][lep, 0], [mass_dict[int(p)] for p in a["pid"][lep]], rtol=1e-6)


def test_dark_reweight_without_eligible_records():
    ds = DarkShower("100GeV", "graphite", 0.010, 0.03, seed=4)
    rec = ParticleBatch.from_arrays(np.array([[3.0, 0.0, 0.0, 3.0]] * 5), 14)  # neutrinos
    d = ds.generate_dark_showers(rec)
    assert d.n == 0


def test_mesons_lighter_than_mV_give_no_vector():
    ds = DarkShower("100GeV", "graphite", 0.010, 0.3, seed=5, active_processes=["TwoBody_BSMDecay"])
    m = mass_dict[111]
    rec = ParticleBatch.from_arrays(np.array([[2.0, 0.0, 0.0, math.sqrt(4.0 - m * m)]] * 100), 111)
    d = ds.generate_dark_showers(rec)
    assert d.n == 0 and d.yields["TwoBody_BSMDecay"] == 0.0
