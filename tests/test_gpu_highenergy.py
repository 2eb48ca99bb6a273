"""High-energy batch on the 400 GeV table set, six primaries: 400 GeV muons
and electrons in lead, cutoff 100 MeV, m_V = 10 MeV.  The
400 GeV set lacks the Brem/PairProd/MuonBrem/DarkBrem/DarkMuonBrem maps, so
the 100 GeV maps stand in (their variables are dimensionless).  Checked on the
device record: per-shower float64 energy budget, muon processes present,
dark entries of every family, and the yield's gradient against central
differences."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from petite_amd.batch import PRIMARY_DTYPE
from petite_amd.constants import mass_dict
from petite_amd.dark_shower import DarkShower


@pytest.fixture(scope="module")
def run():
    ds = DarkShower("400GeV", "lead", 0.100, 0.01, seed=3)
    pids = [13, -13, 13, 11, 11, -11]
    prim = np.zeros(len(pids), dtype=PRIMARY_DTYPE)
    for i, p in enumerate(pids):
        m = mass_dict[p]
        prim[i]["E"] = 400.0
        prim[i]["pz"] = math.sqrt(400.0 ** 2 - m * m)
        prim[i]["mass"] = m
        prim[i]["weight"] = 1.0
        prim[i]["pid"] = p
        prim[i]["shower"] = i
        prim[i]["sid"] = 100 + i
    batch = ds.generate_showers(prim)
    return ds, batch, prim


def test_energy_budget(run):
    ds, batch, prim = run
    e = batch.energy.cpu().numpy()
    m_e = mass_dict[11]
    for s in range(len(prim)):
        dep, fin, drop, natom = e[s]
        bal = prim[s]["E"] + natom * m_e - dep - fin - drop
        assert abs(bal) < 1e-4 * prim[s]["E"], (s, bal)


def test_muon_processes_and_dark_entries(run):
    ds, batch, prim = run
    a = batch.to_numpy()
    proc = (a["meta"] >> 16) & 0xFF
    assert np.any(proc == 7) and np.any(proc == 8)  # MuonE (6+1), MuonBrem (7+1)
    dark = ds.generate_dark_showers(batch)
    codes = set(int(c) for c in dark.process[dark.valid].cpu().numpy())
    assert {8, 9, 10, 11} <= codes, codes
    y = dict(dark.yields)
    assert y.pop("TwoBody_BSMDecay", 0.0) == 0.0  # no meson records in an e/mu/gamma shower
    assert all(v > 0 for v in y.values()), y


def test_yield_gradient(run):
    ds, batch, _ = run
    mV = torch.tensor([0.004, 0.02, 0.05, 0.2], dtype=torch.float64, requires_grad=True)
    Y = ds.dark_yield(batch, mV, 1.0)
    (g,) = torch.autograd.grad(Y.sum(), mV)
    model = ds._yield_model
    h = 1e-5  # weights go as steep powers of m_V between trained masses: 1e-3 leaves ~1e-4 truncation
    mv = mV.detach().numpy()
    Yp, _ = model.evaluate(batch, mv * (1 + h), grad=False)
    Ym, _ = model.evaluate(batch, mv * (1 - h), grad=False)
    fd = (Yp - Ym) / (2 * h * mv)
    np.testing.assert_allclose(g.numpy(), fd, rtol=1e-4)
    assert np.all(np.isfinite(g.numpy()))
