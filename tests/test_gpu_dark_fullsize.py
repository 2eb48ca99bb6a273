"""The full-size dark-vector workload (20,000 e- at 10 GeV, graphite, cutoff
10 MeV, m_V = 1 MeV on the 3 MeV tables, DarkBrem/DarkAnn/DarkComp: about
1.4e7 SM records and 1.8e7 entries) checked through properties that do not
depend on the size: exactly one entry per (SM record, process) with a positive
weight (dark_shower.py:827-848); each entry's process allowed for its parent's
species (GetBSMWeights, 595-647); every entry's weight equal to
GetBSMWeights(E0) times the parent weight; vectors on shell and within the
energy available (the parent's, plus the target electron's mass for DarkAnn
and DarkComp); the yields equal to the per-process weight sums.  At m_V = 1
MeV < 2 m_e the radiative-return integrand is NaN everywhere
(all_processes.py:431), so every DarkAnn entry ends without a vector, as
produce_bsm_particle returns None there; the entries without a vector are
exactly those."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from petite_amd.batch import PRIMARY_DTYPE
from petite_amd.constants import m_electron
from petite_amd.dark_shower import DarkShower

N, E0, MV = 20_000, 10.0, 0.001
PROCS = ("DarkBrem", "DarkAnn", "DarkComp")
PARENT_PIDS = {8: (11, -11), 9: (-11,), 10: (22,)}


@pytest.fixture(scope="module")
def run():
    ds = DarkShower("100GeV", "graphite", 0.010, MV, mode="approx", seed=1234, active_processes=list(PROCS))
    prim = np.zeros(N, dtype=PRIMARY_DTYPE)
    prim["E"] = E0
    prim["pz"] = np.sqrt(E0 * E0 - m_electron ** 2)
    prim["mass"] = m_electron
    prim["weight"] = 1.0
    prim["pid"] = 11
    prim["shower"] = np.arange(N, dtype=np.int32)
    prim["sid"] = np.arange(N, dtype=np.uint64) + (np.uint64(3) << np.uint64(40))
    rec = ds.generate_showers(prim)
    dark = ds.generate_dark_showers(rec)
    a = rec.to_numpy()
    d = {"parent": dark.parent.cpu().numpy().astype(np.int64), "proc": dark.process.cpu().numpy(),
         "valid": dark.valid.cpu().numpy(), "p": dark.p.cpu().numpy().astype(np.float64),
         "w": dark.weight.cpu().numpy().astype(np.float64)}
    return ds, rec, a, dark, d


def test_one_entry_per_record_and_process(run):
    ds, rec, a, dark, d = run
    assert rec.n > 1e7 and dark.n > 1e7
    par, proc = d["parent"], d["proc"]
    assert dark.stats["n_pe_fail"] == 0
    assert np.all(proc[~d["valid"]] == 9) and np.all(d["valid"][proc != 9])
    assert dark.stats["n_no_sample"] == int((~d["valid"]).sum()) > 0
    assert par.min() >= 0 and par.max() < rec.n
    key = par * 4 + (proc - 8)
    assert len(np.unique(key)) == len(key)  # at most one entry per (record, process)
    E = a["p0"][:, 0].astype(np.float64)
    for k, P in enumerate(PROCS):
        wg = ds._dark.bsm_weight(a["pid"], E, P)
        have = np.zeros(rec.n, dtype=bool)
        have[par[proc == 8 + k]] = True
        assert np.array_equal(have, wg > 0.0), P  # one entry exactly where the weight is positive
        sel = proc == 8 + k
        assert np.all(np.isin(a["pid"][par[sel]], PARENT_PIDS[8 + k])), P
        # entries without a vector hold weight 0, and so does DarkComp's
        # below-threshold vector (m_V, 0, 0, 0) (dark_shower.py:777-780)
        wsel = sel & d["valid"] & (d["w"] != 0.0)
        want = wg[par[wsel]] * a["r0"][par[wsel], 3].astype(np.float64)
        np.testing.assert_allclose(d["w"][wsel], want, rtol=1e-5, atol=0.0, err_msg=P)
    z = d["valid"] & (d["w"] == 0.0)
    assert np.all(proc[z] == 10)
    np.testing.assert_allclose(d["p"][z], np.tile([MV, 0.0, 0.0, 0.0], (int(z.sum()), 1)), atol=1e-9)


def test_vectors_on_shell_within_energy(run):
    ds, rec, a, dark, d = run
    ok = d["valid"]
    assert ok.sum() > 1e7
    p = d["p"][ok]
    m2 = p[:, 0] ** 2 - (p[:, 1:] ** 2).sum(1)
    assert np.all(np.abs(m2 - MV ** 2) <= 2e-6 * p[:, 0] ** 2 + 1e-7)
    E_par = a["p0"][d["parent"][ok], 0].astype(np.float64)
    extra = np.where(d["proc"][ok] == 8, 0.0, m_electron)
    assert np.all(p[:, 0] <= (E_par + extra) * (1 + 1e-6))


def test_yields_are_the_weight_sums(run):
    ds, rec, a, dark, d = run
    y = dark.yields
    for k, P in enumerate(PROCS):
        s = d["w"][d["valid"] & (d["proc"] == 8 + k)].sum()
        np.testing.assert_allclose(s, y[P], rtol=1e-5, err_msg=P)
