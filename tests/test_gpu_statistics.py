"""Statistical agreement of GPU showers with the oracle: 10 GeV electrons in
graphite, cutoff 10 MeV.  The oracle's showers are a
committed fixture (tests/golden/make_c1_fixture.py, 320 showers, its own
seed); the GPU runs 2,000 showers from another seed, so the samples are
independent.  Particles of one shower are correlated, so every observable is
compared through per-shower counts: for each bin the difference of the mean
counts per shower over its standard error, |z| < 4.5 for all bins (about 60
bins; 4.5 sigma keeps the false-alarm rate well below 1e-3), and the
per-shower multiplicity distributions by a two-sample KS test (p > 0.01)."""
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from petite_amd import Shower
from petite_amd.batch import PRIMARY_DTYPE
from petite_amd.constants import m_electron

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "c1_graphite_10GeV.npz")
SPECIES = (22, 11, -11)


def _gpu_observables(n=2000, seed=7):
    sh = Shower("100GeV", "graphite", 0.010, seed=seed)
    prim = np.zeros(n, dtype=PRIMARY_DTYPE)
    prim["E"] = 10.0
    prim["pz"] = math.sqrt(100.0 - m_electron ** 2)
    prim["mass"] = m_electron
    prim["weight"] = 1.0
    prim["pid"] = 11
    prim["shower"] = np.arange(n, dtype=np.int32)
    prim["sid"] = np.arange(n, dtype=np.uint64)
    a = sh.generate_showers(prim).to_numpy()
    fx = np.load(FIX)
    keep = a["parent"] >= 0
    pid, sh_i = a["pid"][keep], a["shower"][keep]
    p0 = a["p0"][keep].astype(np.float64)
    z = a["r0"][keep][:, 2].astype(np.float64)
    pn = np.linalg.norm(p0[:, 1:], axis=1)
    theta = np.arccos(np.clip(p0[:, 3] / np.maximum(pn, 1e-300), -1.0, 1.0))
    out = {"n": np.zeros((n, 4))}
    np.add.at(out["n"], (sh_i, 0), 1)
    for k, s in enumerate(SPECIES):
        np.add.at(out["n"], (sh_i[pid == s], k + 1), 1)

    def hist(key, values, mask, edges):
        h = np.zeros((n, len(edges) - 1))
        b = np.digitize(values[mask], edges) - 1
        ok = (b >= 0) & (b < len(edges) - 1)
        np.add.at(h, (sh_i[mask][ok], b[ok]), 1)
        out[key] = h

    lE = np.log10(np.maximum(p0[:, 0], 1e-30))
    for s in SPECIES:
        hist(f"logE_{s}", lE, pid == s, fx["edges_logE"])
    hist("z", z, np.ones_like(z, bool), fx["edges_z"])
    hist("logtheta", np.log10(np.maximum(theta, 1e-30)), np.ones_like(z, bool), fx["edges_logtheta"])
    return out, fx


@pytest.fixture(scope="module")
def obs():
    return _gpu_observables()


def _z(a, b):
    ma, mb = a.mean(0), b.mean(0)
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    ok = se > 0
    return np.abs(ma - mb)[ok] / se[ok]


@pytest.mark.parametrize("key", ["n", "logE_22", "logE_11", "logE_-11", "z", "logtheta"])
def test_per_shower_means_agree(obs, key):
    g, fx = obs
    z = _z(g[key], fx[key].astype(np.float64))
    assert z.size > 0 and z.max() < 4.5, (key, np.round(z, 2))


def test_multiplicity_distribution(obs):
    from scipy.stats import ks_2samp
    g, fx = obs
    for k in range(4):
        p = ks_2samp(g["n"][:, k], fx["n"][:, k]).pvalue
        assert p > 0.01, (k, p)
    # the reference's one-shower anchor (examples/multiple_coulomb_scattering.ipynb:
    # 655 particles) sits inside the spread of the multiplicity
    tot = g["n"][:, 0] + 1
    assert abs(655 - tot.mean()) < 3 * tot.std()
