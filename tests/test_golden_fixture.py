"""The committed oracle fixture of the statistical GPU test is self-consistent:
per-shower histograms count at most the particles of their species, the
species counts add up to the multiplicity, and the sample is the one its
generator (tests/golden/make_c1_fixture.py) describes."""
import os

import numpy as np

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_graphite_10GeV.npz")


def test_fixture_consistent():
    f = np.load(FIX)
    n = f["n"]
    assert n.shape == (320, 4) and int(f["seed"]) == 1000003
    assert np.all(n[:, 1:].sum(1) == n[:, 0])
    for k, s in enumerate((22, 11, -11)):
        h = f[f"logE_{s}"]
        assert np.all(h.sum(1) <= n[:, k + 1])
        assert h.sum() >= 0.95 * n[:, k + 1].sum()  # 10 MeV - 10 GeV covers almost all
    assert np.all(f["z"].sum(1) <= n[:, 0]) and np.all(f["logtheta"].sum(1) <= n[:, 0])
    assert np.all(np.diff(f["edges_logE"]) > 0)
