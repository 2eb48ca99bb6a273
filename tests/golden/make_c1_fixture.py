"""Per-shower observables of 10 GeV electron showers in graphite (cutoff
10 MeV) from the oracle, for the statistical comparison of GPU showers
(tests/test_gpu_statistics.py).

    python tests/golden/make_c1_fixture.py [n_showers] [processes]

Writes tests/golden/c1_graphite_10GeV.npz: for every shower the counts of
non-primary particles by species and in bins of log10 E (GeV), start depth z
(m) and polar angle log10(theta); the oracle's seed and the bin edges.  The
GPU test draws its showers from another seed, so the two samples are
independent.
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

SEED = 1000003
E0, MATERIAL, CUTOFF = 10.0, "graphite", 0.010
EDGES = {
    "logE": np.linspace(-2.0, 1.0, 16),
    "z": np.linspace(0.0, 3.0, 16),
    "logtheta": np.linspace(-4.0, 0.5, 16),
}
SPECIES = (22, 11, -11)


def observables(pid, E, z, theta):
    """Per-shower observables from arrays over the non-primary particles."""
    out = {"n": np.array([len(pid)] + [int(np.sum(pid == s)) for s in SPECIES])}
    lE = np.log10(np.maximum(E, 1e-30))
    for s in SPECIES:
        out[f"logE_{s}"] = np.histogram(lE[pid == s], EDGES["logE"])[0]
    out["z"] = np.histogram(z, EDGES["z"])[0]
    out["logtheta"] = np.histogram(np.log10(np.maximum(theta, 1e-30)), EDGES["logtheta"])[0]
    return out


def _one(s):
    from oracle.constants import m_electron
    from oracle.shower import OracleShower
    from oracle.tables import SMTables
    osh = OracleShower(SMTables(MATERIAL), CUTOFF, seed=SEED)
    parts = osh.generate_shower(11, [E0, 0.0, 0.0, math.sqrt(E0 * E0 - m_electron ** 2)], sid=s, shower=s)
    parts = [p for p in parts if p.parent >= 0]
    pid = np.array([p.pid for p in parts])
    p0 = np.array([p.p0 for p in parts]).reshape(-1, 4)
    z = np.array([p.r0[2] for p in parts])
    pn = np.linalg.norm(p0[:, 1:], axis=1)
    theta = np.arccos(np.clip(p0[:, 3] / np.maximum(pn, 1e-300), -1.0, 1.0))
    return observables(pid, p0[:, 0], z, theta)


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 320
    procs = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    from multiprocessing import Pool
    with Pool(procs) as pool:
        rows = pool.map(_one, range(n))
    out = {k: np.stack([r[k] for r in rows]) for k in rows[0]}
    for k, v in EDGES.items():
        out[f"edges_{k}"] = v
    out["seed"] = np.array(SEED)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_graphite_10GeV.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
