"""Train the VEGAS maps the 400 GeV table set lacks and write them as a pack.

    python -m petite_amd.tablegen.train_missing_maps [--material lead] [--out PATH]

The reference's data_400GeV/ ships no Brem, PairProd or MuonBrem maps, and no
DarkBrem / DarkMuonBrem maps at any mass (.MISSING_LARGE_BLOBS); its own
generator trains them with vegas per (process, energy)
(utilities/generate_integrators.py:49-245).  Without them the 100 GeV maps
stand in, which fit poorly above 100 GeV (a Brem draw at 399 GeV needs ~220
candidates).  This tool, run once on a GPU, puts every missing map on the
energies of its process's 400 GeV cross-section table (0.0016..400 GeV;
DarkBrem / DarkMuonBrem: their dark tables' energies), starting each energy
from the nearest 100 GeV map in log E, and adapts the energies above the
100 GeV set's last map energy with the GPU trainer (petite_amd.vegas_train:
vegas's adaptation rule on (J f)^2, 10 iterations of 2^16 samples per
energy, on the given material); lower energies keep the nearest shipped map,
which the reference's vegas training made (those maps already fit: retraining
them with uniform sampling lowers their efficiency).  A map changes only the
accept/reject efficiency; its end points, kept from the 100 GeV maps, are the
cut.  The pack (default petite_amd/data/maps_400GeV_trained.npz) holds, per
process key (P, or P__<m_V> for the dark processes), the map's energies
(E__), the indices of the trained ones (trained__), their nodes (nodes__,
the layout of the shipped packs) and the node counts (nodecnt__), and a
meta_json record of how it was made.  The other energies keep the nearest
100 GeV map, which the loader (petite_amd.tables.trained_grid) rebuilds the
same way; MaterialTables and DarkTables of the 400 GeV set use these maps in
place of the 100 GeV stand-ins.
"""
import argparse
import json
import os

import numpy as np

from ..tables import DATA_DIR, regrid

SM_MISSING = ["Brem", "PairProd", "MuonBrem"]
DARK_MISSING = ["DarkBrem", "DarkMuonBrem"]
OUT = os.path.join(DATA_DIR, "maps_400GeV_trained.npz")
N_ITER, N_PER_ENERGY = 10, 1 << 16


def _pack(arrays, key, E, nodes, cnt, mask):
    cnt = np.asarray(cnt)
    stride = int(cnt[0].sum())
    idx = np.nonzero(mask)[0]
    arrays[f"E__{key}"] = np.asarray(E, dtype=np.float64)
    arrays[f"trained__{key}"] = idx.astype(np.int32)
    arrays[f"nodes__{key}"] = np.asarray(nodes, dtype=np.float32).reshape(len(E), stride)[idx].ravel()
    arrays[f"nodecnt__{key}"] = cnt


def train_sm(material, arrays, log):
    from ..shower import Shower
    sh = Shower("400GeV", material, 0.010, seed=1)
    t = sh._tables
    E_top = max(float(np.max(t.grids[P][0])) for P in SM_MISSING)
    for P in SM_MISSING:
        regrid(t.grids, P, t.nsigma_tab[P][0])
    ctx = sh.device_context()
    for P in SM_MISSING:
        E, _, cnt = t.grids[P]
        mask = np.asarray(E) > E_top
        before = ctx.sigma_est[P].copy(), ctx.maxf[P].copy()
        nodes = ctx.train_map(P, n_iter=N_ITER, n_per_energy=N_PER_ENERGY, train=mask)
        eff0 = before[0][mask] / before[1][mask]
        eff1 = ctx.sigma_est[P][mask] / ctx.maxf[P][mask]
        log[P] = {"trained_energies": int(mask.sum()), "median_efficiency_before": float(np.median(eff0)),
                  "median_efficiency_after": float(np.median(eff1))}
        print(P, log[P], flush=True)
        _pack(arrays, P, E, nodes, cnt, mask)
    return E_top


def train_dark(material, arrays, log):
    from ..dark_shower import DarkShower
    from ..dark_tables import _load
    _, meta = _load("dark", "400GeV")
    for mV in meta["masses"]:
        ds = DarkShower("400GeV", material, 0.010, mV, seed=1)
        d = ds._dark
        key = repr(d.mV_est)
        E_top = max(float(np.max(d.grids[P][0])) for P in DARK_MISSING)
        for P in DARK_MISSING:
            regrid(d.grids, P, d.xsec[P][:, 0])
        ctx = ds.device_context()
        for P in DARK_MISSING:
            E, _, cnt = d.grids[P]
            mask = np.asarray(E) > E_top
            nodes = ctx.train_map(P, n_iter=N_ITER, n_per_energy=N_PER_ENERGY, train=mask)
            log[f"{P}__{key}"] = {"trained_energies": int(mask.sum())}
            print(P, key, log[f"{P}__{key}"], flush=True)
            _pack(arrays, f"{P}__{key}", E, nodes, cnt, mask)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--material", default="lead")
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--no-dark", action="store_true")
    a = ap.parse_args()
    arrays, log = {}, {}
    E_top = train_sm(a.material, arrays, log)
    if not a.no_dark:
        train_dark(a.material, arrays, log)
    meta = {"table_set": "400GeV", "material": a.material, "n_iter": N_ITER, "n_per_energy": N_PER_ENERGY,
            "trained_above_GeV": E_top, "tool": "petite_amd.tablegen.train_missing_maps", "log": log}
    arrays["meta_json"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(a.out, **arrays)
    print("wrote", a.out, flush=True)


if __name__ == "__main__":
    main()
