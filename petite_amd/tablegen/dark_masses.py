"""Dark tables at vector masses the reference ships none for.

    python -m petite_amd.tablegen.dark_masses [--masses 0.001,0.0015,...] [--set 100GeV]

The reference ships dark maps and cross sections at six masses (3 MeV to
1 GeV); its generator makes a mass by training a VEGAS map per (process,
energy) and tabulating sigma = E[J f] over the map per material
(utilities/script_generateintegrators.py:53-84 for the energy grids,
all_processes.py:1102-1158 for the map domains, 1085-1096 for the iteration
counts, utilities/find_maxes.py:55-119 for the cross sections with the
nuclear target mass).  This tool does the same on the GPU for each new mass:

1. energy grids as the generator's: DarkBrem geomspace(1.25 m_V, E_max),
   DarkMuonBrem geomspace(1.25 (m_V + m_mu), E_max), DarkAnn
   E_R (1 + geomspace(1e-4, (E_hi - E_R)/E_R)) with E_R = (m_V^2 - 2 m_e^2)/
   (2 m_e), DarkComp E_C (1 + geomspace(1e-4, (E_hi - E_C)/E_C)) with
   E_C = m_V (1 + m_V / (2 m_e)), E_hi = max(E_max, 100 E_thr);
2. maps on the generator's domains, uniform at the start (DarkBrem:
   x in [m_V/E, 1 - m_e/E], log10(1 - cos theta) in [-12, log10 2],
   log10(t~) in [-20, 0]; DarkAnn u in [0, 1]; DarkComp cos theta in [-1, 1]),
   1000 increments per dimension as the shipped maps, adapted by the GPU
   trainer (20 iterations of 2^16 samples per energy for the 3-D maps, 10
   for the 1-D ones) on lead;
3. cross sections per material: the mean of J f over 2^17 map points at each
   grid energy, with the material's nuclear mass as the target mass.
The bound-electron DarkAnn / DarkComp cross sections need no table (closed
forms), and the weights and drate tables are built from these cross sections
by DarkTables like any uncached (mass, material) pair.  The pack (default
petite_amd/data/dark_<set>_masses.npz) holds xsec__<P>__<m_V>__<material>
in the layout of the shipped dark pack (with --store-maps also the maps,
E__/nodes__/nodeoff__/nodecnt__<P>__<m_V>; without them DarkTables starts
such a mass's maps flat on the same domains and DarkShower trains them on
the GPU when it needs vectors there); DarkTables(masses="all") and the
m_V-continuous yield use it.
"""
import argparse
import json
import os

import numpy as np

from ..constants import target_information
from ..dark_tables import MAP_ITERATIONS, flat_map, generator_energies
from ..tables import DATA_DIR, MaterialTables

NEW_MASSES = [0.001, 0.0015, 0.002, 0.0025, 0.0055, 0.017, 0.055]
PROCESSES = ["DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"]
SIGMA_SAMPLES = 1 << 17


def _offsets(cnt):
    stride = int(cnt[0].sum())
    offs = np.zeros_like(cnt)
    offs[:, 0] = np.arange(len(cnt)) * stride
    for d in range(1, cnt.shape[1]):
        offs[:, d] = offs[:, d - 1] + cnt[:, d - 1]
    return offs


def make_mass(mV, table_set, materials, train_material, arrays, log, store_maps=False):
    from ..dark_tables import DarkTables
    from ..device import DeviceContext
    sm = MaterialTables(train_material, table_set)
    dt = DarkTables(sm, mV, mode="approx")  # the tables of a trained mass, m_V itself in the integrands
    E_max = float(np.max(sm.nsigma_tab["Brem"][0]))
    key = repr(float(mV))
    grids = {}
    for P in PROCESSES:
        grids[P] = flat_map(P, generator_energies(P, mV, E_max), mV)
        dt.grids[P] = grids[P]
    ctx = DeviceContext(sm, 0.010, dark=dt)
    for P in PROCESSES:
        nodes = ctx.train_map(P, n_iter=MAP_ITERATIONS[P])
        E, _, cnt = grids[P]
        grids[P] = (E, nodes.ravel(), cnt)
        dt.grids[P] = grids[P]
        if store_maps:
            arrays[f"E__{P}__{key}"] = np.asarray(E, dtype=np.float64)
            arrays[f"nodes__{P}__{key}"] = np.asarray(nodes, dtype=np.float32).ravel()
            arrays[f"nodecnt__{P}__{key}"] = cnt
            arrays[f"nodeoff__{P}__{key}"] = _offsets(cnt)
    ctx.close()
    for mat in materials:
        smm = MaterialTables(mat, table_set)
        dtm = DarkTables(smm, mV, mode="approx")
        for P in PROCESSES:
            dtm.grids[P] = grids[P]
        c = DeviceContext(smm, 0.010, dark=dtm, mT=target_information[mat]["mT"])
        for P in PROCESSES:
            _, mean = c.scan_maxf(P, n_samples=SIGMA_SAMPLES, seed=0x51C3A)
            sig = np.where(np.isfinite(mean), mean, 0.0)
            arrays[f"xsec__{P}__{key}__{mat}"] = np.column_stack([grids[P][0], sig])
            eff = c.sigma_est[P] / np.where(c.maxf[P] > 0, c.maxf[P], np.inf)
            log[f"{P}__{key}__{mat}"] = {"sigma_max": float(np.max(sig)),
                                         "median_accept": float(np.median(eff[sig > 0])) if np.any(sig > 0) else 0.0}
        c.close()
        print(key, mat, {P: log[f"{P}__{key}__{mat}"] for P in PROCESSES}, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--masses", default=",".join(str(m) for m in NEW_MASSES))
    ap.add_argument("--set", default="100GeV")
    ap.add_argument("--materials", default=",".join(sorted(target_information)))
    ap.add_argument("--train-material", default="lead")
    ap.add_argument("--out", default=None)
    ap.add_argument("--store-maps", action="store_true")
    a = ap.parse_args()
    masses = [float(m) for m in a.masses.split(",")]
    from ..tables import _load
    _, meta_ship = _load("sm", a.set)
    materials = [m for m in a.materials.split(",") if m in meta_ship["materials"]]
    arrays, log = {}, {}
    for m in masses:
        make_mass(m, a.set, materials, a.train_material, arrays, log, store_maps=a.store_maps)
    meta = {"table_set": a.set, "masses": masses, "materials": materials, "train_material": a.train_material,
            "map_iterations": MAP_ITERATIONS, "sigma_samples": SIGMA_SAMPLES,
            "tool": "petite_amd.tablegen.dark_masses", "log": log}
    arrays["meta_json"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    out = a.out or os.path.join(DATA_DIR, f"dark_{a.set}_masses.npz")
    np.savez_compressed(out, **arrays)
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main()
