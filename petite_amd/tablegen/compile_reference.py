"""Convert the table files PETITE ships into the npz packs this package loads.

Run once, in a checkout where the reference tree is readable::

    python -m petite_amd.tablegen.compile_reference /path/to/PETITE

The reference's ``data/`` (energies up to 100 GeV) and ``data_400GeV/``
(up to 400 GeV) directories hold, per table set (layout documented by
src/PETITE/shower.py:153-175, src/PETITE/dark_shower.py:148-245 and
utilities/find_maxes.py:112-119,203-206,302-305):

* ``sm_xsec.pkl``  {process: {material: [[E, sigma/GeV^-2], ...]}}
* ``<P>/<P>_AdaptiveMaps.npy``  rows [params, AdaptiveMap(grid)] per energy
* ``dark_xsec.pkl``  {mV: {process: {material: [[E, sigma], ...]}}}
* ``Dark<P>/mV_<k>MeV/Dark<P>_AdaptiveMaps.npy``
* ``dark_weights.pkl`` / ``dark_drate.pkl`` caches (partial coverage)

All are read with :mod:`petite_amd.tablegen.static_pickle`, which executes
nothing from the files.  VEGAS grid nodes are stored as float32 (the device
format; a map only changes sampling efficiency, its end points are kept to
float32 rounding); cross sections and weights stay float64.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

from .static_pickle import load_object_npy, load_pickle_file

SM_PROCESSES = ["Brem", "Ann", "PairProd", "Comp", "Moller", "Bhabha", "MuonE", "MuonBrem"]
DARK_PROCESSES = ["DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"]
TABLE_SETS = {"100GeV": "data", "400GeV": "data_400GeV"}


def _pack_grids(rows):
    """rows: object array (N_E, 2) of [params, AdaptiveMapGrid] -> arrays."""
    energies, nodes, offs, counts, params = [], [], [], [], []
    pos = 0
    for prm, amap in rows:
        energies.append(float(prm["E_inc"]))
        o, c = [], []
        for g in amap.grid:
            o.append(pos)
            c.append(len(g))
            nodes.append(np.asarray(g, dtype=np.float32))
            pos += len(g)
        offs.append(o)
        counts.append(c)
        params.append({k: (float(v) if isinstance(v, (int, float)) and not isinstance(v, bool) else v)
                       for k, v in prm.items()
                       if k in ("E_inc", "Eg_min", "Ee_min", "m_lepton", "mV", "Method", "process")})
    return (np.asarray(energies, dtype=np.float64), np.concatenate(nodes),
            np.asarray(offs, dtype=np.int64), np.asarray(counts, dtype=np.int32), params)


def compile_set(ref_root, set_name, out_dir):
    sub = TABLE_SETS[set_name]
    base = os.path.join(ref_root, sub)
    sm = {}
    meta = {"set": set_name, "source_dir": sub, "sm_processes": SM_PROCESSES, "grids": {}, "Eg_min": None, "Ee_min": None}

    xs = load_pickle_file(os.path.join(base, "sm_xsec.pkl"))
    meta["materials"] = sorted(xs["Brem"].keys())
    for P in SM_PROCESSES:
        for mat, tab in xs[P].items():
            sm[f"xsec__{P}__{mat}"] = np.asarray(tab, dtype=np.float64)
        path = os.path.join(base, P, f"{P}_AdaptiveMaps.npy")
        if os.path.exists(path):
            rows = load_object_npy(path)
            E, nodes, offs, counts, params = _pack_grids(rows)
            sm[f"E__{P}"], sm[f"nodes__{P}"], sm[f"nodeoff__{P}"], sm[f"nodecnt__{P}"] = E, nodes, offs, counts
            meta["grids"][P] = {"n_E": int(len(E)), "ndim": int(counts.shape[1]),
                                "Eg_min": params[0].get("Eg_min"), "Ee_min": params[0].get("Ee_min")}
            if P == "Brem":
                meta["Eg_min"] = params[0]["Eg_min"]
                meta["Ee_min"] = params[0]["Ee_min"]
        else:
            meta["grids"][P] = None
    sm["meta_json"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(out_dir, f"sm_{set_name}.npz"), **sm)

    dark = {}
    dmeta = {"set": set_name, "dark_processes": DARK_PROCESSES, "grids": {}, "weights": {}, "drate": {}}
    dx = load_pickle_file(os.path.join(base, "dark_xsec.pkl"))
    dmeta["masses"] = sorted(float(m) for m in dx.keys())
    for mv in dx:
        for P in DARK_PROCESSES:
            for mat, tab in dx[mv][P].items():
                dark[f"xsec__{P}__{mv!r}__{mat}"] = np.asarray(tab, dtype=np.float64)
    for P in DARK_PROCESSES:
        for mv in dmeta["masses"]:
            mdir = f"mV_{int(round(mv * 1000))}MeV"
            path = os.path.join(base, P, mdir, f"{P}_AdaptiveMaps.npy")
            key = f"{P}__{mv!r}"
            if os.path.exists(path):
                rows = load_object_npy(path)
                E, nodes, offs, counts, params = _pack_grids(rows)
                dark[f"E__{key}"], dark[f"nodes__{key}"] = E, nodes
                dark[f"nodeoff__{key}"], dark[f"nodecnt__{key}"] = offs, counts
                dmeta["grids"][key] = {"n_E": int(len(E)), "ndim": int(counts.shape[1]),
                                       "Method": params[0].get("Method")}
            else:
                dmeta["grids"][key] = None
    wpath = os.path.join(base, "dark_weights.pkl")
    if os.path.exists(wpath):
        dw = load_pickle_file(wpath)
        for mv, bym in dw.items():
            for mat, d in bym.items():
                for k, v in d.items():
                    dark[f"weights__{k}__{mv!r}__{mat}"] = np.asarray(v, dtype=np.float64)
                    dmeta["weights"].setdefault(f"{mv!r}__{mat}", []).append(k)
    dpath = os.path.join(base, "dark_drate.pkl")
    if os.path.exists(dpath):
        dd = load_pickle_file(dpath)
        for mv, bym in dd.items():
            for mat, d in bym.items():
                for k, table in d.items():
                    Ei = np.asarray(sorted(table.keys()), dtype=np.float64)
                    rows = [np.asarray(table[e], dtype=np.float64) for e in sorted(table.keys())]
                    nmax = max(r.shape[0] for r in rows)
                    arr = np.zeros((len(rows), nmax, 2))
                    cnt = np.zeros(len(rows), dtype=np.int32)
                    for i, r in enumerate(rows):
                        arr[i, : r.shape[0]] = r
                        cnt[i] = r.shape[0]
                    dark[f"drateE__{k}__{mv!r}__{mat}"] = Ei
                    dark[f"drate__{k}__{mv!r}__{mat}"] = arr
                    dark[f"dratecnt__{k}__{mv!r}__{mat}"] = cnt
                    dmeta["drate"].setdefault(f"{mv!r}__{mat}", []).append(k)
    dark["meta_json"] = np.frombuffer(json.dumps(dmeta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(out_dir, f"dark_{set_name}.npz"), **dark)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    ref_root = argv[0] if argv else "/root/reference"
    out_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data")
    os.makedirs(out_dir, exist_ok=True)
    for s in TABLE_SETS:
        compile_set(ref_root, s, out_dir)
        print("wrote", s)


if __name__ == "__main__":
    main()
