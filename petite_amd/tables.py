"""Host-side staging of the shower tables into the device blobs.

Builds, once per (table set, material), what Shower.__init__ builds in the
reference (src/PETITE/shower.py:189-246, 273-295): the per-process n*sigma
tables [1/cm] (tabulated processes from the shipped sigma tables, Moller,
Bhabha and mu-e from their closed-form totals, all_processes.py:888-1051), the
minimum calculable energies, and the VEGAS grids.  Everything is packed into
one float64 blob (energies, n*sigma) and one float32 blob (grid nodes)
described by a ``TableDesc`` (include/petite_hip.h).
"""
import json
import os

import numpy as np

from . import _lib
from .constants import (GeVsqcm2, all_process_code, m_electron, m_muon, m_proton_grams, process_code,
                        target_information)

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
SM_PROCESSES = ["Brem", "Ann", "PairProd", "Comp", "Moller", "Bhabha", "MuonE", "MuonBrem"]
DARK_PROCESSES = ["DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"]


def _load(kind, table_set):
    path = os.path.join(DATA_DIR, f"{kind}_{table_set}.npz")
    z = np.load(path)
    return z, json.loads(bytes(z["meta_json"]).decode())


def trained_pack(table_set):
    """The pack of maps trained on the GPU for the maps a table set lacks
    (petite_amd.tablegen.train_missing_maps), or None."""
    path = os.path.join(DATA_DIR, f"maps_{table_set}_trained.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path)
    return z, json.loads(bytes(z["meta_json"]).decode())


def trained_grid(pack, key, standin):
    """A map of the trained pack: its energies, the pack's nodes at the
    trained energies and the stand-in map (E, nodes, counts) put on the other
    energies (nearest in log E, as regrid does).  None when the pack lacks it."""
    z = pack[0]
    if f"E__{key}" not in z.files:
        return None
    E = np.array(z[f"E__{key}"], dtype=np.float64)
    g = {"m": standin}
    regrid(g, "m", E)
    _, nodes, cnt = g["m"]
    stride = int(np.asarray(cnt)[0].sum())
    nodes = np.array(nodes, dtype=np.float32).reshape(len(E), stride)
    idx = np.array(z[f"trained__{key}"], dtype=np.int64)
    nodes[idx] = np.array(z[f"nodes__{key}"], dtype=np.float32).reshape(len(idx), stride)
    return E, nodes.ravel(), np.array(z[f"nodecnt__{key}"])


def grid_cuts(meta):
    """(Eg_min, Ee_min) of a table set: the parameters of its Brem grid
    (shower.py:214-215), else of the first grid that records them (the 400 GeV
    set ships no Brem grid; all its grids carry the same generator
    parameters)."""
    if meta.get("Eg_min") is not None and meta.get("Ee_min") is not None:
        return float(meta["Eg_min"]), float(meta["Ee_min"])
    for g in meta["grids"].values():
        if g and g.get("Eg_min") is not None and g.get("Ee_min") is not None:
            return float(g["Eg_min"]), float(g["Ee_min"])
    raise ValueError("no grid of the table set records Eg_min / Ee_min")


def table_set_for(dict_dir):
    """Map the reference's dict_dir argument to a table set name."""
    if dict_dir is None:
        return "100GeV"
    d = str(dict_dir).rstrip("/")
    if d in ("100GeV", "400GeV"):
        return d
    return "400GeV" if d.endswith("data_400GeV") else "100GeV"


# closed-form total cross sections [GeV^-2] (all_processes.py:888-1051)
def sigma_moller(E, Ee_min):
    E = np.asarray(E, dtype=np.float64)
    m = m_electron
    T = Ee_min - m
    thr = 3 * m + 4 * T
    ok = E >= thr
    Ee = np.where(ok, E, thr + 1.0)
    PF = 2 * np.pi * alpha() ** 2 / (m * (Ee ** 2 - m ** 2))
    T1 = Ee - 3 * m - 4 * T + 2 * Ee ** 2 * (-2 / (Ee - 3 * m - 2 * T) + 1 / T + 1 / (-Ee + m + T) + 2 / (Ee + m + 2 * T))
    T2 = 2 * m * (m - 2 * Ee) / (Ee - m) * np.log((-Ee + m + T) * (-Ee + 3 * m + 2 * T) / (T * (Ee + m + 2 * T)))
    return np.where(ok, PF * (T1 + T2), 0.0)


def sigma_bhabha(E, Ee_min):
    E = np.asarray(E, dtype=np.float64)
    m = m_electron
    T = Ee_min - m
    thr = 3 * m + 4 * T
    ok = E >= thr
    Ee = np.where(ok, E, thr + 1.0)
    PF = np.pi * alpha() ** 2 / (12 * (Ee - m) * m * (Ee + m) ** 3 * (Ee - 3 * m - 2 * T) * T)
    T1 = (Ee - 3 * m - 4 * T) * (24 * Ee ** 2 * (Ee + m) ** 2 + (Ee - 3 * m) * (31 * Ee ** 2 + 84 * Ee * m + 57 * m ** 2) * T
                                 - 4 * (16 * Ee ** 2 + 39 * Ee * m + 33 * m ** 2) * T ** 2 + 8 * (Ee - 3 * m) * T ** 3 - 8 * T ** 4)
    T2 = 24 * (Ee + m) * (2 * Ee ** 2 + 4 * Ee * m + m ** 2) * (Ee - 3 * m - 2 * T) * T * np.log(2 * T / (Ee - 3 * m - 2 * T))
    return np.where(ok, PF * (T1 + T2), 0.0)


def muon_e_threshold(Ee_min):
    me, mm = m_electron, m_muon
    return 1.0 / (2.0 * me) * (me * (Ee_min - me) + np.sqrt(me * (Ee_min + me) * (me * (Ee_min - me) + 2 * mm ** 2)))


def sigma_muone(E, Ee_min):
    E = np.asarray(E, dtype=np.float64)
    me, mm = m_electron, m_muon
    s = me ** 2 + mm ** 2 + 2 * me * E
    tmax = 2.0 * me * (me - Ee_min)
    tmin = -4.0 * ((s + me ** 2 - mm ** 2) ** 2 / (4 * s) - me ** 2)
    PF = 16 * np.pi ** 2 * alpha() ** 2 / (8.0 * np.pi * ((s - mm ** 2) ** 2 + me ** 4 - 2 * (s + mm ** 2) * me ** 2))
    T1 = -2.0 * (s ** 2 + mm ** 4 + 5 * me ** 4 - 2 * me ** 2 * (2 * s + mm ** 2)) * (1.0 / tmax - 1.0 / tmin)
    with np.errstate(invalid="ignore", divide="ignore"):
        T2 = 2.0 * (s + 2 * mm ** 2 - 2 * me ** 2) * np.log(tmax / tmin)
    T3 = tmax - tmin
    return np.where(E >= muon_e_threshold(Ee_min), PF * (T1 + T2 + T3), 0.0)


def alpha():
    from .constants import alpha_em
    return alpha_em


class MaterialTables:
    """SM tables of one material in one table set, ready for the device."""

    def __init__(self, target_material, table_set="100GeV"):
        if target_material not in target_information:
            raise Exception("Target Material is not in library")
        self.material = target_material
        self.table_set = table_set
        z, meta = _load("sm", table_set)
        if target_material not in meta["materials"]:
            raise Exception("Target Material is not in library")
        ti = target_information[target_material]
        self.Z, self.A, self.rho, self.mT = float(ti["Z_T"]), float(ti["A_T"]), float(ti["rho"]), float(ti["mT"])
        self.dEdx = ti["dEdx"] * 0.1  # GeV/m (shower.py:649)
        self.nZ = self.rho / m_proton_grams / self.A
        self.ne = self.nZ * self.Z
        self.Eg_min, self.Ee_min = grid_cuts(meta)
        xs = {P: np.array(z[f"xsec__{P}__{target_material}"], dtype=np.float64) for P in SM_PROCESSES}
        mue = xs["MuonE"]
        while mue[0, 1] == 0.0:  # shower.py:238-239
            mue = mue[1:]
        self.xsec = dict(xs, MuonE=mue)
        self.xsec_untrimmed = xs
        self.min_calc = {
            11: min(xs["Brem"][0, 0], xs["Moller"][0, 0]),
            -11: min(xs["Brem"][0, 0], xs["Bhabha"][0, 0], xs["Ann"][0, 0]),
            22: min(xs["PairProd"][0, 0], xs["Comp"][0, 0]),
            13: max(min(xs["MuonBrem"][0, 0], mue[0, 0]), 0.120),
            -13: max(min(xs["MuonBrem"][0, 0], mue[0, 0]), 0.120),
        }
        BS = xs["Brem"]
        ns = {
            "Brem": (BS[:, 0], self.nZ * GeVsqcm2 * BS[:, 1]),
            "PairProd": (xs["PairProd"][:, 0], self.nZ * GeVsqcm2 * xs["PairProd"][:, 1]),
            "Ann": (xs["Ann"][:, 0], self.ne * GeVsqcm2 * xs["Ann"][:, 1]),
            "Comp": (xs["Comp"][:, 0], self.ne * GeVsqcm2 * xs["Comp"][:, 1]),
            "MuonBrem": (xs["MuonBrem"][:, 0], self.nZ * GeVsqcm2 * xs["MuonBrem"][:, 1]),
        }
        bm = np.geomspace(3.0 * m_electron + self.Ee_min, BS[-1, 0], len(BS))  # shower.py:285
        ns["Moller"] = (bm, self.ne * GeVsqcm2 * sigma_moller(bm, self.Ee_min))
        ns["Bhabha"] = (bm, self.ne * GeVsqcm2 * sigma_bhabha(bm, self.Ee_min))
        emin = muon_e_threshold(self.Ee_min)
        if emin < mue[0, 0]:
            emin = mue[0, 0]
        me_E = np.geomspace(emin, bm[-1], len(BS))  # shower.py:289-293
        ns["MuonE"] = (me_E, self.ne * GeVsqcm2 * sigma_muone(me_E, self.Ee_min))
        self.nsigma_tab = ns
        # grids.  The 400 GeV set lacks the Brem/PairProd/MuonBrem maps: they
        # come from the pack of GPU-trained maps when it is there, else the
        # 100 GeV maps stand in (a map only changes sampling efficiency)
        self.grids = {}
        self.map_source = {}
        fb = None
        tp = trained_pack(table_set)
        for P in SM_PROCESSES:
            if f"E__{P}" in z.files:
                self.grids[P] = (np.array(z[f"E__{P}"]), np.array(z[f"nodes__{P}"], dtype=np.float32),
                                 np.array(z[f"nodecnt__{P}"]))
                self.map_source[P] = "shipped"
                continue
            if fb is None:
                fb, _ = _load("sm", "100GeV")
            standin = (np.array(fb[f"E__{P}"]), np.array(fb[f"nodes__{P}"], dtype=np.float32),
                       np.array(fb[f"nodecnt__{P}"]))
            g = trained_grid(tp, P, standin) if tp is not None else None
            self.grids[P] = g if g is not None else standin
            self.map_source[P] = "trained" if g is not None else "100GeV stand-in"

    def regrid_map(self, process, E_new):
        regrid(self.grids, process, E_new)

    def nsigma(self, process, E):
        xg, yg = self.nsigma_tab[process]
        E = np.asarray(E, dtype=np.float64)
        return np.where((E < xg[0]) | (E > xg[-1]), 0.0, np.interp(E, xg, yg))


def regrid(grids, process, E_new):
    """Put a process's map (grids[process] = (E, nodes, counts)) on the
    energies E_new, each new energy taking the nodes of the nearest old one in
    log E: the starting point for retraining (Shower.train_maps) at energies
    the shipped maps lack."""
    E, nodes, cnt = grids[process]
    E_new = np.asarray(E_new, dtype=np.float64)
    cnt = np.asarray(cnt)
    stride = int(cnt[0].sum())
    nodes = np.asarray(nodes, dtype=np.float32).reshape(len(E), stride)
    idx = np.abs(np.log(E_new)[:, None] - np.log(np.asarray(E, dtype=np.float64))[None, :]).argmin(axis=1)
    grids[process] = (E_new, nodes[idx].ravel().copy(), np.repeat(cnt[:1], len(E_new), axis=0))


def geometric(E):
    """(ln E0, (n-1)/ln(E_last/E0)) for a geometric grid, else (0, 0)."""
    E = np.asarray(E, dtype=np.float64)
    if len(E) < 2 or np.any(E <= 0):
        return 0.0, 0.0
    r = np.log(E[1:] / E[:-1])
    if np.max(np.abs(r / r.mean() - 1.0)) > 1e-9:
        return 0.0, 0.0
    return float(np.log(E[0])), float((len(E) - 1) / np.log(E[-1] / E[0]))


def build_blobs(mt, dark=None):
    """Pack MaterialTables (and optional dark tables) into (TableDesc, blob64, blob32)."""
    desc = _lib.TableDesc()
    b64, b32 = [], []
    n64 = 0
    n32 = 0

    def put64(arr):
        nonlocal n64
        a = np.ascontiguousarray(arr, dtype=np.float64)
        off = n64
        b64.append(a)
        n64 += a.size
        return off

    def put32(arr):
        nonlocal n32
        a = np.ascontiguousarray(arr, dtype=np.float32)
        off = n32
        b32.append(a)
        n32 += a.size
        return off

    grids = dict(mt.grids)
    sigmas = dict(mt.nsigma_tab)
    if dark is not None:
        grids.update(dark.grids)
        sigmas.update(dark.nsigma_tab)
    for P, (E, nodes, cnt) in grids.items():
        code = all_process_code[P]
        g = desc.grid[code]
        cnt = np.asarray(cnt)
        if not np.all(cnt == cnt[0]):
            raise ValueError(f"{P}: node counts vary with energy")
        g.n_E = len(E)
        g.ndim = cnt.shape[1]
        for d in range(g.ndim):
            g.ninc[d] = int(cnt[0, d]) - 1
        g.stride = int(cnt[0].sum())
        if nodes.size != g.n_E * g.stride:
            raise ValueError(f"{P}: node blob size mismatch")
        g.E_off = put64(E)
        g.node_off = put32(nodes)
        g.lnE0, g.inv_dlog = geometric(E)
    for P, (xg, yg) in sigmas.items():
        code = all_process_code[P]
        s = desc.sigma[code]
        s.n = len(xg)
        s.E_off = put64(xg)
        s.v_off = put64(yg)
        s.lnE0, s.inv_dlog = geometric(xg)
    desc.Z, desc.A, desc.rho, desc.dEdx = mt.Z, mt.A, mt.rho, mt.dEdx
    desc.mT = mt.A  # event_info 'mT': A_T (shower.py:435, dark_shower.py:672)
    desc.Eg_min, desc.Ee_min = mt.Eg_min, mt.Ee_min
    for i, pid in enumerate((11, -11, 22, 13, -13)):
        desc.min_calc[i] = mt.min_calc[pid]
    blob64 = np.concatenate(b64) if b64 else np.zeros(1)
    blob32 = np.concatenate(b32) if b32 else np.zeros(1, dtype=np.float32)
    return desc, blob64, blob32
