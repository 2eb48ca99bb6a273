"""Table loading and n*sigma interpolation (oracle restatement).

Follows Shower.set_material_properties / set_n_targets / set_cross_sections /
set_NSigmas / get_mfp, src/PETITE/shower.py:189-246, 273-295, 357-389, and
reads the same npz packs the product uses (converted from the reference's
``sm_xsec.pkl`` and ``<P>_AdaptiveMaps.npy`` files).

interp1d semantics restated: linear between nodes, value 0 outside
[E_first, E_last] (``fill_value=0.0, bounds_error=False``).
"""
import json
import os

import numpy as np

from .constants import (GeVsqcm2, cmtom, m_electron, m_muon, m_proton_grams, target_information)
from . import integrands as I

DATA_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "petite_amd", "data")
SM_PROCESSES = ["Brem", "Ann", "PairProd", "Comp", "Moller", "Bhabha", "MuonE", "MuonBrem"]


def load_pack(kind, table_set):
    z = np.load(os.path.join(DATA_DIR, f"{kind}_{table_set}.npz"))
    meta = json.loads(bytes(z["meta_json"]).decode())
    return z, meta


class MergedPack:
    """npz packs read as one: the union of their files, each item from the
    first pack that holds it."""

    def __init__(self, packs):
        self.packs = list(packs)
        self.files = [f for z in self.packs for f in z.files]

    def __getitem__(self, key):
        for z in self.packs:
            if key in z.files:
                return z[key]
        raise KeyError(key)


def dark_packs(table_set, masses="shipped"):
    """(pack, masses) of the dark tables: the shipped pack, and with
    masses="all" also the cross sections of the generated mass pack
    (dark_<set>_masses.npz, made on the GPU by the product's tablegen tool)."""
    z, meta = load_pack("dark", table_set)
    mv = [float(m) for m in meta["masses"]]
    path = os.path.join(DATA_DIR, f"dark_{table_set}_masses.npz")
    if masses == "all" and os.path.exists(path):
        g = np.load(path)
        gm = json.loads(bytes(g["meta_json"]).decode())["masses"]
        return MergedPack([z, g]), sorted(set(mv) | {float(m) for m in gm})
    return z, mv


def grid_cuts(meta):
    """(Eg_min, Ee_min) of a table set: the Brem grid's, else the first grid's
    that records them."""
    if meta.get("Eg_min") is not None and meta.get("Ee_min") is not None:
        return float(meta["Eg_min"]), float(meta["Ee_min"])
    for g in meta["grids"].values():
        if g and g.get("Eg_min") is not None and g.get("Ee_min") is not None:
            return float(g["Eg_min"]), float(g["Ee_min"])
    raise ValueError("no grid of the table set records Eg_min / Ee_min")


def interp0(xg, yg, x):
    """scipy.interpolate.interp1d(xg, yg, fill_value=0, bounds_error=False)(x)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.interp(x, xg, yg)
    return np.where((x < xg[0]) | (x > xg[-1]), 0.0, out)


class Grid:
    """VEGAS grid nodes of one process: per E-index, per-dimension node arrays."""

    def __init__(self, E, nodes, offs, cnts):
        self.E = np.asarray(E, dtype=np.float64)
        self.nodes = np.asarray(nodes, dtype=np.float64)  # float32 values promoted
        self.offs = np.asarray(offs)
        self.cnts = np.asarray(cnts)
        self.ndim = self.cnts.shape[1]

    def dim_nodes(self, ie, d):
        o = int(self.offs[ie, d])
        return self.nodes[o:o + int(self.cnts[ie, d])]

    def lu_key(self, E):
        """LU_Key = argmin|E_i - E| + 1, clamped to N-1 (shower.py:416-426)."""
        k = int(np.argmin(np.abs(self.E - E))) + 1
        return min(max(k, 0), len(self.E) - 1)


def grid_from(E, nodes, cnt):
    """A Grid from (energies, nodes of every energy back to back, node counts)."""
    cnt = np.asarray(cnt)
    stride = int(cnt[0].sum())
    offs = np.zeros_like(cnt)
    offs[:, 0] = np.arange(len(E)) * stride
    for d in range(1, cnt.shape[1]):
        offs[:, d] = offs[:, d - 1] + cnt[:, d - 1]
    return Grid(E, np.asarray(nodes, dtype=np.float64), offs, cnt)


class SMTables:
    """Per-material SM tables: n*sigma interpolators, thresholds, grids."""

    def __init__(self, material, table_set="100GeV", grids=None):
        """grids: {process: (E, nodes, node counts)} to use instead of the
        packs' maps (a map sets only the draws' efficiency; tests pass the
        product's maps so that both sides draw from the same one)."""
        z, meta = load_pack("sm", table_set)
        self.meta = meta
        self.material = material
        ti = target_information[material]
        self.Z, self.A, self.rho, self.dEdx_MeVcm = ti["Z_T"], ti["A_T"], ti["rho"], ti["dEdx"]
        self.mT = ti["mT"]
        self.dEdx = self.dEdx_MeVcm * 0.1  # GeV/m (shower.py:649)
        self.nZ = self.rho / m_proton_grams / self.A  # shower.py:202-208
        self.ne = self.nZ * self.Z
        # the cuts come from the Brem grid's parameters (shower.py:214-215); the
        # 400 GeV set ships no Brem grid, so they come from another grid of
        # the set (all carry the same generator parameters)
        self.Eg_min, self.Ee_min = grid_cuts(meta)
        xs = {P: np.array(z[f"xsec__{P}__{material}"]) for P in SM_PROCESSES}
        self.xsec_raw = xs
        mue = xs["MuonE"]
        while mue[0, 1] == 0.0:  # shower.py:238-239
            mue = mue[1:]
        self.xsec = dict(xs)
        self.xsec["MuonE"] = mue
        self.min_calc = {  # shower.py:242-246
            11: min(xs["Brem"][0, 0], xs["Moller"][0, 0]),
            -11: min(xs["Brem"][0, 0], xs["Bhabha"][0, 0], xs["Ann"][0, 0]),
            13: max(min(xs["MuonBrem"][0, 0], mue[0, 0]), 0.120),
            -13: max(min(xs["MuonBrem"][0, 0], mue[0, 0]), 0.120),
            22: min(xs["PairProd"][0, 0], xs["Comp"][0, 0]),
        }
        # n*sigma tables [1/cm] (shower.py:273-295)
        BS = xs["Brem"]
        ns = {}
        ns["Brem"] = (BS[:, 0], self.nZ * GeVsqcm2 * BS[:, 1])
        ns["PairProd"] = (xs["PairProd"][:, 0], self.nZ * GeVsqcm2 * xs["PairProd"][:, 1])
        ns["Ann"] = (xs["Ann"][:, 0], self.ne * GeVsqcm2 * xs["Ann"][:, 1])
        ns["Comp"] = (xs["Comp"][:, 0], self.ne * GeVsqcm2 * xs["Comp"][:, 1])
        bm = np.geomspace(3.0 * m_electron + self.Ee_min, BS[-1, 0], len(BS))
        ns["Moller"] = (bm, self.ne * GeVsqcm2 * I.sigma_moller(bm, self.Ee_min))
        ns["Bhabha"] = (bm, self.ne * GeVsqcm2 * I.sigma_bhabha(bm, self.Ee_min))
        mu_min = I.muon_e_minimum(self.Ee_min)
        if mu_min < mue[0, 0]:
            mu_min = mue[0, 0]
        me_E = np.geomspace(mu_min, bm[-1], len(BS))
        ns["MuonE"] = (me_E, self.ne * GeVsqcm2 * I.sigma_muone(me_E, self.Ee_min))
        ns["MuonBrem"] = (xs["MuonBrem"][:, 0], self.nZ * GeVsqcm2 * xs["MuonBrem"][:, 1])
        self.nsigma_tab = ns
        # grids: given ones, else the set's, else (maps the set lacks) the 100 GeV set's
        self.grids = {}
        fallback = None
        for P in SM_PROCESSES:
            if grids is not None and P in grids:
                self.grids[P] = grid_from(*grids[P])
                continue
            src = z
            if f"E__{P}" not in z.files:
                if fallback is None:
                    fallback, _ = load_pack("sm", "100GeV")
                src = fallback
            self.grids[P] = Grid(src[f"E__{P}"], src[f"nodes__{P}"], src[f"nodeoff__{P}"], src[f"nodecnt__{P}"])

    def nsigma(self, process, E):
        xg, yg = self.nsigma_tab[process]
        return interp0(xg, yg, E)

    # species channel lists, order as in shower.py:665-698
    CHANNELS = {11: ("Brem", "Moller"), -11: ("Brem", "Ann", "Bhabha"), 22: ("PairProd", "Comp"),
                13: ("MuonE", "MuonBrem"), -13: ("MuonE", "MuonBrem")}

    # summation order of _NSigmaElectron/Photon/Positron/Muon (shower.py:355-366)
    SUM_ORDER = {11: ("Brem", "Moller"), -11: ("Brem", "Bhabha", "Ann"), 22: ("PairProd", "Comp"),
                 13: ("MuonBrem", "MuonE"), -13: ("MuonBrem", "MuonE")}

    def nsigma_total(self, pid, E):
        # _NSigmaElectron/Photon/Positron/Muon, shower.py:357-368
        return sum(self.nsigma(p, E) for p in self.SUM_ORDER[pid])

    def mfp(self, pid, E):
        """get_mfp, shower.py:370-389: metres; 1e12 when n*sigma <= 0."""
        ns = float(self.nsigma_total(pid, E))
        if ns <= 0.0:
            return 1.0e12
        return cmtom / ns

    def event_info(self, process, E, mV=0.0):
        """event_info dict of draw_sample (shower.py:435-439): note mT = A_T."""
        ev = {"E_inc": E, "Z_T": self.Z, "A_T": self.A, "mT": self.A, "mV": mV,
              "Eg_min": self.Eg_min, "Ee_min": self.Ee_min, "m_lepton": m_electron}
        if process in ("MuonBrem", "DarkMuonBrem"):
            ev["m_lepton"] = m_muon
        return ev
