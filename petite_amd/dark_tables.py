"""Host-side dark-sector tables (DarkShower initialisation).

Builds, once per (table set, material, m_V), what DarkShower.__init__ builds in
the reference (src/PETITE/dark_shower.py:85-593):

* mass selection (closest_lesser_value, 158-176),
* dark cross-section tables and thresholds (219-245),
* bound-electron annihilation and Compton cross sections on their 200-point
  grids (254-292, src/PETITE/atomic_annihilation.py, atomic_compton.py),
* log-log n*sigma interpolants (294-309, interpolate1d 31-46),
* track-length weights W(E_i) and the 10-bin interaction-energy rates
  (drate) (311-593), from the survival factor of the SM interaction
  integrals (src/PETITE/shower.py:298-354).

The weights and rates are integrals the reference evaluates with
scipy.integrate.quad.  Here they use Gauss-Legendre panels that follow the
exponential survival factor (width ~ dE/dx x mean free path) and break at
every table node; caches shipped with the reference (dark_weights.pkl,
dark_drate.pkl) are used when they cover the (mass, material) pair, as the
reference does.
"""
import json
import os

import numpy as np

from .constants import (TWO_BODY_CODE, GeVsqcm2, alpha_em, cmtom, m_electron, m_muon,
                        meson_twobody_branchingratios)
from .tables import DATA_DIR, MaterialTables, _load, geometric, trained_grid, trained_pack

DARK_PROCESSES = ["DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"]
# survival channels per species (shower.py:322-354) and the summation order of
# the mean free path (get_mfp with _NSigmaElectron/Positron/Muon, shower.py:355-389)
SURVIVAL = {11: ("Brem", "Moller"), -11: ("Brem", "Ann", "Bhabha"), 13: ("MuonBrem", "MuonE"),
            -13: ("MuonBrem", "MuonE")}
MFP_ORDER = {11: ("Brem", "Moller"), -11: ("Brem", "Bhabha", "Ann"), 13: ("MuonBrem", "MuonE"),
             -13: ("MuonBrem", "MuonE")}
WEIGHT_FAMILIES = (("brem_elec", "DarkBrem", 11), ("brem_positron", "DarkBrem", -11),
                   ("muon_brem", "DarkMuonBrem", 13), ("annihilation", "DarkAnn", -11))


class _Packs:
    """Several npz packs read as one: ``files`` is their union and an item
    comes from the first pack that holds it."""

    def __init__(self, packs):
        self._packs = list(packs)
        self.files = [f for z in self._packs for f in z.files]

    def __getitem__(self, key):
        for z in self._packs:
            if key in z.files:
                return z[key]
        raise KeyError(key)


def generator_energies(process, mV, E_max=100.0, n=100):
    """The energies at which the reference's generator tabulates a dark
    process at mass mV (utilities/script_generateintegrators.py:62-83)."""
    if process == "DarkBrem":
        return np.geomspace(1.25 * mV, E_max, n)
    if process == "DarkMuonBrem":
        return np.geomspace(1.25 * (mV + m_muon), E_max, n)
    if process == "DarkAnn":
        E0 = (mV ** 2 - 2 * m_electron ** 2) / (2 * m_electron)
    else:
        E0 = mV * (1 + mV / (2 * m_electron))
    E_hi = max(E_max, 100 * E0)
    return E0 * (1 + np.geomspace(1e-4, (E_hi - E0) / E0, n))


MAP_INCREMENTS = 1000
MAP_ITERATIONS = {"DarkBrem": 20, "DarkMuonBrem": 20, "DarkAnn": 10, "DarkComp": 10}  # all_processes.py:1085-1096


def flat_map(process, E, mV):
    """Uniform maps on the generator's domains (all_processes.py:1102-1158) at
    the energies E, 1000 increments per dimension as the shipped maps:
    DarkBrem / DarkMuonBrem x in [m_V/E, 1 - m_e/E], log10(1 - cos theta) in
    [-12, log10 2], log10(t~) in [-20, 0]; DarkAnn u in [0, 1]; DarkComp
    cos theta in [-1, 1].  Returns (E, nodes float32, node counts)."""
    rows = []
    for e in E:
        if process in ("DarkBrem", "DarkMuonBrem"):
            dom = [(max(0.0, mV / e), 1.0 - m_electron / e), (-12.0, np.log10(2.0)), (-20.0, 0.0)]
        elif process == "DarkAnn":
            dom = [(0.0, 1.0)]
        else:
            dom = [(-1.0, 1.0)]
        rows.append(np.concatenate([np.linspace(a, b, MAP_INCREMENTS + 1) for a, b in dom]))
    ndim = 3 if process in ("DarkBrem", "DarkMuonBrem") else 1
    return (np.asarray(E, dtype=np.float64), np.asarray(rows, dtype=np.float32).ravel(),
            np.full((len(E), ndim), MAP_INCREMENTS + 1))


def mass_pack(table_set):
    """The pack of dark tables at the masses made by
    petite_amd.tablegen.dark_masses (dark cross sections per material and
    maps at masses the reference ships none for), or None."""
    path = os.path.join(DATA_DIR, f"dark_{table_set}_masses.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path)
    return z, json.loads(bytes(z["meta_json"]).decode())


def dark_pack(table_set, masses="shipped"):
    """(pack, trained masses) of a table set: the shipped dark tables, and with
    masses="all" also those of the generated mass pack."""
    z, meta = _load("dark", table_set)
    mv = [float(m) for m in meta["masses"]]
    if masses == "all":
        extra = mass_pack(table_set)
        if extra is not None:
            mv = sorted(set(mv) | {float(m) for m in extra[1]["masses"]})
            return _Packs([z, extra[0]]), mv
    elif masses != "shipped":
        raise ValueError('masses: "shipped" or "all"')
    return z, mv


def closest_lesser_value(values, x, below="clamp"):
    """dark_shower.py:158-165.  For x below the lightest mass the reference
    indexes [-1] and wraps to the heaviest (below='wrap'); below='clamp'
    takes the lightest instead."""
    arr = np.asarray(values, dtype=np.float64)
    i = int(np.argmin(np.abs(arr - x)))
    if arr[i] <= x:
        return float(arr[i])
    if i == 0 and below == "clamp":
        return float(arr[0])
    return float(arr[i - 1])


# ---------------------------------------------------------------- bound e- --
def _atomic_ann(k, mV, Zeff):
    """sigma of e+ e-(bound) -> V, tree + radiative tail (atomic_annihilation.py:70-104).
    That module uses its own alpha = 1/137 and m_e = 0.511 MeV.  The closed form
    cancels strongly near threshold, so each term is formed in the reference's
    order (scalar float64) to reproduce its values there."""
    return np.array([_atomic_ann1(np.float64(kk), mV, Zeff) for kk in np.atleast_1d(k)])


def _atomic_ann1(k, mV, Zeff):
    al, me = 1 / 137, 0.511e-3
    lam = Zeff * al * me
    a = me / lam
    b = mV ** 2 / (2 * k * lam)
    pre = (4.0 * np.pi * al) * ((mV ** 2 + 2 * me ** 2) * (2 / 3 / lam / me) * (1 / k ** 2) * (1 / ((a - b) ** 2 + 1) ** 3))
    s = 2 * k * me + me ** 2
    beta = 2 * al / np.pi * (np.log(s / me ** 2) - 1.0)
    tail = (4.0 * np.pi * al) * ((mV ** 2 + 2 * me ** 2) * (2 / 3 / lam / me) * (1 / k ** 2) * (beta / 2) * _fancy1(a, b))
    return pre + tail


def _fancy1(a, b):
    # closed-form radiative-tail integral of atomic_annihilation.py:10-68 (series for b >= 1000)
    if b >= 1000:
        return (3081 / 280 - (3081 * a ** 2) / 40) / b ** 8 - (709 * a) / (35 * b ** 7) - 91 / (30 * b ** 6)
    p1 = (1 + a ** 2)
    p3 = (1 + a ** 2) ** 3
    c5 = 5 + 3 * a ** 2
    p2 = (1 + a ** 2) ** 2
    e1 = -1 + 15 * a ** 2 + 12 * a ** 4
    t1 = p1 * (a * p3 * c5 - p2 * e1 * b + a * (1 + a ** 2) * (-23 + 38 * a ** 2 + 21 * a ** 4) * b ** 2
               + 4 * (5 + a ** 2 - 18 * a ** 4 - 6 * a ** 6) * b ** 3 + 3 * a * (3 + a ** 2) * (-1 + 7 * a ** 2) * b ** 4
               + (7 - 29 * a ** 2 - 12 * a ** 4) * b ** 5 + a * (7 + 3 * a ** 2) * b ** 6)
    poly = (3 * (1 + a ** 2) ** 6 - a * (1 + a ** 2) ** 3 * (3 + 26 * a ** 2 + 15 * a ** 4) * b
            + 3 * (1 + a ** 2) ** 2 * (-7 + 21 * a ** 2 + 23 * a ** 4 + 11 * a ** 6) * b ** 2
            - 3 * a * (1 + a ** 2) * (7 + 69 * a ** 2 + 45 * a ** 4 + 15 * a ** 6) * b ** 3
            + (-11 + 132 * a ** 2 + 9 * a ** 4 * (38 + 5 * a ** 2 * (4 + a ** 2))) * b ** 4
            - 3 * a * (11 + 73 * a ** 2 + 41 * a ** 4 + 11 * a ** 6) * b ** 5
            + 3 * (-1 + 27 * a ** 2 + 17 * a ** 4 + 5 * a ** 6) * b ** 6 - a * (15 + 10 * a ** 2 + 3 * a ** 4) * b ** 7)
    t4 = (1 + a ** 2) ** 3 + 6 * a * (1 + a ** 2) ** 2 * b - 3 * (1 + 6 * a ** 2 + 5 * a ** 4) * b ** 2 + \
        4 * a * (3 + 5 * a ** 2) * b ** 3 - 3 * (1 + 5 * a ** 2) * b ** 4 + 6 * a * b ** 5 - b ** 6
    num = (t1 + poly * np.arctan2(a, 1) + poly * np.arctan2(1 + a ** 2 - a * b, b)
           + 4 * b * (t4 * (np.log(1 + (a - b) ** 2) - 2 * np.log(b))))
    return num / (8 * (1 + a ** 2) ** 3 * (1 + (a - b) ** 2) ** 3 * b)


def _atomic_comp(k, mV, Zeff):
    """sigma of gamma e-(bound) -> e- V (atomic_compton.py:94-109 with combine/expr1/expr2),
    term order as in the reference (see _atomic_ann)."""
    return np.array([_atomic_comp1(np.float64(kk), mV, Zeff) for kk in np.atleast_1d(k)])


def _atomic_comp1(k, mV, Zeff):
    me, al = m_electron, alpha_em
    lam = Zeff * al * me
    a = me / lam
    b = mV ** 2 / (2 * k * lam)
    s = 2 * k * me + me ** 2
    beta = 2 * al / np.pi * (np.log(s / me ** 2) - 1.0)
    if beta < 0.0:
        beta = 0.0
    return (4.0 * np.pi * al) * ((mV ** 2 + 2 * me ** 2) * (2 / 3 / lam / me) * (1 / k ** 2) * (beta / 4) * _combine1(a, b))


def _inner(a, b):
    return (-a * (a ** 2 + 1) ** 3 * (3 * a ** 2 + 5) + 6 * (a ** 4 + 4 * a ** 2 - 5) * b ** 5
            - 4 * a * (6 * a ** 4 + 23 * a ** 2 - 31) * b ** 4 + (a - 1) * (a + 1) * (39 * a ** 4 + 190 * a ** 2 + 55) * b ** 3
            - a * (a ** 2 + 1) * (33 * a ** 4 + 98 * a ** 2 - 111) * b ** 2 + (a ** 2 + 1) ** 2 * (15 * a ** 4 + 32 * a ** 2 - 23) * b)


def _coeff(a, b):
    return (3 * (a ** 2 + 1) ** 4 - 2 * a * (3 * a ** 4 + 10 * a ** 2 + 15) * (a ** 2 + 1) * b
            + 6 * (a ** 6 + 5 * a ** 4 + 15 * a ** 2 - 5) * b ** 2)


def _combine1(a, b):
    d = (a - b) ** 2 + 1
    lt = np.log(d) - 2 * np.log(b)
    if a ** 2 + 1 < a * b:
        # region above the kinematic knee: expr1, with its large-b series
        num = (8 * b * (a ** 2 - 6 * a * b + 1) * lt + (a ** 2 + 1) * _inner(a, b) / d ** 2
               + _coeff(a, b) * (np.arctan(b / (a ** 2 - a * b + 1)) + np.arctan(1 / a)))
        e1 = -num / (8 * (a ** 2 + 1) ** 4 * b)
        if e1 < 0.0 or (a ** 2 + 1 < 0.01 * a * b):
            e1 = (-145 + 1015 * a ** 2 + 330 * a * b + 64 * b ** 2) / (420 * b ** 8)
        return e1
    c = _coeff(a, b)
    t1 = (a ** 2 + 1) * -_inner(a, b)
    t2 = np.pi * c * d ** 2
    t3 = c * d ** 2 * np.arctan(1 / a)
    t4 = d ** 2 * (8 * b * (a ** 2 - 6 * a * b + 1) * lt + c * np.arctan(b / (a ** 2 - a * b + 1)))
    return (t1 + t2 - t3 - t4) / (8 * (a ** 2 + 1) ** 4 * b * d ** 2)


# ------------------------------------------------------------ interpolants --
class LogLog:
    """interpolate1d(xspace='log', yspace='log', fill_value=-20) (dark_shower.py:31-46)."""

    def __init__(self, x, y):
        self.lx = np.log10(np.asarray(x, dtype=np.float64) + 1e-20)
        self.ly = np.log10(np.asarray(y, dtype=np.float64) + 1e-20)

    def __call__(self, E):
        l = np.log10(np.asarray(E, dtype=np.float64))
        out = np.interp(l, self.lx, self.ly)
        out = np.where((l < self.lx[0]) | (l > self.lx[-1]), -20.0, out)
        return 10.0 ** out


def _lin0(xg, yg, x):
    x = np.asarray(x, dtype=np.float64)
    return np.where((x < xg[0]) | (x > xg[-1]), 0.0, np.interp(x, xg, yg))


def quadrature_loss_tables(sm):
    """{channel: (x, y, grid)}: the n sigma table of each SM interaction
    integral and the grid its running integral is kept on -- its own, except
    MuonBrem, whose integral is the one of MuonE on the MuonBrem grid
    (shower.py:316)."""
    out = {P: (sm.nsigma_tab[P][0], sm.nsigma_tab[P][1], sm.nsigma_tab[P][0])
           for P in ("Brem", "Ann", "Moller", "Bhabha", "MuonE")}
    out["MuonBrem"] = (sm.nsigma_tab["MuonE"][0], sm.nsigma_tab["MuonE"][1], sm.nsigma_tab["MuonBrem"][0])
    return out


def dark_quadrature(sm, specs):
    """Weights (mode 0) and drate rows (mode 1) of the given jobs on the
    device (petite_dark_quadrature).  Each spec: lx, ly (log-log n sigma of the
    process), zero_small, pid (survival and mean-free-path channels), nodes,
    energies, lo, mode.  Returns per spec the (n,) weights or (n, 10, 2) rows."""
    from . import _lib
    parts, n = [], 0
    cache = {}

    def put(*arrays):
        nonlocal n
        off = n
        for a in arrays:
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            parts.append(a)
            n += a.size
        return off

    def table(key, x, y=None):
        if key not in cache:
            t = _lib.QTable()
            t.off = put(x) if y is None else put(x, y)
            t.n = len(x)
            cache[key] = t
        return cache[key]

    loss = quadrature_loss_tables(sm)
    jobs = (_lib.DarkQuadJob * len(specs))()
    out_n = 0
    for i, sp in enumerate(specs):
        J = jobs[i]
        pid = int(sp["pid"])
        J.ns = table(("ns", i), sp["lx"], sp["ly"])
        J.n_loss = len(SURVIVAL[pid])
        for c, P in enumerate(SURVIVAL[pid]):
            x, y, g = loss[P]
            J.loss[c] = table(("loss", P), x, y)
            J.grid[c] = table(("grid", P), g)
        J.n_mfp = len(MFP_ORDER[pid])
        for c, P in enumerate(MFP_ORDER[pid]):
            J.mfp[c] = table(("mfp", P), *sm.nsigma_tab[P])
        J.nodes = table(("nodes", i), sp["nodes"])
        J.energies = table(("E", i), sp["energies"])
        J.mode = int(sp["mode"])
        J.zero_small = 1 if sp["zero_small"] else 0
        J.lo = float(sp["lo"])
        J.dEdx = float(sm.dEdx)
        J.out_off = out_n
        out_n += len(sp["energies"]) * (20 if J.mode else 1)
    blob = np.concatenate(parts) if parts else np.zeros(1)
    out = np.zeros(max(out_n, 1))
    lib = _lib.load()
    _lib.check(lib.petite_dark_quadrature(jobs, len(specs), blob.ctypes.data, blob.size, out.ctypes.data, out.size),
               "petite_dark_quadrature")
    res = []
    for i, sp in enumerate(specs):
        m = len(sp["energies"])
        o = jobs[i].out_off
        res.append(out[o:o + m].copy() if sp["mode"] == 0 else out[o:o + 20 * m].reshape(m, 10, 2).copy())
    return res


class DarkTables:
    """Dark tables of one material, table set and mass."""

    def __init__(self, sm: MaterialTables, mV, mode="exact", Zeff=29.508, bound_electron=True, kinetic_mixing=1.0,
                 below="clamp", masses="shipped"):
        if not bound_electron:
            raise NotImplementedError("only bound_electron=True (the reference's default) is implemented")
        z, self.mV_list = dark_pack(sm.table_set, masses)
        shipped = [float(m) for m in _load("dark", sm.table_set)[1]["masses"]]
        self.sm = sm
        self.material = sm.material
        if mode == "exact":
            self.mV = closest_lesser_value(self.mV_list, mV, below)
            self.mV_est = self.mV
        elif mode == "approx":
            self.mV = float(mV)
            self.mV_est = closest_lesser_value(self.mV_list, mV, below)
        else:
            raise Exception("Mode not valid. Chose exact or approx.")
        self.Zeff = Zeff
        self.kinetic_mixing = kinetic_mixing
        self.eps2 = kinetic_mixing ** 2  # g_e^2 / (4 pi alpha) (dark_shower.py:109-111)
        key = repr(self.mV_est)
        mat = self.material
        self._generated = float(self.mV_est) not in shipped

        def xs(P):
            return np.array(z[f"xsec__{P}__{key}__{mat}"], dtype=np.float64)

        dbs = xs("DarkBrem")
        while dbs[0, 1] == 0.0:
            dbs = dbs[1:]
        dmb = xs("DarkMuonBrem")
        while dmb[0, 1] == 0.0:
            dmb = dmb[1:]
        self.xsec = {"DarkBrem": dbs, "DarkAnn_free": xs("DarkAnn"), "DarkComp_free": xs("DarkComp"), "DarkMuonBrem": dmb}
        mV_ = self.mV
        self.E_res = (mV_ ** 2 - 2 * m_electron ** 2) / (2 * m_electron)
        self.E_C = mV_ ** 2 / (2 * m_electron) + mV_
        self.min_calc_dark = {11: {"DarkBrem": dbs[0, 0]},
                              -11: {"DarkBrem": dbs[0, 0], "DarkAnn": self.E_res / 1000.0},
                              22: {"DarkComp": self.E_C / 1000.0},
                              13: {"DarkMuonBrem": dmb[0, 0]}, -13: {"DarkMuonBrem": dmb[0, 0]}}
        # bound-electron grids (dark_shower.py:254-292)
        Ea = np.logspace(np.log10(max(sm.min_calc[-11], 0.001 * self.E_res)), np.log10(dbs[-1, 0]), 200)
        self.xsec["DarkAnn"] = np.column_stack([Ea, _atomic_ann(Ea, mV_, Zeff)])
        Ec = np.logspace(np.log10(max(sm.min_calc[22], 0.001 * self.E_C)), np.log10(xs("DarkComp")[-1, 0]), 200)
        self.xsec["DarkComp"] = np.column_stack([Ec, _atomic_comp(Ec, mV_, Zeff)])
        nZ, ne = sm.nZ, sm.ne
        self.ns = {"DarkBrem": LogLog(dbs[:, 0], nZ * GeVsqcm2 * dbs[:, 1]),
                   "DarkAnn": LogLog(Ea, ne * GeVsqcm2 * self.xsec["DarkAnn"][:, 1]),
                   "DarkComp": LogLog(Ec, ne * GeVsqcm2 * self.xsec["DarkComp"][:, 1]),
                   "DarkMuonBrem": LogLog(dmb[:, 0], nZ * GeVsqcm2 * dmb[:, 1])}
        self._z, self._key = z, key
        self.from_cache = f"weights__brem_elec_weights__{key}__{self.material}" in z.files
        self._W = self._D = None
        self.grids = {}
        self.map_source = {}
        fallback = None
        tp = trained_pack(sm.table_set) if sm.table_set != "100GeV" else None
        for P in DARK_PROCESSES:
            gk = f"{P}__{key}"
            if f"E__{gk}" in z.files:
                self.grids[P] = (np.array(z[f"E__{gk}"]), np.array(z[f"nodes__{gk}"], dtype=np.float32),
                                 np.array(z[f"nodecnt__{gk}"]))
                self.map_source[P] = "shipped"
            elif self._generated:
                # a generated mass: its pack holds cross sections only; the
                # maps start flat on the generator's domains at the table
                # energies and DarkShower trains them on the GPU
                self.grids[P] = flat_map(P, np.array(z[f"xsec__{P}__{key}__{mat}"])[:, 0], self.mV_est)
                self.map_source[P] = "untrained"
            elif sm.table_set != "100GeV":
                # the 400 GeV set ships no DarkBrem/DarkMuonBrem maps: the
                # pack of GPU-trained maps when it has them, else the 100 GeV
                # maps of the same mass (their variables are dimensionless;
                # energies above the last grid energy take the last map, as
                # LU_Key clamps)
                if fallback is None:
                    fallback, _ = _load("dark", "100GeV")
                if f"E__{gk}" not in fallback.files:
                    continue
                standin = (np.array(fallback[f"E__{gk}"]), np.array(fallback[f"nodes__{gk}"], dtype=np.float32),
                           np.array(fallback[f"nodecnt__{gk}"]))
                g = trained_grid(tp, gk, standin) if tp is not None else None
                self.grids[P] = g if g is not None else standin
                self.map_source[P] = "trained" if g is not None else "100GeV stand-in"
        # n*sigma of the dark processes for the device (log-log -> tabulate as given)
        self.nsigma_tab = {}

    # -- weights and rates (dark_shower.py:311-593) -----------------------------
    # Track-length integrals over the survival factor of the SM interaction
    # integrals (shower.py:298-354), computed on the device by
    # petite_dark_quadrature on first use.
    @property
    def W(self):
        if self._W is None:
            self._build_weights()
        return self._W

    @property
    def D(self):
        if self._D is None:
            self._build_weights()
        return self._D

    def _energies(self, proc):
        return {"DarkAnn": self.xsec["DarkAnn"][:, 0], "DarkMuonBrem": self.xsec["DarkMuonBrem"][:, 0]}.get(
            proc, self.xsec["DarkBrem"][:, 0])

    def _nodes(self, proc, pid):
        """Panel nodes: the grids of the species' survival integrals and the
        nodes of the process's log-log n sigma table."""
        loss = quadrature_loss_tables(self.sm)
        pts = [loss[P][2] for P in SURVIVAL[pid]]
        pts.append(10.0 ** self.ns[proc].lx)
        return np.unique(np.concatenate(pts))

    def _build_weights(self):
        z, key, mat = self._z, self._key, self.material
        cached = self.from_cache
        W, D, jobs = {}, {}, []
        for name, proc, pid in WEIGHT_FAMILIES:
            Es = self._energies(proc)
            ll = self.ns[proc]
            spec = dict(lx=ll.lx, ly=ll.ly, zero_small=proc != "DarkAnn", pid=pid, nodes=self._nodes(proc, pid),
                        energies=Es, lo=Es[0])
            if cached and name != "annihilation":
                W[name] = np.array(z[f"weights__{name}_weights__{key}__{mat}"])
            else:
                jobs.append((("W", name), dict(spec, mode=0)))
            ck = f"drate__{name}_drate__{key}__{mat}"
            if name != "annihilation" and ck in z.files:
                arr = np.array(z[ck])
                D[name] = (np.array(z[f"drateE__{name}_drate__{key}__{mat}"]), arr[:, :10, :],
                           np.array(z[f"dratecnt__{name}_drate__{key}__{mat}"]))
            else:
                jobs.append((("D", name), dict(spec, mode=1)))
        res = dark_quadrature(self.sm, [j for _, j in jobs])
        for ((kind, name), spec), out in zip(jobs, res):
            Es = spec["energies"]
            if kind == "W":
                W[name] = np.column_stack([Es, out])
            else:
                D[name] = (Es, out, np.full(len(Es), 10))
        self._W, self._D = W, D

    def bsm_weight(self, pid, E0, process):
        """GetBSMWeights (dark_shower.py:595-647) for arrays of (pid, E0)."""
        pid = np.asarray(pid)
        E0 = np.asarray(E0, dtype=np.float64)
        out = np.zeros(np.broadcast(pid, E0).shape)
        sm = self.sm
        if process == "DarkComp":
            m = (pid == 22) & (E0 >= self.min_calc_dark[22]["DarkComp"]) & (E0 >= sm.min_calc[22])
            den = sm.nsigma("PairProd", E0) + sm.nsigma("Comp", E0)
            with np.errstate(all="ignore"):
                v = self.ns["DarkComp"](E0) / den
            out = np.where(m, v, out)
        elif process == "DarkBrem":
            for p, name in ((11, "brem_elec"), (-11, "brem_positron")):
                m = (pid == p) & (E0 >= self.min_calc_dark[p]["DarkBrem"])
                W = self.W[name]
                out = np.where(m, _lin0(W[:, 0], W[:, 1], E0), out)
        elif process == "DarkAnn":
            m = (pid == -11) & (E0 >= self.min_calc_dark[-11]["DarkAnn"])
            W = self.W["annihilation"]
            out = np.where(m, _lin0(W[:, 0], W[:, 1], E0), out)
        elif process == "DarkMuonBrem":
            m = (np.abs(pid) == 13) & (E0 >= self.min_calc_dark[13]["DarkMuonBrem"])
            W = self.W["muon_brem"]
            out = np.where(m, _lin0(W[:, 0], W[:, 1], E0), out)
        return self.eps2 * out


def build_dark_blob(dt: DarkTables, active_codes):
    """Pack the weight arrays, drate tables and the DarkComp n*sigma of a
    DarkTables into (DarkDesc, float64 blob) for petite_dark_setup."""
    from . import _lib
    desc = _lib.DarkDesc()
    parts, n = [], 0

    def put(a):
        nonlocal n
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        off = n
        parts.append(a)
        n += a.size
        return off

    nb = _lib.DARK_DRATE_BINS
    for f, name in enumerate(_lib.DARK_FAMILIES):
        W = np.asarray(dt.W[name], dtype=np.float64)
        if len(W) > 1 and np.any(np.diff(W[:, 0]) <= 0):
            raise ValueError(f"{name}: weight energies must ascend")
        desc.w_n[f] = len(W)
        desc.w_off[f] = put(np.concatenate([W[:, 0], W[:, 1]])) if len(W) else 0
        desc.w_lnE0[f], desc.w_inv_dlog[f] = geometric(W[:, 0]) if len(W) > 1 else (0.0, 0.0)
        Es, rows, cnt = dt.D[name]
        Es = np.asarray(Es, dtype=np.float64)
        order = np.argsort(Es, kind="stable")
        rows = np.asarray(rows, dtype=np.float64)[order]
        cnt = np.asarray(cnt)[order]
        pad = np.zeros((len(Es), nb, 2))
        k = min(nb, rows.shape[1])
        pad[:, :k, :] = rows[:, :k, :]
        desc.d_n[f] = len(Es)
        desc.d_off[f] = put(np.concatenate([Es[order], np.minimum(cnt, nb).astype(np.float64), pad.ravel()]))
    ll = dt.ns["DarkComp"]
    desc.comp_n = len(ll.lx)
    desc.comp_off = put(np.concatenate([ll.lx, ll.ly]))
    d_lx = np.diff(ll.lx)
    if len(ll.lx) > 1 and np.all(d_lx > 0) and np.max(np.abs(d_lx / d_lx.mean() - 1.0)) < 1e-9:
        desc.comp_l0, desc.comp_inv_dl = float(ll.lx[0]), float((len(ll.lx) - 1) / (ll.lx[-1] - ll.lx[0]))
    codes = list(active_codes)
    if len(codes) > 8:
        raise ValueError("at most 8 active dark processes")
    desc.n_active = len(codes)
    for i, c in enumerate(codes):
        desc.active[i] = int(c)
    desc.eps2 = float(dt.eps2)
    if TWO_BODY_CODE in codes:
        # GetBSMWeights of a meson (dark_shower.py:633-640): 2 kappa^2 (1 - r^2)^3 BR(M -> gamma gamma)
        for i, pid in enumerate((111, 221, 331)):
            desc.meson_w[i] = 2.0 * float(dt.kinetic_mixing) ** 2 * meson_twobody_branchingratios[pid]
    desc.E_res, desc.E_C = float(dt.E_res), float(dt.E_C)
    desc.min_brem = float(dt.min_calc_dark[11]["DarkBrem"])
    desc.min_ann = float(dt.min_calc_dark[-11]["DarkAnn"])
    desc.min_comp = float(dt.min_calc_dark[22]["DarkComp"])
    desc.min_muon_brem = float(dt.min_calc_dark[13]["DarkMuonBrem"])
    blob = np.concatenate(parts) if parts else np.zeros(1)
    return desc, blob
