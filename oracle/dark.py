"""Dark-sector reweighting (oracle, float64).

Restates DarkShower (src/PETITE/dark_shower.py): mass selection (158-176),
dark cross sections and thresholds (219-245), bound-electron cross sections
(254-292 with src/PETITE/atomic_annihilation.py and atomic_compton.py),
log-log n*sigma (294-309), the survival factors (src/PETITE/shower.py:298-354),
weight arrays and drate tables with scipy.integrate.quad (311-593),
GetBSMWeights (595-647), draw_dark_sample (649-704), draw_pe_sample (706-719)
and produce_bsm_particle (721-804).  Random numbers come from the dark stream
of each (SM particle, dark process) pair (oracle.philox, PURPOSE_DARK).

Pinned by: the bound-electron closed forms, evaluated in 60-digit arithmetic
(the `m` namespace argument), against the large-b expansions the reference
carries beside them, continuity at the expr1/expr2 branch point and the 1s
momentum-tail integral of the tree level; the stored dark_weights.pkl /
dark_drate.pkl at the level of the SM tables' MC noise (the caches were made
from SM tables other than the shipped sm_xsec.pkl).  The weight/drate quad
integrals here check the product's panel quadrature on the same tables
(tests/test_dark_tables.py).  Sampling (draw_dark_sample, produce_bsm_particle)
reuses the pinned VEGAS and kinematics restatements.
"""
import math

import numpy as np
from scipy.integrate import quad

from . import philox, vegas
from .constants import GeVsqcm2, alpha_em, cmtom, m_electron, m_muon
from .integrands import INTEGRAND
from .kinematics import compton_bound, l_to_lV, radiative_return
from .moliere import scatter
from .shower import lose_energy, rotation_matrix
from .tables import Grid, dark_packs, grid_from, interp0, load_pack

DARK_CODE = {"DarkBrem": 8, "DarkAnn": 9, "DarkComp": 10, "DarkMuonBrem": 11}


def closest_lesser_value(input_list, input_value, below="clamp"):
    arr = np.asarray(input_list)
    index = (np.abs(arr - input_value)).argmin()
    if arr[index] <= input_value:
        return float(arr[index])
    if index == 0 and below == "clamp":
        return float(arr[0])
    return float(arr[index - 1])


class _NumpyMath:
    log, atan, atan2, pi = np.log, np.arctan, np.arctan2, np.pi


def f_ab(a, b):
    return 1 / ((a - b) ** 2 + 1) ** 3


def fancy_integral(a, b, m=_NumpyMath, series=True):
    """atomic_annihilation.py:10-68.  `m` supplies log/atan/atan2/pi (float64
    numpy by default; an mpmath namespace evaluates the same expression in
    extended precision); series=False skips the b >= 1000 expansion."""
    A_ = 1 + a ** 2
    B = (1 + a ** 2) ** 3
    C = 5 + 3 * a ** 2
    D = (1 + a ** 2) ** 2
    E = -1 + 15 * a ** 2 + 12 * a ** 4
    term1 = (1 + a ** 2) * (a * B * C - D * E * b + a * (1 + a ** 2) * (-23 + 38 * a ** 2 + 21 * a ** 4) * b ** 2
                            + 4 * (5 + a ** 2 - 18 * a ** 4 - 6 * a ** 6) * b ** 3 + 3 * a * (3 + a ** 2) * (-1 + 7 * a ** 2) * b ** 4
                            + (7 - 29 * a ** 2 - 12 * a ** 4) * b ** 5 + a * (7 + 3 * a ** 2) * b ** 6)
    term2 = (3 * (1 + a ** 2) ** 6 - a * (1 + a ** 2) ** 3 * (3 + 26 * a ** 2 + 15 * a ** 4) * b
             + 3 * (1 + a ** 2) ** 2 * (-7 + 21 * a ** 2 + 23 * a ** 4 + 11 * a ** 6) * b ** 2
             - 3 * a * (1 + a ** 2) * (7 + 69 * a ** 2 + 45 * a ** 4 + 15 * a ** 6) * b ** 3
             + (-11 + 132 * a ** 2 + 9 * a ** 4 * (38 + 5 * a ** 2 * (4 + a ** 2))) * b ** 4
             - 3 * a * (11 + 73 * a ** 2 + 41 * a ** 4 + 11 * a ** 6) * b ** 5
             + 3 * (-1 + 27 * a ** 2 + 17 * a ** 4 + 5 * a ** 6) * b ** 6 - a * (15 + 10 * a ** 2 + 3 * a ** 4) * b ** 7)
    term3 = term2
    term4 = ((1 + a ** 2) ** 3 + 6 * a * (1 + a ** 2) ** 2 * b - 3 * (1 + 6 * a ** 2 + 5 * a ** 4) * b ** 2
             + 4 * a * (3 + 5 * a ** 2) * b ** 3 - 3 * (1 + 5 * a ** 2) * b ** 4 + 6 * a * b ** 5 - b ** 6)
    numerator = (term1 + term2 * m.atan2(a, 1) + term3 * m.atan2(1 + a ** 2 - a * b, b)
                 + 4 * b * (term4 * (m.log(1 + (a - b) ** 2) - 2 * m.log(b))))
    denominator = 8 * (1 + a ** 2) ** 3 * (1 + (a - b) ** 2) ** 3 * b
    if b < 1000 or not series:
        return numerator / denominator
    return (3081 / 280 - (3081 * a ** 2) / 40) / b ** 8 - (709 * a) / (35 * b ** 7) - 91 / (30 * b ** 6)


def sigma_atomic(k, mV, Zeff):
    """atomic_annihilation.py:70-104 (local alpha = 1/137, m_e = 0.511 MeV)."""
    al, me = 1 / 137, 0.511e-3
    Lambda = Zeff * al * me
    a = me / Lambda
    b = mV ** 2 / (2 * k * Lambda)
    s0 = (4.0 * np.pi * al) * ((mV ** 2 + 2 * me ** 2) * (2 / 3 / Lambda / me) * (1 / k ** 2) * f_ab(a, b))
    s = 2 * k * me + me ** 2
    beta = 2 * al / np.pi * (np.log(s / me ** 2) - 1.0)
    s1 = (4.0 * np.pi * al) * ((mV ** 2 + 2 * me ** 2) * (2 / 3 / Lambda / me) * (1 / k ** 2) * (beta / 2) * fancy_integral(a, b))
    return s0 + s1


def fancy_series(a, b):
    """Large-b expansion used by fancy_integral (atomic_annihilation.py:66-67)."""
    return (3081 / 280 - (3081 * a ** 2) / 40) / b ** 8 - (709 * a) / (35 * b ** 7) - 91 / (30 * b ** 6)


def _expr1(a, b, m=_NumpyMath):
    """atomic_compton.py:6-44."""
    delta = (a - b) ** 2 + 1
    log_term = m.log(delta) - 2 * m.log(b)
    num_part1 = 8 * b * (a ** 2 - 6 * a * b + 1) * log_term
    inner = (-a * (a ** 2 + 1) ** 3 * (3 * a ** 2 + 5) + 6 * (a ** 4 + 4 * a ** 2 - 5) * b ** 5
             - 4 * a * (6 * a ** 4 + 23 * a ** 2 - 31) * b ** 4 + (a - 1) * (a + 1) * (39 * a ** 4 + 190 * a ** 2 + 55) * b ** 3
             - a * (a ** 2 + 1) * (33 * a ** 4 + 98 * a ** 2 - 111) * b ** 2 + (a ** 2 + 1) ** 2 * (15 * a ** 4 + 32 * a ** 2 - 23) * b)
    num_part2 = (a ** 2 + 1) * inner / delta ** 2
    coeff = (3 * (a ** 2 + 1) ** 4 - 2 * a * (3 * a ** 4 + 10 * a ** 2 + 15) * (a ** 2 + 1) * b
             + 6 * (a ** 6 + 5 * a ** 4 + 15 * a ** 2 - 5) * b ** 2)
    num_part3 = coeff * (m.atan(b / (a ** 2 - a * b + 1)) + m.atan(1 / a))
    return -(num_part1 + num_part2 + num_part3) / (8 * (a ** 2 + 1) ** 4 * b)


def comp_series(a, b):
    """Large-b form combine() falls back to (atomic_compton.py:83-86)."""
    return (-145 + 1015 * a ** 2 + 330 * a * b + 64 * b ** 2) / (420 * b ** 8)


def _expr2(a, b, m=_NumpyMath):
    """atomic_compton.py:46-79."""
    delta = (a - b) ** 2 + 1
    denominator = 8 * (a ** 2 + 1) ** 4 * b * delta ** 2
    term1_inner = (a * (a ** 2 + 1) ** 3 * (3 * a ** 2 + 5) - 6 * (a ** 4 + 4 * a ** 2 - 5) * b ** 5
                   + 4 * a * (6 * a ** 4 + 23 * a ** 2 - 31) * b ** 4 - (a - 1) * (a + 1) * (39 * a ** 4 + 190 * a ** 2 + 55) * b ** 3
                   + a * (a ** 2 + 1) * (33 * a ** 4 + 98 * a ** 2 - 111) * b ** 2 - (a ** 2 + 1) ** 2 * (15 * a ** 4 + 32 * a ** 2 - 23) * b)
    term1 = (a ** 2 + 1) * term1_inner
    coeff = (3 * (a ** 2 + 1) ** 4 - 2 * a * (3 * a ** 4 + 10 * a ** 2 + 15) * (a ** 2 + 1) * b
             + 6 * (a ** 6 + 5 * a ** 4 + 15 * a ** 2 - 5) * b ** 2)
    term2 = m.pi * coeff * delta ** 2
    term3 = coeff * delta ** 2 * m.atan(1 / a)
    log_term = m.log(delta) - 2 * m.log(b)
    term4 = delta ** 2 * (8 * b * (a ** 2 - 6 * a * b + 1) * log_term + coeff * m.atan(b / (a ** 2 - a * b + 1)))
    return (term1 + term2 - term3 - term4) / denominator


def sigma_atomic_comp(k, mV, Zeff):
    """atomic_compton.py:187-214."""
    m_e = m_electron
    Lambda = Zeff * alpha_em * m_e
    a = m_e / Lambda
    b = mV ** 2 / (2 * k * Lambda)
    s = 2 * k * m_e + m_e ** 2
    beta = 2 * alpha_em / np.pi * (np.log(s / m_e ** 2) - 1.0)
    if beta < 0.0:
        beta = 0.0
    if a ** 2 + 1 < a * b:
        c = _expr1(a, b)
        if c < 0.0 or (a ** 2 + 1 < 0.01 * a * b):
            c = comp_series(a, b)
    else:
        c = _expr2(a, b)
    return (4.0 * np.pi * alpha_em) * ((mV ** 2 + 2 * m_e ** 2) * (2 / 3 / Lambda / m_e) * (1 / k ** 2) * (beta / 4) * c)


class LogInterp:
    """interpolate1d(xspace='log', yspace='log', fill_value=-20) (dark_shower.py:31-46)."""

    def __init__(self, x, y):
        self.lx = np.log10(np.asarray(x) + 1e-20)
        self.ly = np.log10(np.asarray(y) + 1e-20)

    def __call__(self, x):
        l = np.log10(x)
        v = np.interp(l, self.lx, self.ly)
        return 10 ** np.where((l < self.lx[0]) | (l > self.lx[-1]), -20.0, v)


class OracleDark:
    def __init__(self, sm_tables, mV_in_GeV, mode="exact", Zeff=29.508, kinetic_mixing=1.0, table_set="100GeV",
                 below="clamp", use_cache=True, with_integrals=True, grids=None, masses="shipped"):
        """grids: {process: (E, nodes, node counts)} to use instead of the
        packs' maps (tests pass the product's maps, so both sides draw from
        the same map)."""
        z, self.mV_list = dark_packs(table_set, masses)
        self.T = sm_tables
        T = sm_tables
        if mode == "exact":
            self.mV = closest_lesser_value(self.mV_list, mV_in_GeV, below)
            self.mV_est = self.mV
        else:
            self.mV = mV_in_GeV
            self.mV_est = closest_lesser_value(self.mV_list, mV_in_GeV, below)
        self.Zeff = Zeff
        self.eps2 = kinetic_mixing ** 2
        key, mat = repr(self.mV_est), T.material

        def xs(P):
            return np.array(z[f"xsec__{P}__{key}__{mat}"])

        DBS = xs("DarkBrem")
        while DBS[0, 1] == 0.0:
            DBS = DBS[1:]
        DMB = xs("DarkMuonBrem")
        while DMB[0, 1] == 0.0:
            DMB = DMB[1:]
        self.DBS, self.DMB = DBS, DMB
        self.E_res = (self.mV ** 2 - 2 * m_electron ** 2) / (2 * m_electron)
        self.E_C = self.mV ** 2 / (2 * m_electron) + self.mV
        self.min_dark = {11: {"DarkBrem": DBS[0, 0]}, -11: {"DarkBrem": DBS[0, 0], "DarkAnn": self.E_res / 1000.0},
                         22: {"DarkComp": self.E_C / 1000.0}, 13: {"DarkMuonBrem": DMB[0, 0]},
                         -13: {"DarkMuonBrem": DMB[0, 0]}}
        Emin = max(T.min_calc[-11], 0.001 * self.E_res)
        Ea = np.logspace(np.log10(Emin), np.log10(DBS[-1, 0]), 200)
        self.DAS = np.column_stack([Ea, [sigma_atomic(e, self.mV, Zeff) for e in Ea]])
        Emin = max(T.min_calc[22], 0.001 * self.E_C)
        Ec = np.logspace(np.log10(Emin), np.log10(xs("DarkComp")[-1, 0]), 200)
        self.DCS = np.column_stack([Ec, [sigma_atomic_comp(e, self.mV, Zeff) for e in Ec]])
        nZ, ne = T.nZ, T.ne
        self.NS = {"DarkBrem": LogInterp(DBS[:, 0], nZ * GeVsqcm2 * DBS[:, 1]),
                   "DarkAnn": LogInterp(Ea, ne * GeVsqcm2 * self.DAS[:, 1]),
                   "DarkComp": LogInterp(Ec, ne * GeVsqcm2 * self.DCS[:, 1]),
                   "DarkMuonBrem": LogInterp(DMB[:, 0], nZ * GeVsqcm2 * DMB[:, 1])}
        # interaction integrals via quad of the interpolants (shower.py:298-320)
        self.II = {}
        for P in (("Brem", "Ann", "Moller", "Bhabha", "MuonE") if with_integrals else ()):
            xg = T.nsigma_tab[P][0]
            f = lambda e, P=P: float(T.nsigma(P, e))
            vals = np.array([quad(f, xg[0], xg[i], full_output=1, limit=200, points=xg[:i + 1][1:-1] if i > 1 else None)[0]
                             if i > 0 else 0.0 for i in range(len(xg))])
            self.II[P] = (xg, vals)
        xg = T.nsigma_tab["MuonBrem"][0]
        f = lambda e: float(T.nsigma("MuonE", e))
        mx = T.nsigma_tab["MuonE"][0]
        if with_integrals:
            self.II["MuonBrem"] = (xg, np.array([quad(f, xg[0], xg[i], full_output=1, limit=200,
                                                      points=mx[(mx > xg[0]) & (mx < xg[i])])[0] if i > 0 else 0.0
                                                 for i in range(len(xg))]))
        self.dEdxT = T.dEdx
        self.z, self.key = z, key
        self.use_cache = use_cache
        self._grids = {}
        for P in DARK_CODE:
            gk = f"{P}__{key}"
            if grids is not None and P in grids:
                self._grids[P] = grid_from(*grids[P])
                continue
            src = z
            if f"E__{gk}" not in z.files and table_set != "100GeV":
                src, _ = load_pack("dark", "100GeV")  # maps missing from the 400 GeV set
            if f"E__{gk}" in src.files:
                self._grids[P] = Grid(src[f"E__{gk}"], src[f"nodes__{gk}"], src[f"nodeoff__{gk}"], src[f"nodecnt__{gk}"])

    def _ii(self, P, E):
        xg, yg = self.II[P]
        return interp0(xg, yg, E)

    def exp_factor(self, pid, E, Ei):
        chans = {11: ("Brem", "Moller"), -11: ("Brem", "Ann", "Bhabha"), 13: ("MuonBrem", "MuonE")}[pid if pid != -13 else 13]
        d = sum(float(self._ii(P, Ei) - self._ii(P, E)) for P in chans)
        if d < 0.0 or E > Ei:
            return 0.0
        return math.exp(-d / self.dEdxT / cmtom)

    def integrand(self, proc, pid, E, Ei):
        k = self.T.dEdx * cmtom
        n = float(self.NS[proc](E))
        if proc != "DarkAnn" and n < 1.0e-18:
            return 0.0
        return n / k * self.exp_factor(pid, E, Ei)

    def weight(self, proc, pid, Ei, E0):
        mfp = self.T.mfp(pid if pid != -13 else 13, Ei)
        eb = max(Ei - 10 * mfp * self.dEdxT, E0)
        f = lambda e: self.integrand(proc, pid, e, Ei)
        return (quad(f, E0, eb, full_output=1)[0] + quad(f, eb, Ei, full_output=1)[0])

    def drate(self, proc, pid, Ei, lo_table):
        mfp = self.T.mfp(pid if pid != -13 else 13, Ei)
        e = np.linspace(max(Ei - 10 * mfp * self.dEdxT, lo_table), Ei, 11)
        f = lambda x: self.integrand(proc, pid, x, Ei)
        w = np.array([quad(f, e[i], e[i + 1], full_output=1)[0] for i in range(10)])
        return np.column_stack([0.5 * (e[1:] + e[:-1]), w])


# ------------------------------------------------------------------ sampling --
# Random numbers of one (SM particle, dark process) pair come from its dark
# stream dsid = derive_sid(seed, sid, code, PURPOSE_DARK):
#   INTERACT block 0: x drate bin, y MCS radius, z MCS azimuth (sign: bit 0 of z), w kinematic azimuth
#   INTERACT block 1: x bound-electron cos(theta_e) (2u - 1), y rotation azimuth of the bound-electron frame
#   BOUND_E block k:  (x, y) and (z, w) accept/reject pairs of the bound-electron momentum
#   VEGAS_A: the accept/reject candidates of draw_dark_sample
W_FAMILY = {("DarkBrem", 11): "brem_elec", ("DarkBrem", -11): "brem_positron", ("DarkAnn", -11): "annihilation",
            ("DarkMuonBrem", 13): "muon_brem", ("DarkMuonBrem", -13): "muon_brem"}
PE_MAX_PAIRS = 1 << 16


def electron_wave_function(pe, Zeff):
    """dark_shower.py:706-708."""
    lam = alpha_em * Zeff * m_electron
    return 32 / np.pi * lam ** 5 * pe ** 2 / (pe ** 2 + lam ** 2) ** 4


def draw_pe(seed, dsid, Zeff):
    """draw_pe_sample (dark_shower.py:710-719): uniform pe in [0, 1 MeV), accepted
    against the 1s momentum density normalised at its maximum."""
    c = electron_wave_function(alpha_em * Zeff * m_electron / math.sqrt(3), Zeff)
    for k in range(PE_MAX_PAIRS // 2):
        u = philox.to_uniform(philox.words(seed, dsid, k, philox.PURPOSE_BOUND_E))
        for x, a in ((u[0], u[1]), (u[2], u[3])):
            pe = 1e-3 * x
            if a < electron_wave_function(pe, Zeff) / c:
                return pe
    return None


def drate_pick(Es, rows, cnt, E0, u):
    """Interaction energy of produce_bsm_particle (dark_shower.py:741-750): the
    saved energy Ei (largest <= E0, clamped to the table), a bin chosen with
    probability rate/sum (numpy.random.choice: first cdf entry above u),
    shifted by E0 - Ei.  None when the rates sum to zero."""
    Es = np.asarray(Es)
    if E0 < Es.min():
        i = int(np.argmin(Es))
    elif E0 > Es.max():
        i = int(np.argmax(Es))
    else:
        i = int(np.nonzero(Es == np.max(Es[Es <= E0]))[0][0])
    c = int(cnt[i])
    e_c, rate = rows[i, :c, 0], rows[i, :c, 1]
    tot = np.sum(rate)
    if tot == 0.0:
        return None
    cdf = np.cumsum(rate / tot)
    cdf = cdf / cdf[-1]
    k = min(int(np.searchsorted(cdf, u, side="right")), c - 1)
    return e_c[k] + (E0 - Es[i])


class OracleDarkSampler:
    """GetBSMWeights (dark_shower.py:595-647), produce_bsm_particle (721-804) and
    the loop of generate_dark_shower (806-849) over a fixed SM record.

    Inputs: an OracleDark (cross sections and thresholds), the weight arrays
    W[name] = (n, 2) and drate tables D[name] = (energies, rows (n, 10, 2),
    counts) that DarkShower builds at initialisation, and the max_F arrays of
    the dark maps."""

    def __init__(self, od, W, D, maxf, rescale_MCS=1.0, maxF_fudge=1.0, max_trials=int(1e4) * 300):
        self.od, self.T = od, od.T
        self.W, self.D, self.maxf = W, D, maxf
        self.rescale, self.fudge, self.max_trials = rescale_MCS, maxF_fudge, max_trials

    def weight(self, pid, E0, process):
        od, T = self.od, self.T
        if pid not in (11, -11, 13, -13, 22):
            return 0.0
        thr = od.min_dark[pid]
        if process not in thr or E0 < thr[process]:
            return 0.0
        if pid == 22:
            if process != "DarkComp" or E0 < T.min_calc[22]:
                return 0.0
            return od.eps2 * float(od.NS["DarkComp"](E0)) / float(T.nsigma("PairProd", E0) + T.nsigma("Comp", E0))
        name = W_FAMILY.get((process, pid))
        if name is None:
            return 0.0
        W = self.W[name]
        return od.eps2 * float(interp0(W[:, 0], W[:, 1], E0))

    def produce(self, rec, process, wg, seed):
        """rec: dict with pid, p0, pf, rf, weight, sid.  Returns the V record or None."""
        od, T = self.od, self.T
        code = DARK_CODE[process]
        pid = rec["pid"]
        dsid = philox.derive_sid(seed, rec["sid"], code, philox.PURPOSE_DARK)
        w0 = philox.words(seed, dsid, 0, philox.PURPOSE_INTERACT)
        u0 = philox.to_uniform(w0)
        p0 = np.asarray(rec["p0"], dtype=np.float64)
        pf = np.asarray(rec["pf"], dtype=np.float64)
        name = W_FAMILY.get((process, pid)) if process != "DarkComp" else None
        if name is not None:
            E0 = p0[0]
            E_int = drate_pick(*self.D[name], E0, u0[0])
            if E_int is None:
                return None
            dist = (E0 - E_int) / T.dEdx
            sign = 1.0 if int(w0[2]) & 1 else -1.0
            # beta from the on-shell momentum of the record's energy and mass
            pn = float(np.linalg.norm(p0[1:]))
            pon = math.sqrt(max(E0 * E0 - rec["mass"] ** 2, 0.0))
            pin = np.concatenate([[E0], p0[1:] * (pon / pn)]) if pn > 0 else p0
            ps = scatter(pin, T.rho * (dist / cmtom), T.A, T.Z, self.rescale, m_electron, u0[1], u0[2], sign)
            pf = lose_energy(ps, rec["mass"], E0 - E_int)
        E = pf[0]
        RM = rotation_matrix(pf)
        bound = process in ("DarkAnn", "DarkComp")
        mV = od.mV
        if process == "DarkAnn" and E <= od.E_res:
            V = np.array([math.sqrt(E ** 2 - m_electron ** 2 + mV ** 2), 0.0, 0.0, math.sqrt(E ** 2 - m_electron ** 2)])
        elif process == "DarkComp" and E <= od.E_C:
            V = np.array([math.sqrt(E ** 2 + mV ** 2), 0.0, 0.0, E])
        elif process == "DarkAnn" and not mV > 2 * m_electron:
            # the radiative-return integrand carries sqrt(1 - 4 m_e^2/m_V^2)
            # (all_processes.py:431): NaN at every point, so draw_dark_sample's
            # loop ends without a sample
            return None
        else:
            grid = od._grids[process]
            ie = grid.lu_key(E)
            ev = T.event_info(process, E, mV)
            ev["Zeff"] = od.Zeff
            x, trials, _ = vegas.draw(grid, process, E, ev, self.maxf[process][ie] * self.fudge, seed, dsid,
                                      max_trials=self.max_trials)
            if x is None:
                return None
            if process == "DarkComp":
                u1 = philox.to_uniform(philox.words(seed, dsid, 1, philox.PURPOSE_INTERACT))
                pe = draw_pe(seed, dsid, od.Zeff)
                if pe is None:
                    return None
                c0 = -1.0 + 2.0 * u1[0]
                s = m_electron ** 2 + 2 * E * (math.sqrt(m_electron ** 2 + pe ** 2) - c0 * pe)
                Ee = (s - mV ** 2 + m_electron ** 2) / (2 * math.sqrt(s))
                if Ee < m_electron:
                    V = np.array([mV, 0.0, 0.0, 0.0])
                    wg = 0.0
                else:
                    V = compton_bound(E, x, mV, pe, c0, u0[3], u1[1])
            elif process == "DarkAnn":
                V = radiative_return(E, x, mV)
            else:
                V = l_to_lV(E, rec["mass"], x, u0[3], mV)[1]
        p4 = np.concatenate([[V[0]], RM @ V[1:]])
        return {"p": p4, "r": np.asarray(rec["rf"], dtype=np.float64)[:3], "weight": wg * rec["weight"],
                "process": process + ("_bound" if bound else ""), "code": code}

    def generate(self, records, processes, seed):
        """generate_dark_shower's loop: one V per (SM particle, active process) with weight > 0."""
        out = []
        for i, rec in enumerate(records):
            for P in processes:
                wg = self.weight(rec["pid"], rec["p0"][0], P)
                if wg > 0.0:
                    v = self.produce(rec, P, wg, seed)
                    if v is not None:
                        v["parent"] = i
                        out.append(v)
        return out


# ------------------------------------------------------- m_V-continuous yield --
def sampler_weights(s, pid, E0, process):
    """OracleDarkSampler.weight over arrays of (pid, E0)."""
    pid = np.asarray(pid)
    E0 = np.asarray(E0, dtype=np.float64)
    od, T = s.od, s.T
    out = np.zeros(E0.shape)
    if process == "DarkComp":
        m = (pid == 22) & (E0 >= od.min_dark[22]["DarkComp"]) & (E0 >= T.min_calc[22])
        if m.any():
            e = E0[m]
            out[m] = od.eps2 * od.NS["DarkComp"](e) / (T.nsigma("PairProd", e) + T.nsigma("Comp", e))
        return out
    for (P, p), name in W_FAMILY.items():
        if P != process:
            continue
        m = (pid == p) & (E0 >= od.min_dark[p][process])
        if m.any():
            W = s.W[name]
            out[m] = od.eps2 * interp0(W[:, 0], W[:, 1], E0[m])
    return out


BOUND_PROCESSES = ("DarkAnn", "DarkComp")


def bound_tables(T, mV, table_set="100GeV", Zeff=29.508, masses="all"):
    """The bound-electron tables DarkShower builds for its own mass
    (dark_shower.py:254-292, mode "approx": cross sections at mV itself, grid
    tops from the trained mass at or below it) and the DarkAnn track-length
    weights by the panel rule (oracle.dark_panels)."""
    from .dark_panels import Panels
    od = OracleDark(T, mV, mode="approx", Zeff=Zeff, table_set=table_set, with_integrals=False, masses=masses)
    return od, Panels(od).weight_array("DarkAnn", -11)


def bound_weights(od, Wa, pid, E0, process):
    """GetBSMWeights of DarkAnn / DarkComp (595-647) with eps^2 = 1 from bound_tables."""
    pid = np.asarray(pid)
    E0 = np.asarray(E0, dtype=np.float64)
    T = od.T
    out = np.zeros(E0.shape)
    if process == "DarkComp":
        m = (pid == 22) & (E0 >= od.min_dark[22]["DarkComp"]) & (E0 >= T.min_calc[22])
        if m.any():
            e = E0[m]
            out[m] = od.NS["DarkComp"](e) / (T.nsigma("PairProd", e) + T.nsigma("Comp", e))
        return out
    m = (pid == -11) & (E0 >= od.min_dark[-11]["DarkAnn"])
    if m.any():
        out[m] = interp0(Wa[:, 0], Wa[:, 1], E0[m])
    return out


def interp_with_derivative(x, y, dx, dy, v):
    """numpy.interp(v, x, y), 0 outside [x0, x_n-1], and its derivative in a
    parameter that moves the nodes (dx) and values (dy):
    y'_lo + s (y'_hi - y'_lo) + (y_hi - y_lo) s', s' = -[x'_lo (1-s) + x'_hi s]/(x_hi - x_lo)."""
    v = np.asarray(v, dtype=np.float64)
    inside = (v >= x[0]) & (v <= x[-1])
    lo = np.clip(np.searchsorted(x, v, side="right") - 1, 0, len(x) - 2)
    xl, xh, yl, yh = x[lo], x[lo + 1], y[lo], y[lo + 1]
    val = (yh - yl) / (xh - xl) * (v - xl) + yl
    sw = (v - xl) / (xh - xl)
    sd = -(dx[lo] * (1.0 - sw) + dx[lo + 1] * sw) / (xh - xl)
    der = dy[lo] + sw * (dy[lo + 1] - dy[lo]) + (yh - yl) * sd
    top = v == x[-1]
    val = np.where(top, y[-1], val)
    der = np.where(top, dy[-1], der)
    return np.where(inside, val, 0.0), np.where(inside, der, 0.0)


def bound_yield(T, processes, pid, E0, w, mV, table_set="100GeV", rel_step=1e-5, Zeff=29.508, masses="all"):
    """Y and dY/dm_V of the bound-electron processes with tables at each m_V
    itself.  The derivative is that of the tables' interpolants at m_V: node
    positions and values differentiated by central differences of the tables
    at m_V (1 +- rel_step), the interval weights analytically
    (interp_with_derivative; DarkComp in log10 space, W = 10^L / den)."""
    procs = [P for P in processes if P in BOUND_PROCESSES]
    pid = np.asarray(pid)
    E0 = np.asarray(E0, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    Y, dY = [], []
    for m in np.atleast_1d(mV):
        (od, Wa), (op, Wp), (om, Wm) = (bound_tables(T, mm, table_set, Zeff, masses)
                                        for mm in (m, m * (1 + rel_step), m * (1 - rel_step)))
        f = 1.0 / (2 * rel_step * m)
        y = dy = 0.0
        if "DarkAnn" in procs:
            sel = (pid == -11) & (E0 >= od.min_dark[-11]["DarkAnn"])
            val, der = interp_with_derivative(Wa[:, 0], Wa[:, 1], (Wp[:, 0] - Wm[:, 0]) * f, (Wp[:, 1] - Wm[:, 1]) * f,
                                              E0[sel])
            y += float(np.sum(val * w[sel]))
            dy += float(np.sum(der * w[sel]))
        if "DarkComp" in procs:
            sel = (pid == 22) & (E0 >= od.min_dark[22]["DarkComp"]) & (E0 >= T.min_calc[22])
            e = E0[sel]
            den = T.nsigma("PairProd", e) + T.nsigma("Comp", e)
            ns, nsp, nsm = od.NS["DarkComp"], op.NS["DarkComp"], om.NS["DarkComp"]
            l = np.log10(e)
            inside = (l >= ns.lx[0]) & (l <= ns.lx[-1])
            L, dL = interp_with_derivative(ns.lx, ns.ly, (nsp.lx - nsm.lx) * f, (nsp.ly - nsm.ly) * f, l)
            L = np.where(inside, L, -20.0)
            Wc = 10.0 ** L / den
            y += float(np.sum(Wc * w[sel]))
            dy += float(np.sum(np.where(inside, Wc * math.log(10.0) * dL, 0.0) * w[sel]))
        Y.append(y)
        dY.append(dy)
    return np.array(Y), np.array(dY)


def yield_and_grad(samplers, masses, processes, pid, E0, w, mV, bound_T=None, table_set="100GeV"):
    """Y(m_V) = sum_i w_i sum_P W_P(E0_i; m_V) and dY/dm_V with each particle's
    weight interpolated between the trained masses: geometric in
    t = ln(m_V/m_j)/ln(m_{j+1}/m_j) where both weights are positive, linear
    otherwise; t clamped to [0, 1] (zero derivative) outside the masses.
    samplers[j] evaluates the weights at masses[j] with eps^2 = 1.  With
    bound_T (the SM tables), DarkAnn and DarkComp use tables at each m_V
    instead (bound_yield)."""
    if bound_T is not None:
        rest = [P for P in processes if P not in BOUND_PROCESSES]
        Yb, dYb = bound_yield(bound_T, processes, pid, E0, w, mV, table_set)
        Yr, dYr = yield_and_grad(samplers, masses, rest, pid, E0, w, mV) if rest else (0.0, 0.0)
        return Yb + Yr, dYb + dYr
    masses = np.asarray(masses, dtype=np.float64)
    wts = np.array([[sampler_weights(samplers[j], pid, E0, P) for P in processes] for j in range(len(masses))])
    w = np.asarray(w, dtype=np.float64)
    Y, dY = [], []
    for m in np.atleast_1d(mV):
        j = 0
        while j + 1 < len(masses) - 1 and masses[j + 1] <= m:
            j += 1
        if len(masses) == 1:
            lo = hi = wts[0]
            t = dtdm = 0.0
        else:
            dl = math.log(masses[j + 1] / masses[j])
            t, dtdm = math.log(m / masses[j]) / dl, 1.0 / (m * dl)
            if t <= 0.0:
                t = 0.0
                dtdm = 0.0 if m < masses[0] else dtdm
            if t >= 1.0:
                t = 1.0
                dtdm = 0.0 if m > masses[-1] else dtdm
            lo, hi = wts[j], wts[j + 1]
        both = (lo > 0) & (hi > 0)
        with np.errstate(all="ignore"):
            r = np.where(both, np.log(hi / lo), 0.0)
            Wg = lo * np.exp(t * r)
        W = np.where(both, Wg, lo + t * (hi - lo))
        dW = np.where(both, Wg * r * dtdm, (hi - lo) * dtdm)
        Y.append(float(np.sum(W.sum(0) * w)))
        dY.append(float(np.sum(dW.sum(0) * w)))
    return np.array(Y), np.array(dY)


# ------------------------------------------------------------- meson decays --
MESON_MASS = {111: 134.9768e-3, 221: 547.862e-3, 331: 957.78e-3}
MESON_BR_GG = {111: 0.98823, 221: 0.3936, 331: 0.02307}   # particle.py:30-48
TWO_BODY_CODE = 12


def meson_weight(pid, m_meson, mV, kinetic_mixing=1.0):
    """GetBSMWeights of a meson for TwoBody_BSMDecay (dark_shower.py:633-640)."""
    r = mV / m_meson
    if r >= 1.0:
        return 0.0
    return 2 * kinetic_mixing ** 2 * (1.0 - r ** 2) ** 3 * MESON_BR_GG[pid]


def boost_matrix(pf, m0):
    """Particle.boost_matrix (particle.py:187-207): rest frame -> lab along pf."""
    E0, px0, py0, pz0 = (float(v) for v in pf)
    gamma = E0 / m0
    beta = 1.0 if gamma == 1.0 else np.sqrt(1.0 - 1.0 / gamma ** 2)
    pmag = np.linalg.norm([px0, py0, pz0])
    if pmag == 0.0:
        return np.identity(4)
    bx, by, bz = beta * np.array([px0, py0, pz0]) / pmag
    b2 = beta ** 2
    return np.array([[gamma, gamma * bx, gamma * by, gamma * bz],
                     [gamma * bx, 1 + (gamma - 1) * bx ** 2 / b2, (gamma - 1) * bx * by / b2, (gamma - 1) * bx * bz / b2],
                     [gamma * by, (gamma - 1) * by * bx / b2, 1 + (gamma - 1) * by ** 2 / b2, (gamma - 1) * by * bz / b2],
                     [gamma * bz, (gamma - 1) * bz * bx / b2, (gamma - 1) * bz * by / b2, 1 + (gamma - 1) * bz ** 2 / b2]])


def two_body_vector(pf, m_meson, mV, cos_theta, phi):
    """Lab 4-vector of V in M -> gamma V (two_body_decay, particle.py:209-254,
    gamma first, V second), for given rest-frame angles."""
    E2 = (m_meson ** 2 - 0.0 + mV ** 2) / (2 * m_meson)
    E1 = (m_meson ** 2 - mV ** 2 + 0.0) / (2 * m_meson)
    pF = np.sqrt(E1 ** 2)
    sin_theta = np.sqrt(1 - cos_theta ** 2)
    p2_rf = [E2, pF * sin_theta * np.sin(phi), pF * sin_theta * np.cos(phi), pF * cos_theta]
    return np.dot(boost_matrix(pf, m_meson), p2_rf)


def produce_meson(rec, mV, seed, kinetic_mixing=1.0):
    """The dark vector of one meson record (dark_shower.py:836-844) with the
    angles from its dark stream dsid = derive_sid(seed, sid, 12, DARK):
    INTERACT block 0, cos theta = 2x - 1, phi = 2 pi y.  Returns None when
    m_V >= m_M, else {"p", "r", "weight"}.  The meson's mass is its PDG value."""
    pid = rec["pid"]
    m = MESON_MASS[pid]
    wg = meson_weight(pid, m, mV, kinetic_mixing)
    if wg <= 0.0:
        return None
    dsid = philox.derive_sid(seed, rec["sid"], TWO_BODY_CODE, philox.PURPOSE_DARK)
    u = philox.block(seed, dsid, 0, philox.PURPOSE_INTERACT)
    p = two_body_vector(rec["pf"], m, mV, -1.0 + 2.0 * u[0], 2.0 * np.pi * u[1])
    return {"p": p, "r": np.asarray(rec["rf"], dtype=np.float64), "weight": wg * rec["weight"]}
