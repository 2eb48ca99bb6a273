"""Differential cross sections in the VEGAS map variables (oracle, float64).

Each function takes ``x`` of shape (n, ndim) and an ``ev`` dict of event
information (keys as in src/PETITE/shower.py:435-439 and
dark_shower.py:672-676: E_inc, Z_T, A_T, mT, mV, Eg_min, Ee_min, m_lepton,
Zeff) and returns dsigma of shape (n,), in GeV^-2 per unit map volume, such
that the integral over a process's map domain is the tabulated total cross
section (utilities/find_maxes.py:102-110).  The formulas are restated from
src/PETITE/all_processes.py; line numbers are given per function.

Pinned by: the sigma-table integral identity (tests/test_oracle_tables.py),
closed-form total cross sections for Moller/Bhabha/mu-e and Klein-Nishina
for Compton (tests/test_oracle_closed_forms.py).
"""
import numpy as np

from .constants import alpha_em, m_electron, m_muon, m_proton, GeV

PI = np.pi


def _aa(Z):
    # all_processes.py:92-94
    return 184.15 * (2.718) ** -0.5 * Z ** (-1.0 / 3.0) / m_electron


def g2_elastic(Z, t):
    # all_processes.py:99-103
    a0 = _aa(Z)
    return Z ** 2 * a0 ** 4 * t ** 2 / (1 + a0 ** 2 * t) ** 2


def brem_vars(x, ev):
    """The Brem / MuonBrem integrand's kinematic variables of the map point x
    (all_processes.py:139-172): photon energy w, the scaled angles d, d' of
    photon and lepton, the azimuth ph between them, the final lepton energy
    epp and the momentum transfer q^2 to the nucleus."""
    ep = ev["E_inc"]
    egmin = ev["Eg_min"]
    m = ev["m_lepton"]
    x1, x2, x3, x4 = x.T
    w = egmin + x1 * (ep - m - egmin)
    d = ep / (2 * m) * (x2 + x3)
    dp = ep / (2 * m) * (x2 - x3)
    ph = (x4 - 0.5) * 2 * PI
    epp = ep - w
    with np.errstate(all="ignore"):
        qsq = m ** 2 * ((d ** 2 + dp ** 2 - 2 * d * dp * np.cos(ph))
                        + m ** 2 * ((1 + d ** 2) / (2 * ep) - (1 + dp ** 2) / (2 * epp)) ** 2)
    return w, d, dp, ph, epp, qsq


def brem(x, ev):
    """Brem / MuonBrem, all_processes.py:139-206."""
    ep = ev["E_inc"]
    egmin = ev["Eg_min"]
    m = ev["m_lepton"]
    w, d, dp, ph, epp, qsq = brem_vars(x, ev)
    ok = (egmin < w) & (w < ep - m) & (m < epp) & (epp < ep) & (d > 0.0) & (dp > 0.0)
    with np.errstate(all="ignore"):
        PF = 8.0 / PI * alpha_em * (alpha_em / m) ** 2 * (epp * m ** 4) / (w * ep * qsq ** 2) * d * dp
        jac = PI * ep ** 2 * (ep - m - egmin) / m ** 2
        FF = g2_elastic(ev["Z_T"], qsq)
        T1 = d ** 2 / (1 + d ** 2) ** 2
        T2 = dp ** 2 / (1 + dp ** 2) ** 2
        T3 = w ** 2 / (2 * ep * epp) * (d ** 2 + dp ** 2) / ((1 + d ** 2) * (1 + dp ** 2))
        T4 = -(epp / ep + ep / epp) * (d * dp * np.cos(ph)) / ((1 + d ** 2) * (1 + dp ** 2))
        out = PF * (T1 + T2 + T3 + T4) * jac * FF
    return np.where(ok, out, 0.0)


def pairprod_vars(x, ev):
    """The PairProd integrand's kinematic variables of the map point x
    (all_processes.py:531-580): positron and electron energies epp, epm, their
    scaled angles dp, dm, the azimuth ph between them and the momentum
    transfer q^2 / m_e^2 (qq) to the nucleus."""
    w = ev["E_inc"]
    m = m_electron
    x1, x2, x3, x4 = x.T
    epp = m + x1 * (w - 2 * m)
    dp = w / (2 * m) * (x2 + x3)
    dm = w / (2 * m) * (x2 - x3)
    ph = x4 * 2 * PI
    epm = w - epp
    with np.errstate(all="ignore"):
        qq = (dp ** 2 + dm ** 2 + 2.0 * dp * dm * np.cos(ph)) + m ** 2 * (
            (1.0 + dp ** 2) / (2.0 * epp) + (1.0 + dm ** 2) / (2.0 * epm)) ** 2
    return epp, epm, dp, dm, ph, qq


def pairprod(x, ev):
    """PairProd, all_processes.py:531-622."""
    w = ev["E_inc"]
    m = m_electron
    epp, epm, dp, dm, ph, qq = pairprod_vars(x, ev)
    ok = (m < epm) & (epm < w) & (m < epp) & (epp < w) & (dm > 0.0) & (dp > 0.0)
    with np.errstate(all="ignore"):
        PF = 8.0 / PI * alpha_em * (alpha_em / m) ** 2 * epp * epm / (w ** 3 * qq ** 2) * dp * dm
        jac = PI * w ** 2 * (w - 2 * m) / m ** 2
        FF = g2_elastic(ev["Z_T"], m ** 2 * qq)
        T1 = -1.0 * dp ** 2 / (1.0 + dp ** 2) ** 2
        T2 = -1.0 * dm ** 2 / (1.0 + dm ** 2) ** 2
        T3 = w ** 2 / (2.0 * epp * epm) * (dp ** 2 + dm ** 2) / ((1.0 + dp ** 2) * (1.0 + dm ** 2))
        T4 = (epp / epm + epm / epp) * (dp * dm * np.cos(ph)) / ((1.0 + dp ** 2) * (1.0 + dm ** 2))
        out = PF * (T1 + T2 + T3 + T4) * jac * FF
    return np.where(ok, out, 0.0)


def compton_t(x, ev):
    """The Mandelstam t the Compton / DarkComp integrand assigns to cos(theta)
    = x[:, 0] (all_processes.py:660-668)."""
    Eg = ev["E_inc"]
    mV = ev.get("mV", 0.0)
    m = m_electron
    ct = x[:, 0]
    s = m ** 2 + 2 * Eg * m
    R = np.sqrt(m ** 4 + (mV ** 2 - s) ** 2 - 2 * m ** 2 * (mV ** 2 + s))
    return -0.5 * (m ** 4 + s * (-(mV ** 2) + s + ct * R) - m ** 2 * (mV ** 2 + 2 * s + ct * R)) / s


def compton(x, ev):
    """Comp / DarkComp (free electron at rest), all_processes.py:624-742."""
    Eg = ev["E_inc"]
    mV = ev.get("mV", 0.0)
    m = m_electron
    ct = x[:, 0]
    s = m ** 2 + 2 * Eg * m
    if s < (m + mV) ** 2:
        return np.zeros(len(ct))
    jac = (s - m ** 2) / (2 * s) * np.sqrt((s - mV ** 2) ** 2 - 2 * m ** 2 * (s + mV ** 2) + m ** 4)
    t = compton_t(x, ev)
    PF = 2.0 * PI * alpha_em ** 2 / (s - m ** 2) ** 2
    if mV == 0.0:
        T1 = (6.0 * m ** 2 * s + 3.0 * m ** 4 - s ** 2) / ((m ** 2 - s) * (-(m ** 2) + s + t))
        T2 = 4 * m ** 4 / (s + t - m ** 2) ** 2
        T3 = (t * (s - m ** 2) + (s + m ** 2) ** 2) / (s - m ** 2) ** 2
    else:
        T1 = (2.0 * m ** 2 * (mV ** 2 - 3 * s) - 3 * m ** 4 - 2 * mV ** 2 * s + 2 * mV ** 4 + s ** 2) / (
            (m ** 2 - s) * (m ** 2 + mV ** 2 - s - t))
        T2 = (2 * m ** 2 * (2 * m ** 2 + mV ** 2)) / (m ** 2 + mV ** 2 - s - t) ** 2
        T3 = ((m ** 2 + s) * (m ** 2 + mV ** 2 + s) + t * (s - m ** 2)) / (m ** 2 - s) ** 2
    return PF * jac * (T1 + T2 + T3)


def annihilation(x, ev):
    """Ann (e+ e- -> gamma gamma/V on an electron at rest), all_processes.py:468-529."""
    Ee = ev["E_inc"]
    mV = ev.get("mV", 0.0)
    m = m_electron
    s = 2.0 * m * (Ee + m)
    egmin = ev.get("Eg_min", 0.0)
    ctMax = np.sqrt((Ee + m) / (Ee - m)) * (2 * m * (Ee - 2 * egmin + m) - mV ** 2) / (2 * m * (Ee + m) - mV ** 2)
    ct = x[:, 0]
    if s < mV ** 2:
        return np.zeros(len(ct))
    b = np.sqrt(1.0 - 4.0 * m ** 2 / s)
    val = 4.0 * PI * alpha_em ** 2 / (s * (1 - b ** 2 * ct ** 2)) * (
        (s - mV ** 2) / (2 * s) * (1 + ct ** 2) + 2.0 * mV ** 2 / (s - mV ** 2))
    return np.where(ct > ctMax, 0.0, val)


def moller(x, ev):
    """Moller, all_processes.py:786-840."""
    Ee = ev["E_inc"]
    DE = ev.get("Ee_min", 0.010)
    m = m_electron
    dct = 2.0 * DE / (Ee - m)
    ct = x[:, 0]
    ok = (ct > -1 + dct) & (ct < 1.0 - dct)
    s = m ** 2 + 2 * Ee * m
    with np.errstate(all="ignore"):
        val = (16 * PI ** 2 * alpha_em ** 2 * (s ** 2 * (3 + ct ** 2) ** 2 - 8 * m ** 2 * s * (7 + ct ** 4)
                                               + 16 * m ** 4 * (6 - 3 * ct ** 2 + ct ** 4))
               / (8 * PI * s * (s - 4 * m ** 2) ** 2 * (1 - ct) ** 2 * (1 + ct) ** 2))
    return np.where(ok, val, 0.0)


def bhabha(x, ev):
    """Bhabha, all_processes.py:947-1007."""
    Ee = ev["E_inc"]
    DE = ev.get("Ee_min", 0.010)
    m = m_electron
    dct = 2.0 * DE / (Ee - m)
    ct = x[:, 0]
    ok = (ct > -1 + dct) & (ct < 1.0 - dct)
    s = m ** 2 + 2 * Ee * m
    with np.errstate(all="ignore"):
        num = (256 * (-1 + ct) ** 2 * ct ** 2 * m ** 8
               - 128 * (-1 + ct) * (1 + ct * (1 + ct) * (-3 + 2 * ct)) * m ** 6 * s
               + 16 * (7 + ct * (2 + ct * (-5 + 6 * (-1 + ct) * ct))) * m ** 4 * s ** 2
               - 8 * (7 + ct * (-3 + ct * (3 + ct * (-1 + 2 * ct)))) * m ** 2 * s ** 3
               + (3 + ct ** 2) ** 2 * s ** 4)
        val = alpha_em ** 2 * PI * num / (2 * (-1 + ct) ** 2 * s ** 3 * (-4 * m ** 2 + s) ** 2)
    return np.where(ok, val, 0.0)


def muone_t(x, ev):
    """The Mandelstam t (electron side) the mu e integrand assigns to
    cos(theta) = x[:, 0] (all_processes.py:858-862)."""
    Emu = ev["E_inc"]
    s = m_electron ** 2 + m_muon ** 2 + 2 * m_electron * Emu
    ct = x[:, 0]
    return -2.0 * (1 - ct) * ((s + m_electron ** 2 - m_muon ** 2) ** 2 / (4.0 * s) - m_electron ** 2)


def muone(x, ev):
    """mu e -> mu e on an electron at rest, all_processes.py:842-886."""
    Emu = ev["E_inc"]
    s = m_electron ** 2 + m_muon ** 2 + 2 * m_electron * Emu
    DE = ev.get("Ee_min", 0.010)
    t_limit = 2.0 * m_electron * (m_electron - DE)
    t = muone_t(x, ev)
    ok = t < t_limit
    with np.errstate(all="ignore"):
        val = (16 * PI ** 2 * alpha_em ** 2 * (s ** 2 + 2 * (m_electron ** 2 + m_muon ** 2) * (2 * t + m_muon ** 2 - 3 * m_electron ** 2)
                                               + (s + t - 4 * m_electron ** 2) ** 2) / (16 * PI * s * t ** 2))
    return np.where(ok, val, 0.0)


def G_el_inel_over_t2(ev, t):
    # all_processes.py:111-133
    Z = ev["Z_T"]
    A = ev["A_T"]
    mu_p = 2.79
    c1 = (111 * Z ** (-1 / 3) / m_electron) ** 2
    c2 = 0.164 * GeV ** 2 * A ** (-2 / 3)
    Gel = (1.0 / (1.0 + c1 * t)) ** 2 * (1 + t / c2) ** (-2)
    ap2 = (773.0 * Z ** (-2.0 / 3) / m_electron) ** 2
    Ginel = (Z / (c1 ** 2 * Z ** 2) * np.power(ap2 / (1.0 + ap2 * t), 2.0)
             * ((1.0 + (mu_p ** 2 - 1.0) * t / (4.0 * m_proton ** 2)) / (1.0 + t / 0.71) ** 4))
    return Z ** 2 * c1 ** 2 * (Gel + Ginel)


def dark_brem_vars(x, ev):
    """The DarkBrem / DarkMuonBrem integrand's vector kinematics of the map
    point x (all_processes.py:225-236): 1 - cos(theta_V), cos(theta_V) and
    |k_V| for the energy fraction x[:, 0] and log10(1 - cos theta_V) = x[:, 1]."""
    xx, l1mct = x[:, 0], x[:, 1]
    one_minus_costheta = 10 ** l1mct
    k = np.sqrt(np.fabs((xx * ev["E_inc"]) ** 2 - ev["mV"] ** 2))
    return one_minus_costheta, 1.0 - one_minus_costheta, k


def dark_brem(x, ev):
    """DarkBrem / DarkMuonBrem exact tree level, 'Log' method, all_processes.py:208-372."""
    m_lepton = ev["m_lepton"]
    mV = ev["mV"]
    Ebeam = ev["E_inc"]
    MT = ev["mT"]
    xx, l1mct, lttilde = x.T
    with np.errstate(all="ignore"):
        one_minus_costheta, costheta, k = dark_brem_vars(x, ev)
        ttilde = 10 ** lttilde
        Jacobian = one_minus_costheta * ttilde * np.log(10.0) ** 2
        p = np.sqrt(Ebeam ** 2 - m_lepton ** 2)
        V = np.sqrt(p ** 2 + k ** 2 - 2 * p * k * costheta)
        utilde = -2 * (xx * Ebeam ** 2 - k * p * costheta) + mV ** 2
        discr = utilde ** 2 + 4 * MT * utilde * ((1 - xx) * Ebeam + MT) + 4 * MT ** 2 * V ** 2
        Qplus = V * (utilde + 2 * MT * ((1 - xx) * Ebeam + MT)) + ((1 - xx) * Ebeam + MT) * np.sqrt(np.abs(discr))
        Qplus = Qplus / (2 * ((1 - xx) * Ebeam + MT) ** 2 - 2 * V ** 2)
        Qminus = V * (utilde + 2 * MT * ((1 - xx) * Ebeam + MT)) - ((1 - xx) * Ebeam + MT) * np.sqrt(np.abs(discr))
        Qminus = Qminus / (2 * ((1 - xx) * Ebeam + MT) ** 2 - 2 * V ** 2)
        Qplus = np.fabs(Qplus)
        Qminus = np.fabs(Qminus)
        tplus = 2 * MT * (np.sqrt(MT ** 2 + Qplus ** 2) - MT)
        tminus = 2 * MT * (np.sqrt(MT ** 2 + Qminus ** 2) - MT)
        tconv = (2 * MT * (MT + Ebeam) * np.sqrt(Ebeam ** 2 + m_lepton ** 2)
                 / (MT * (MT + 2 * Ebeam) + m_lepton ** 2)) ** 2
        t = ttilde * tconv
        q0 = -t / (2 * MT)
        q = np.sqrt(t ** 2 / (4 * MT ** 2) + t)
        costhetaq = -(V ** 2 + q ** 2 + m_lepton ** 2 - (Ebeam + q0 - xx * Ebeam) ** 2) / (2 * V * q)
        mVsq2mlepsq = mV ** 2 + 2 * m_lepton ** 2
        Am2 = -8 * MT * (4 * Ebeam ** 2 * MT - t * (2 * Ebeam + MT)) * mVsq2mlepsq
        A1 = 8 * MT ** 2 / utilde
        Am1 = (8 / utilde) * (MT ** 2 * (2 * t * utilde + utilde ** 2 + 4 * Ebeam ** 2 * (2 * (xx - 1) * mVsq2mlepsq - t * ((xx - 2) * xx + 2))
                                         + 2 * t * (-(mV ** 2) + 2 * m_lepton ** 2 + t))
                              - 2 * Ebeam * MT * t * ((1 - xx) * utilde + (xx - 2) * (mVsq2mlepsq + t))
                              + t ** 2 * (utilde - mV ** 2))
        A0 = (8 / utilde ** 2) * (MT ** 2 * (2 * t * utilde + (t - 4 * Ebeam ** 2 * (xx - 1) ** 2) * mVsq2mlepsq)
                                  + 2 * Ebeam * MT * t * (utilde - (xx - 1) * mVsq2mlepsq))
        Y = -t + 2 * q0 * Ebeam - 2 * q * p * (p - k * costheta) * costhetaq / V
        W = np.fabs(Y ** 2 - 4 * q ** 2 * p ** 2 * k ** 2 * (1 - costheta ** 2) * (1 - costhetaq ** 2) / V ** 2)
        ok = ((xx * Ebeam >= mV) & (discr >= 0) & (tplus > tminus) & (t > tminus) & (t < tplus)
              & (np.fabs(costhetaq) <= 1.0) & (W > 0))
        phi_integral = np.where(ok, (A0 + Y * A1 + Am1 / np.sqrt(W) + Y * Am2 / W ** 1.5) / (8 * MT ** 2), 0.0)
        ff = G_el_inel_over_t2(ev, t)
        ans = ff * np.power(alpha_em, 3) * k * Ebeam * phi_integral / (p * np.sqrt(k ** 2 + p ** 2 - 2 * p * k * costheta))
        out = np.where(ok, ans * tconv * Jacobian, 0.0)
    return np.nan_to_num(out, nan=0.0, posinf=0.0, neginf=0.0) * ok


def _fl_kf(x, s):
    # radiative_return.py:26-34
    beta = (2.0 * alpha_em / PI) * (np.log(s / m_electron ** 2) - 1.0)
    x = np.where(x >= 1.0, 1.0 - 1e-10, x)
    return (beta / 16.0) * ((8.0 + 3.0 * beta) * np.power(1.0 - x, beta / 2.0 - 1.0) - 4.0 * (1.0 + x))


def _fl_kf_scaled(x, s):
    # radiative_return.py:36-44
    beta = (2.0 * alpha_em / PI) * (np.log(s / m_electron ** 2) - 1.0)
    return (beta / 16.0) * ((8.0 + 3.0 * beta) - 4.0 * (1.0 + x) * np.power(1.0 - x, 1.0 - beta / 2.0))


def _transformed_lumi(s, y, u):
    # radiative_return.py:61-79
    beta = (2.0 * alpha_em / PI) * (np.log(s / m_electron ** 2) - 1.0)
    x = 1.0 - np.power(u, 2.0 / beta)
    return _fl_kf(y / x, s) * _fl_kf_scaled(x, s) * (1.0 / x) * (2.0 / beta)


def dark_ann(x, ev):
    """DarkAnn radiative return in u, all_processes.py:400-466."""
    mV = ev["mV"]
    Ee = ev["E_inc"]
    s = 2.0 * m_electron * (Ee + m_electron)
    u0 = x[:, 0]
    if s < mV ** 2:
        return np.zeros(len(u0))
    beta = (2.0 * alpha_em / PI) * (np.log(s / m_electron ** 2) - 1.0)
    umax = np.power(1.0 - mV ** 2 / s, beta / 2.0)
    betaf = np.sqrt(1.0 - 4.0 * (m_electron ** 2) / (mV ** 2))
    prefac = (4.0 * PI ** 2) * alpha_em * betaf * (3.0 / 2.0 - betaf ** 2 / 2.0) / s * umax
    with np.errstate(all="ignore"):
        x1 = 1.0 - np.power(u0 * umax, 2.0 / beta)
        x2 = mV ** 2 / (x1 * s)
        ok = (x2 < 1.0) & (x1 > 0.0) & (u0 < 1.0)
        val = 2.0 * prefac * _transformed_lumi(s, mV ** 2 / s, u0 * umax)
    return np.where(ok, val, 0.0)


def sigma_moller(E, Ee_min):
    """Closed-form Moller total cross section, all_processes.py:907-944."""
    Ee = np.asarray(E, dtype=np.float64)
    TeMIN = Ee_min - m_electron
    m = m_electron
    threshold = 3 * m + 4 * TeMIN
    with np.errstate(all="ignore"):
        PF = 2 * PI * alpha_em ** 2 / (m * (Ee ** 2 - m ** 2))
        T1 = (Ee - 3 * m - 4 * TeMIN + 2 * Ee ** 2 * (-2 / (Ee - 3 * m - 2 * TeMIN) + 1 / TeMIN
                                                        + 1 / (-Ee + m + TeMIN) + 2 / (Ee + m + 2 * TeMIN)))
        T2 = (2 * m * (m - 2 * Ee) / (Ee - m) * np.log(
            ((-Ee + m + TeMIN) * (-Ee + 3 * m + 2 * TeMIN) / (TeMIN * (Ee + m + 2 * TeMIN))) * np.heaviside(Ee - threshold, 1)
            + np.heaviside(threshold - Ee, 1)))
        out = PF * (T1 + T2) * np.heaviside(Ee - threshold, 1)
    return np.where(Ee - threshold >= 0, out, 0.0)


def sigma_bhabha(E, Ee_min):
    """Closed-form Bhabha total cross section, all_processes.py:1010-1051."""
    Ee = np.asarray(E, dtype=np.float64)
    TeMIN = Ee_min - m_electron
    m = m_electron
    threshold = 3 * m + 4 * TeMIN
    with np.errstate(all="ignore"):
        PF = PI * alpha_em ** 2 / (12 * (Ee - m) * m * (Ee + m) ** 3 * (Ee - 3 * m - 2 * TeMIN) * TeMIN)
        T1 = (Ee - 3 * m - 4 * TeMIN) * (24 * Ee ** 2 * (Ee + m) ** 2
                                         + (Ee - 3 * m) * (31 * Ee ** 2 + 84 * Ee * m + 57 * m ** 2) * TeMIN
                                         - 4 * (16 * Ee ** 2 + 39 * Ee * m + 33 * m ** 2) * TeMIN ** 2
                                         + 8 * (Ee - 3 * m) * TeMIN ** 3 - 8 * TeMIN ** 4)
        T2 = (24 * (Ee + m) * (2 * Ee ** 2 + 4 * Ee * m + m ** 2) * (Ee - 3 * m - 2 * TeMIN) * TeMIN
              * np.log((2 * TeMIN / (Ee - 3 * m - 2 * TeMIN)) * np.heaviside(Ee - threshold, 1) + np.heaviside(threshold - Ee, 1)))
        out = PF * (T1 + T2) * np.heaviside(Ee - threshold, 1)
    return np.where(Ee - threshold >= 0, out, 0.0)


def sigma_muone(E, Ee_min):
    """Closed-form mu-e total cross section, all_processes.py:888-905."""
    Emu = np.asarray(E, dtype=np.float64)
    me, mm = m_electron, m_muon
    threshold = 1.0 / (2.0 * me) * (me * (Ee_min - me) + np.sqrt(me * (Ee_min + me) * (me * (Ee_min - me) + 2 * mm ** 2)))
    s = me ** 2 + mm ** 2 + 2 * me * Emu
    t_max = 2.0 * me * (me - Ee_min)
    t_min = -4.0 * ((s + me ** 2 - mm ** 2) ** 2 / (4 * s) - me ** 2)
    with np.errstate(all="ignore"):
        PF = 16 * PI ** 2 * alpha_em ** 2 / (8.0 * PI * ((s - mm ** 2) ** 2 + me ** 4 - 2 * (s + mm ** 2) * me ** 2))
        T1 = -2.0 * (s ** 2 + mm ** 4 + 5 * me ** 4 - 2 * me ** 2 * (2 * s + mm ** 2)) * (1.0 / t_max - 1.0 / t_min)
        T2 = 2.0 * (s + 2 * mm ** 2 - 2 * me ** 2) * np.log(t_max / t_min)
        T3 = t_max - t_min
        out = PF * (T1 + T2 + T3) * np.heaviside(Emu - threshold, 1)
    return np.where(Emu - threshold >= 0, out, 0.0)


def muon_e_minimum(Ee_min):
    # shower.py:289
    me, mm = m_electron, m_muon
    return 1.0 / (2.0 * me) * (me * (Ee_min - me) + np.sqrt(me * (Ee_min + me) * (me * (Ee_min - me) + 2 * mm ** 2)))


INTEGRAND = {"Brem": brem, "MuonBrem": brem, "PairProd": pairprod, "Comp": compton, "Ann": annihilation,
             "Moller": moller, "Bhabha": bhabha, "MuonE": muone, "DarkBrem": dark_brem,
             "DarkMuonBrem": dark_brem, "DarkAnn": dark_ann, "DarkComp": compton}
