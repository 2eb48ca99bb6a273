"""Final-state four-vectors in the parent frame (oracle, float64).

Restates src/PETITE/kinematics.py.  The parent moves along +z with energy E
and mass m; ``az`` is the azimuth uniform that the reference draws with
np.random.uniform(0, 2 pi) inside each function (here 2 pi * u).  Each
function returns (first, second) four-vectors in the reference's order.

Pinned by: energy conservation and on-shell invariants per vertex
(tests/test_oracle_physics.py), and the momentum transfer, Mandelstam t,
annihilation cut and dark-vector angle recomputed from the four-vectors
against the integrands' variables (tests/test_oracle_kinematics_pins.py).
"""
import math

import numpy as np

from .constants import m_electron, m_muon, alpha_em

EGAMMA_MIN = 0.001  # kinematics.py:9 (module constant, independent of the tables)


def brem(E, m, x, u_az):
    """e_to_egamma_fourvecs, kinematics.py:10-41 -> (lepton, photon)."""
    ep = E
    x1, x2, x3, x4 = x[:4]
    w = EGAMMA_MIN + x1 * (ep - m - EGAMMA_MIN)
    ct = math.cos((x2 + x3) / 2)
    ctp = math.cos((x2 - x3) * ep / (2 * (ep - w)))
    ph = (x4 - 0.5) * 2.0 * math.pi
    epp = ep - w
    pp = math.sqrt(epp ** 2 - m ** 2)
    al = 2.0 * math.pi * u_az
    cal, sal = math.cos(al), math.sin(al)
    st, stp = math.sqrt(1.0 - ct ** 2), math.sqrt(1.0 - ctp ** 2)
    sp, cp = math.sin(ph), math.cos(ph)
    g = [w, w * cal * st, w * sal * st, w * ct]
    l = [epp, pp * (sal * sp * stp + cal * (ctp * st - cp * ct * stp)),
         pp * (ctp * sal * st - (cp * ct * sal + cal * sp) * stp), pp * (ct * ctp + cp * st * stp)]
    return np.array(l), np.array(g)


def pairprod(E, x, u_az):
    """gamma_to_epem_fourvecs, kinematics.py:70-102 -> (positron, electron)."""
    w = E
    m = m_electron
    x1, x2, x3, x4 = x[:4]
    epp = m + x1 * (w - 2 * m)
    ctp = math.cos(w * (x2 + x3) / (2 * epp))
    ctm = math.cos(w * (x2 - x3) / (2 * (w - epp)))
    ph = x4 * 2 * math.pi
    epm = w - epp
    pm, pp = math.sqrt(epm ** 2 - m ** 2), math.sqrt(epp ** 2 - m ** 2)
    al = 2.0 * math.pi * u_az
    cal, sal = math.cos(al), math.sin(al)
    stp, stm = math.sqrt(1.0 - ctp ** 2), math.sqrt(1.0 - ctm ** 2)
    spal, cpal = math.sin(ph + al), math.cos(ph + al)
    return (np.array([epp, pp * stp * cal, pp * stp * sal, pp * ctp]),
            np.array([epm, pm * stm * cpal, pm * stm * spal, pm * ctm]))


def compton(E, x, u_az, mV=0.0):
    """compton_fourvecs, kinematics.py:104-132 -> (electron, photon/V)."""
    m = m_electron
    Eg = E
    ct = x[0]
    s = m ** 2 + 2 * Eg * m
    rs = math.sqrt(s)
    Ee0 = (s + m ** 2) / (2.0 * rs)
    Ee = (s - mV ** 2 + m ** 2) / (2 * rs)
    EV = (s + mV ** 2 - m ** 2) / (2 * rs)
    pF = math.sqrt(Ee ** 2 - m ** 2)
    g0 = Ee0 / m
    b0 = 1.0 / g0 * math.sqrt(g0 ** 2 - 1.0)
    ph = 2.0 * math.pi * u_az
    st = math.sqrt(1 - ct ** 2)
    pe = [g0 * Ee + b0 * g0 * pF * ct, -pF * st * math.sin(ph), -pF * st * math.cos(ph), b0 * g0 * Ee + g0 * pF * ct]
    pV = [g0 * EV - b0 * g0 * pF * ct, pF * st * math.sin(ph), pF * st * math.cos(ph), b0 * g0 * EV - g0 * pF * ct]
    return np.array(pe), np.array(pV)


def ee_to_ee(E, x, u_az):
    """ee_to_ee_fourvecs (Moller/Bhabha), kinematics.py:213-237 -> (scattered, new electron)."""
    m = m_electron
    ct = x[0]
    s = 2 * m ** 2 + 2 * E * m
    Ee0 = math.sqrt(s) / 2.0
    pF = math.sqrt(Ee0 ** 2 - m ** 2)
    g0 = Ee0 / m
    b0 = 1.0 / g0 * math.sqrt(g0 ** 2 - 1.0)
    ph = 2.0 * math.pi * u_az
    st = math.sqrt(1 - ct ** 2)
    a = [g0 * Ee0 + b0 * g0 * pF * ct, -pF * st * math.sin(ph), -pF * st * math.cos(ph), b0 * g0 * Ee0 + g0 * pF * ct]
    b = [g0 * Ee0 - b0 * g0 * pF * ct, pF * st * math.sin(ph), pF * st * math.cos(ph), b0 * g0 * Ee0 - g0 * pF * ct]
    return np.array(a), np.array(b)


def mue_to_mue(E, x, u_az):
    """mue_to_mue_fourvecs, kinematics.py:239-265 -> (muon, electron)."""
    me, mm = m_electron, m_muon
    ct = x[0]
    s = me ** 2 + mm ** 2 + 2 * E * me
    rs = math.sqrt(s)
    Ee0 = (s + me ** 2 - mm ** 2) / (2.0 * rs)
    Em0 = (s + mm ** 2 - me ** 2) / (2.0 * rs)
    pe = math.sqrt(Ee0 ** 2 - me ** 2)
    pm = math.sqrt(Em0 ** 2 - mm ** 2)
    g0 = Ee0 / me
    b0 = 1.0 / g0 * math.sqrt(g0 ** 2 - 1.0)
    ph = 2.0 * math.pi * u_az
    st = math.sqrt(1 - ct ** 2)
    mu = [g0 * Em0 + b0 * g0 * pm * ct, pm * st * math.sin(ph), pm * st * math.cos(ph), b0 * g0 * Em0 + g0 * pm * ct]
    el = [g0 * Ee0 - b0 * g0 * pe * ct, -pe * st * math.sin(ph), -pe * st * math.cos(ph), b0 * g0 * Ee0 - g0 * pe * ct]
    return np.array(mu), np.array(el)


def annihilation(E, x, u_az, mV=0.0):
    """annihilation_fourvecs, kinematics.py:301-334 -> (photon, photon/V)."""
    m = m_electron
    ct = x[0]
    s = 2 * m * (E + m)
    rs = math.sqrt(s)
    EeCM = rs / 2.0
    Eg = (s - mV ** 2) / (2 * rs)
    EV = (s + mV ** 2) / (2 * rs)
    pF = Eg
    g0 = EeCM / m
    b0 = 1.0 / g0 * math.sqrt(g0 ** 2 - 1.0)
    ph = 2.0 * math.pi * u_az
    st = math.sqrt(1 - ct ** 2)
    pg = [g0 * Eg - b0 * g0 * pF * ct, -pF * st * math.sin(ph), -pF * st * math.cos(ph), b0 * g0 * Eg - g0 * pF * ct]
    pV = [g0 * EV + b0 * g0 * pF * ct, pF * st * math.sin(ph), pF * st * math.cos(ph), b0 * g0 * EV + g0 * pF * ct]
    return np.array(pg), np.array(pV)


def l_to_lV(E, m, x, u_az, mV):
    """l_to_lV_fourvecs, kinematics.py:43-68 -> (lepton placeholder, V)."""
    w = x[0] * E
    ct = 1 - 10 ** x[1]
    k = math.sqrt(w ** 2 - mV ** 2)
    al = 2.0 * math.pi * u_az
    st = math.sqrt(1.0 - ct ** 2)
    return np.array([E, 0, 0, math.sqrt(E ** 2 - m ** 2)]), np.array([w, k * math.cos(al) * st, k * math.sin(al) * st, k * ct])


def boost(p, v):
    """radiative_return.py:18-24: boost v into the rest frame of p."""
    lp = lambda a, b: a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3]
    rsq = math.sqrt(lp(p, p))
    v0 = lp(p, v) / rsq
    c1 = (v[0] + v0) / (rsq + p[0])
    return np.array([v0, v[1] - c1 * p[1], v[2] - c1 * p[2], v[3] - c1 * p[3]])


def radiative_return(E, x, mV):
    """radiative_return_fourvecs, kinematics.py:267-299 -> V four-vector."""
    m = m_electron
    s = 2.0 * m * (m + E)
    beta = (2.0 * alpha_em / math.pi) * (math.log(s / m ** 2) - 1.0)
    umax = (1.0 - mV ** 2 / s) ** (beta / 2.0)
    x1 = 1.0 - (x[0] * umax) ** (2.0 / beta)
    x2 = mV ** 2 / (x1 * s)
    E1 = x1 * math.sqrt(s) / 2.0
    E2 = x2 * math.sqrt(s) / 2.0
    pV = np.array([E1 + E2, 0.0, 0.0, E1 - E2])
    return boost(np.array([math.sqrt(s) / 2.0, 0.0, 0.0, -math.sqrt(s / 4.0 - m ** 2)]), pV)


def compton_bound(E, x, mV, Pe, cte, u_ph, u_phie):
    """compton_fourvecs_boundelectron, kinematics.py:134-183 -> V four-vector."""
    m = m_electron
    Eg = E
    ct = x[0]
    s = m ** 2 + 2 * Eg * (math.sqrt(m ** 2 + Pe ** 2) - cte * Pe)
    rs = math.sqrt(s)
    Ee = (s - mV ** 2 + m ** 2) / (2 * rs)
    EV = (s + mV ** 2 - m ** 2) / (2 * rs)
    pF = math.sqrt(Ee ** 2 - m ** 2)
    bnum = math.sqrt(Eg ** 2 + 2 * cte * Eg * Pe + Pe ** 2)
    bden = Eg + math.sqrt(m ** 2 + Pe ** 2)
    b0 = bnum / bden
    g0 = 1.0 / math.sqrt(1.0 - b0 ** 2)
    ph = 2.0 * math.pi * u_ph
    st = math.sqrt(1 - ct ** 2)
    EVLab = g0 * EV - b0 * g0 * pF * ct
    pV3 = np.array([pF * st * math.sin(ph), pF * st * math.cos(ph), b0 * g0 * EV - g0 * pF * ct])
    ctz = (Eg + cte * Pe) / math.sqrt(Eg ** 2 + 2 * cte * Eg * Pe + Pe ** 2)
    stz = math.sqrt(1.0 - ctz ** 2)
    phie = 2.0 * math.pi * u_phie
    R = np.array([[ctz * math.cos(phie), -math.sin(phie), stz * math.cos(phie)],
                  [ctz * math.sin(phie), math.cos(phie), stz * math.sin(phie)],
                  [-stz, 0, ctz]])
    return np.concatenate([[EVLab], R @ pV3])


SM_KINEMATICS = {"Brem": "brem", "MuonBrem": "brem", "PairProd": "pairprod", "Comp": "compton",
                 "Ann": "annihilation", "Moller": "ee_to_ee", "Bhabha": "ee_to_ee", "MuonE": "mue_to_mue"}


def sm_final_state(process, E, m, x, u_az):
    name = SM_KINEMATICS[process]
    if name == "brem":
        return brem(E, m, x, u_az)
    if name == "pairprod":
        return pairprod(E, x, u_az)
    if name == "compton":
        return compton(E, x, u_az)
    if name == "annihilation":
        return annihilation(E, x, u_az)
    if name == "ee_to_ee":
        return ee_to_ee(E, x, u_az)
    return mue_to_mue(E, x, u_az)
