"""Multiple Coulomb scattering, fast (Lynch-Dahl) mode (oracle, float64).

Restates src/PETITE/moliere.py:196-219 (chi_c^2, chi_a^2), 265-284
(generate_moliere_angle_simplified_alt), 287-348 (get_rotation_matrix) and
350-400 (get_scattered_momentum_fast).  Random numbers: the reference's
random.choice([-1,1]) * sqrt(gauss^2 + gauss^2) is a signed Rayleigh(1)
variate; it is drawn here as s * sqrt(-2 ln u_r), with the azimuth 2 pi u_a
and the sign s = +1 when bit 0 of the azimuth word is set (words of the
substep's or the closing increment's Philox block, see oracle.shower).

Pinned by: the width theta0 against Highland's formula
(tests/test_oracle_moliere_pin.py), the Rayleigh law of the space angle (KS)
and |p| preservation (tests/test_oracle_physics.py).
"""
import math

import numpy as np

from .constants import MeV, alpha_em, m_electron

F_LD = 0.98


def chic2(t, beta, A, Z, m_lepton):
    p = (m_lepton / MeV) * beta / math.sqrt(1.0 - beta ** 2)
    return 0.157 * Z * (Z + 1) * (t / A) * (1.0 / (p * beta)) ** 2


def chia2(beta, A, Z, m_lepton):
    p = (m_lepton / MeV) * beta / math.sqrt(1.0 - beta ** 2)
    return 2.007e-5 * Z ** (2.0 / 3.0) * (1.0 + 3.34 * (Z * alpha_em / beta) ** 2) / p ** 2


def theta0(t, beta, A, Z, m_lepton):
    c2 = chic2(t, beta, A, Z, m_lepton)
    a2 = chia2(beta, A, Z, m_lepton)
    omega = c2 / a2
    v = 0.5 * omega / (1.0 - F_LD)
    return math.sqrt(c2 * ((1.0 + v) * math.log(1.0 + v) / v - 1) / (1.0 + F_LD ** 2))


def rotation_to_z(v):
    """get_rotation_matrix (moliere.py:287-348): R with R v = |v| z-hat."""
    vx, vy, vz = v
    if abs(vx) > 0.0 and abs(vy) > 0.0:
        a = math.atan(abs(vy / vx))
        if vx > 0.0 and vy > 0.0:
            a = -a
        if vx < 0.0 and vy > 0.0:
            a = -(math.pi - a)
        if vx < 0.0 and vy < 0.0:
            a = -(math.pi + a)
        if vx > 0.0 and vy < 0.0:
            a = -(2.0 * math.pi - a)
        ca, sa = math.cos(a), math.sin(a)
    elif abs(vy) > 0.0:
        ca, sa = 0.0, 1.0
    else:
        ca, sa = 1.0, 0.0
    Ra = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    vxp = vx * ca - vy * sa
    if abs(vz) > 0.0 and abs(vxp) > 0.0:
        b = math.atan(abs(vxp / vz))
        if vz > 0.0 and vxp > 0.0:
            b = -b
        if vz < 0.0 and vxp > 0.0:
            b = -(math.pi - b)
        if vz < 0.0 and vxp < 0.0:
            b = -(math.pi + b)
        if vz > 0.0 and vxp > 0.0:  # the reference repeats this case (moliere.py:331)
            b = -(2.0 * math.pi - b)
        cb, sb = math.cos(b), math.sin(b)
    elif vxp > 0.0:
        cb, sb = 0.0, -1.0
    elif vxp < 0.0:
        cb, sb = 0.0, 1.0
    else:
        cb, sb = 1.0, 0.0
    Rb = np.array([[cb, 0.0, sb], [0.0, 1.0, 0.0], [-sb, 0.0, cb]])
    return Rb @ Ra


def scatter(p4, t, A, Z, rescale, m_lepton, u_r, u_a, sign):
    """get_scattered_momentum_fast (moliere.py:350-400) with radius/azimuth uniforms and a sign."""
    p3 = np.asarray(p4[1:], dtype=np.float64)
    pn = float(np.linalg.norm(p3))
    if not pn > 0:
        return np.array(p4, dtype=np.float64)
    beta = pn / p4[0]
    Rz = rotation_to_z(p3).T
    th0 = theta0(t, beta, A, Z, m_lepton)
    theta = sign * math.sqrt(-2.0 * math.log(u_r)) * th0 * rescale
    phi = 2.0 * math.pi * u_a
    cth, sth, cph, sph = math.cos(theta), math.sin(theta), math.cos(phi), math.sin(phi)
    local = pn * np.array([sph * sth, -cph * sth, cth])
    out = np.empty(4)
    out[0] = p4[0]
    out[1:] = Rz @ local
    return out
