"""Pin the 1-D oracle integrands and closed-form sigmas against each other and
against textbook results (Klein-Nishina), src/PETITE/all_processes.py:624-1051."""
import numpy as np
import pytest
from scipy.integrate import quad

from oracle import integrands as I
from oracle.constants import alpha_em, m_electron

EE_MIN = 0.005


def _q(fn, ev, lo, hi, pts=None):
    return quad(lambda c: fn(np.array([[c]]), ev)[0], lo, hi, limit=400, points=pts)[0]


@pytest.mark.parametrize("E", [0.03, 0.5, 10.0, 100.0])
def test_compton_klein_nishina(E):
    k = E / m_electron
    re = alpha_em / m_electron
    kn = 2 * np.pi * re ** 2 * ((1 + k) / k ** 2 * (2 * (1 + k) / (1 + 2 * k) - np.log(1 + 2 * k) / k)
                                + np.log(1 + 2 * k) / (2 * k) - (1 + 3 * k) / (1 + 2 * k) ** 2)
    assert abs(_q(I.compton, {"E_inc": E}, -1, 1, [1 - 1e-3, 1 - 1e-5]) / kn - 1) < 1e-8


@pytest.mark.parametrize("E", [0.3, 1.0, 10.0, 100.0])
def test_muone_closed_form(E):
    q = _q(I.muone, {"E_inc": E, "Ee_min": EE_MIN}, -1, 1, [1 - 1e-4, 1 - 1e-6])
    assert abs(q / I.sigma_muone(E, EE_MIN) - 1) < 1e-6


@pytest.mark.parametrize("E", [10.0, 100.0])
@pytest.mark.parametrize("fn,cf", [(I.moller, I.sigma_moller), (I.bhabha, I.sigma_bhabha)])
def test_moller_bhabha_closed_form_high_energy(E, fn, cf):
    # the closed forms cut on kinetic energy T >= Ee_min - m_e; at high energy
    # that is the symmetric cos-theta cut of the integrand with DE = Ee_min - m_e
    de = EE_MIN - m_electron
    d = 2 * de / (E - m_electron)
    q = _q(fn, {"E_inc": E, "Ee_min": de}, -1 + d, 1 - d, [-1 + 2 * d, 1 - 2 * d])
    assert abs(q / cf(E, EE_MIN) - 1) < 1e-3


def test_annihilation_heitler_limit():
    # Eg_min=0 and mV=0: Heitler total = pi r_e^2/(g+1) [ (g^2+4g+1)/(g^2-1) ln(g+sqrt(g^2-1))
    # - (g+3)/sqrt(g^2-1) ]; it counts each photon pair once, so the integral over
    # ct in [-1,1] is twice it.  The reference's integrand drops electron-mass
    # terms of the matrix element, so agreement is asymptotic (1e-4 at 50 GeV).
    for E in (50.0, 1000.0):
        g = E / m_electron
        re = alpha_em / m_electron
        h = np.pi * re ** 2 / (g + 1) * ((g * g + 4 * g + 1) / (g * g - 1) * np.log(g + np.sqrt(g * g - 1))
                                         - (g + 3) / np.sqrt(g * g - 1))
        q = _q(I.annihilation, {"E_inc": E, "mV": 0.0, "Eg_min": 0.0}, -1, 1, [-1 + 1e-4, 1 - 1e-4])
        assert abs(q / (2 * h) - 1) < 1e-4, (E, q, h)
