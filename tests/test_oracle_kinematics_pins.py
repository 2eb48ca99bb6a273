"""Pin the oracle's final-state kinematics (src/PETITE/kinematics.py:10-334)
against the kinematic variables its integrands assign to the same map point
(src/PETITE/all_processes.py:139-886).

The draw samples x with density dsigma(x); the kinematics turn x into
four-vectors.  Both must describe the same final state, so an invariant
recomputed from the four-vectors has to equal the one the integrand uses:
* Brem / MuonBrem, PairProd: the momentum transfer to the nucleus,
  q = p_in - sum p_out, against the integrand's q^2 (its small-angle form, so
  to 5e-3 of the q^2 scale m^2 (d^2 + d'^2) + the squared longitudinal terms,
  with every final lepton at >= 50 m; measured at most 1.7e-3), and the
  outgoing energies exactly;
* Compton / DarkComp, mu e: the Mandelstam t of the outgoing vector (or
  electron) against the integrand's t(cos theta);
* Moller / Bhabha: t of the scattered beam particle (t -> 0 at cos theta -> 1,
  the integrand's forward pole) and u of the other, against the centre-of-mass
  relations t = -(s - 4 m^2)(1 -+ cos theta)/2;
* annihilation: at the integrand's cut cos theta = ctMax the first photon has
  exactly the minimum photon energy Eg_min;
* DarkBrem: the vector's energy, |k| and 1 - cos theta against the
  integrand's x E, k and 10^x1.
A flipped cos theta, a relative azimuth off by pi, a missing angle factor
(theta' = (x2 - x3)/2 instead of (x2 - x3) E/(2 E')) or a swapped output
order fails these (checked by planting each in a copy of the oracle).
"""
import math

import numpy as np
import pytest

from oracle import integrands as I
from oracle import kinematics as K
from oracle.constants import m_electron as ME, m_muon as MMU

RNG = np.random.default_rng(11)


def _mink(a):
    return a[0] ** 2 - a[1] ** 2 - a[2] ** 2 - a[3] ** 2


def _small_angle_points(E, m, n, xmin, xmax):
    """x with scaled angles d, d' in (0, 3) (x2 +- x3 = 2 m d / E) and energy
    fractions x1 in [xmin, xmax]."""
    out = []
    for _ in range(n):
        d, dp = RNG.uniform(0.02, 3.0, 2)
        s_, d_ = m * d / E, m * dp / E  # (x2 + x3)/2, (x2 - x3)/2
        out.append([RNG.uniform(xmin, xmax), s_ + d_, s_ - d_, RNG.uniform()])
    return np.array(out)


@pytest.mark.parametrize("E,m", [(0.1, ME), (1.0, ME), (30.0, ME), (30.0, MMU), (100.0, MMU)])
def test_brem_momentum_transfer(E, m):
    ev = {"E_inc": E, "Eg_min": K.EGAMMA_MIN, "m_lepton": m}
    x = _small_angle_points(E, m, 300, 0.02, 1.0 - 50 * m / (E - m))
    w, d, dp, ph, epp, qsq = I.brem_vars(x, ev)
    p_in = np.array([0.0, 0.0, math.sqrt(E * E - m * m)])
    for i in range(len(x)):
        lep, gam = K.brem(E, m, x[i], RNG.uniform())
        assert abs(gam[0] - w[i]) < 1e-12 * E and abs(lep[0] - epp[i]) < 1e-12 * E
        q = p_in - lep[1:] - gam[1:]
        scale = m * m * (d[i] ** 2 + dp[i] ** 2) + m ** 4 * (
            ((1 + d[i] ** 2) / (2 * E)) ** 2 + ((1 + dp[i] ** 2) / (2 * epp[i])) ** 2)
        assert abs(q @ q - qsq[i]) <= 5e-3 * scale, (E, i, q @ q, qsq[i], scale)


@pytest.mark.parametrize("E", [0.1, 1.0, 30.0])
def test_pairprod_momentum_transfer(E):
    ev = {"E_inc": E}
    f = 50 * ME / (E - 2 * ME)
    x = _small_angle_points(E, ME, 300, f, 1.0 - f)
    epp, epm, dp, dm, ph, qq = I.pairprod_vars(x, ev)
    for i in range(len(x)):
        pos, ele = K.pairprod(E, x[i], RNG.uniform())
        assert abs(pos[0] - epp[i]) < 1e-12 * E and abs(ele[0] - epm[i]) < 1e-12 * E
        q = np.array([0.0, 0.0, E]) - pos[1:] - ele[1:]
        m2 = ME * ME
        scale = m2 * (dp[i] ** 2 + dm[i] ** 2) + m2 * m2 * (
            ((1 + dp[i] ** 2) / (2 * epp[i])) ** 2 + ((1 + dm[i] ** 2) / (2 * epm[i])) ** 2)
        assert abs(q @ q - m2 * qq[i]) <= 5e-3 * scale, (E, i, q @ q, m2 * qq[i], scale)


@pytest.mark.parametrize("mV", [0.0, 0.003])
@pytest.mark.parametrize("E", [0.05, 1.0, 30.0])
def test_compton_t(E, mV):
    for ct in RNG.uniform(-1, 1, 100):
        e, V = K.compton(E, [ct], RNG.uniform(), mV=mV)
        t = _mink(np.array([E, 0.0, 0.0, E]) - V)
        tI = I.compton_t(np.array([[ct]]), {"E_inc": E, "mV": mV})[0]
        assert abs(t - tI) <= 1e-12 * E * E + 1e-9 * abs(tI), (E, mV, ct, t, tI)
        assert abs(_mink(V) - mV * mV) <= 1e-12 * E * E


@pytest.mark.parametrize("E", [0.5, 5.0, 50.0])
def test_muone_t(E):
    for ct in RNG.uniform(-1, 1, 100):
        mu, el = K.mue_to_mue(E, [ct], RNG.uniform())
        t = _mink(np.array([ME, 0.0, 0.0, 0.0]) - el)
        tI = I.muone_t(np.array([[ct]]), {"E_inc": E})[0]
        assert abs(t - tI) <= 1e-12 * E * E + 1e-9 * abs(tI), (E, ct, t, tI)


@pytest.mark.parametrize("E", [0.05, 5.0, 50.0])
def test_ee_t_and_u(E):
    p_in = np.array([E, 0.0, 0.0, math.sqrt(E * E - ME * ME)])
    s = 2 * ME * ME + 2 * E * ME
    for ct in RNG.uniform(-1, 1, 100):
        a, b = K.ee_to_ee(E, [ct], RNG.uniform())
        t, u = _mink(p_in - a), _mink(p_in - b)
        tI, uI = -(s - 4 * ME * ME) * (1 - ct) / 2, -(s - 4 * ME * ME) * (1 + ct) / 2
        assert abs(t - tI) <= 1e-12 * E * E + 1e-9 * abs(tI), (E, ct, t, tI)
        assert abs(u - uI) <= 1e-12 * E * E + 1e-9 * abs(uI), (E, ct, u, uI)
    # the Bhabha integrand's forward pole (cos theta -> 1) is t -> 0 of the scattered positron
    fwd = I.bhabha(np.array([[1 - 1e-3]]), {"E_inc": E, "Ee_min": 0.0})[0]
    bwd = I.bhabha(np.array([[-1 + 1e-3]]), {"E_inc": E, "Ee_min": 0.0})[0]
    assert fwd > 10 * bwd


@pytest.mark.parametrize("E", [0.02, 0.5, 5.0, 50.0])
def test_annihilation_cut_is_the_minimum_photon_energy(E):
    egmin = K.EGAMMA_MIN
    ev = {"E_inc": E, "Eg_min": egmin}
    m = ME
    ctmax = math.sqrt((E + m) / (E - m)) * (2 * m * (E - 2 * egmin + m)) / (2 * m * (E + m))
    lo = I.annihilation(np.array([[ctmax * (1 - 1e-12)]]), ev)[0]
    hi = I.annihilation(np.array([[min(ctmax * (1 + 1e-9), 1.0)]]), ev)[0]
    assert lo > 0.0 and (hi == 0.0 or ctmax >= 1.0)
    g, g2 = K.annihilation(E, [ctmax], RNG.uniform())
    assert abs(g[0] - egmin) < 1e-9 * E
    assert abs(g[0] + g2[0] - (E + m)) < 1e-12 * E


@pytest.mark.parametrize("m", [ME, MMU])
@pytest.mark.parametrize("mV", [0.003, 0.3])
def test_dark_brem_vector(m, mV):
    E = 10.0
    ev = {"E_inc": E, "mV": mV, "m_lepton": m}
    x = np.column_stack([RNG.uniform(1.01 * mV / E, 0.999, 100), RNG.uniform(-8, 0.3, 100), RNG.uniform(-10, 0, 100)])
    omc, ct, k = I.dark_brem_vars(x, ev)
    for i in range(len(x)):
        _, V = K.l_to_lV(E, m, x[i], RNG.uniform(), mV)
        kk = math.sqrt(V[1] ** 2 + V[2] ** 2 + V[3] ** 2)
        assert abs(V[0] - x[i, 0] * E) < 1e-12 * E
        assert abs(kk - k[i]) < 1e-10 * E
        if ct[i] > -1.0:
            assert abs((1.0 - V[3] / kk) - omc[i]) < 1e-9 * max(omc[i], 1e-6)
