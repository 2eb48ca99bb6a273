"""Pin oracle kinematics (vertex invariants), MCS (Rayleigh law) and the
transport (exponential free path, per-shower energy budget, multiplicity)."""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import kinematics as K, moliere, philox
from oracle.constants import m_electron, m_muon
from oracle.shower import OracleShower, lose_energy
from oracle.tables import SMTables

RNG = np.random.default_rng(5)


def _mass2(p):
    return p[0] ** 2 - p[1] ** 2 - p[2] ** 2 - p[3] ** 2


@pytest.mark.parametrize("E", [0.02, 1.0, 40.0])
def test_two_to_two_invariants(E):
    for _ in range(50):
        ct = RNG.uniform(-0.999, 0.999)
        u = RNG.uniform()
        a, b = K.compton(E, [ct], u)
        assert abs(a[0] + b[0] - (E + m_electron)) < 1e-12 * (E + 1)
        assert np.allclose(a[1:] + b[1:], [0, 0, E], atol=1e-9 * (E + 1))
        assert abs(_mass2(a) - m_electron ** 2) < 1e-9 * E ** 2 and abs(_mass2(b)) < 1e-9 * E ** 2
        a, b = K.ee_to_ee(E, [ct], u)
        p = math.sqrt(E ** 2 - m_electron ** 2)
        assert abs(a[0] + b[0] - (E + m_electron)) < 1e-12 * (E + 1)
        assert np.allclose(a[1:] + b[1:], [0, 0, p], atol=1e-9 * (E + 1))
        a, b = K.annihilation(E, [ct], u)
        assert abs(a[0] + b[0] - (E + m_electron)) < 1e-12 * (E + 1)
        assert np.allclose(a[1:] + b[1:], [0, 0, p], atol=1e-9 * (E + 1))
    Em = E + 0.2
    a, b = K.mue_to_mue(Em, [0.3], 0.1)
    assert abs(a[0] + b[0] - (Em + m_electron)) < 1e-12 * Em
    assert abs(_mass2(a) - m_muon ** 2) < 1e-9 * Em ** 2


def test_brem_and_pairprod_energy_split():
    for _ in range(100):
        E = 10 ** RNG.uniform(-1, 2)
        x = RNG.uniform(0, 1, 4) * np.array([1, 0.001, 0.001, 1])
        l, g = K.brem(E, m_electron, x, RNG.uniform())
        assert abs(l[0] + g[0] - E) < 1e-12 * E
        assert abs(_mass2(l) - m_electron ** 2) < 1e-9 * E ** 2
        pp, pm = K.pairprod(E, x, RNG.uniform())
        assert abs(pp[0] + pm[0] - E) < 1e-12 * E


def test_mcs_preserves_p_and_rayleigh_law():
    p4 = np.array([1.0, 0.1, -0.2, math.sqrt(1 - m_electron ** 2 - 0.05)])
    t, A, Z = 0.5, 12, 6
    n = 4000
    w = philox.words(3, np.arange(n, dtype=np.uint64), 0, philox.PURPOSE_STEP)
    u = philox.to_uniform(w)
    beta = np.linalg.norm(p4[1:]) / p4[0]
    th0 = moliere.theta0(t, beta, A, Z, m_electron)
    ang = []
    for i in range(n):
        q = moliere.scatter(p4, t, A, Z, 1.0, m_electron, u[i, 2], u[i, 3], 1.0 if int(w[i, 3]) & 1 else -1.0)
        assert abs(np.linalg.norm(q[1:]) - np.linalg.norm(p4[1:])) < 1e-12
        c = np.dot(q[1:], p4[1:]) / np.dot(p4[1:], p4[1:])
        ang.append(math.acos(min(1.0, c)))
    # space angle |theta| is Rayleigh(theta0)
    assert stats.kstest(np.array(ang) / th0, "rayleigh").pvalue > 1e-3


def test_rotation_to_z():
    for _ in range(200):
        v = RNG.normal(size=3)
        R = moliere.rotation_to_z(v)
        assert np.allclose(R @ v, [0, 0, np.linalg.norm(v)], atol=1e-12)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)


def test_lose_energy_on_shell():
    pf = np.array([2.0, 0.3, 0.4, math.sqrt(4 - 0.25 - m_electron ** 2)])
    q = lose_energy(pf, m_electron, 0.5)
    assert abs(q[0] - 1.5) < 1e-15 and abs(_mass2(q) - m_electron ** 2) < 1e-12
    q = lose_energy(pf, m_electron, 5.0)
    assert q[0] == m_electron and np.all(q[1:] == 0)


@pytest.fixture(scope="module")
def graphite():
    return SMTables("graphite")


def test_photon_free_path_exponential(graphite):
    sh = OracleShower(graphite, 0.010, seed=11)
    E = 1.0
    mfp = graphite.mfp(22, E)
    from oracle.shower import OParticle
    d = []
    for s in range(400):
        p = OParticle(22, 0.0, [E, 0, 0, E], [0, 0, 0], sid=s)
        sh.propagate(p)
        d.append(p.rf[2])
    assert stats.kstest(np.array(d) / mfp, "expon").pvalue > 1e-3


@pytest.mark.slow
def test_shower_energy_budget_and_multiplicity(graphite):
    sh = OracleShower(graphite, 0.010, seed=0)
    n = []
    for s in range(2):
        parts = sh.generate_shower(11, [10.0, 0, 0, math.sqrt(100 - m_electron ** 2)], sid=s, shower=s)
        st = sh.stats
        bal = 10.0 + st.atomic.get(s, 0) - st.deposit.get(s, 0) - st.final.get(s, 0) - st.dropped.get(s, 0)
        assert abs(bal) < 1e-10
        n.append(len(parts))
    # reference anchor: 655 particles (examples/multiple_coulomb_scattering.ipynb:258-259)
    assert 450 < np.mean(n) < 900
