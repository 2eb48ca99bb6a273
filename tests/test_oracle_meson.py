"""Meson decays M -> gamma V (TwoBody_BSMDecay, dark_shower.py:633-640,
836-844; particle.py:187-254) in the oracle, pinned by kinematics and known
distributions: V on shell and the complement p_M - p_V light-like (the
photon), isotropy in the rest frame, the flat lab-energy spectrum of an
isotropic two-body decay between gamma(E* -+ beta p*), and the weight's
limits (2 kappa^2 BR(gamma gamma) as m_V -> 0, the (1 - r^2)^3 phase-space
factor of Batell, Pospelov and Ritz, Phys. Rev. D 80 (2009) 095024)."""
import numpy as np
from scipy import stats

from oracle import dark as od


def _pf(E, m, direction):
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    p = np.sqrt(E * E - m * m)
    return np.array([E, *(p * d)])


def test_vector_on_shell_and_photon_massless():
    rng = np.random.default_rng(1)
    for pid, mV in ((111, 0.05), (221, 0.3), (331, 0.01)):
        m = od.MESON_MASS[pid]
        for _ in range(200):
            pf = _pf(rng.uniform(1.0, 300.0) * m, m, rng.normal(size=3))
            v = od.two_body_vector(pf, m, mV, rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi))
            m2v = v[0] ** 2 - v[1:] @ v[1:]
            np.testing.assert_allclose(m2v, mV ** 2, rtol=1e-6, atol=1e-9 * v[0] ** 2)
            g = pf - v
            assert abs(g[0] ** 2 - g[1:] @ g[1:]) <= 1e-9 * pf[0] ** 2
            assert g[0] > 0


def test_rest_frame_isotropy():
    m, mV = od.MESON_MASS[221], 0.2
    rng = np.random.default_rng(2)
    ct = []
    for _ in range(4000):
        v = od.two_body_vector([m, 0.0, 0.0, 0.0], m, mV, rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi))
        ct.append(v[3] / np.linalg.norm(v[1:]))
    assert stats.kstest(ct, "uniform", args=(-1, 2)).pvalue > 1e-3


def test_flat_lab_energy_spectrum():
    m, mV, E = od.MESON_MASS[111], 0.06, 20.0
    pf = _pf(E, m, [0.0, 0.0, 1.0])
    g = E / m
    b = np.sqrt(1 - 1 / g ** 2)
    Es = (m * m + mV * mV) / (2 * m)
    ps = (m * m - mV * mV) / (2 * m)
    lo, hi = g * (Es - b * ps), g * (Es + b * ps)
    rng = np.random.default_rng(3)
    EV = [od.two_body_vector(pf, m, mV, rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi))[0] for _ in range(4000)]
    assert lo * (1 - 1e-9) <= min(EV) and max(EV) <= hi * (1 + 1e-9)
    assert stats.kstest(EV, "uniform", args=(lo, hi - lo)).pvalue > 1e-3


def test_weight_limits():
    for pid in (111, 221, 331):
        m = od.MESON_MASS[pid]
        np.testing.assert_allclose(od.meson_weight(pid, m, 1e-9, 0.5), 2 * 0.25 * od.MESON_BR_GG[pid], rtol=1e-12)
        assert od.meson_weight(pid, m, m, 1.0) == 0.0
        assert od.meson_weight(pid, m, 1.5 * m, 1.0) == 0.0
        np.testing.assert_allclose(od.meson_weight(pid, m, m / 2, 1.0), 2 * od.MESON_BR_GG[pid] * 0.75 ** 3, rtol=1e-12)


def test_produce_uses_the_dark_stream():
    rec = {"pid": 111, "pf": _pf(5.0, od.MESON_MASS[111], [0.1, 0.0, 1.0]), "rf": [0.0, 0.0, 0.3], "weight": 2.0,
           "sid": np.uint64(77)}
    a = od.produce_meson(rec, 0.02, seed=5)
    b = od.produce_meson(rec, 0.02, seed=5)
    c = od.produce_meson(dict(rec, sid=np.uint64(78)), 0.02, seed=5)
    assert np.array_equal(a["p"], b["p"]) and not np.array_equal(a["p"], c["p"])
    np.testing.assert_allclose(a["weight"], 2.0 * od.meson_weight(111, od.MESON_MASS[111], 0.02), rtol=1e-15)
    assert od.produce_meson(rec, 0.2, seed=5) is None


def test_mesons_and_neutrinos_end_where_they_start():
    # generate_shower propagates photons, e+- and mu+- only (shower.py:640-653);
    # a neutrino is ended in place and a meson's decay_particle returns no
    # products (particle.py:391-423), so the shower is the primary alone
    from oracle.shower import OracleShower
    from oracle.tables import SMTables
    sh = OracleShower(SMTables("graphite"), 0.010, seed=3)
    for s, pid in enumerate((111, 221, 331, 14, -14)):
        m = od.MESON_MASS.get(pid, 0.0)
        p0 = _pf(7.0, m, [0.1, -0.2, 1.0])
        out = sh.generate_shower(pid, p0, r0=(0.01, 0.02, 0.3), sid=s, shower=s)
        assert len(out) == 1 and out[0].ended and out[0].steps == 0
        assert np.array_equal(out[0].pf, out[0].p0) and np.array_equal(out[0].rf, out[0].r0)
        assert sh.stats.final[s] == 7.0 and s not in sh.stats.deposit
