"""Particle decay methods (particle.py:187-423) pinned by kinematics and
closed forms: the boost is a Lorentz transformation taking (m, 0) to pf;
products are on shell and conserve four-momentum; isotropy in the rest frame
(KS); the flat lab-energy spectrum of an isotropic two-body decay; uniform
occupation of the Dalitz region (the m12^2 marginal proportional to the
width of the m23^2 range) and its reweighting by a matrix element; the
truncated exponential of the decay-or-interaction distance; the branching
fractions and bookkeeping of decay_particle."""
import numpy as np
import pytest
from scipy import integrate, stats

from petite_amd.constants import mass_dict, meson_decay_dict
from petite_amd.particle import Particle

G = np.diag([1.0, -1.0, -1.0, -1.0])


def _meson(pid, E, direction=(0.2, -0.1, 1.0), **ids):
    m = mass_dict[pid]
    d = np.asarray(direction, dtype=np.float64)
    p = np.sqrt(E * E - m * m) * d / np.linalg.norm(d)
    return Particle(np.concatenate([[E], p]), [0.01, -0.02, 0.5], dict({"PID": pid, "mass": m}, **ids))


def _m2(k):
    k = np.asarray(k, dtype=np.float64)
    return k[0] ** 2 - k[1:] @ k[1:]


def test_boost_is_lorentz_and_takes_rest_to_pf():
    q = _meson(221, 7.0)
    L = np.asarray(q.boost_matrix())
    np.testing.assert_allclose(L.T @ G @ L, G, atol=1e-12)
    np.testing.assert_allclose(L @ [mass_dict[221], 0, 0, 0], q.get_pf(), rtol=1e-12, atol=1e-12)
    rest = Particle([mass_dict[111], 0.0, 0.0, 0.0], [0, 0, 0], {"PID": 111, "mass": mass_dict[111]})
    assert np.array_equal(rest.boost_matrix(), np.identity(4))


def test_two_body_conservation_and_masses():
    np.random.seed(1)
    for pid, E in ((111, 0.2), (221, 5.0), (331, 80.0)):
        q = _meson(pid, E)
        for _ in range(50):
            a, b = q.two_body_decay({"PID": 22}, {"PID": 4900022, "mass": 0.05})
            np.testing.assert_allclose(np.asarray(a.get_p0()) + np.asarray(b.get_p0()), q.get_pf(), rtol=1e-12,
                                       atol=1e-12 * E)
            assert abs(_m2(a.get_p0())) <= 1e-10 * E * E
            np.testing.assert_allclose(_m2(b.get_p0()), 0.05 ** 2, rtol=1e-7, atol=1e-10 * E * E)
            assert np.array_equal(a.get_r0(), q.get_rf()) and a.get_ids()["mass"] == 0.0


def test_two_body_isotropic_at_rest_and_flat_in_the_lab():
    np.random.seed(2)
    m = mass_dict[221]
    rest = Particle([m, 0.0, 0.0, 0.0], [0, 0, 0], {"PID": 221, "mass": m})
    ct = []
    for _ in range(3000):
        _, v = rest.two_body_decay({"PID": 22}, {"PID": 4900022, "mass": 0.1})
        p = np.asarray(v.get_p0())
        ct.append(p[3] / np.linalg.norm(p[1:]))
    assert stats.kstest(ct, "uniform", args=(-1, 2)).pvalue > 1e-3
    # moving meson: E_V uniform between gamma (E* -+ beta p*)
    E, mV = 12.0, 0.1
    q = _meson(221, E)
    g = E / m
    b = np.sqrt(1 - 1 / g ** 2)
    Es, ps = (m * m + mV * mV) / (2 * m), (m * m - mV * mV) / (2 * m)
    EV = [q.two_body_decay({"PID": 22}, {"PID": 4900022, "mass": mV})[1].get_p0()[0] for _ in range(3000)]
    lo, hi = g * (Es - b * ps), g * (Es + b * ps)
    assert stats.kstest(EV, "uniform", args=(lo, hi - lo)).pvalue > 1e-3


def test_two_body_angular_information():
    np.random.seed(3)
    m = mass_dict[111]
    rest = Particle([m, 0.0, 0.0, 0.0], [0, 0, 0], {"PID": 111, "mass": m})
    # product 2 along (sin t sin f, sin t cos f, cos t): cos t = 1 -> +z
    a, b = rest.two_body_decay({"PID": 22}, {"PID": 22}, angular_information=(lambda u: 1.0, lambda u: 0.3))
    np.testing.assert_allclose(b.get_p0(), [m / 2, 0, 0, m / 2], atol=1e-15)
    np.testing.assert_allclose(a.get_p0(), [m / 2, 0, 0, -m / 2], atol=1e-15)
    # cos t = 0, f = 0 -> +y
    _, b = rest.two_body_decay({"PID": 22}, {"PID": 22}, angular_information=(lambda u: 0.0, lambda u: 0.0))
    np.testing.assert_allclose(b.get_p0(), [m / 2, 0, m / 2, 0], atol=1e-15)
    # one inverse CDF: cos t only
    cts = [rest.two_body_decay({"PID": 22}, {"PID": 22}, angular_information=lambda u: 2 * u ** 0.5 - 1)[1]
           .get_p0()[3] / (m / 2) for _ in range(2000)]
    assert stats.kstest((np.array(cts) + 1) / 2, lambda x: x ** 2).pvalue > 1e-3


def test_missing_masses_raise():
    q = _meson(111, 1.0)
    with pytest.raises(ValueError):
        q.two_body_decay({"weight": 1.0}, {"PID": 22})
    with pytest.raises(ValueError):
        q.two_body_decay({"PID": 999999}, {"PID": 22})
    d = {"PID": 22}
    q.two_body_decay(d, {"PID": 22})
    assert d["mass"] == 0.0  # filled in, as the reference does


def _dalitz_width(q, s12, m1, m2, m3, M):
    a, b = q.dalitz_range(s12, m1, m2, m3, M)
    return b - a


def test_three_body_conservation_and_flat_dalitz():
    np.random.seed(4)
    pid = 221
    M = mass_dict[pid]
    m1 = m2 = m3 = mass_dict[111]
    q = _meson(pid, 4.0)
    s12 = []
    for _ in range(2000):
        ks = q.three_body_decay({"PID": 111}, {"PID": 111}, {"PID": 111})
        p = [np.asarray(k.get_p0()) for k in ks]
        np.testing.assert_allclose(sum(p), q.get_pf(), rtol=1e-12, atol=1e-12 * 4.0)
        for k in p:
            np.testing.assert_allclose(_m2(k), m1 * m1, rtol=1e-7)
        s12.append(_m2(p[0] + p[1]))
    lo, hi = (m1 + m2) ** 2, (M - m3) ** 2
    tot = integrate.quad(lambda s: _dalitz_width(q, s, m1, m2, m3, M), lo, hi)[0]
    cdf = np.vectorize(lambda x: integrate.quad(lambda s: _dalitz_width(q, s, m1, m2, m3, M), lo, x)[0] / tot)
    assert stats.kstest(s12, cdf).pvalue > 1e-3


def test_three_body_matrix_element_and_rest_isotropy():
    np.random.seed(5)
    M = mass_dict[331]
    m1, m2, m3 = mass_dict[111], mass_dict[111], mass_dict[221]
    rest = Particle([M, 0.0, 0.0, 0.0], [0, 0, 0], {"PID": 331, "mass": M})
    lo, hi = (m1 + m2) ** 2, (M - m3) ** 2
    f = lambda s12, s23, *a: (s12 - lo) / (hi - lo)
    s12, ct = [], []
    for _ in range(2000):
        ks = rest.three_body_decay({"PID": 111}, {"PID": 111}, {"PID": 221}, dalitz_information=(1.0, f))
        p = [np.asarray(k.get_p0()) for k in ks]
        s12.append(_m2(p[0] + p[1]))
        ct.append(p[0][3] / np.linalg.norm(p[0][1:]))
    w = lambda s: _dalitz_width(rest, s, m1, m2, m3, M) * (s - lo) / (hi - lo)
    tot = integrate.quad(w, lo, hi)[0]
    cdf = np.vectorize(lambda x: integrate.quad(w, lo, x)[0] / tot)
    assert stats.kstest(s12, cdf).pvalue > 1e-3
    assert stats.kstest(ct, "uniform", args=(-1, 2)).pvalue > 1e-3
    with pytest.raises(ValueError):
        rest.three_body_decay({"PID": 111}, {"PID": 111}, {"PID": 221}, dalitz_information="not flat")

    def broken(*args):
        raise ZeroDivisionError("inside the matrix element")
    with pytest.raises(ZeroDivisionError):
        rest.three_body_decay({"PID": 111}, {"PID": 111}, {"PID": 221}, dalitz_information=(1.0, broken))


def test_decay_distance_law():
    np.random.seed(6)
    q = _meson(211, 5.0)
    L_int, ctau0 = 0.1796, 7.8045
    m = mass_dict[211]
    g = 5.0 / m
    R = 1 / (ctau0 * g * np.sqrt(1 - 1 / g ** 2)) + 1 / L_int
    r0, dens = q.decay_int_prob(0.0, L_int, ctau0)
    np.testing.assert_allclose([r0, dens], [R, R], rtol=1e-12)
    np.testing.assert_allclose(q.decay_int_prob(0.3, L_int, ctau0)[1], R * np.exp(-0.3 * R), rtol=1e-12)
    x = np.array([q.draw_x_sample(L_int, ctau0) for _ in range(3000)])
    assert x.max() <= 4 / R
    trunc = lambda t: (1 - np.exp(-R * t)) / (1 - np.exp(-4.0))
    assert stats.kstest(x, trunc).pvalue > 1e-3
    ctau = ctau0 * g * np.sqrt(1 - 1 / g ** 2)
    P = [q.prob_decay_b_int(L_int, ctau0) for _ in range(200)]
    assert min(P) >= 0 and max(P) <= (1 - np.exp(-4.0)) / (1 + ctau / L_int) + 1e-15


def test_decay_particle_bookkeeping():
    np.random.seed(7)
    q = _meson(111, 3.0, stability="short-lived", ID=5, weight=2.0, generation_number=3, decay_time=1.5)
    out = q.decay_particle()
    assert q.get_ended() and len(out) == 2
    np.testing.assert_allclose(np.asarray(out[0].get_p0()) + np.asarray(out[1].get_p0()), q.get_pf(), rtol=1e-12)
    for k, d in enumerate(o.get_ids() for o in out):
        assert d["PID"] == 22 and d["ID"] == 10 + k and d["generation_number"] == 4
        assert d["generation_process"] == "SMDecay" and d["production_time"] == 1.5
        np.testing.assert_allclose(d["weight"], 2.0 * 0.98823, rtol=1e-15)
    # long-lived pi+: moved along pf to its decay point, mu+ nu_mu, weights x P(decay before interaction)
    pi = _meson(211, 5.0, stability="long-lived")
    r_before = np.asarray(pi.get_rf(), dtype=np.float64).copy()
    mu, nu = pi.decay_particle()
    step = np.asarray(pi.get_rf()) - r_before
    pf = np.asarray(pi.get_pf())
    np.testing.assert_allclose(np.cross(step, pf[1:]), 0.0, atol=1e-12)
    assert step @ pf[1:] > 0
    assert (mu.get_ids()["PID"], nu.get_ids()["PID"]) == (-13, 14)
    assert 0 < mu.get_ids()["weight"] < 0.9998 and 0 < nu.get_ids()["weight"] < 0.9998
    np.testing.assert_allclose(mu.get_r0(), pi.get_rf())
    # stable: nothing, ended
    s = _meson(111, 3.0)
    assert s.decay_particle() == [] and s.get_ended()
    with pytest.raises(ValueError):
        _meson(13, 3.0).decay_particle()  # no decay channels listed


def test_eta_channel_fractions():
    # eta: gamma gamma with probability 0.3936 / (0.3936 + 0.3257), else the
    # three-pi0 channel, which raises (three-body SM decays are not implemented)
    np.random.seed(8)
    n, ok = 3000, 0
    for _ in range(n):
        q = _meson(221, 2.0, stability="short-lived")
        try:
            out = q.decay_particle()
            ok += 1
            np.testing.assert_allclose(out[0].get_ids()["weight"], 0.3936 + 0.3257, rtol=1e-15)
        except ValueError:
            pass
    p = meson_decay_dict[221][0][0] / (meson_decay_dict[221][0][0] + meson_decay_dict[221][1][0])
    assert stats.binomtest(ok, n, p).pvalue > 1e-3
