"""Oracle of the m_V-continuous dark yield (no counterpart in the reference):
consistency of its value and derivative.

* at a trained mass the yield is the plain weight sum of GetBSMWeights
  (src/PETITE/dark_shower.py:595-647) times the parent weights;
* its analytic dY/dm_V matches central finite differences inside a segment;
* the geometric interpolation is exact for weights that follow a power law in
  m_V, and eps^2 enters linearly by construction of the caller.
"""
import numpy as np
import pytest

from oracle.dark import OracleDark, OracleDarkSampler, sampler_weights, yield_and_grad
from oracle.tables import SMTables
from oracle.dark_panels import Panels

MASSES = [0.01, 0.03, 0.1]
PROCS = ["DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"]


@pytest.fixture(scope="module")
def setup():
    T = SMTables("graphite")
    samplers = []
    for m in MASSES:
        od = OracleDark(T, m)
        samplers.append(OracleDarkSampler(od, *Panels(od).tables(), None))
    rng = np.random.default_rng(8)
    n = 400
    pid = rng.choice([11, -11, 22, 13], size=n)
    E0 = np.exp(rng.uniform(np.log(0.02), np.log(50.0), n))
    w = rng.uniform(0.5, 1.5, n)
    return samplers, pid, E0, w


def test_trained_masses_give_weight_sums(setup):
    samplers, pid, E0, w = setup
    Y, _ = yield_and_grad(samplers, MASSES, PROCS, pid, E0, w, MASSES)
    for j, s in enumerate(samplers):
        direct = sum(s.weight(int(p), float(e), P) * ww for p, e, ww in zip(pid, E0, w) for P in PROCS)
        assert abs(Y[j] - direct) <= 1e-12 * abs(direct)
        vec = sum(np.sum(sampler_weights(s, pid, E0, P) * w) for P in PROCS)
        assert abs(vec - direct) <= 1e-12 * abs(direct)


def test_derivative_matches_finite_differences(setup):
    samplers, pid, E0, w = setup
    mv = np.array([0.012, 0.02, 0.05, 0.08])
    Y, dY = yield_and_grad(samplers, MASSES, PROCS, pid, E0, w, mv)
    h = 1e-4
    Yp, _ = yield_and_grad(samplers, MASSES, PROCS, pid, E0, w, mv * (1 + h))
    Ym, _ = yield_and_grad(samplers, MASSES, PROCS, pid, E0, w, mv * (1 - h))
    fd = (Yp - Ym) / (2 * h * mv)
    np.testing.assert_allclose(dY, fd, rtol=1e-5)
    # clamped outside the trained masses
    Yo, dYo = yield_and_grad(samplers, MASSES, PROCS, pid, E0, w, [0.005, 0.2])
    Ye, _ = yield_and_grad(samplers, MASSES, PROCS, pid, E0, w, [MASSES[0], MASSES[-1]])
    np.testing.assert_allclose(Yo, Ye, rtol=1e-14)
    assert np.all(dYo == 0)


class _PowerLaw:
    """Stand-in sampler whose weights are c(E) m^-2 exactly."""

    def __init__(self, m):
        self.m = m


def test_geometric_interpolation_exact_for_power_law(monkeypatch):
    import oracle.dark as od
    monkeypatch.setattr(od, "sampler_weights", lambda s, pid, E0, P: np.asarray(E0) / s.m ** 2)
    E0 = np.linspace(1, 2, 7)
    Y, dY = od.yield_and_grad([_PowerLaw(m) for m in MASSES], MASSES, ["DarkBrem"], np.zeros(7), E0, np.ones(7),
                              [0.017, 0.06])
    np.testing.assert_allclose(Y, E0.sum() / np.array([0.017, 0.06]) ** 2, rtol=1e-12)
    np.testing.assert_allclose(dY, -2 * E0.sum() / np.array([0.017, 0.06]) ** 3, rtol=1e-12)


def test_bound_yield_at_trained_mass_equals_weight_sum(setup):
    # tables built at a trained mass are that mass's own tables
    from oracle.dark import bound_yield
    samplers, pid, E0, w = setup
    T = samplers[0].T
    for j, m in enumerate(MASSES):
        Yb, _ = bound_yield(T, ["DarkAnn", "DarkComp"], pid, E0, w, [m])
        want = sum(np.sum(sampler_weights(samplers[j], pid, E0, P) * w) for P in ("DarkAnn", "DarkComp"))
        assert want > 0 and abs(Yb[0] / want - 1) < 1e-12, (m, Yb[0], want)


def test_bound_yield_derivative_is_a_consistent_central_difference(setup):
    from oracle.dark import bound_yield
    samplers, pid, E0, w = setup
    T = samplers[0].T
    m = 0.013
    Y, dY = bound_yield(T, ["DarkAnn", "DarkComp"], pid, E0, w, [m])
    h = 1e-4
    Yp, _ = bound_yield(T, ["DarkAnn", "DarkComp"], pid, E0, w, [m * (1 + h)])
    Ym, _ = bound_yield(T, ["DarkAnn", "DarkComp"], pid, E0, w, [m * (1 - h)])
    assert abs(dY[0] / ((Yp[0] - Ym[0]) / (2 * h * m)) - 1) < 1e-5


def test_interpolant_derivative_exact_for_linear_tables():
    # nodes x(m) = x0 (1 + m), values y = a(m) x + b(m): the interpolant is
    # the line itself at any m, so its derivative at fixed v is a'(m) v + b'(m)
    from oracle.dark import interp_with_derivative
    m = 0.3
    x0 = np.array([1.0, 1.7, 2.9, 4.0, 6.5])
    a, da, b, db = 2.0 + m * m, 2 * m, 0.5 - m, -1.0
    x, dx = x0 * (1 + m), x0
    y, dy = a * x + b, da * x + a * dx + db
    v = np.array([1.3, 2.0, 3.3, 7.4, 8.45, 0.5, 9.0])
    val, der = interp_with_derivative(x, y, dx, dy, v)
    ins = (v >= x[0]) & (v <= x[-1])
    np.testing.assert_allclose(val[ins], a * v[ins] + b, rtol=1e-13)
    np.testing.assert_allclose(der[ins], da * v[ins] + db, rtol=1e-12, atol=1e-12)
    assert np.all(val[~ins] == 0) and np.all(der[~ins] == 0)
