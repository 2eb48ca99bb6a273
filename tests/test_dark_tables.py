"""Dark-sector tables: bound-electron cross sections, track-length weights and
interaction-energy rates (DarkShower initialisation, src/PETITE/dark_shower.py:85-593).

The oracle's bound-electron closed forms (atomic_annihilation.py, atomic_compton.py)
are pinned in 60-digit arithmetic against the large-b expansions the reference
carries beside them and against continuity at the expr1/expr2 branch point.  The
product reproduces the oracle's float64 values exactly (same evaluation order:
above b ~ 100 both are dominated by cancellation, as the reference is).  The
panel rule of the weight and rate integrals (oracle.dark_panels, which the
device quadrature follows: tests/test_gpu_dark_quadrature.py) is compared with
the oracle's scipy quad integrals on the same tables, and with the caches
shipped with the reference (made from other SM tables: ~1 % level).
"""
import mpmath as mp
import numpy as np
import pytest

from oracle.dark import (OracleDark, _expr1, _expr2, comp_series, fancy_integral, fancy_series,
                         sigma_atomic, sigma_atomic_comp)
from oracle.dark_panels import Panels
from oracle.tables import SMTables
from petite_amd.dark_tables import DarkTables, closest_lesser_value
from petite_amd.tables import MaterialTables


class _MP:
    log = staticmethod(mp.log)
    atan = staticmethod(mp.atan)
    atan2 = staticmethod(mp.atan2)
    pi = mp.pi


A_VALUES = [mp.mpf(137) / z for z in (6, 29.508, 82)]


@pytest.fixture(autouse=True)
def _dps():
    old = mp.mp.dps
    mp.mp.dps = 60
    yield
    mp.mp.dps = old


@pytest.mark.parametrize("a", A_VALUES)
def test_annihilation_tail_matches_large_b_series(a):
    # the expansion is exact to O(b^-9): relative gap falls as b^-3
    gaps = []
    for b in (1000, 3000, 10000):
        F = fancy_integral(a, mp.mpf(b), _MP, series=False)
        gaps.append(abs(fancy_series(a, mp.mpf(b)) / F - 1))
    assert gaps[0] < 2e-3
    assert gaps[1] < gaps[0] / 20 and gaps[2] < gaps[1] / 20


@pytest.mark.parametrize("a", A_VALUES)
def test_compton_expr1_matches_large_b_series(a):
    gaps = []
    for b in (1000, 3000, 10000):
        E1 = _expr1(a, mp.mpf(b), _MP)
        gaps.append(abs(comp_series(a, mp.mpf(b)) / E1 - 1))
    assert gaps[0] < 1e-3
    assert gaps[1] < gaps[0] / 20 and gaps[2] < gaps[1] / 20


@pytest.mark.parametrize("a", A_VALUES)
def test_compton_branches_meet_at_knee(a):
    bk = (a * a + 1) / a
    lo = _expr2(a, bk * (1 - mp.mpf("1e-12")), _MP)
    hi = _expr1(a, bk * (1 + mp.mpf("1e-12")), _MP)
    assert lo > 0 and abs(hi / lo - 1) < 1e-9


def test_float64_forms_accurate_where_well_conditioned():
    a = 137 / 29.508
    for b in (0.1, 1.0, 3.0, 10.0):
        am, bm = mp.mpf(137) / mp.mpf(29.508), mp.mpf(b)
        F = fancy_integral(am, bm, _MP, series=False)
        assert abs(fancy_integral(a, b) / float(F) - 1) < 1e-11
        C = _expr1(am, bm, _MP) if am * am + 1 < am * bm else _expr2(am, bm, _MP)
        c = _expr1(a, b) if a * a + 1 < a * b else _expr2(a, b)
        assert abs(c / float(C) - 1) < 1e-11


def test_tree_level_is_hydrogenic_momentum_tail():
    # 1/((a-b)^2+1)^3 = 6 * int_{a-b}^inf p dp / (1+p^2)^4  (1s momentum density)
    from scipy.integrate import quad
    for y in (-3.0, 0.0, 0.7, 4.0):
        v = 6 * quad(lambda p: p / (1 + p * p) ** 4, y, np.inf)[0]
        assert abs(v * (y * y + 1) ** 3 - 1) < 1e-9


def test_closest_lesser_value():
    m = [0.003, 0.01, 0.03, 0.1]
    assert closest_lesser_value(m, 0.02) == 0.01
    assert closest_lesser_value(m, 0.03) == 0.03
    assert closest_lesser_value(m, 0.5) == 0.1
    assert closest_lesser_value(m, 0.001) == 0.003
    assert closest_lesser_value(m, 0.001, below="wrap") == 0.1


@pytest.fixture(scope="module", params=[("lead", 0.01), ("graphite", 0.003)])
def pair(request):
    mat, mv = request.param
    return DarkTables(MaterialTables(mat), mv), OracleDark(SMTables(mat), mv)


@pytest.fixture(scope="module", params=[("lead", 0.01), ("graphite", 0.003)])
def panels(request):
    mat, mv = request.param
    od = OracleDark(SMTables(mat), mv)
    return Panels(od), od


def test_bound_sigma_grids_identical(pair):
    d, o = pair
    np.testing.assert_array_equal(d.xsec["DarkAnn"], o.DAS)
    np.testing.assert_array_equal(d.xsec["DarkComp"], o.DCS)
    assert np.all(d.xsec["DarkAnn"][:, 1] > 0) and np.all(d.xsec["DarkComp"][:, 1] >= 0)


def test_bound_sigma_scalar_forms(pair):
    d, _ = pair
    Ea = d.xsec["DarkAnn"][::37, 0]
    assert np.array_equal([sigma_atomic(e, d.mV, d.Zeff) for e in Ea], d.xsec["DarkAnn"][::37, 1])
    Ec = d.xsec["DarkComp"][::37, 0]
    assert np.array_equal([sigma_atomic_comp(e, d.mV, d.Zeff) for e in Ec], d.xsec["DarkComp"][::37, 1])


@pytest.mark.parametrize("proc,pid", [("DarkBrem", 11), ("DarkBrem", -11), ("DarkMuonBrem", 13), ("DarkAnn", -11)])
def test_panel_weights_match_quad(panels, proc, pid):
    P, o = panels
    W = P.weight_array(proc, pid)
    E0 = W[0, 0]
    for i in range(1, len(W), max(1, len(W) // 6)):
        wo = o.weight(proc, pid, W[i, 0], E0)
        if wo > 0:
            assert abs(W[i, 1] / wo - 1) < 5e-4, (proc, pid, W[i, 0])
        else:
            assert W[i, 1] <= 1e-30


def test_panel_drate_matches_quad(panels):
    P, o = panels
    for proc, pid in (("DarkBrem", 11), ("DarkAnn", -11)):
        Es = P.energies(proc)
        for i in range(3, len(Es), max(1, len(Es) // 5)):
            mine = P.drate(proc, pid, Es[i], Es[0])
            r = o.drate(proc, pid, Es[i], Es[0])
            np.testing.assert_allclose(mine[:, 0], r[:, 0], rtol=1e-12)
            ok = r[:, 1] > 1e-300
            assert np.max(np.abs(mine[ok, 1] / r[ok, 1] - 1)) < 2e-4, (proc, Es[i])


def test_panel_integrals_close_to_shipped_caches():
    od = OracleDark(SMTables("graphite"), 0.03)
    P = Panels(od)
    W, _ = P.tables()
    # MuonBrem's SM tables carry the largest MC noise (several % in graphite)
    for name, proc, pid, tol in (("brem_elec", "DarkBrem", 11, 0.05), ("brem_positron", "DarkBrem", -11, 0.05),
                                 ("muon_brem", "DarkMuonBrem", 13, 0.15)):
        Wc = W[name]
        idx = np.arange(len(Wc) // 4, len(Wc), len(Wc) // 5)
        mine = np.array([P.integrate(proc, pid, Wc[0, 0], Wc[i, 0], Wc[i, 0], P.nodes(proc, pid)) for i in idx])
        assert np.max(np.abs(mine / Wc[idx, 1] - 1)) < tol, name


def test_cumint_exact_for_piecewise_linear():
    # the running integral of a piecewise-linear table on another grid: a
    # quadratic's integral where the table is linear, zero outside it
    from oracle.dark_panels import cumint
    xg = np.array([1.0, 2.0, 4.0, 7.0])
    yg = np.array([2.0, 3.0, 1.0, 1.0])
    nodes = np.array([0.5, 1.5, 3.0, 5.0, 8.0])
    # int_1^1.5 (1 + x) = 1.125, to 2: 2.5, to 3: 5.0, to 4: 6.5, to 5: 7.5, to 7: 9.5, then 0
    np.testing.assert_allclose(cumint(xg, yg, nodes), [0.0, 1.125, 5.0, 7.5, 9.5], rtol=1e-15)
    np.testing.assert_allclose(cumint(xg, yg, xg), [0.0, 2.5, 6.5, 9.5], rtol=1e-15)
