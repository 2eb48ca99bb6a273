"""Dark-sector tables: bound-electron cross sections, track-length weights and
interaction-energy rates (DarkShower initialisation, src/PETITE/dark_shower.py:85-593).

The oracle's bound-electron closed forms (atomic_annihilation.py, atomic_compton.py)
are pinned in 60-digit arithmetic against the large-b expansions the reference
carries beside them and against continuity at the expr1/expr2 branch point.  The
product reproduces the oracle's float64 values exactly (same evaluation order:
above b ~ 100 both are dominated by cancellation, as the reference is).  The
weight and rate integrals of the product (Gauss-Legendre panels) are compared
with the oracle's scipy quad integrals on the same tables, and with the caches
shipped with the reference (made from other SM tables: ~1 % level).
"""
import mpmath as mp
import numpy as np
import pytest

from oracle.dark import (OracleDark, _expr1, _expr2, comp_series, fancy_integral, fancy_series,
                         sigma_atomic, sigma_atomic_comp)
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


@pytest.mark.parametrize("name,proc,pid", [("brem_elec", "DarkBrem", 11), ("brem_positron", "DarkBrem", -11),
                                           ("muon_brem", "DarkMuonBrem", 13), ("annihilation", "DarkAnn", -11)])
def test_weights_match_quad(pair, name, proc, pid):
    d, o = pair
    W = d.W[name]
    E0 = W[0, 0]
    for i in range(1, len(W), max(1, len(W) // 6)):
        wo = o.weight(proc, pid, W[i, 0], E0)
        if wo > 0:
            assert abs(W[i, 1] / wo - 1) < 5e-4, (name, W[i, 0])
        else:
            assert W[i, 1] <= 1e-30


def test_drate_matches_quad(pair):
    d, o = pair
    for name, proc, pid in (("brem_elec", "DarkBrem", 11), ("annihilation", "DarkAnn", -11)):
        Es, arr, _ = d.D[name]
        for i in range(3, len(Es), max(1, len(Es) // 5)):
            r = o.drate(proc, pid, Es[i], Es[0])
            np.testing.assert_allclose(arr[i, :, 0], r[:, 0], rtol=1e-12)
            ok = r[:, 1] > 1e-300
            assert np.max(np.abs(arr[i, ok, 1] / r[ok, 1] - 1)) < 2e-4, (name, Es[i])


def test_integrals_close_to_shipped_caches():
    d = DarkTables(MaterialTables("graphite"), 0.03)
    assert d.from_cache
    # MuonBrem's SM tables carry the largest MC noise (several % in graphite)
    for name, proc, pid, tol in (("brem_elec", "DarkBrem", 11, 0.05), ("brem_positron", "DarkBrem", -11, 0.05),
                                 ("muon_brem", "DarkMuonBrem", 13, 0.15)):
        W = d.W[name]
        idx = np.arange(len(W) // 4, len(W), len(W) // 5)
        mine = np.array([d.integrate(proc, pid, W[0, 0], W[i, 0], W[i, 0], d._nodes(proc, pid)) for i in idx])
        assert np.max(np.abs(mine / W[idx, 1] - 1)) < tol, name


def test_bsm_weight_selection():
    d = DarkTables(MaterialTables("graphite"), 0.03)
    E = np.array([0.01, 1.0, 10.0])
    w = d.bsm_weight(np.array([11, 11, 11]), E, "DarkBrem")
    assert w[0] == 0.0 or E[0] >= d.min_calc_dark[11]["DarkBrem"]
    assert np.all(np.diff(w[1:]) > 0)
    assert np.all(d.bsm_weight(np.array([22, 22, 22]), E, "DarkBrem") == 0)
    assert np.all(d.bsm_weight(np.array([11, 11, 11]), E, "DarkAnn") == 0)
