"""Pin the oracle's accept/reject draw (draw_sample, src/PETITE/shower.py:401-465)
by the law of its accepted points.

With candidates x = map(y), y uniform, density 1/J(x) on the map's domain,
and acceptance when max_F * u < J(x) f(x), the accepted points have density
proportional to f on the map's domain, whatever the map, as long as max_F
bounds J f (no clipping).  So:
* 1-D processes: the histogram of accepted x against f itself, normalised on
  a 2e5-point grid of the domain (chi-square over 40 bins of equal expected
  mass), at energies between grid energies, where f is evaluated at the true
  energy and the map is the next grid energy's (LU_Key);
* Brem and PairProd (4-D): each marginal of the accepted points against the
  J f-weighted histogram of 10^6 map points (the VEGAS estimate of the same
  marginal), 20 bins of equal weighted mass per dimension.
max_F is set above the largest J f on the fine grid (1-D) or in the 10^6
points (4-D), so the law is f's and not the clipped one.  A draw that
inverts the test, drops J, evaluates f at the grid energy, takes the accept
uniform from a map coordinate or forgets max_F fails these (checked by
planting each mistake in a copy of the oracle).
"""
import numpy as np
import pytest
from scipy import stats

from oracle import vegas
from oracle.integrands import INTEGRAND
from oracle.tables import SMTables

SEED = 77


@pytest.fixture(scope="module")
def graphite():
    return SMTables("graphite")


def _between(grid, E):
    """An energy halfway (in log) between the grid energies around E."""
    i = int(np.searchsorted(grid.E, E))
    return float(np.sqrt(grid.E[i - 1] * grid.E[i]))


def _accepted(grid, P, E, ev, max_f, n, ndim):
    xs = np.empty((n, ndim))
    for s in range(n):
        x, trials, w = vegas.draw(grid, P, E, ev, max_f, SEED, np.uint64(s), max_trials=100_000, chunk=64)
        assert x is not None
        xs[s] = x[:ndim]
    return xs


@pytest.mark.parametrize("P,E0", [("Comp", 0.004), ("Comp", 3.0), ("Ann", 0.05), ("Moller", 0.5), ("Bhabha", 0.5),
                                  ("MuonE", 5.0)])
def test_one_dim_accepted_points_follow_f(graphite, P, E0):
    g = graphite.grids[P]
    E = _between(g, E0)
    ev = graphite.event_info(P, E)
    ie = g.lu_key(E)
    nodes = g.dim_nodes(ie, 0)
    lo, hi = float(nodes[0]), float(nodes[-1])
    # f on a fine grid of the map domain (midpoints), its CDF, and sup J f
    m = 200_000
    xf = lo + (np.arange(m) + 0.5) * (hi - lo) / m
    f = INTEGRAND[P](xf[:, None], ev)
    N = len(nodes) - 1
    inc = np.clip(np.searchsorted(nodes, xf, side="right") - 1, 0, N - 1)
    J = N * (nodes[inc + 1] - nodes[inc])
    max_f = 1.3 * float(np.max(J * f))
    cdf = np.cumsum(f)
    cdf /= cdf[-1]
    nb = 40
    edges = np.interp(np.linspace(0, 1, nb + 1), np.concatenate([[0.0], cdf]),
                      np.concatenate([[lo], xf + 0.5 * (hi - lo) / m]))
    n = 6000
    xs = _accepted(g, P, E, ev, max_f, n, 1)[:, 0]
    cnt, _ = np.histogram(xs, bins=edges)
    assert cnt.sum() == n
    chi2 = float(np.sum((cnt - n / nb) ** 2 / (n / nb)))
    assert stats.chi2.sf(chi2, nb - 1) > 1e-4, (P, E, chi2)


@pytest.mark.parametrize("P,E0", [("Brem", 0.3), ("Brem", 20.0), ("PairProd", 2.0)])
def test_four_dim_marginals_follow_vegas_estimate(graphite, P, E0):
    g = graphite.grids[P]
    E = _between(g, E0)
    ev = graphite.event_info(P, E)
    ie = g.lu_key(E)
    rng = np.random.default_rng(3)
    y = rng.uniform(size=(1_000_000, 4))
    xr, J = vegas.map_points(g, ie, y)
    wr = J * INTEGRAND[P](xr, ev)
    max_f = 2.0 * float(np.max(wr))
    n = 4000
    xs = _accepted(g, P, E, ev, max_f, n, 4)
    nb = 20
    for d in range(4):
        order = np.argsort(xr[:, d])
        cw = np.cumsum(wr[order])
        cw /= cw[-1]
        cut = np.searchsorted(cw, np.linspace(0, 1, nb + 1)[1:-1])
        edges = np.concatenate([[-np.inf], xr[order][cut, d], [np.inf]])
        cnt, _ = np.histogram(xs[:, d], bins=edges)
        chi2 = float(np.sum((cnt - n / nb) ** 2 / (n / nb)))
        assert stats.chi2.sf(chi2, nb - 1) > 1e-4, (P, E, d, chi2)
