"""Per-candidate parity of the device map transform and integrands (J*f)
against the float64 oracle on identical Philox words."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import philox, vegas as ovegas
from oracle.integrands import INTEGRAND
from oracle.tables import SMTables
from petite_amd.constants import all_process_code, process_code
from petite_amd.device import DeviceContext
from petite_amd.tables import MaterialTables

SM = ["Brem", "PairProd", "Comp", "Ann", "Moller", "Bhabha", "MuonE", "MuonBrem"]
# float32 bodies (Brem, MuonBrem, PairProd) vs float64 bodies
TOL = {"Brem": 2e-5, "MuonBrem": 2e-5, "PairProd": 2e-5}


@pytest.fixture(scope="module", params=["graphite", "lead"])
def ctx(request):
    return DeviceContext(MaterialTables(request.param), 0.010), SMTables(request.param)


@pytest.mark.parametrize("P", SM)
def test_points(ctx, P):
    dctx, T = ctx
    g = T.grids[P]
    xg, yg = T.nsigma_tab[P]
    lo = xg[np.argmax(yg > 0)]
    rng = np.random.default_rng(3)
    n = 3000
    E = np.exp(rng.uniform(np.log(lo * 1.01), np.log(xg[-1] * 0.999), n))
    ctr = np.zeros((n, 4), dtype=np.uint32)
    ctr[:, 0] = np.arange(n, dtype=np.uint32)
    ctr[:, 3] = philox.PURPOSE_VEGAS_A
    words = philox.philox4x32_10(ctr, np.array([5, 6], dtype=np.uint32))
    x, w = dctx.eval_points(np.full(n, process_code[P], np.int32), E, words)
    y = philox.to_uniform(words)
    wo = np.empty(n)
    xo = np.empty((n, g.ndim))
    for i in range(n):
        xx, J = ovegas.map_points(g, g.lu_key(E[i]), y[i:i + 1, : g.ndim])
        xo[i] = xx[0]
        wo[i] = J[0] * INTEGRAND[P](xx, T.event_info(P, E[i]))[0]
    assert np.allclose(x[:, : g.ndim], xo, rtol=1e-12, atol=1e-12)
    big = wo > 1e-6 * wo.max()
    rel = np.abs(w - wo)[big] / wo[big]
    tol = TOL.get(P, 1e-9)
    assert np.quantile(rel, 0.999) < tol, (P, np.quantile(rel, [0.5, 0.99, 0.999]), rel.max())
    assert rel.max() < 100 * tol, (P, rel.max())
    assert np.all(np.abs(w[~big]) <= 1e-5 * wo.max() + 1e-300)


DARK = ["DarkBrem", "DarkMuonBrem", "DarkAnn", "DarkComp"]
# (0.999 quantile, max) of the relative J*f difference: the exact tree-level
# dark bremsstrahlung integrand (all_processes.py:208-372) cancels between its
# A0/A1/Am1/Am2 terms near the t and cos(theta_q) edges, where float64 orders
# of evaluation differ by up to ~5e-6 relative
DARK_TOL = {"DarkBrem": (1e-5, 5e-5), "DarkMuonBrem": (1e-5, 5e-5)}


@pytest.fixture(scope="module", params=[("graphite", 0.03), ("lead", 0.003)])
def dark_ctx(request):
    from oracle.dark import OracleDark
    from petite_amd.dark_tables import DarkTables
    mat, mv = request.param
    mt = MaterialTables(mat)
    dt = DarkTables(mt, mv)
    return DeviceContext(mt, 0.010, dark=dt), OracleDark(SMTables(mat), mv)


@pytest.mark.parametrize("P", DARK)
def test_dark_points(dark_ctx, P):
    dctx, od = dark_ctx
    T = od.T
    g = od._grids[P]
    rng = np.random.default_rng(4)
    n = 3000
    E = np.exp(rng.uniform(np.log(g.E[0] * 1.01), np.log(g.E[-1] * 0.999), n))
    ctr = np.zeros((n, 4), dtype=np.uint32)
    ctr[:, 0] = np.arange(n, dtype=np.uint32)
    ctr[:, 3] = philox.PURPOSE_VEGAS_A
    words = philox.philox4x32_10(ctr, np.array([7, 8], dtype=np.uint32))
    x, w = dctx.eval_points(np.full(n, all_process_code[P], np.int32), E, words)
    y = philox.to_uniform(words)
    wo = np.empty(n)
    xo = np.empty((n, g.ndim))
    for i in range(n):
        xx, J = ovegas.map_points(g, g.lu_key(E[i]), y[i:i + 1, : g.ndim])
        xo[i] = xx[0]
        wo[i] = J[0] * INTEGRAND[P](xx, T.event_info(P, E[i], od.mV))[0]
    assert np.allclose(x[:, : g.ndim], xo, rtol=1e-12, atol=1e-12)
    big = wo > 1e-6 * wo.max()
    assert big.sum() > n // 10
    rel = np.abs(w - wo)[big] / wo[big]
    q_tol, max_tol = DARK_TOL.get(P, (1e-9, 1e-7))
    assert np.quantile(rel, 0.5) < 1e-9, (P, np.quantile(rel, [0.5, 0.99, 0.999]))
    assert np.quantile(rel, 0.999) < q_tol, (P, np.quantile(rel, [0.5, 0.99, 0.999]), rel.max())
    assert rel.max() < max_tol, (P, rel.max())
    assert np.all(np.abs(w[~big]) <= 1e-5 * wo.max() + 1e-300)
