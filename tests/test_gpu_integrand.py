"""Per-candidate parity of the device map transform and integrands (J*f)
against the float64 oracle on identical Philox words."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import philox, vegas as ovegas
from oracle.integrands import INTEGRAND
from oracle.tables import SMTables
from petite_amd.constants import process_code
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
