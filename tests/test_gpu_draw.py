"""GPU parity of the accept/reject draw and the max_F scan against the oracle.

Identical Philox streams on both sides: the accepted candidate index (trial
count) is index work and must agree in every draw.  The kernel's float32
4-D integrands decide alone only outside a 4e-4 relative band around the
threshold max_F * u; inside it the float64 reference formula decides, as the
oracle does.  The accepted points agree to float32 rounding of the stored x.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import vegas as ovegas
from oracle.tables import SMTables
from petite_amd.constants import process_code
from petite_amd.device import DeviceContext
from petite_amd.tables import MaterialTables

SM = ["Brem", "PairProd", "Comp", "Ann", "Moller", "Bhabha", "MuonE", "MuonBrem"]
SEED = 1234


@pytest.fixture(scope="module")
def ctxs():
    out = {}
    for mat in ("graphite", "lead"):
        out[mat] = (DeviceContext(MaterialTables(mat), 0.010), SMTables(mat))
    return out


@pytest.mark.parametrize("mat", ["graphite", "lead"])
@pytest.mark.parametrize("P", SM)
def test_maxf_scan_matches_oracle(ctxs, mat, P):
    ctx, T = ctxs[mat]
    mx = ctx.maxf[P]
    mean = ctx.sigma_est[P]
    for ie in (5, 37, 80, 99):
        ev = T.event_info(P, T.grids[P].E[ie])
        omx, omean = ovegas.scan_maxf(T, P, ie, ev)
        assert abs(mx[ie] - omx) <= 2e-5 * abs(omx) + 1e-300, (P, ie, mx[ie], omx)
        assert abs(mean[ie] - omean) <= 2e-5 * abs(omean) + 1e-300, (P, ie, mean[ie], omean)


def _energies(T, P, n, rng):
    xg, yg = T.nsigma_tab[P]
    lo = xg[np.argmax(yg > 0)]
    hi = xg[-1]
    return np.exp(rng.uniform(np.log(lo * 1.02), np.log(hi * 0.999), n))


@pytest.mark.parametrize("mat", ["graphite", "lead"])
@pytest.mark.parametrize("P", SM)
def test_draw_matches_oracle(ctxs, mat, P):
    ctx, T = ctxs[mat]
    rng = np.random.default_rng(7 if mat == "graphite" else 8)
    n = 600 if P in ("Brem", "PairProd", "MuonBrem") else 1000
    E = _energies(T, P, n, rng)
    sid = np.arange(n, dtype=np.uint64) + 1000
    x, trials, w, rc = ctx.draw(np.full(n, process_code[P], np.int32), E, sid, SEED)
    assert rc == 0
    x = x.cpu().numpy().astype(np.float64)
    trials = trials.cpu().numpy()
    g = T.grids[P]
    same = 0
    for i in range(n):
        ie = g.lu_key(E[i])
        mf = ctx.maxf[P][ie]
        xo, to, wo = ovegas.draw(g, P, E[i], T.event_info(P, E[i]), mf, SEED, sid[i])
        if to == trials[i]:
            same += 1
            d = g.ndim
            assert np.allclose(x[i, :d], xo[:d], rtol=2e-6, atol=2e-6), (P, i, x[i, :d], xo)
    assert same == n, (P, mat, same, n)
