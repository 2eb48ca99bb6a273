"""m_V-continuous dark yield on the GPU (k_dark_yield) against the float64
oracle, central finite differences and torch autograd.  Record: lead, e-
primaries with E log-uniform in [1, 20] GeV, cutoff 0.010 GeV; masses
geomspace(1, 100) MeV.  Tolerances: 1e-9 relative against the oracle (both
float64 on the same float32 record values; DarkAnn and DarkComp from tables
at each m_V, the other processes interpolated between trained masses); 1e-4 against central finite
differences with step 1e-5 m_V at all 32 masses, one-sided (second order)
at the trained ones, where the interpolation has a kink (between two
trained masses a particle's weight goes as a power of m_V, up to m_V^-30
where a resonance moves past it, so a step of 1e-3 m_V would leave a
truncation error near 1e-4)."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle.dark import OracleDark, OracleDarkSampler, yield_and_grad
from oracle.tables import SMTables
from petite_amd import Particle
from petite_amd.constants import mass_dict
from petite_amd.dark_shower import DarkShower

PROCS = ["DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"]
MV = np.geomspace(1e-3, 0.1, 32)


@pytest.fixture(scope="module")
def run():
    ds = DarkShower("100GeV", "lead", 0.010, 0.01, seed=4)
    rng = np.random.default_rng(6)
    m = mass_dict[11]
    prims = []
    for E in np.exp(rng.uniform(0.0, math.log(20.0), 24)):
        prims.append(Particle([E, 0.0, 0.0, math.sqrt(E * E - m * m)], [0, 0, 0], {"PID": 11}))
    batch = ds.generate_showers(prims)
    mV = torch.tensor(MV, dtype=torch.float64, requires_grad=True)
    ds.dark_yield(batch, mV[:1], 1.0)  # builds the yield model
    model = ds._yield_model
    T = SMTables("lead")
    samplers = [OracleDarkSampler(OracleDark(T, m, with_integrals=False, masses="all"), model.dark[m].W,
                                  model.dark[m].D, None) for m in model.masses]
    a = batch.to_numpy()
    return ds, batch, model, samplers, a


def test_yield_matches_oracle(run):
    # every fourth mass: the oracle builds the bound-electron tables of each
    # mass (and of m_V (1 +- 1e-5)) on the CPU
    ds, batch, model, samplers, a = run
    mv = MV[::4]
    Y, dY = model.evaluate(batch, mv)
    Yo, dYo = yield_and_grad(samplers, model.masses, PROCS, a["pid"], a["p0"][:, 0].astype(np.float64),
                             a["r0"][:, 3].astype(np.float64), mv, bound_T=SMTables("lead"))
    assert batch.n > 2000
    np.testing.assert_allclose(Y, Yo, rtol=1e-9, atol=0)
    # the bound-electron part of dY/dm_V is a central difference of step
    # 1e-5 m_V: table differences at 1e-14 show up near 1e-9 there
    np.testing.assert_allclose(dY, dYo, rtol=1e-7, atol=1e-12 * np.abs(dYo).max())


def _fd(model, batch, masses, h=1e-5):
    # central between trained masses; at a trained mass (a kink of the
    # interpolation) second-order one-sided on the side the model
    # differentiates: the segment starting there (ending there for the heaviest)
    out = []
    for m in masses:
        if not any(abs(m / t - 1) < 1e-12 for t in model.masses):
            Yp, _ = model.evaluate(batch, [m * (1 + h)], grad=False)
            Ym, _ = model.evaluate(batch, [m * (1 - h)], grad=False)
            out.append((Yp[0] - Ym[0]) / (2 * h * m))
            continue
        sgn = -1.0 if m >= model.masses[-1] * (1 - 1e-12) else 1.0
        Y = [model.evaluate(batch, [m * (1 + sgn * i * h)], grad=False)[0][0] for i in range(3)]
        out.append(sgn * (-3 * Y[0] + 4 * Y[1] - Y[2]) / (2 * h * m))
    return np.array(out)


def test_gradient_matches_finite_differences(run):
    # every mass of the scan, the trained ones at 1 MeV and 100 MeV included
    ds, batch, model, _, _ = run
    Y, dY = model.evaluate(batch, MV)
    fd = _fd(model, batch, MV)
    np.testing.assert_allclose(dY, fd, rtol=1e-4, atol=1e-9 * np.abs(dY).max())


def test_autograd(run):
    ds, batch, model, _, _ = run
    mV = torch.tensor(MV, dtype=torch.float64, requires_grad=True)
    eps2 = torch.tensor(3e-8, dtype=torch.float64, requires_grad=True)
    out = ds.dark_yield(batch, mV, eps2)
    gm, ge = torch.autograd.grad(out.sum(), (mV, eps2))
    Y, dY = model.evaluate(batch, MV)
    np.testing.assert_allclose(out.detach().numpy(), 3e-8 * Y, rtol=1e-14)
    np.testing.assert_allclose(gm.numpy(), 3e-8 * dY, rtol=1e-14)
    assert abs(float(ge) - Y.sum()) <= 1e-12 * Y.sum()
