"""The full-size shower workload (100,000 primaries e-/e+/gamma, E
log-uniform in [0.1, 100] GeV within 1 mrad, lead, cutoff 10 MeV: about 2.2e8
records) checked through properties that do not depend on the size, on the
device: finite and on-shell records; parent links, generations, shower ids
and child species consistent with each vertex's process (shower.py:87-96,
486-507); kept children above the cutoff (shower.py:703-705); energy
conservation at every vertex whose two children were kept (the atomic
electron's mass added for Compton, Moller, Bhabha, annihilation and mu e);
every shower's float64 energy budget closed; unique stream ids; and the
record's step, particle and dropped-child counts equal to the run
statistics."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import record_checks
from petite_amd import Shower
from petite_amd.batch import PRIMARY_DTYPE
from petite_amd.constants import mass_dict

N, CUT, ME = 100_000, 0.010, mass_dict[11]


def _primaries():
    rng = np.random.default_rng(1234)
    prim = np.zeros(N, dtype=PRIMARY_DTYPE)
    pid = rng.choice(np.array([11, -11, 22], dtype=np.int32), N)
    E = np.exp(rng.uniform(np.log(0.1), np.log(100.0), N))
    th, ph = rng.uniform(0, 1e-3, N), rng.uniform(0, 2 * np.pi, N)
    m = np.where(pid == 22, 0.0, ME)
    p = np.sqrt(E * E - m * m)
    prim["E"], prim["px"], prim["py"], prim["pz"] = E, p * np.sin(th) * np.cos(ph), p * np.sin(th) * np.sin(ph), \
        p * np.cos(th)
    prim["mass"], prim["weight"], prim["pid"] = m, 1.0, pid
    prim["shower"] = np.arange(N, dtype=np.int32)
    prim["sid"] = np.arange(N, dtype=np.uint64)
    return prim


@pytest.fixture(scope="module")
def run():
    prim = _primaries()
    sh = Shower("100GeV", "lead", CUT, seed=1234)
    b = sh.generate_showers(prim, seed=1234)
    return b, prim


def test_records_finite_and_on_shell(run):
    b, _ = run
    assert b.n > 1e8
    record_checks.finite_and_on_shell(b)


def test_links_generations_and_species(run):
    b, _ = run
    record_checks.links_generations_and_species(b, N, CUT)


def test_vertex_energy_conservation(run):
    b, _ = run
    record_checks.vertex_energy_conservation(b, 10_000_000)


def test_energy_budget_and_counts(run):
    b, prim = run
    record_checks.budget_and_counts(b, prim)
