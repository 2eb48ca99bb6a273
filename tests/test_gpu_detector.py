"""GPU detector acceptance (petite_detector_cut) against the oracle
(transverse_position / detector_cut, src/PETITE/shower.py:815-865), and the
shower-record checkpoint feeding the dark reweighting.

The kernel works in float64 without contraction on the float32 records, as
the oracle does on the same values: pass masks agree bit for bit, weight sums
to 1e-12 relative (summation order differs)."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import detector as od
from petite_amd import DarkShower, Particle, Shower, detector_cut
from petite_amd.batch import ParticleBatch
from petite_amd.constants import mass_dict
from petite_amd.detector import cut_masks

Z = [0.5, 2.0, 10.0]
R, R_IN, ECUT = 0.4, 0.01, (0.05, 5.0)


def random_records(n, seed=1):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 4), dtype=np.float32)
    r = np.zeros((n, 4), dtype=np.float32)
    p[:, 0] = np.exp(rng.uniform(np.log(0.01), np.log(20.0), n))
    th = np.abs(rng.normal(0.0, 0.1, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    p[:, 1] = p[:, 0] * np.sin(th) * np.cos(ph)
    p[:, 2] = p[:, 0] * np.sin(th) * np.sin(ph)
    p[:, 3] = p[:, 0] * np.cos(th)
    r[:, :3] = rng.normal(0.0, 0.05, (n, 3))
    r[:, 3] = rng.uniform(0.0, 2.0, n)
    # edge cases: pz = 0, backwards, on axis, exactly at the outer radius of plane 1
    p[0, 3] = 0.0
    p[1, 3] = -p[1, 3]
    p[2, 1:3] = 0.0
    r[2, :2] = 0.0
    p[3] = [1.0, 0.2, 0.0, 1.0]
    r[3] = [0.0, 0.0, 0.0, 1.0]
    return p, r


@pytest.mark.parametrize("planes", [Z, list(np.linspace(0.25, 12.0, 12))])  # register sums (<= 8) and LDS sums
def test_kernel_matches_oracle_bitwise(planes):
    p, r = random_records(200_000)
    dev = torch.device("cuda")
    mask, wsum = cut_masks(torch.from_numpy(p).to(dev), torch.from_numpy(r).to(dev), planes, R, ECUT, R_IN)
    win, ok = od.pass_matrix(p.astype(np.float64), r.astype(np.float64), planes, R, R_IN, ECUT)
    m = mask.cpu().numpy().view(np.uint64)
    for d in range(len(planes)):
        got = ((m >> np.uint64(d)) & np.uint64(1)).astype(bool)
        assert np.array_equal(got, ok[d]), d
    assert ok.sum() > 1000 and (~ok).sum() > 1000
    w = r[:, 3].astype(np.float64)
    want = [np.sum(w[ok[d]]) for d in range(len(planes))] + [np.sum(w[win])]
    np.testing.assert_allclose(wsum.cpu().numpy(), want, rtol=1e-12)


def test_empty_input_and_limits():
    dev = torch.device("cuda")
    e = torch.zeros((0, 4), device=dev)
    mask, wsum = cut_masks(e, e, [1.0], 1.0)
    assert mask.numel() == 0 and wsum.cpu().tolist() == [0.0, 0.0]
    with pytest.raises(ValueError):
        cut_masks(e, e, list(range(65)), 1.0)


def _particles(p, r):
    return [Particle(p[i].astype(np.float64), r[i, :3].astype(np.float64), {"PID": 22, "weight": float(r[i, 3])})
            for i in range(len(p))]


@pytest.mark.filterwarnings("ignore:invalid value encountered in sqrt")  # Particle's mass of off-shell test records
def test_particle_list_methods_match_oracle():
    p, r = random_records(3000, seed=4)
    parts = _particles(p, r)
    p64, r64 = p.astype(np.float64), r.astype(np.float64)
    samp = detector_cut(parts, Z, R, "Sample", ECUT, R_IN)
    want = od.detector_cut(p64, r64, Z, R, "Sample", ECUT, R_IN)
    E = p64[:, 0]
    kept = np.nonzero((E < ECUT[1]) & (E > ECUT[0]))[0]
    for d in range(len(Z)):
        assert [id(q) for q in samp[d]] == [id(parts[kept[i]]) for i in want[d]]
    assert np.array_equal(detector_cut(parts, Z, R, "SampleW", ECUT, R_IN),
                          od.detector_cut(p64, r64, Z, R, "SampleW", ECUT, R_IN))
    for meth in ("Efficiency", "TotalWeight"):
        np.testing.assert_allclose(detector_cut(parts, Z, R, meth, ECUT, R_IN),
                                   od.detector_cut(p64, r64, Z, R, meth, ECUT, R_IN), rtol=1e-12)
    assert detector_cut(parts, Z, R, "Sample", (100.0, 200.0)) == [[], [], []]
    assert detector_cut(parts, Z, R, "Unknown") is None


def _prim(E, pid):
    m = mass_dict[pid]
    return Particle([E, 0.0, 0.0, math.sqrt(E * E - m * m)], [0, 0, 0], {"PID": pid})


def test_batches_and_record_checkpoint(tmp_path):
    ds = DarkShower("100GeV", "graphite", 0.010, 0.03, seed=9)
    prims = [_prim(E, pid) for E, pid in ((3.0, 11), (5.0, 22), (2.0, -11), (4.0, 11))]
    batch = ds.generate_showers(prims)
    a = batch.to_numpy()
    # SM record on the device
    got = detector_cut(batch, Z, 5.0, "TotalWeight")
    want = od.detector_cut(a["p0"].astype(np.float64), a["r0"].astype(np.float64), Z, 5.0, "TotalWeight")
    np.testing.assert_allclose(got, want, rtol=1e-12)
    # dark vectors: the valid entries in the reference's order
    dark = ds.generate_dark_showers(batch)
    idx = dark.order()
    pv = dark.p.cpu().numpy()[idx].astype(np.float64)
    rv = dark.tensor("r").cpu().numpy()[idx].astype(np.float64)
    got = detector_cut(dark, [1.0, 3.0], 0.2, "Sample")
    want = od.detector_cut(pv, rv, [1.0, 3.0], 0.2, "Sample")
    assert [g.cpu().tolist() for g in got] == [w.tolist() for w in want]
    # checkpoint: the reloaded record reweights to the same dark vectors
    path = str(tmp_path / "record.npz")
    batch.save(path)
    back = ParticleBatch.load(path)
    dark2 = ds.generate_dark_showers(back)
    assert dark2.n == dark.n
    idx2 = dark2.order()  # entries are appended in completion order: compare in the reference's order
    assert np.array_equal(dark2.p.cpu().numpy()[idx2], dark.p.cpu().numpy()[idx])
    assert np.array_equal(dark2.tensor("r").cpu().numpy()[idx2], dark.tensor("r").cpu().numpy()[idx])
    sm, vecs = ds.generate_dark_shower(ExDir=path)
    assert len(sm) == batch.n and len(vecs) == len(idx)
    with pytest.raises(NotImplementedError):
        ds.generate_dark_shower(ExDir=str(tmp_path / "shower.p"))
