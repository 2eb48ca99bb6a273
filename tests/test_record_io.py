"""Shower-record checkpoint (ParticleBatch.save / load): a plain-array .npz,
read back without unpickling; it replaces the reference's ExDir files for
DarkShower.generate_dark_shower (dark_shower.py:822-823)."""
import json

import numpy as np
import pytest
import torch

from petite_amd.batch import PRIMARY_DTYPE, STORE_KEYS, ParticleBatch


def make_batch(n=7, cap=11):
    g = torch.Generator().manual_seed(3)
    store = {k: torch.randn((cap, 4), generator=g) for k in ("p0", "r0", "pf", "rf")}
    for k in ("pid", "parent", "meta", "shower"):
        store[k] = torch.randint(-20, 20, (cap,), generator=g, dtype=torch.int32)
    store["sid"] = torch.randint(-2 ** 62, 2 ** 62, (cap,), generator=g, dtype=torch.int64)
    prim = np.zeros(2, dtype=PRIMARY_DTYPE)
    prim["E"] = [1.0, 2.0]
    prim["sid"] = [2 ** 63 + 5, 9]
    energy = torch.arange(8, dtype=torch.float64).reshape(2, 4)
    return ParticleBatch(store, n, {"n_steps": 123, "ms_transport": 1.5}, energy, prim)


def test_round_trip_keeps_every_array(tmp_path):
    b = make_batch()
    path = tmp_path / "rec.npz"
    b.save(path)
    c = ParticleBatch.load(path, device="cpu")
    assert c.n == b.n
    for k in STORE_KEYS:
        assert torch.equal(c.tensor(k), b.tensor(k)), k
        assert c._store[k].shape[0] == b.n  # only the record's rows are written
    assert torch.equal(c.energy, b.energy)
    assert c.stats == b.stats
    assert np.array_equal(c.primaries, b.primaries)
    with np.load(path, allow_pickle=False) as z:  # plain arrays only
        assert json.loads(str(z["stats"]))["n_steps"] == 123


def test_load_refuses_other_files(tmp_path):
    path = tmp_path / "other.npz"
    np.savez(path, a=np.zeros(3))
    with pytest.raises(ValueError, match="not a petite_amd shower record"):
        ParticleBatch.load(path, device="cpu")


def test_from_arrays_checks_shapes_before_touching_the_device():
    with pytest.raises(ValueError):
        ParticleBatch.from_arrays(np.zeros((3, 3)), 111)
    with pytest.raises(ValueError):
        ParticleBatch.from_arrays(np.zeros((3, 4)), 111, r=np.zeros((2, 3)))
