"""The launch-size autotuner on the live device: a small search returns a
valid point no slower than the defaults by more than timing noise, the point
survives the per-device table, and runs under it give the same records as
under the defaults (matched by stream id, bit for bit)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from petite_amd import Shower, autotune
from petite_amd.workloads import batch_primaries


def _by_sid(b):
    a = b.to_numpy()
    o = np.argsort(a["sid"].astype(np.uint64), kind="stable")
    return {k: a[k][o] for k in ("sid", "pid", "p0", "pf", "rf", "meta")}


def test_autotune_small_search(tmp_path):
    sh = Shower("100GeV", "lead", 0.010, seed=1234)
    prim = batch_primaries(3000, seed=5)
    space = {"group_requests": [1, 4], "batch_lanes": [40, 48]}
    point, ms, dms, hist = autotune.autotune(sh, prim, seed=7, space=space, repeats=2, rounds=1, log=lambda m: None)
    assert len(hist) >= 3 and ms <= dms * 1.0001
    assert point["group_requests"] in (1, 4) and point["batch_lanes"] in (40, 44, 48)
    path = str(tmp_path / "t.json")
    autotune.save(point, ms, dms, "test", path=path, key="dev")
    assert autotune.load(path, key="dev") == autotune.validate(point)
    ctx = sh.device_context()
    ctx.set_tuning(None)
    a = _by_sid(sh.generate_showers(prim, seed=7))
    ctx.set_tuning(point)
    b = _by_sid(sh.generate_showers(prim, seed=7))
    ctx.set_tuning(None)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
