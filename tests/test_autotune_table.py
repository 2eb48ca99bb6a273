"""The per-device tuning table of petite_amd.autotune: points round-trip
through the JSON file, keyed by device; unknown fields are rejected; the
coordinate moves touch one field (or one draw_waves entry) at a time."""
import json

import pytest

from petite_amd import autotune


POINT = {"batch_lanes": 40, "yield_steps": 32, "transport_waves": 16, "emit_waves": 24,
         "draw_waves": [12, 2, 8, 2, 2, 2, 2, 2], "min_group_lanes": 4, "sweep_group": 16, "group_requests": 4, "small_sweep": 16384, "small_yield_steps": 40, "graph": 2,
         "sweep_bound": 1 << 23, "resume_cap": 1 << 20,
         "tail_sweep": 16384, "tail_waves": 8, "tail_depth": 128, "tail_budget": 16}


def test_save_and_load_round_trip(tmp_path):
    path = str(tmp_path / "tuning.json")
    autotune.save(POINT, 200.0, 245.0, "100k lead", path=path, key="gfx950:256")
    autotune.save(dict(POINT, batch_lanes=48), 100.0, 110.0, "other", path=path, key="gfx942:304")
    assert autotune.load(path, key="gfx950:256") == POINT
    assert autotune.load(path, key="gfx942:304")["batch_lanes"] == 48
    assert autotune.load(path, key="gfx1100:48") is None
    with open(path) as f:
        t = json.load(f)
    assert t["gfx950:256"]["ms"] == 200.0 and t["gfx950:256"]["default_ms"] == 245.0


def test_missing_table_is_none(tmp_path):
    assert autotune.load(str(tmp_path / "none.json"), key="gfx950:256") is None


def test_unknown_field_rejected():
    with pytest.raises(ValueError):
        autotune.validate(dict(POINT, warp_size=32))


def test_coordinate_moves():
    q = autotune._set(POINT, "draw_waves.2", 6)
    assert q["draw_waves"][2] == 6 and POINT["draw_waves"][2] == 8
    assert autotune._get(q, "draw_waves.2") == 6
    q = autotune._set(POINT, "batch_lanes", 52)
    assert q["batch_lanes"] == 52 and autotune._get(q, "batch_lanes") == 52
    for name in autotune.SPACE:
        autotune._get(POINT, name)
