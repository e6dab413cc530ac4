"""Host logic of the launch-shape tuner (pybrown_amd.autotune): grid to
per-order table, the choice rule with its noise margin, and the saved
format.  No GPU."""
import pytest

from pybrown_amd import autotune


def test_table_from_takes_the_nearest_grid_point_in_log_order():
    res = {"grid": [2, 8, 64], "codes": [4, 61, 62]}
    t = autotune.table_from(res, t_max=200)
    assert len(t) == 201 and t[0] == 0
    assert t[1] == 4 and t[2] == 4 and t[3] == 4
    assert t[4] == 4              # log 4 is as near to log 2 as to log 8: the smaller point
    assert t[5] == 61 and t[8] == 61 and t[22] == 61
    assert t[23] == 62 and t[64] == 62 and t[200] == 62


def test_choose_keeps_the_default_within_the_margin():
    assert autotune.default_code(64) == 61 and autotune.default_code(65) == 62
    assert autotune.choose({4: 0.99, 61: 1.0, 62: 1.2}, 5, margin=0.02) == 61
    assert autotune.choose({4: 0.95, 61: 1.0, 62: 1.2}, 5, margin=0.02) == 4
    assert autotune.choose({4: 3.0, 61: 2.5, 62: 2.0}, 100, margin=0.02) == 62
    assert autotune.choose({4: 3.0, 61: 1.5, 62: 2.0}, 100, margin=0.02) == 61


def test_saved_result_round_trips_and_is_validated(tmp_path):
    res = {"grid": [2, 16], "codes": [61, 4], "ms": {"2": {"4": 1.0}}, "reps": 1}
    p = tmp_path / "t.json"
    autotune.save(res, str(p))
    assert autotune.load(str(p))["codes"] == [61, 4]
    (tmp_path / "bad.json").write_text('{"grid": [2], "codes": [7]}')
    with pytest.raises(ValueError, match="unknown shape code"):
        autotune.load(str(tmp_path / "bad.json"))
    with pytest.raises(ValueError, match="without a grid"):
        autotune.table_from({"grid": [], "codes": []})
