"""The oracle's drain_over_bed (src/fluid_core.rs:5123-5214): the pour moves,
and the volume the dry cut wipes is put back, so the water volume is kept to
the rounding of the reference's f32 volume sums (a 40 000-cell f32 sum is
good to about 1e-5 relative, and each step's restore adds that error back)."""
import numpy as np

import pyoracle as O
from silt_amd import scenes


def test_drain_keeps_volume_and_settles():
    sc = scenes.window_scene(seed=3)
    o = O.Core((0.0, 0.0))
    o.map_water_bed(sc["bed"], sc["water"])
    o.pour_water(-3.0, 3.0, 4.0, 8.0, 0.05)
    before = float(np.maximum(o.planes()["water"], 0).astype(np.float64).sum()) * 0.04
    w0 = o.planes()["water"]
    v = o.drain_over_bed(1.0 / 60.0, 200)
    p = o.planes()
    after = float(np.maximum(p["water"], 0).astype(np.float64).sum()) * 0.04
    assert abs(after - before) <= 1e-3 * before
    assert abs(v - after) <= 1e-4 * after
    assert np.abs(p["water"] - w0).max() > 0.01  # the pour spread


def test_drain_with_no_steps_only_measures():
    sc = scenes.window_scene(seed=1)
    o = O.Core((0.0, 0.0))
    o.map_water_bed(sc["bed"], sc["water"])
    p0 = o.planes()
    v = o.drain_over_bed(1.0 / 60.0, 0)
    assert v == np.float32(O.column_sum(p0["water"], 0.2))
    for k in ("water", "uw"):
        assert np.array_equal(o.planes()[k], p0[k])
