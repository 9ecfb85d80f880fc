"""The oracle's recenter (src/fluid_core.rs:5409-5422, shift_coarse
:9004-9226): moving the window away and back restores every cell, the
parked set empties again, and the ledger counts parked volume meanwhile."""
import numpy as np

import pyoracle as O
from silt_amd import scenes


def test_round_trip_restores_the_window():
    sc = scenes.window_scene(seed=1)
    o = O.Core((0.0, 0.0))
    o.map_water_bed(sc["bed"], sc["water"])
    o.lay_slurry(sc["slurry"], "clay", 0.6)
    p0, l0 = o.planes(), o.ledger()
    o.recenter(3.4, -1.0)
    assert tuple(o.origin()) == (np.float32(-16.6), np.float32(-21.0))
    assert o.parked_cells() > 0
    o.recenter(63.0, 5.0)  # the whole window leaves: every wet cell parked
    assert o.planes()["water"].max() == 0.0
    total = o.ledger()
    # (the window volume is the reference's f32 sum, good to ~1e-5 relative)
    assert abs(total["frozen_water_m3"] - l0["water_m3"]) <= 1e-4 * l0["water_m3"]
    o.recenter(0.0, 0.0)
    assert o.parked_cells() == 0
    p1 = o.planes()
    for k in ("water", "slurry", "uw", "vw", "moisture"):
        assert np.array_equal(p1[k], p0[k]), k
    assert abs(o.ledger()["total_m3"] - l0["total_m3"]) <= 1e-6 * l0["total_m3"]


def test_a_move_under_one_cell_is_no_move():
    o = O.Core((0.0, 0.0))
    o.pour_water(-1.0, 1.0, -1.0, 1.0, 0.05)
    o.recenter(0.05, 0.05)
    assert tuple(o.origin()) == (np.float32(-20.0), np.float32(-20.0))
    assert o.parked_cells() == 0
