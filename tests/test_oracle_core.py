"""Pins of the FluidCore restatement (oracle/core.c) against the reference's
own FluidCore and near-field tests, with their thresholds."""
import math

import numpy as np
import pytest

import pyoracle as O


def tire(x, z=0.0, radius=0.38, width=0.2, speed=3.0, slip=0.3):
    return dict(x=x, z=z, forward_x=0.0, forward_z=1.0, radius=radius, width=width, speed=speed, slip=slip)


def test_floor_scale_keeps_cells_non_negative():
    """floor_scale_keeps_cells_non_negative (src/fluid_core/near.rs:354-364)."""
    r = np.zeros(100, np.float32)
    r[0], r[1] = -0.04, 0.04
    assert O.floor_scale(0.05, r) == 1.0
    a = O.floor_scale(0.01, r)
    assert abs(0.01 + a * r[0]) < 1e-7
    assert O.floor_scale(0.0, r) == 0.0


def test_recentre_sums_to_zero():
    """recentre_sums_to_zero (src/fluid_core/near.rs:366-375)."""
    r = np.array([math.sin(n * 0.37) * 0.01 + 0.003 for n in range(100)], np.float32)
    out = O.recentre(r)
    assert abs(out.astype(np.float64).sum()) < 1e-6


def test_fine_patch_leaves_no_hole_in_deep_water():
    """fine_patch_leaves_no_hole_in_deep_water (src/fluid_core.rs:11615-11660):
    a still tire in 40 cm of water for 30 steps at 1/30 s keeps the volume
    within 1e-3 and leaves no hole deeper than 3 cm around the patch."""
    core = O.Core((0.0, 0.0))
    core.pour_water(-6.0, 6.0, -6.0, 6.0, 0.40)
    before = core.ledger()["total_m3"]
    for _ in range(30):
        core.push_tire(**tire(0.0, radius=0.01, width=0.01, speed=0.0, slip=0.0))
        core.step(1.0 / 30.0)
    after = core.ledger()["total_m3"]
    assert abs(after - before) < 1e-3 * before
    lowest = min(core.sample_water(-2.0 + 0.1 * a, -2.0 + 0.1 * b)
                 for a in range(41) for b in range(41)
                 if math.hypot(-2.0 + 0.1 * a, -2.0 + 0.1 * b) >= 0.5)
    assert lowest > 0.37, lowest


def test_detail_survives_refine_and_keeps_volume():
    """detail_survives_refine_and_keeps_volume (src/fluid_core/near.rs:384-410):
    after six steps behind a tire the kept residual re-lays a patch with more
    than 1 mm of spread, and refine / restrict move under 1e-6 m^3."""
    dt = 1.0 / 30.0
    core = O.Core((0.0, 0.0))
    core.pour_water(-2.0, 2.0, -2.0, 2.0, 0.08)
    for _ in range(6):
        core.push_tire(**tire(0.7))
        core.step(dt)
    assert core.near_tiles() > 0
    before = core.ledger()["water_m3"]
    core.push_tire(**tire(0.7))
    core.refine(dt)
    fw, _ = core.fine_water()
    F = fw.reshape(160, 160)
    spread = 0.0
    for j0 in range(0, 160, 10):
        for i0 in range(0, 160, 10):
            blk = F[j0:j0 + 10, i0:i0 + 10]
            if blk.max() > 0:
                spread = max(spread, float(blk.max() - blk.min()))
    assert spread > 1e-3, spread
    laid = core.ledger()["water_m3"]
    core.restrict()
    after = core.ledger()["water_m3"]
    assert abs(laid - before) < 1e-6 and abs(after - before) < 1e-6
    assert (fw >= 0).all()


def test_unheld_tiles_fade_out():
    """unheld_tiles_fade_out (src/fluid_core/near.rs:413-427)."""
    dt = 1.0 / 30.0
    core = O.Core((0.0, 0.0))
    core.pour_water(-2.0, 2.0, -2.0, 2.0, 0.08)
    for _ in range(6):
        core.push_tire(**tire(0.7))
        core.step(dt)
    assert core.near_tiles() > 0
    for _ in range(int(12.0 * 1.5 / dt)):
        core.step(dt)
    assert core.near_tiles() == 0


def test_spray_is_booked_and_the_ledger_closes():
    """Spray peeled by a fast tire leaves the water sheet and enters spray_m3
    (peel_spray :8241-8314, step :6101-6103): the ledger total stays within
    the reference's 1e-3 relative band while the shoved water moves."""
    core = O.Core((0.0, 0.0))
    core.pour_water(-5.0, 5.0, -5.0, 5.0, 0.06)
    t0 = core.ledger()["total_m3"]
    for k in range(20):
        for x in (-0.7, 0.7):
            for z in (-1.2, 1.2):
                core.push_tire(**tire(x, z + 0.1 * k, speed=3.0))
        core.step(1.0 / 30.0)
    led = core.ledger()
    assert led["spray_m3"] > 0.0
    assert abs(led["total_m3"] - t0) < 1e-3 * t0
    p = core.planes()
    assert np.abs(p["uw"]).max() > 0.1  # the tires shoved the water


def test_publish_film_and_flow_layout():
    """coarse_cell / publish (src/fluid_core.rs:6998-7034,
    src/render_contract.rs:468-496): film is (water, slurry), flow is the
    thickness-weighted (u, v), its speed, and 0 turbidity with no solids; dry
    cells publish zero flow and nothing is NaN."""
    core = O.Core((0.0, 0.0))
    core.pour_water(-3.0, 3.0, -3.0, 3.0, 0.05)
    core.pour_slurry(-1.0, 1.0, -1.0, 1.0, 0.02, "clay", 0.6)
    for _ in range(5):
        core.push_tire(**tire(0.0, speed=2.0))
        core.step(1.0 / 30.0)
    film, flow = core.publish()
    p = core.planes()
    assert np.array_equal(film[0::2], p["water"]) and np.array_equal(film[1::2], p["slurry"])
    h = p["water"] + p["slurry"]
    wet = h >= 1e-5
    u = np.where(wet, (p["uw"] * p["water"] + p["us"] * p["slurry"]) / np.where(wet, h, 1), 0)
    assert np.allclose(flow[0::4], u, rtol=1e-6, atol=1e-9)
    assert np.allclose(flow[2::4], np.hypot(flow[0::4], flow[1::4]), rtol=1e-6, atol=1e-9)
    assert not flow[0::4][~wet].any() and np.isfinite(flow).all()


def test_publish_surface_edges_and_crest():
    """wet_edge / foam_crest / moisture (src/render_contract.rs:457-500,
    :526-569) on a square pool: the edge ring is the pool's inner and outer
    perimeter, the crest excess of a corner cell is its depth minus the mean
    of its two wet and two dry neighbours, placed soil adds to the crest,
    and the soil's wetness can make dry cells wet (no edge inside a wet
    patch)."""
    core = O.Core((0.0, 0.0))
    core.pour_water(-3.0, 3.0, -3.0, 3.0, 0.05)  # 30 x 30 cells
    p = core.planes()
    wet = p["water"].reshape(200, 200) > 0
    js, is_ = np.nonzero(wet)
    side = js.max() - js.min() + 1
    assert side == 30 and wet.sum() == 900
    moist, edge, foam = core.publish_surface()
    assert edge.sum() == (4 * side - 4) + 4 * side  # inner ring + outer ring
    crest = foam[1::2].reshape(200, 200)
    corner = (js.min(), is_.min())
    assert abs(crest[corner] - 0.05 * 0.5) < 1e-7  # two of four neighbours wet
    assert crest[js.min() + 5, is_.min() + 5] == 0.0  # flat inside
    assert not foam[0::2].any() and not moist.any()
    placed = np.zeros(200 * 200, np.float32)
    placed[0] = 0.04
    _, _, foam2 = core.publish_surface(placed=placed)
    assert foam2[1] == np.float32(0.04)
    wetness = np.zeros(200 * 200, np.float32)
    wetness.reshape(200, 200)[js.min() - 5:js.max() + 6, is_.min() - 5:is_.max() + 6] = 0.9
    m2, edge2, _ = core.publish_surface(wetness=wetness)
    assert m2.max() == np.float32(0.9)
    assert edge2.sum() == (4 * (side + 10) - 4) + 4 * (side + 10)  # the edge moved out to the wet soil
