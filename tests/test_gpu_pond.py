"""GPU parity of the pond passes (pond.hip: pond_ruts, seat_rut_water and
shove_rut_slurry, src/fluid_core.rs:5509-5792) against the FP32 restatement
(oracle/core.c), value for value: the passes are f32 arithmetic in the
reference's order, their totals summed in cell order, so water, slurry,
grain, the moved volume and the pond surface must be equal.

Covered: the reference's own seat case (src/fluid_core.rs:11505-11537); a
random pond with scattered cuts; cuts and banks standing above a pond of
uneven depth, with slurry scraped off cuts whose rings overlap (the shares
land in the order the reference walks its list); and a vehicle cutting ruts
through a puddle, each step followed by pond_ruts and seat_rut_sheet as
ThroughDrive runs them (src/world.rs:4686-4694), every pass compared with the
oracle's on the same planes and the FP64 ledger closed per step."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from silt_amd import drive as D

pytestmark = pytest.mark.gpu

N = 200 * 200
DT = 1.0 / 30.0


def _eq(a, b, what):
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, f"{what}: {bad.size} cells differ, first {bad[:5]} {a[bad[:5]]} vs {b[bad[:5]]}"


def _pair(bed, water, slurry=None, grain="clay"):
    g, o = S.FluidCore((0.0, 0.0), path="exact"), O.Core((0.0, 0.0))
    for c in (g, o):
        c.map_water_bed(np.asarray(bed, np.float32).ravel(), np.asarray(water, np.float32).ravel())
        if slurry is not None:
            c.lay_slurry(np.asarray(slurry, np.float32).ravel(), grain, 0.6)
    return g, o


def _same(g, o, what):
    gp, op = g.planes(), o.planes()
    for k in ("water", "slurry", "uw", "vw", "us", "vs", "bed"):
        _eq(gp[k], op[k], f"{k} {what}")
    _eq(g.grain(), o.grain(), f"grain {what}")


def test_reference_seat_case_matches_the_oracle():
    g, o = S.FluidCore((0.0, 0.0), path="exact"), O.Core((0.0, 0.0))
    x, z = o.cell_centres()
    rut = ((np.abs(x) < 0.3) & (np.abs(z) < 1.0)).ravel()
    for c in (g, o):
        c.pour_water(-3.0, 3.0, -3.0, 3.0, 0.05)
        c.lay_slurry(np.where(rut, 0.08, 0.0).astype(np.float32), "clay", 0.6)
    pressed = np.where(rut, -0.05, 0.0).astype(np.float32)
    eg = g.seat_rut_water(pressed, rut)
    eo = o.seat_rut_water(pressed, rut)
    assert eg == eo and abs(eg - 0.05) < 1e-4
    _same(g, o, "after the seat")
    s = g.planes()["slurry"]
    assert np.all(np.abs(s[rut] - 0.05) < 1e-4) and s[~rut].max() > 0.01


def test_pond_on_a_random_layout_matches_the_oracle():
    rng = np.random.default_rng(5)
    bed = rng.uniform(-0.007, 0.029, (200, 200)).astype(np.float32)
    water = np.where(rng.random((200, 200)) < 0.6, rng.uniform(0.019, 0.08, (200, 200)), 0.0).astype(np.float32)
    cuts = rng.random((200, 200)) < 0.01
    bed[cuts] = rng.uniform(-0.3, -0.02, cuts.sum()).astype(np.float32)
    water[cuts] = rng.uniform(0.0, 0.05, cuts.sum()).astype(np.float32)
    g, o = _pair(bed, water)
    mg, mo = g.pond_ruts(), o.pond_ruts()
    assert mg == np.float32(mo) and mg > 0.0
    _same(g, o, "after the pond")
    # nothing to do: no cut left deeper than the pond's surface needs
    g2, o2 = _pair(np.zeros(N, np.float32), water)
    assert g2.pond_ruts() == 0.0 == o2.pond_ruts()
    _same(g2, o2, "uncut window")


def _seat_layout(seed):
    rng = np.random.default_rng(seed)
    water = np.zeros((200, 200), np.float32)
    water[50:150, 50:150] = rng.uniform(0.01, 0.09, (100, 100)).astype(np.float32)
    slurry = np.zeros((200, 200), np.float32)
    pressed = np.zeros((200, 200), np.float32)
    rut = np.zeros((200, 200), bool)
    for j in (80, 84, 120):  # three tracks; the two close ones share rings
        rut[j:j + 2, 60:140] = True
    rut[100:104, 100:104] = True
    pressed[rut] = -rng.uniform(0.02, 0.08, rut.sum()).astype(np.float32)
    water[rut] = rng.uniform(0.0, 0.3, rut.sum()).astype(np.float32)
    slurry[rut] = rng.uniform(0.0, 0.2, rut.sum()).astype(np.float32)
    slurry[50:150, 50:150] += np.where(rng.random((100, 100)) < 0.3, 0.01, 0.0).astype(np.float32)
    bank = np.zeros((200, 200), bool)
    bank[48:50, 50:150] = True
    pressed[bank] = 0.045
    water[bank] = 0.2
    slurry[bank] = 0.1
    return water, slurry, pressed.ravel(), rut.ravel()


@pytest.mark.parametrize("seed", [7, 8])
def test_seat_caps_shoves_and_pours_like_the_oracle(seed):
    water, slurry, pressed, rut = _seat_layout(seed)
    g, o = _pair(np.zeros(N, np.float32), water, slurry)
    eg, eo = g.seat_rut_water(pressed, rut), o.seat_rut_water(pressed, rut)
    assert eg is not None and eg == np.float32(eo)
    _same(g, o, f"seed {seed}")
    # too few floor cells: None on both, nothing changes
    w = np.zeros(N, np.float32)
    w[:7] = 0.05
    g2, o2 = _pair(np.zeros(N, np.float32), w)
    assert g2.seat_rut_water(np.zeros(N, np.float32), np.zeros(N, bool)) is None
    assert o2.seat_rut_water(np.zeros(N, np.float32), np.zeros(N, bool)) is None
    _same(g2, o2, "no floor")


def test_drive_through_a_puddle_with_ponds_matches_the_oracle_pass_by_pass():
    g = S.FluidCore((0.0, 0.0), path="exact")
    bed = np.zeros(N, np.float32)
    g.set_ground(bed)
    water = np.zeros((200, 200), np.float32)
    water[60:140, 60:140] = 0.05
    slurry = np.zeros((200, 200), np.float32)
    slurry[90:110, 60:140] = 0.02
    g.map_water_bed(bed, water.ravel())
    g.lay_slurry(slurry.ravel(), "clay", 0.6)
    o = O.Core((0.0, 0.0))
    o.map_water_bed(bed, water.ravel())
    o.lay_slurry(slurry.ravel(), "clay", 0.6)
    drv = D.Drive(3, speed=3.0, half=6.0)
    first = g.ledger_lines()
    seated = 0
    for k in range(24):
        g.stamp_ruts(drv.contacts(), DT)
        for p in drv.pushes():
            g.push_tire(**p)
        g.step(DT)
        o.set_planes(g.planes())
        mg, mo = g.pond_ruts(), o.pond_ruts()
        assert mg == np.float32(mo), (k, mg, mo)
        _same(g, o, f"pond step {k}")
        gb, gpl, grut = g.ground()
        pressed, cut = gb + gpl, grut > 0.012
        ok = g.seat_rut_sheet()
        eo = o.seat_rut_water(pressed, cut)
        assert ok == (eo is not None), k
        seated += ok
        _same(g, o, f"seat step {k}")
        L = g.ledger_lines()
        for l in (0, 1):
            assert abs(L["gap_m3"][l]) <= 1e-9 * L["total_m3"][l], (k, l, L["gap_m3"][l])
        drv.advance(DT)
    assert seated > 0 and grut.max() > 0.012
    assert abs(L["handoff_m3"][0] - first["handoff_m3"][0]) > 0.0


def test_seat_rut_sheet_skips_a_dry_uncut_window_and_needs_a_ground():
    g = S.FluidCore((0.0, 0.0), path="exact")
    with pytest.raises(Exception):
        g.seat_rut_sheet()
    g.set_ground(np.zeros(N, np.float32))
    assert g.seat_rut_sheet() is False
    assert np.all(g.planes()["water"] == 0.0)
