"""Pins of the suspended-solids restatement (oracle/core.c: entrain_wake and
advect_suspended, src/fluid_core.rs:8750-9003) by closed forms and known
answers:

* fall-out of a still column: s' = s exp(-dt w_s / h) with w_s = 4 mm/s, in
  FP64 against the closed form, in f32 against the same expression;
* a dry cell and an over-full cell settle their whole load;
* a column with v of a quarter cell per step hands a quarter of its load to
  the next row (down for v > 0, up for v < 0), a receiver already at the
  donor's concentration takes nothing, the give is capped at 0.45 of the
  load, and a donor on the last row loses its give off the window;
* the lift behind a tire against the footprint formula evaluated directly
  (cells ahead of the tire, beside it, or under a slow tire get none);
* the account closes over a drive: column + settled + parked = entrained."""
import math

import numpy as np
import pytest

import pyoracle as O

N = 200 * 200
DT = 1.0 / 30.0
CELL = 0.2
AREA = float(np.float32(0.2)) ** 2
# The FP64 oracle widens the reference's f32 constants: these are its values.
K_CELL, K_FALL, K_W = float(np.float32(0.2)), float(np.float32(0.004)), float(np.float32(0.1))
KEEP = math.exp(-DT * K_FALL / K_W)


def _still(water=0.1, v=0.0, precision=32):
    c = O.Core((0.0, 0.0), precision=precision)
    c.map_water_bed(np.zeros(N, np.float32), np.full(N, water, np.float32))
    if v:
        p = c.planes()
        p["vw"][:] = v
        c.set_planes(p)
    return c


def test_still_column_falls_out_at_the_closed_form():
    c64 = _still(0.1, precision=64)
    s = np.zeros(N)
    s[100 * 200 + 100] = 0.01
    c64.set_suspended(s)
    c64.suspended_step(DT)
    got, acc = c64.suspended()
    assert got[100 * 200 + 100] == pytest.approx(0.01 * KEEP, rel=1e-14)
    assert acc["settled_m3"] == pytest.approx(0.01 * (1.0 - KEEP) * AREA, rel=1e-10)
    c = _still(0.1)
    s32 = np.zeros(N, np.float32)
    s32[100 * 200 + 100] = 0.01
    c.set_suspended(s32)
    c.suspended_step(DT)
    got, _ = c.suspended()
    w32 = np.float32(0.1)
    k32 = np.float32(math.exp(np.float32(-np.float32(DT) * np.float32(0.004)) / w32))
    assert got[100 * 200 + 100] == pytest.approx(float(np.float32(0.01) - np.float32(0.01) * (np.float32(1) - k32)),
                                                 rel=2e-7)


def test_dry_and_overfull_cells_settle_everything():
    c = _still(0.1)
    p = c.planes()
    p["water"][5] = 0.0
    p["water"][7] = 0.001
    c.set_planes(p)
    s = np.zeros(N, np.float32)
    s[5] = 0.002
    s[7] = 0.003  # more solids than water
    c.set_suspended(s)
    c.suspended_step(DT)
    got, acc = c.suspended()
    assert got[5] == 0.0 and got[7] == 0.0
    assert acc["settled_m3"] == pytest.approx(0.005 * AREA, rel=1e-6)


@pytest.mark.parametrize("sign", [1, -1])
def test_quarter_cell_per_step_hands_a_quarter_downstream(sign):
    v = sign * 0.25 * CELL / DT
    c = _still(0.1, v=v, precision=64)
    s = np.zeros(N)
    j, i = 100, 100
    s[j * 200 + i] = 0.02
    c.set_suspended(s)
    c.suspended_step(DT)
    got, acc = c.suspended()
    fv = abs(v) * DT / K_CELL
    left = 0.02 * KEEP
    assert got[j * 200 + i] == pytest.approx(left * (1.0 - fv), rel=1e-12)
    assert got[(j + sign) * 200 + i] == pytest.approx(left * fv, rel=1e-12)
    assert got[(j - sign) * 200 + i] == 0.0 and got[j * 200 + i + 1] == 0.0
    assert got.sum() == pytest.approx(left, rel=1e-12)


def test_receiver_at_the_donors_concentration_takes_nothing_and_the_give_is_capped():
    v = 0.9 * CELL / DT  # 0.9 of a cell: capped at 0.45
    c = _still(0.1, v=v, precision=64)
    s = np.zeros(N)
    s[100 * 200 + 100] = 0.02
    s[101 * 200 + 100] = 0.02 * KEEP * 1.5  # denser than the donor
    c.set_suspended(s)
    before = s.copy()
    c.suspended_step(DT)
    got, _ = c.suspended()
    # the donor's give finds no room below it, so it keeps its load
    assert got[100 * 200 + 100] == pytest.approx(before[100 * 200 + 100] * KEEP, rel=1e-12)
    c2 = _still(0.1, v=v, precision=64)
    s2 = np.zeros(N)
    s2[100 * 200 + 100] = 0.02
    c2.set_suspended(s2)
    c2.suspended_step(DT)
    g2, _ = c2.suspended()
    assert g2[101 * 200 + 100] == pytest.approx(0.02 * KEEP * 0.45, rel=1e-7)


def test_last_row_gives_off_the_window():
    v = 0.25 * CELL / DT
    c = _still(0.1, v=v, precision=64)
    s = np.zeros(N)
    s[199 * 200 + 50] = 0.01
    c.set_suspended(s)
    c.suspended_step(DT)
    got, acc = c.suspended()
    fv = v * DT / K_CELL
    assert got[199 * 200 + 50] == pytest.approx(0.01 * KEEP * (1 - fv), rel=1e-12)
    # what fell out plus the give that left the window
    assert acc["settled_m3"] == pytest.approx((0.01 * (1 - KEEP) + 0.01 * KEEP * fv) * AREA, rel=1e-10)


def _footprint(px, pz, fx, fz, radius, width, speed, water, dt, susp=0.0):
    """entrain_wake's lift per cell, evaluated directly in float64."""
    fl = max(math.hypot(fx, fz), 1e-4)
    fx, fz = fx / fl, fz / fl
    rx, rz = fz, -fx
    aft, half = radius * 3.2, width * 0.55 + 0.08
    out = {}
    for j in range(200):
        for i in range(200):
            x, z = -20.0 + (i + 0.5) * CELL, -20.0 + (j + 0.5) * CELL
            dx, dz = x - px, z - pz
            along = dx * fx + dz * fz
            if along > -0.05 or along < -aft:
                continue
            lat = abs(dx * rx + dz * rz)
            if lat > half:
                continue
            core = (1 - min(max(-along / aft, 0), 1)) * (1 - min(max(lat / half, 0), 1))
            add = min(4.0 * dt * core * water, max(water - susp, 0.0))
            if add > 1e-8:
                out[j * 200 + i] = add
    return out


def test_lift_behind_a_tire_matches_the_footprint_formula():
    c = _still(0.05)
    c.push_tire(0.1, 0.1, 0.6, 0.8, 0.38, 0.28, 2.0, 0.1)
    c.push_tire(5.1, 0.1, 1.0, 0.0, 0.38, 0.28, 0.2, 0.1)  # too slow to lift
    c.entrain(DT)
    got, acc = c.suspended()
    want = _footprint(0.1, 0.1, 0.6, 0.8, 0.38, 0.28, 2.0, 0.05, DT)
    assert set(np.flatnonzero(got)) == set(want)
    # f32 cell centres move the footprint's edges by ~1e-6 m: the error is
    # absolute, a few parts in 1e5 of the largest lift
    top = max(want.values())
    for k, v in want.items():
        assert abs(got[k] - v) <= 2e-5 * top
    assert acc["entrained_m3"] == pytest.approx(sum(want.values()) * AREA, rel=2e-5)
    assert acc["column_m3"] == pytest.approx(acc["entrained_m3"], rel=1e-5)


@pytest.mark.parametrize("precision,tol", [(32, 1e-5), (64, 1e-12)])
def test_the_account_closes_over_a_drive(precision, tol):
    c = O.Core((0.0, 0.0), precision=precision)
    water = np.zeros((200, 200), np.float32)
    water[60:140, 80:120] = 0.06
    c.map_water_bed(np.zeros(N, np.float32), water.ravel())
    for k in range(20):
        z0 = -6.0 + 0.25 * k
        for x in (-0.7, 0.7):
            for dz in (-1.2, 1.2):
                c.push_tire(x, z0 + dz, 0.0, 1.0, 0.38, 0.28, 7.5, 0.2)
        c.step(DT)
    for _ in range(10):
        c.step(DT)
    _, acc = c.suspended()
    assert acc["entrained_m3"] > 1e-4
    gap = acc["column_m3"] + acc["settled_m3"] + acc["frozen_m3"] - acc["entrained_m3"]
    assert abs(gap) <= tol * acc["entrained_m3"]
    c.recenter(30.0, 0.0)
    _, acc2 = c.suspended()
    assert acc2["frozen_m3"] > 0.0
    gap2 = acc2["column_m3"] + acc2["settled_m3"] + acc2["frozen_m3"] - acc2["entrained_m3"]
    assert abs(gap2) <= tol * acc2["entrained_m3"]
