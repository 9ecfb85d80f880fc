"""Pins of the pond restatement (oracle/core.c: pond_ruts, seat_rut_water and
shove_rut_slurry, src/fluid_core.rs:5509-5792, with prefix_plane /
plane_window :1502-1525) by the reference's own test, closed forms and brute
force:

* seat_shoves_rut_slurry_beside_the_cut (src/fluid_core.rs:11505-11537), the
  reference's known answer: pond surface 0.05, rut slurry 0.05, slurry kept,
  nothing thrown farther than a metre, a lip above 1 cm;
* a flat pond and a cut: the cut fills flush with the pond's free surface,
  every donor gives the same depth, the volume is kept; with too little room
  every donor drains to exactly 12 mm;
* the window means of the summed-area planes against a brute-force box sum
  (a transposed index or an off-by-one in plane_window fails it);
* a cut farther than 10 cells from any still cell takes nothing; a berm is
  left alone; a window with no cut moves nothing;
* seat_rut_water's median against a sort, its None below 8 floor cells, the
  rut column capped at the pond surface with the excess poured onto the
  floor in proportion to depth (volume kept), a bank column cut to the
  surface, and slurry on a cut walled in by cuts staying where it is."""
import numpy as np
import pytest

import pyoracle as O

N = 200 * 200
AREA = np.float32(0.2) * np.float32(0.2)


def _core(bed, water, slurry=None, precision=32):
    c = O.Core((0.0, 0.0), precision=precision)
    c.map_water_bed(np.asarray(bed, np.float32).ravel(), np.asarray(water, np.float32).ravel())
    if slurry is not None:
        p = c.planes()
        p["slurry"][:] = np.asarray(slurry, np.float32).ravel()
        c.set_planes(p)
    return c


def _volume(a):
    return float(np.sum(np.asarray(a, np.float64))) * float(AREA)


def test_reference_seat_shoves_rut_slurry_beside_the_cut():
    # src/fluid_core.rs:11505-11537
    c = O.Core((0.0, 0.0))
    c.pour_water(-3.0, 3.0, -3.0, 3.0, 0.05)
    x, z = c.cell_centres()
    rut = ((np.abs(x) < 0.3) & (np.abs(z) < 1.0)).ravel()
    c.lay_slurry(np.where(rut, 0.08, 0.0).astype(np.float32), "clay", 0.6)
    before = c.ledger()["slurry_m3"]
    eta = c.seat_rut_water(np.where(rut, -0.05, 0.0).astype(np.float32), rut)
    assert eta is not None and abs(eta - 0.05) < 1e-4
    after = c.ledger()["slurry_m3"]
    assert abs(after - before) <= 1e-5 * before
    s = c.planes()["slurry"]
    assert np.all(np.abs(s[rut] - 0.05) < 1e-4)
    xr, zr = x.ravel(), z.ravel()
    thrown = (~rut) & (s > 0.0)
    off = np.maximum(np.abs(xr[thrown]) - 0.3, np.abs(zr[thrown]) - 1.0)
    assert np.all(off < 1.0)
    assert s[thrown].max() > 0.01
    # a cell that had no slurry takes the cut's grain (clay = 0)
    assert np.all(c.grain()[thrown] == O.GRAIN["clay"])


def _flat_pond(depth=0.06, cut_rows=slice(98, 102), cut_cols=slice(60, 140), cut_bed=-0.05, cut_water=0.01):
    bed = np.zeros((200, 200), np.float32)
    water = np.zeros((200, 200), np.float32)
    water[80:120, 40:160] = depth
    bed[cut_rows, cut_cols] = cut_bed
    water[cut_rows, cut_cols] = cut_water
    return bed, water


def test_pond_fills_the_cut_flush_and_keeps_the_volume():
    bed, water = _flat_pond()
    c = _core(bed, water)
    v0 = _volume(c.planes()["water"])
    moved = c.pond_ruts()
    w = c.planes()["water"].reshape(200, 200)
    cut = bed < -0.012
    # the cut's free surface reaches the pond's: bed + water = 0.06
    surf = (bed + w)[cut]
    assert np.allclose(surf, 0.06, atol=2e-7)
    need = float(np.sum((np.float32(0.06) - bed[cut] - water[cut]).astype(np.float64))) * float(AREA)
    assert moved == pytest.approx(need, rel=1e-5)
    donors = (~cut) & (water > 0.0)
    d = w[donors]
    assert np.ptp(d) <= 2e-8, "every donor gives the same depth"
    assert d[0] == pytest.approx(0.06 - need / (donors.sum() * float(AREA)), rel=1e-5)
    # The reference sums the needs and the rooms in f32, so the donors give
    # the moved volume to the rounding of a 4480-term sum; in FP64 the same
    # algorithm keeps the volume to the last digits.
    assert abs(_volume(w) - v0) <= 1e-4 * moved
    c64 = _core(bed, water, precision=64)
    c64.pond_ruts()
    assert _volume(c64.planes()["water"]) == pytest.approx(_volume(water.astype(np.float64)), rel=1e-12)


def test_pond_with_little_room_drains_every_donor_to_twelve_mm():
    # shallow pond (room 3 mm per cell over 12 mm) and a wide deep cut
    bed, water = _flat_pond(depth=0.015, cut_rows=slice(90, 110), cut_bed=-0.2, cut_water=0.0)
    c = _core(bed, water)
    moved = c.pond_ruts()
    w = c.planes()["water"].reshape(200, 200)
    donors = (bed > -0.008) & (water > 0.018)
    assert not donors.any(), "15 mm is below the still threshold of 18 mm"
    assert moved == 0.0
    bed, water = _flat_pond(depth=0.020, cut_rows=slice(90, 110), cut_bed=-0.2, cut_water=0.0)
    c = _core(bed, water)
    v0 = _volume(water)
    moved = c.pond_ruts()
    w = c.planes()["water"].reshape(200, 200)
    donors = (bed > -0.008) & (water > 0.018)
    room = float(np.sum((water[donors] - np.float32(0.012)).astype(np.float64))) * float(AREA)
    # f32 sums of 4800 rooms in the reference's order: to 1e-4
    assert moved == pytest.approx(room, rel=1e-4)
    assert np.allclose(w[donors], 0.012, atol=1e-8)
    assert abs(_volume(w) - v0) <= 1e-4 * moved
    c64 = _core(bed, water, precision=64)
    moved64 = c64.pond_ruts()
    w64 = c64.planes()["water"].reshape(200, 200)
    area64 = float(np.float32(0.2)) ** 2
    room64 = float(np.sum(water[donors].astype(np.float64) - float(np.float32(0.012)))) * area64
    assert moved64 == pytest.approx(room64, rel=1e-12)
    assert np.allclose(w64[donors], float(np.float32(0.012)), rtol=0, atol=1e-15)
    assert _volume(w64) == pytest.approx(_volume(water), rel=1e-12)


def test_pond_window_means_match_a_brute_force_box_sum():
    rng = np.random.default_rng(5)
    bed = rng.uniform(-0.007, 0.029, (200, 200)).astype(np.float32)
    water = np.where(rng.random((200, 200)) < 0.6, rng.uniform(0.019, 0.08, (200, 200)), 0.0).astype(np.float32)
    cuts = rng.random((200, 200)) < 0.01
    bed[cuts] = rng.uniform(-0.3, -0.02, cuts.sum()).astype(np.float32)
    water[cuts] = 0.0
    c = _core(bed, water)
    moved = c.pond_ruts()
    w = c.planes()["water"].reshape(200, 200)
    still = (bed > -0.008) & (bed < 0.03) & (water > 0.018)
    eta = np.where(still, bed.astype(np.float64) + water, 0.0)
    need = np.zeros((200, 200))
    for j, i in zip(*np.nonzero(cuts)):
        j0, j1, i0, i1 = max(j - 10, 0), min(j + 10, 199), max(i - 10, 0), min(i + 10, 199)
        cnt = still[j0:j1 + 1, i0:i1 + 1].sum()
        if cnt < 1:
            continue
        extra = max(eta[j0:j1 + 1, i0:i1 + 1].sum() / cnt - bed[j, i] - water[j, i], 0.0)
        if extra > 1e-4:
            need[j, i] = extra
    total_need = need.sum() * float(AREA)
    room = np.where(still, np.maximum(water - 0.012, 0.0), 0.0)
    room = np.where(room > 1e-4, room, 0.0)
    scale = min(room.sum() * float(AREA) / total_need, 1.0)
    assert moved == pytest.approx(total_need * scale, rel=1e-4)
    got = (w - water)[cuts]
    assert np.allclose(got, (need * scale)[cuts], atol=2e-6)


def test_pond_leaves_far_cuts_berms_and_uncut_windows_alone():
    bed, water = _flat_pond(cut_rows=slice(98, 100), cut_cols=slice(60, 62))
    bed[10:12, 10:12] = -0.05  # a cut 30 cells from the pond
    water[10:12, 10:12] = 0.01
    bed[150:152, 150:152] = 0.2  # a berm
    water[150:152, 150:152] = 0.0
    c = _core(bed, water)
    c.pond_ruts()
    w = c.planes()["water"].reshape(200, 200)
    assert np.all(w[10:12, 10:12] == np.float32(0.01))
    assert np.all(w[150:152, 150:152] == 0.0)
    assert np.all(w[98:100, 60:62] > 0.1)
    bed2, water2 = _flat_pond(cut_bed=-0.01)  # not deeper than 12 mm: no cut
    c = _core(bed2, water2)
    assert c.pond_ruts() == 0.0
    assert np.array_equal(c.planes()["water"], water2.ravel())


def test_seat_median_floor_count_and_rut_cap():
    rng = np.random.default_rng(7)
    bed = np.zeros((200, 200), np.float32)
    water = np.zeros((200, 200), np.float32)
    water[60:140, 60:140] = rng.uniform(0.01, 0.09, (80, 80)).astype(np.float32)
    pressed = np.zeros((200, 200), np.float32)
    rut = np.zeros((200, 200), bool)
    rut[99:101, 70:130] = True
    pressed[rut] = -0.06
    water[rut] = 0.3  # stands far above the pond
    bank = np.zeros((200, 200), bool)
    bank[58:60, 60:140] = True
    pressed[bank] = 0.045  # above the floor's band
    water[bank] = 0.2  # pushed up the bank
    c = _core(bed, water)
    v0 = _volume(water)
    eta = c.seat_rut_water(pressed.ravel(), rut.ravel())
    floor = (~rut) & (pressed >= -0.02) & (pressed < 0.04) & (water > 0.008)
    etas = np.sort((pressed + water)[floor].astype(np.float32))
    assert eta == etas[len(etas) // 2]
    w = c.planes()["water"].reshape(200, 200)
    assert np.allclose(w[rut], max(eta + 0.06, 0.0), atol=1e-7)
    assert np.allclose(w[bank], max(eta - 0.045, 0.0), atol=1e-7)
    # the excess is poured onto the floor in proportion to depth
    gain = (w - water)[floor]
    ratio = gain / water[floor]
    assert np.ptp(ratio) < 1e-5 * ratio.mean() + 1e-7
    assert _volume(w) == pytest.approx(v0, rel=1e-6)
    # fewer than 8 floor cells: None, nothing changes
    water2 = np.zeros((200, 200), np.float32)
    water2[0, :7] = 0.05
    c = _core(bed, water2)
    assert c.seat_rut_water(np.zeros(N, np.float32), np.zeros(N, bool)) is None
    assert np.array_equal(c.planes()["water"], water2.ravel())


def test_seat_slurry_on_a_walled_in_cut_stays():
    bed = np.zeros((200, 200), np.float32)
    water = np.zeros((200, 200), np.float32)
    water[50:150, 50:150] = 0.05
    rut = np.zeros((200, 200), bool)
    rut[90:99, 90:99] = True  # 9 x 9 cuts: the centre's rings 1..4 are all cuts
    slurry = np.zeros((200, 200), np.float32)
    slurry[94, 94] = 0.2
    pressed = np.where(rut, -0.05, 0.0).astype(np.float32)
    c = _core(bed, water, slurry)
    c.seat_rut_water(pressed.ravel(), rut.ravel())
    s = c.planes()["slurry"].reshape(200, 200)
    # room on the centre cut: eta - pressed - water = 0.05 + 0.05 - 0.05
    assert s[94, 94] == np.float32(np.float32(0.05) + (np.float32(0.2) - np.float32(0.05)))
    assert np.count_nonzero(s) == 1
