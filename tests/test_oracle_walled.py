"""Pins of the walled-sheet restatement (oracle/sheet.c, oracle/core.c) by
known answers, closed forms and invariants the mathematics fixes:

* the surface lift, the no-slip ghost, the no-penetration filter, the
  displacement and the redirect on hand-worked cells
  (src/fluid_core.rs:4225-4563);
* one wall face of the walled transport against its closed form, and a wall
  across a channel that no water crosses (:3894-4223);
* a lake at rest around a wall stays at rest (the lift keeps the pressure
  flat at the wall);
* the wall mask against a brute-force rasterisation of the rotated
  rectangle (:7288-7362, :4584-4604);
* hydro_load's hydrostatic force (rho g h^2 / 2 per metre of wall) and the
  half-submerged wheel's buoyancy (pi r^2 w / 2) (:7164-7249, :4606-4615);
* the scour rate on one cell beside a contact (:7364-7421)."""
import math

import numpy as np
import pytest

import pyoracle as O

G = np.float32(9.81)


def _grid(res, fill=0.0, dtype=np.float32):
    return np.full(res * res, fill, dtype)


@pytest.mark.parametrize("precision", [32, 64])
def test_lift_takes_the_mean_of_wet_open_cells_within_two(precision):
    dt_ = np.float32 if precision == 32 else np.float64
    res = 7
    eta = np.arange(res * res, dtype=dt_) * dt_(0.01)
    h = _grid(res, 0.0, dt_)
    solid = np.zeros(res * res, np.uint8)
    solid[3 * res + 3] = 1
    solid[3 * res + 4] = 1  # a second wall cell: never averaged in
    wet = [(1, 1), (5, 3), (3, 5), (2, 4)]  # (i, j); (5, 3) is two cells right
    for i, j in wet:
        h[j * res + i] = 0.1
    h[3 * res + 4] = 0.5  # wet but solid: skipped
    h[0] = 0.2  # wet but three cells away: outside the window
    before = eta.copy()
    O.lift_solid_surface(eta, h, solid, res, precision=precision)
    want = np.mean([before[j * res + i] for i, j in wet])
    assert eta[3 * res + 3] == pytest.approx(want, rel=1e-6)
    # The other wall cell (4, 3) sees (2, 4)? no: |di| = 2, |dj| = 1 -> yes,
    # (5, 3) yes, (3, 5) yes, (1, 1) no (|di| = 3).
    want4 = np.mean([before[j * res + i] for i, j in ((5, 3), (3, 5), (2, 4))])
    assert eta[3 * res + 4] == pytest.approx(want4, rel=1e-6)
    others = np.ones(res * res, bool)
    others[[3 * res + 3, 3 * res + 4]] = False
    assert np.array_equal(eta[others], before[others])


@pytest.mark.parametrize("precision", [32, 64])
def test_ghost_is_minus_the_mean_of_wet_open_neighbours(precision):
    dt_ = np.float32 if precision == 32 else np.float64
    res = 5
    u, v, h = _grid(res, 0, dt_), _grid(res, 0, dt_), _grid(res, 0, dt_)
    solid = np.zeros(res * res, np.uint8)
    c = 2 * res + 2
    solid[c] = 1
    u[c - 1], v[c - 1], h[c - 1] = 1.0, 0.5, 0.1
    u[c + 1], v[c + 1], h[c + 1] = 3.0, -1.5, 0.1
    u[c - res], h[c - res] = 9.0, 0.0  # dry: skipped
    u[c], v[c] = 7.0, 7.0
    O.noslip_ghost(u, v, h, solid, res, precision=precision)
    assert u[c] == pytest.approx(-2.0) and v[c] == pytest.approx(0.5)
    # A wall cell with no wet open neighbour gets zero velocity.
    u2, v2, h2 = _grid(res, 4.0, dt_), _grid(res, 4.0, dt_), _grid(res, 0, dt_)
    O.noslip_ghost(u2, v2, h2, solid, res, precision=precision)
    assert u2[c] == 0.0 and v2[c] == 0.0 and u2[c - 1] == 4.0


def test_block_drops_only_the_into_wall_part():
    res = 5
    u, v, h = _grid(res, 0), _grid(res, 0), _grid(res, 0.1)
    solid = np.zeros(res * res, np.uint8)
    solid[2 * res + 3] = 1  # the wall is at +x of cell (2, 2)
    c = 2 * res + 2
    u[c], v[c] = 2.0, 1.0
    u[c - 2], v[c - 2] = -2.0, 1.0  # cell (0, 2): no wall beside it
    u[2 * res + 4], v[2 * res + 4] = -3.0, 0.5  # cell (4, 2): wall at -x, flow into it
    O.block_into_solid(u, v, h, solid, res)
    assert (u[c], v[c]) == (0.0, 1.0)
    assert (u[2 * res + 4], v[2 * res + 4]) == (0.0, 0.5)
    # Flow away from the wall stays.
    u[c] = -1.0
    O.block_into_solid(u, v, h, solid, res)
    assert u[c] == -1.0


def test_displacement_splits_the_column_over_open_neighbours():
    res = 9
    h, u, v = _grid(res, 0), _grid(res, 0.3), _grid(res, 0.3)
    solid = np.zeros(res * res, np.uint8)
    c = 4 * res + 4
    solid[c] = 1
    solid[c - 1] = 1  # one 4-neighbour is a wall: three share
    h[c] = 0.12
    O.displace_occupied(h, u, v, solid, res)
    assert h[c] == 0.0 and u[c] == 0.0 and v[c] == 0.0
    for q in (c + 1, c - res, c + res):
        assert h[q] == np.float32(0.12) / np.float32(3.0)
    # A cell walled in on all four sides spreads over the open cells within two.
    h, u, v = _grid(res, 0), _grid(res, 0), _grid(res, 0)
    solid = np.zeros(res * res, np.uint8)
    for q in (c, c - 1, c + 1, c - res, c + res):
        solid[q] = 1
    h[c] = 0.2
    O.displace_occupied(h, u, v, solid, res)
    open_near = [(j * res + i) for j in range(2, 7) for i in range(2, 7) if not solid[j * res + i]]
    assert len(open_near) == 20
    assert all(h[q] == np.float32(0.2) / np.float32(20.0) for q in open_near)
    # Enclosed within two cells: the column stays.
    h, u, v = _grid(res, 0), _grid(res, 0), _grid(res, 0)
    solid = np.ones(res * res, np.uint8)
    h[c] = 0.2
    O.displace_occupied(h, u, v, solid, res)
    assert h[c] == np.float32(0.2)


def test_redirect_moves_the_excess_to_the_flanks():
    """Cell (3, 3) runs at u = 0.5 into a wall at (4, 3); its flanks (3, 2)
    and (3, 4) hold 0.02: dh = min(h - 0.02, h min(u dt / cell, 0.45)), half
    to each flank."""
    res, cell, dt = 7, np.float32(0.2), np.float32(1.0 / 60.0)
    h, u, v = _grid(res, 0), _grid(res, 0), _grid(res, 0)
    bed = _grid(res, 0)
    solid = np.zeros(res * res, np.uint8)
    c = 3 * res + 3
    solid[c + 1] = 1
    h[c], u[c] = 0.1, 0.5
    h[c - res] = h[c + res] = 0.02
    O.redirect_into_walls(h, u, v, solid, bed, res, cell, dt)
    frac = min(np.float32(0.5) * dt / cell, np.float32(0.45))
    dh = min(np.float32(0.1) - np.float32(0.02), np.float32(0.1) * frac)
    share = dh / np.float32(2.0)
    assert h[c] == np.float32(0.1) - dh
    assert h[c - res] == np.float32(0.02) + share and h[c + res] == np.float32(0.02) + share
    # Deep flanks: only the water above their mean turns aside (0.1 - 0.095).
    h3, u3 = _grid(res, 0), _grid(res, 0)
    h3[c], u3[c] = 0.1, 2.0
    h3[c - res] = h3[c + res] = 0.095
    O.redirect_into_walls(h3, u3, v, solid, bed, res, cell, dt)
    ex = np.float32(0.1) - (np.float32(0.095) + np.float32(0.095)) / np.float32(2.0)
    assert ex < np.float32(0.1) * min(np.float32(2.0) * dt / cell, np.float32(0.45))
    assert h3[c] == np.float32(0.1) - ex
    # Below the 0.05 m/s threshold nothing moves.
    h2 = h.copy()
    u[c] = 0.04
    O.redirect_into_walls(h2, u, v, solid, bed, res, cell, dt)
    assert np.array_equal(h2, h)


@pytest.mark.parametrize("precision", [32, 64])
def test_walled_transport_face_closed_form_and_no_flux_into_walls(precision):
    dt_ = np.float32 if precision == 32 else np.float64
    res, cell, dt = 6, dt_(0.2), dt_(1.0 / 30.0)
    h, u, v = _grid(res, 0, dt_), _grid(res, 0, dt_), _grid(res, 0, dt_)
    gate = np.ones(res * res, np.uint8)
    bed = _grid(res, 0, dt_)
    solid = np.zeros(res * res, np.uint8)
    a = 2 * res + 1
    h[a], h[a + 1] = 0.1, 0.05
    u[a], u[a + 1] = 0.6, 0.2  # face speed 0.4
    u[a + 2] = -0.2  # the next face is still
    u[a - 1] = -0.6  # and so is the one before
    O.transport_walled(h, u, v, gate, solid, bed, res, cell, dt, True, precision=precision)
    frac = min(dt_(0.4) * dt / cell, dt_(0.45))
    dh = dt_(0.1) * frac
    assert h[a] == dt_(0.1) - dh and h[a + 1] == dt_(0.05) + dh
    # The same face beside a wall cell carries nothing.
    h[a], h[a + 1] = 0.1, 0.05
    solid[a + 1] = 1
    O.transport_walled(h, u, v, gate, solid, bed, res, cell, dt, True, precision=precision)
    assert h[a] == dt_(0.1) and h[a + 1] == dt_(0.05)


def _channel(res=40, wall_col=20):
    bed = np.zeros((res, res), np.float32)
    for i in range(res):
        bed[:, i] = 0.002 * (res - i)  # slopes down toward +x
    h = np.zeros((res, res), np.float32)
    h[4:36, 4:18] = 0.08
    solid = np.zeros((res, res), np.uint8)
    solid[:, wall_col] = 1  # a wall across the whole channel
    return bed.ravel(), h.ravel(), solid.ravel()


@pytest.mark.parametrize("precision", [32, 64])
def test_a_wall_across_the_channel_holds_the_water(precision):
    dt_ = np.float32 if precision == 32 else np.float64
    res = 40
    bed, h, solid = (a.astype(dt_) if a.dtype != np.uint8 else a for a in _channel(res))
    u, v = np.zeros_like(h), np.zeros_like(h)
    nu = np.full(res * res, O.dynamic_viscosity(3, 1.0), dt_)
    gate = np.ones(res * res, np.uint8)
    v0 = h.astype(np.float64).sum()
    for _ in range(60):
        O.step_sheet_walled(h, u, v, bed, nu, gate, solid, res, 0.2, 1.0 / 60.0, ("gravity", 0.08),
                            precision=precision)
    H = h.reshape(res, res)
    assert H[:, 20].max() == 0.0, "water stands in the wall"
    assert H[:, 21:].max() == 0.0, "water crossed the wall"
    assert H[:, 17:20].max() > 0.0, "water reached the wall"
    assert abs(h.astype(np.float64).sum() - v0) <= 1e-4 * v0


@pytest.mark.parametrize("precision", [32, 64])
def test_a_lake_at_rest_around_a_wall_stays_at_rest(precision):
    """Flat free surface over a bumpy bed with a block in it: the lift gives
    the wall cells the lake's level, so no cell feels a pressure gradient
    and nothing moves."""
    dt_ = np.float32 if precision == 32 else np.float64
    res = 32
    rng = np.random.default_rng(5)
    bed = (0.05 * rng.random((res, res))).astype(dt_)
    level = dt_(0.3)
    h = (level - bed).astype(dt_).ravel()
    bed = bed.ravel()
    solid = np.zeros((res, res), np.uint8)
    solid[12:18, 10:14] = 1
    solid = solid.ravel()
    h[solid == 1] = 0.0
    u, v = np.zeros_like(h), np.zeros_like(h)
    nu = np.full(res * res, O.dynamic_viscosity(3, 1.0), dt_)
    gate = np.ones(res * res, np.uint8)
    h0 = h.copy()
    for _ in range(5):
        O.step_sheet_walled(h, u, v, bed, nu, gate, solid, res, 0.2, 1.0 / 60.0, ("gravity", 0.08),
                            precision=precision)
    tol = 1e-5 if precision == 32 else 1e-12
    assert np.abs(u).max() <= tol and np.abs(v).max() <= tol
    assert np.abs(h - h0).max() <= tol


def test_solid_mask_matches_a_brute_force_rasterisation():
    c = O.Core((0.0, 0.0))
    x0 = c.origin().astype(np.float64)
    X, Z = c.cell_centres()
    rng = np.random.default_rng(3)
    bed = (0.02 * rng.random(200 * 200)).astype(np.float32)
    water = np.zeros(200 * 200, np.float32)
    water.reshape(200, 200)[80:120, 80:120] = 0.15
    c.map_water_bed(bed, water)
    belly = dict(x=0.3, z=-0.2, yaw=0.7, half_length=1.6, half_width=0.8, bottom_y=0.1, top_y=1.0, contact=0)
    wheel = dict(x=1.1, z=0.4, yaw=0.7, half_length=0.35, half_width=0.14, bottom_y=-1.0, top_y=1.0, contact=1,
                 radius=0.38, width=0.28)
    c.set_hydro_solids([belly, wheel])
    got = c.solid_mask().reshape(200, 200)
    p = c.planes()
    eta = (p["bed"] + p["water"]).reshape(200, 200).astype(np.float64)
    pad = np.pad(eta, 1, constant_values=-np.inf)
    eta_n = np.max(np.stack([eta, pad[:-2, 1:-1], pad[2:, 1:-1], pad[1:-1, :-2], pad[1:-1, 2:]]), axis=0)
    bedp = p["bed"].reshape(200, 200).astype(np.float64)
    want = np.zeros((200, 200), np.uint8)
    for s, val in ((belly, 2), (wheel, 1)):
        fx, fz = math.sin(s["yaw"]), math.cos(s["yaw"])
        dx, dz = X.astype(np.float64) - s["x"], Z.astype(np.float64) - s["z"]
        along = dx * fx + dz * fz
        lat = dx * fz - dz * fx
        inside = (np.abs(along) <= s["half_length"]) & (np.abs(lat) <= s["half_width"])
        wall = inside & (eta_n > s["bottom_y"] + 1e-4) & (bedp < s["top_y"])
        want[wall] = val
    diff = got != want
    # FP32 rounding may flip a cell whose centre lies on the rectangle's edge.
    assert diff.sum() <= 4, int(diff.sum())
    assert (got == 1).sum() >= 3 and (got == 2).sum() > 50
    del x0


def test_hydrostatic_force_and_wheel_buoyancy_closed_forms():
    c = O.Core((0.0, 0.0))
    bed = np.zeros(200 * 200, np.float32)
    water = np.zeros((200, 200), np.float32)
    water[:, :100] = 0.25  # water west of x = 0, dry east
    c.map_water_bed(bed, water.ravel())
    # A wall two cells thick right at the shore, 4 m long (20 cells).
    wall = dict(x=0.2, z=0.0, yaw=0.0, half_length=2.0, half_width=0.2, bottom_y=-1.0, top_y=2.0, contact=0)
    c.set_hydro_solids([wall])
    m = c.solid_mask().reshape(200, 200)
    rows = np.flatnonzero(m.any(axis=1))
    cols = np.flatnonzero(m.any(axis=0))
    assert cols.min() == 100 and len(rows) == 20
    load = c.hydro_load()
    # rho g h^2 / 2 per metre on the wetted face, times the face length.
    want = 1000.0 * 9.81 * 0.25 ** 2 / 2.0 * (len(rows) * 0.2)
    assert load["force_x"] == pytest.approx(want, rel=1e-5)
    assert abs(load["force_z"]) <= 1e-3 * want
    assert load["belly_blocks"] == m.astype(bool).sum() and load["wheel_blocks"] == 0
    # A wheel in water as deep as its radius displaces half its cylinder.
    c2 = O.Core((0.0, 0.0))
    r, w = 0.38, 0.28
    c2.map_water_bed(bed, np.full(200 * 200, r, np.float32))
    c2.set_hydro_solids([dict(x=0.0, z=0.0, yaw=0.0, half_length=0.3, half_width=0.1, bottom_y=-1.0, top_y=1.0,
                              contact=1, radius=r, width=w)])
    ld = c2.hydro_load()
    assert ld["displaced_m3"] == pytest.approx(math.pi * r * r * w / 2.0, rel=1e-5)
    assert ld["wheel_buoyancy_n"] == pytest.approx(1000.0 * 9.81 * ld["displaced_m3"], rel=1e-6)
    # At half the radius the segment is r^2 (pi / 3 - sqrt(3) / 4).
    c3 = O.Core((0.0, 0.0))
    c3.map_water_bed(bed, np.full(200 * 200, r / 2.0, np.float32))
    c3.set_hydro_solids([dict(x=0.0, z=0.0, yaw=0.0, half_length=0.3, half_width=0.1, bottom_y=-1.0, top_y=1.0,
                              contact=1, radius=r, width=w)])
    want = r * r * (math.pi / 3.0 - math.sqrt(3.0) / 4.0) * w
    assert c3.hydro_load()["displaced_m3"] == pytest.approx(want, rel=1e-5)


def test_scour_rate_beside_a_contact():
    """A wet cell next to a contact running at 0.3 m/s: tau = 1000 * 0.03 *
    0.09 = 2.7 Pa over tau_c = 1.5, rate = min(0.0035 (1.8 - 1), 0.012) =
    0.0028 m/s, cut = rate dt; at 0.6 m/s the rate is capped at 0.012."""
    c = O.Core((0.0, 0.0))
    bed = np.zeros(200 * 200, np.float32)
    water = np.full(200 * 200, 0.1, np.float32)
    c.map_water_bed(bed, water)
    c.set_hydro_solids([dict(x=0.1, z=0.1, yaw=0.0, half_length=0.09, half_width=0.09, bottom_y=-1.0, top_y=1.0,
                             contact=1, radius=0.38, width=0.28)])
    m = c.solid_mask().reshape(200, 200)
    assert m.sum() == 1
    j, i = map(int, np.argwhere(m == 1)[0])
    p = c.planes()
    p["uw"].reshape(200, 200)[j, i - 1] = 0.3
    p["uw"].reshape(200, 200)[j, i + 1] = 0.6
    c.set_planes(p)
    dt = 1.0 / 30.0
    c.scour(dt)
    rate = 0.0035 * (1000.0 * 0.03 * 0.09 / 1.5 - 1.0)
    assert drop_capped(c, j, i + 1) == pytest.approx(0.012 * dt, rel=1e-5)
    drop = c.bed_drop().reshape(200, 200)
    assert drop[j, i - 1] == pytest.approx(rate * dt, rel=1e-5)
    assert np.count_nonzero(drop) == 2
    q = c.planes()
    assert q["slurry"].reshape(200, 200)[j, i - 1] == pytest.approx(rate * dt, rel=1e-5)
    assert c.hydro_load()["eroded_m3"] == pytest.approx((rate + 0.012) * dt * 0.04, rel=1e-5)


def drop_capped(c, j, i):
    return float(c.bed_drop().reshape(200, 200)[j, i])
