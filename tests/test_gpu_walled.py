"""GPU parity of the walled sheet (silt_step_sheet_walled, include/silt_sheet.h)
and of the FluidCore with hydro solids (include/silt_core.h) against the FP32
restatement (oracle/sheet.c, oracle/core.c).

The walled stage kernels keep the reference's operation order (no fused
multiply-add, IEEE division; uniform viscosity, so no device transcendental
enters), so every case must equal the oracle value for value after every
step: a wall across a channel, a rotated block in a flowing puddle under both
of the reference's kick forms (in-place sweep, and the parallel kicks of a
large box with row splits), a wind-driven walled pond, the 2 cm grid, and the
body displacement.  The core with a belly and four tire contacts moving with
the vehicle compares water value for value and slurry (per-cell viscosity,
device exp) to 1e-5 of its scale, the wall masks exactly, hydro_load to 1e-4,
and holds the FP64 ledger gap to 1e-9 of the total per step with the scour
booked as a source."""
import math

import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def _eq(a, b, what):
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, f"{what}: {bad.size} cells differ, first {bad[:5]} {a[bad[:5]]} vs {b[bad[:5]]}"


def _close(a, b, what, tol=1e-5):
    scale = max(float(np.abs(b).max()), 1e-12)
    err = float(np.abs(a.astype(np.float64) - b).max())
    assert err <= tol * scale, f"{what}: max err {err:.3e} vs scale {scale:.3e}"


def _both(state, bed, nu, gate, solid, res, cell, dt, drive, steps, row_split=True):
    g = [a.copy() for a in state]
    o = [a.copy() for a in state]
    for k in range(steps):
        ng = S.step_sheet_walled(*g, bed, nu, gate, solid, res, cell, dt, drive, row_split=int(row_split))
        no = O.step_sheet_walled(*o, bed, nu, gate, solid, res, cell, dt, drive, inner_par=row_split)
        assert ng == no, f"substeps differ at step {k}: gpu {ng} oracle {no}"
        for a, b, name in zip(g, o, "huv"):
            _eq(a, b, f"{name} step {k}")
    return g


def _rect_mask(res, cell, cx, cz, yaw, hl, hw):
    c = (np.arange(res) + 0.5) * cell - res * cell / 2.0
    X, Z = np.meshgrid(c, c)
    fx, fz = math.sin(yaw), math.cos(yaw)
    dx, dz = X - cx, Z - cz
    along, lat = dx * fx + dz * fz, dx * fz - dz * fx
    return ((np.abs(along) <= hl) & (np.abs(lat) <= hw)).astype(np.uint8).ravel()


def test_wall_across_a_channel_equals_the_oracle():
    res, cell = 40, 0.2
    n = res * res
    bed = np.repeat(-0.01 * np.arange(res, dtype=np.float32)[None, :], res, axis=0).ravel()
    h = np.zeros((res, res), np.float32)
    h[4:36, 4:18] = 0.08
    solid = np.zeros((res, res), np.uint8)
    solid[:, 20] = 1
    nu = np.full(n, O.dynamic_viscosity(3, 1.0), np.float32)
    gate = np.ones(n, np.uint8)
    g = _both([h.ravel(), np.zeros(n, np.float32), np.zeros(n, np.float32)], bed, nu, gate, solid.ravel(), res,
              cell, 1.0 / 60.0, ("gravity", 0.08), 25)
    H = g[0].reshape(res, res)
    assert H[:, 20:].max() == 0.0 and H[:, 17:20].max() > 0.0


@pytest.mark.parametrize("row_split", [False, True])
def test_block_in_a_flowing_puddle_equals_the_oracle(row_split):
    """A 120^2 puddle flowing at 0.4 m/s onto a rotated block: the redirect,
    the no-penetration filter and the ghosts all act; the wet box is large
    enough that row splits select the parallel kicks."""
    res, cell = 120, 0.2
    n = res * res
    c = (np.arange(res) + 0.5) * cell - res * cell / 2.0
    X, Z = np.meshgrid(c, c)
    h = (0.12 * np.exp(-(X ** 2 + Z ** 2) / (2 * 4.0 ** 2))).astype(np.float32).ravel()
    h[h < 1e-6] = 0.0
    u = np.where(h > 0, np.float32(0.4), np.float32(0.0)).astype(np.float32)
    v = np.where(h > 0, np.float32(-0.1), np.float32(0.0)).astype(np.float32)
    bed = (0.01 * np.sin(X / 3.0) * np.cos(Z / 5.0)).astype(np.float32).ravel()
    solid = _rect_mask(res, cell, 1.5, 0.3, 0.6, 1.2, 0.5)
    h[solid == 1] = 0.0
    nu = np.full(n, O.dynamic_viscosity(3, 1.0), np.float32)
    gate = np.ones(n, np.uint8)
    _both([h, u, v], bed, nu, gate, solid, res, cell, 1.0 / 30.0, ("gravity", 0.08), 10, row_split=row_split)


def test_wind_on_a_walled_pond_equals_the_oracle():
    res, cell = 48, 0.2
    n = res * res
    h = np.zeros((res, res), np.float32)
    h[6:42, 6:42] = 0.05
    solid = np.zeros((res, res), np.uint8)
    solid[20:28, 30:33] = 1
    h[solid == 1] = 0.0
    nu = np.full(n, O.dynamic_viscosity(3, 1.0), np.float32)
    gate = np.ones(n, np.uint8)
    _both([h.ravel(), np.zeros(n, np.float32), np.zeros(n, np.float32)], np.zeros(n, np.float32), nu, gate,
          solid.ravel(), res, cell, 1.0 / 60.0, ("wind", 1.5, 0.3, 0.05), 12)


def test_two_centimetre_walled_grid_equals_the_oracle():
    res, cell = 160, 0.02
    n = res * res
    rng = np.random.default_rng(8)
    X, Z = np.meshgrid(np.arange(res), np.arange(res))
    h = np.zeros((res, res), np.float32)
    for _ in range(6):
        ci, cj = rng.integers(20, res - 20, 2)
        h += np.where((X - ci) ** 2 + (Z - cj) ** 2 <= rng.integers(30, 200), np.float32(0.03), np.float32(0.0))
    solid = np.zeros((res, res), np.uint8)
    solid[70:90, 60:66] = 1
    h[solid == 1] = 0.0
    bed = (0.004 * np.sin(X / 9.0) * np.cos(Z / 7.0)).astype(np.float32).ravel()
    nu = np.full(n, O.dynamic_viscosity(3, 1.0), np.float32)
    gate = np.ones(n, np.uint8)
    _both([h.ravel(), np.zeros(n, np.float32), np.zeros(n, np.float32)], bed, nu, gate, solid.ravel(), res, cell,
          1.0 / 60.0, ("gravity", 0.08), 5, row_split=False)


def test_displacement_equals_the_oracle():
    res = 64
    n = res * res
    rng = np.random.default_rng(4)
    h = (0.05 * rng.random(n)).astype(np.float32)
    u = rng.standard_normal(n).astype(np.float32)
    v = rng.standard_normal(n).astype(np.float32)
    solid = (rng.random(n) < 0.3).astype(np.uint8)
    solid.reshape(res, res)[20:30, 20:30] = 1  # cells walled in on all sides
    solid.reshape(res, res)[40:50, 40:50] = 1
    h.reshape(res, res)[45, 45] = 1e-9  # dries without moving
    g = [h.copy(), u.copy(), v.copy()]
    o = [h.copy(), u.copy(), v.copy()]
    S.displace_occupied(*g, solid, res)
    O.displace_occupied(*o, solid, res)
    for a, b, name in zip(g, o, "huv"):
        _eq(a, b, name)
    assert g[0].reshape(res, res)[45, 45] == 0.0 and g[0].reshape(res, res)[25, 25] > 0.0


def test_walled_device_step_equals_the_host_step():
    res, cell = 40, 0.2
    n = res * res
    h = np.zeros((res, res), np.float32)
    h[5:30, 5:15] = 0.06
    solid = np.zeros((res, res), np.uint8)
    solid[10:20, 16:18] = 1
    bed = np.zeros(n, np.float32)
    nu = np.full(n, O.dynamic_viscosity(3, 1.0), np.float32)
    gate = np.ones(n, np.uint8)
    host = [h.ravel().copy(), np.full(n, 0.3, np.float32), np.zeros(n, np.float32)]
    dev = [DeviceArray(a.copy()) for a in host]
    dbed, dsolid = DeviceArray(bed), DeviceArray(solid.ravel())
    for _ in range(5):
        S.step_sheet_walled(*host, bed, nu, gate, solid.ravel(), res, cell, 1.0 / 60.0, ("gravity", 0.08))
        S.step_sheet_walled_device(*(d.ptr for d in dev), dbed.ptr, 0, 0, dsolid.ptr, res, cell, 1.0 / 60.0,
                                   ("gravity", 0.08), nu_uniform=float(nu[0]))
    synchronize()
    for a, d, name in zip(host, dev, "huv"):
        _eq(d.numpy(), a, name)


def _vehicle(k):
    """A belly block and four tire contacts under the four pushes of a
    vehicle rolling 8 cm a step along +z."""
    z0 = -2.0 + 0.08 * k
    pushes = [dict(x=x, z=z0 + dz, forward_x=0.0, forward_z=1.0, radius=0.38, width=0.28, speed=3.0, slip=0.2)
              for x in (-0.7, 0.7) for dz in (-1.2, 1.2)]
    solids = [dict(x=0.0, z=z0, yaw=0.0, half_length=1.6, half_width=0.55, bottom_y=0.02, top_y=1.2, contact=0)]
    solids += [dict(x=p["x"], z=p["z"], yaw=0.0, half_length=0.3, half_width=0.14, bottom_y=-1.0, top_y=0.8,
                    contact=1, radius=0.38, width=0.28) for p in pushes]
    return pushes, solids


def test_core_with_walls_matches_the_oracle_and_closes_its_ledger():
    from silt_amd import scenes
    sc = scenes.window_scene(seed=1)
    g, o = S.FluidCore((0.0, 0.0), path="exact"), O.Core((0.0, 0.0))
    try:
        water = sc["water"].copy()
        water.reshape(200, 200)[80:120, 85:115] += 0.08  # a ford the vehicle sits in
        for c in (g, o):
            c.map_water_bed(sc["bed"], water)
            c.lay_slurry(sc["slurry"], "clay", 0.6)
        prev = g.ledger_lines()
        for k in range(10):
            pushes, solids = _vehicle(k)
            for c in (g, o):
                c.set_hydro_solids(solids)
                for p in pushes:
                    c.push_tire(**p)
            _eq(g.solid_mask(), o.solid_mask(), f"window mask before step {k}")
            g.step(1.0 / 30.0)
            o.step(1.0 / 30.0)
            pg, po = g.planes(), o.planes()
            for key in ("water", "uw", "vw"):
                _eq(pg[key], po[key], f"{key} step {k}")
            for key in ("slurry", "us", "vs"):
                _close(pg[key], po[key], f"{key} step {k}")
            L = g.ledger_lines()
            for l in (0, 1):
                step_gap = L["gap_m3"][l] - prev["gap_m3"][l]
                assert abs(step_gap) <= 1e-9 * L["total_m3"][l], (k, l, step_gap)
            prev = L
        lg, lo = g.hydro_load(), o.hydro_load()
        for key in ("force_x", "force_z", "buoyancy_n", "hull_buoyancy_n", "wheel_buoyancy_n", "displaced_m3"):
            assert abs(lg[key] - lo[key]) <= 1e-4 * max(abs(lo[key]), 1.0), (key, lg[key], lo[key])
        assert lg["wheel_blocks"] == lo["wheel_blocks"] > 0 and lg["belly_blocks"] == lo["belly_blocks"] > 0
        assert lo["eroded_m3"] > 0.0, "no scour beside the contacts"
        assert abs(lg["eroded_m3"] - lo["eroded_m3"]) <= 1e-5 * max(lo["eroded_m3"], 1e-9)
        assert prev["source_m3"][1] == pytest.approx(lg["eroded_m3"], rel=1e-5, abs=1e-12)
    finally:
        g.close()
