"""GPU checks of the FluidCore entries beside the step (include/silt_core.h):
drain_over_bed against the oracle's restatement (exact path, value for
value), map_water_velocity, non_finite_count, coarse_cell / sample and the
full ledger against the planes they read."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from silt_amd import scenes

pytestmark = pytest.mark.gpu

N = 200 * 200


def _eq(a, b, what):
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, f"{what}: {bad.size} cells differ, first {bad[:5]} {a[bad[:5]]} vs {b[bad[:5]]}"


def test_drain_over_bed_matches_the_oracle():
    sc = scenes.window_scene(seed=3)
    g, o = S.FluidCore((0.0, 0.0), path="exact"), O.Core((0.0, 0.0))
    try:
        for c in (g, o):
            c.map_water_bed(sc["bed"], sc["water"])
            c.lay_slurry(sc["slurry"], "clay", 0.6)
            c.pour_water(-3.0, 3.0, 4.0, 8.0, 0.05)
        before = float(np.maximum(g.planes()["water"], 0).astype(np.float64).sum()) * 0.04
        vg = g.drain_over_bed(1.0 / 60.0, 40)
        vo = o.drain_over_bed(1.0 / 60.0, 40)
        assert vg == np.float32(vo), (vg, vo)
        pg, po = g.planes(), o.planes()
        for k in ("water", "uw", "vw"):
            _eq(pg[k], po[k], k)
        for k in ("slurry", "us", "vs"):
            scale = max(float(np.abs(po[k]).max()), 1e-12)
            assert float(np.abs(pg[k] - po[k]).max()) <= 1e-5 * scale, k
        # Dry cut losses went back: the water volume is kept (to the f32 sum's
        # rounding).
        assert abs(vg - before) <= 1e-4 * before, (vg, before)
    finally:
        g.close()


def test_map_water_velocity_sets_wet_cells_only():
    sc = scenes.window_scene(seed=1)
    g = S.FluidCore((0.0, 0.0))
    try:
        g.map_water_bed(sc["bed"], sc["water"])
        rng = np.random.default_rng(2)
        u = rng.normal(0, 1, N).astype(np.float32)
        v = rng.normal(0, 1, N).astype(np.float32)
        u[5] = np.nan
        g.map_water_velocity(u, v)
        p = g.planes()
        wet = p["water"] > 1e-6
        exp_u = np.where(wet, np.where(np.isfinite(u), u, 0.0), 0.0).astype(np.float32)
        exp_v = np.where(wet, v, 0.0).astype(np.float32)
        _eq(p["uw"], exp_u, "uw")
        _eq(p["vw"], exp_v, "vw")
        assert g.non_finite_count() == 0
        p["uw"][7] = np.inf
        g.set_planes({"uw": p["uw"]})
        assert g.non_finite_count() == 1
    finally:
        g.close()


def test_coarse_cell_sample_center_and_full_ledger():
    sc = scenes.window_scene(seed=1)
    g = S.FluidCore((0.0, 0.0))
    try:
        g.map_water_bed(sc["bed"], sc["water"])
        g.lay_slurry(sc["slurry"], "clay", 0.6)
        p = g.planes()
        o = g.origin()
        i, j = 100, 97
        k = j * 200 + i
        cc = g.coarse_cell(i, j)
        assert cc["x"] == np.float32(o[0] + (i + 0.5) * np.float32(0.2))
        assert cc["water_m"] == p["water"][k] and cc["slurry_m"] == p["slurry"][k] and cc["bed_m"] == p["bed"][k]
        assert g.coarse_cell(200, 3) is None
        s = g.sample(float(cc["x"]), float(cc["z"]))
        assert s["water_m"] == p["water"][k] and s["u"] == p["us"][k] and s["moisture"] == p["moisture"][k]
        assert g.sample(-500.0, 0.0)["water_m"] == 0.0
        g.set_center(3.0, -2.0)
        assert tuple(g.center()) == (3.0, -2.0)
        assert tuple(g.origin()) == tuple(o)
        full, led = g.ledger_full(), g.ledger()
        for key in ("water_m3", "slurry_m3", "spray_m3", "total_m3"):
            assert full[key] == led[key]
        assert full["dust_air_m3"] == 0.0 and full["reach_m3"] == 0.0
    finally:
        g.close()


def test_recenter_matches_the_oracle_and_restores_parked_cells():
    """Scroll shifts (a few cells, then past the whole window) and back, with
    tire steps in between, exact path: planes, origin, parked set and
    ledger equal the oracle's; the first window's cells come back."""
    sc = scenes.window_scene(seed=2)
    g, o = S.FluidCore((0.0, 0.0), path="exact"), O.Core((0.0, 0.0))
    try:
        for c in (g, o):
            c.map_water_bed(sc["bed"], sc["water"])
            c.lay_slurry(sc["slurry"], "clay", 0.6)
        moist = np.full(N, 0.3, np.float32)
        grain = np.full(N, 4, np.uint8)
        for c in (g, o):
            c.read_substrate(sc["bed"], moist, grain)
        moves = [(3.4, -1.0), (1.0, 2.2), (63.0, 5.0), (0.0, 0.0)]
        for k, (x, z) in enumerate(moves):
            for c in (g, o):
                c.recenter(x, z)
                c.push_tire(x=x + 0.5, z=z, forward_x=1.0, forward_z=0.0, radius=0.38, width=0.28, speed=3.0,
                            slip=0.2)
                c.step(1.0 / 60.0)
            assert tuple(g.origin()) == tuple(o.origin()), k
            assert g.parked_cells() == o.parked_cells(), k
            pg, po = g.planes(), o.planes()
            for name in ("water", "uw", "vw", "moisture"):
                _eq(pg[name], po[name], f"{name} move {k}")
        # The oracle sums the window in f32 as the reference does (good to about
        # 2e-5 relative over 40 000 cells); the library sums in FP64.
        lg, lo = g.ledger(), o.ledger()
        for key in ("water_m3", "frozen_water_m3", "frozen_slurry_m3", "total_m3"):
            assert abs(lg[key] - lo[key]) <= 1e-4 * max(abs(lo[key]), 1e-9), (key, lg[key], lo[key])
    finally:
        g.close()


def test_lay_slurry_on_water_and_window_bounds():
    sc = scenes.window_scene(seed=1)
    g = S.FluidCore((1.0, -2.0))
    try:
        g.map_water_bed(sc["bed"], sc["water"])
        p0 = g.planes()
        vol = g.lay_slurry_on_water(0.01, "silt")
        p = g.planes()
        deep = p0["water"] > np.float32(0.012)
        assert np.array_equal(p["slurry"], np.where(deep, np.float32(0.01), p0["slurry"]).astype(np.float32))
        assert np.array_equal(p["moisture"], np.where(deep, np.maximum(p0["moisture"], np.float32(0.9)),
                                                      p0["moisture"]).astype(np.float32))
        assert abs(vol - deep.sum() * 0.01 * 0.04) <= 1e-4 * vol
        x0, z0, ext = g.window_bounds()
        o = g.origin()
        assert (x0, z0) == (float(o[0]), float(o[1])) and ext == np.float32(40.0)
    finally:
        g.close()


def test_device_publish_equals_the_host_publish():
    """silt_core_publish_device / _surface_device write the same maps as the
    host calls, straight into device buffers (the renderer's textures in
    ContractGpu::upload, src/render_contract.rs:1146-1166), with soil inputs
    taken from device buffers or from the core's own ground."""
    from hipmem import DeviceArray, synchronize
    from silt_amd import drive as D

    g = D.nested_scene_core(S, scenes, seed=2)
    drv = D.Drive(seed=4)
    try:
        for _ in range(6):
            D.drive_step(g, drv, 1.0 / 30.0)
        n = 200 * 200
        ground, film, flow = g.publish()
        outs = [DeviceArray(np.zeros(k * n, np.float32)) for k in (4, 2, 4)]
        g.publish_device(outs[0].ptr, outs[1].ptr, outs[2].ptr)
        synchronize()
        for a, b, name in zip(outs, (ground, film, flow), ("ground_rut", "film", "flow")):
            assert np.array_equal(a.numpy(), b), name
        # Soil planes given by the caller, on the device.
        rng = np.random.default_rng(1)
        soil = [rng.random(n).astype(np.float32) for _ in range(3)]
        want = g.publish(*soil)
        dsoil = [DeviceArray(a) for a in soil]
        g.publish_device(outs[0].ptr, outs[1].ptr, 0, *(d.ptr for d in dsoil))
        synchronize()
        assert np.array_equal(outs[0].numpy(), want[0]) and np.array_equal(outs[1].numpy(), want[1])
        moist, edge, foam = g.publish_surface(soil_wetness=soil[0])
        souts = [DeviceArray(np.zeros(k * n, np.float32)) for k in (1, 1, 2)]
        g.publish_surface_device(souts[0].ptr, souts[1].ptr, souts[2].ptr, soil_wetness=dsoil[0].ptr)
        synchronize()
        for a, b, name in zip(souts, (moist, edge, foam), ("moisture", "wet_edge", "foam_crest")):
            assert np.array_equal(a.numpy(), b), name
    finally:
        g.close()
