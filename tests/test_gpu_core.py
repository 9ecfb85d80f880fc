"""GPU parity of the FluidCore handle (include/silt_core.h) against the
FluidCore restatement (oracle/core.c), step by step.

The water planes follow the reference order exactly (uniform water viscosity,
exact sheet path, the coarse sheets' in-place confinement sweep, the same
per-cell arithmetic in refine / restrict / near field / shoves / spray), so
they must be equal value for value.  The slurry's damping factor comes from
the per-cell viscosity through the device exp/pow, so the slurry is compared
to 1e-5 of its scale.  Ledgers are compared to the reference's 1e-3 band and
the GPU's own FP64 ledger must close to 1e-6 relative.
"""
import math

import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from silt_amd import scenes

pytestmark = pytest.mark.gpu


def _eq(a, b, what):
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, f"{what}: {bad.size} cells differ, first {bad[:5]} {a[bad[:5]]} vs {b[bad[:5]]}"


def _close(a, b, what, tol=1e-5):
    scale = max(float(np.abs(b).max()), 1e-12)
    err = float(np.abs(a.astype(np.float64) - b).max())
    assert err <= tol * scale, f"{what}: max err {err:.3e} vs scale {scale:.3e}"


def _pair(center=(0.0, 0.0), path="exact"):
    return S.FluidCore(center, path=path), O.Core(center)


def _tires(k, speed=3.0):
    """Four tires (hubs at +-0.7 m x +-1.2 m) rolling forward 8 cm per step."""
    z0 = -2.0 + 0.08 * k
    return [dict(x=x, z=z0 + dz, forward_x=0.0, forward_z=1.0, radius=0.38, width=0.28, speed=speed, slip=0.2)
            for x in (-0.7, 0.7) for dz in (-1.2, 1.2)]


def _compare(g, o, step, slurry_tol=1e-5):
    pg, po = g.planes(), o.planes()
    for k in ("water", "uw", "vw"):
        _eq(pg[k], po[k], f"{k} step {step}")
    for k in ("slurry", "us", "vs"):
        _close(pg[k], po[k], f"{k} step {step}", slurry_tol)


@pytest.mark.parametrize("near", [False, True])
def test_core_steps_match_oracle_with_four_tires(near):
    g, o = _pair()
    sc = scenes.window_scene(seed=1)
    for c in (g, o):
        c.map_water_bed(sc["bed"], sc["water"])
        c.lay_slurry(sc["slurry"], "clay", 0.6)
        c.set_near_field(near)
    for k in range(12):
        for t in _tires(k):
            g.push_tire(**t)
            o.push_tire(**t)
        ev, n = g.step(1.0 / 30.0)
        no, vol_o = o.step(1.0 / 30.0)
        assert n == no, f"spray events step {k}: gpu {n} oracle {no}"
        _compare(g, o, k)
    if near:
        assert g.near_tiles() == o.near_tiles() > 0
    lg, lo = g.ledger(), o.ledger()
    for key in ("water_m3", "slurry_m3", "spray_m3", "total_m3"):
        assert abs(lg[key] - lo[key]) <= 1e-3 * max(abs(lo[key]), 1e-9), (key, lg[key], lo[key])


def test_core_fine_pace3_and_still_tire_in_deep_water():
    """The acceptance scene's fine pace (3) and the deep-water still tire of
    fine_patch_leaves_no_hole_in_deep_water (src/fluid_core.rs:11615-11660)."""
    g, o = _pair()
    for c in (g, o):
        c.pour_water(-6.0, 6.0, -6.0, 6.0, 0.40)
        c.set_fine_pace(3)
    t = dict(x=0.0, z=0.0, forward_x=0.0, forward_z=1.0, radius=0.01, width=0.01, speed=0.0, slip=0.0)
    before = g.ledger()["total_m3"]
    for k in range(10):
        g.push_tire(**t)
        o.push_tire(**t)
        g.step(1.0 / 30.0)
        o.step(1.0 / 30.0)
        _compare(g, o, k)
    after = g.ledger()["total_m3"]
    assert abs(after - before) < 1e-3 * before


def test_fine_patch_after_near_apply_matches_oracle():
    """fine_water_debug (src/fluid_core/near.rs:341-346) after steps with a
    tire at the world origin: the patch laid from the coarse sheet plus the
    near-field residuals equals the oracle's, including the columns whose
    world keys are (-1, -1) and (-1, -2)."""
    g, o = _pair()
    for c in (g, o):
        c.pour_water(-6.0, 6.0, -6.0, 6.0, 0.40)
    for k in range(3):
        t = dict(x=0.0, z=0.08 * k, forward_x=0.0, forward_z=1.0, radius=0.38, width=0.28, speed=3.0, slip=0.2)
        g.push_tire(**t)
        o.push_tire(**t)
        g.step(1.0 / 30.0)
        o.step(1.0 / 30.0)
        assert g.near_tiles() == o.near_tiles()
        fg, fo = g.fine_water_debug([t]), o.fine_water_debug([t])
        for a, b, name in zip(fg, fo, ("water", "uw", "vw", "origin")):
            _eq(a, b, f"fine {name} after step {k}")
        _compare(g, o, k)


def test_core_publish_matches_oracle():
    g, o = _pair()
    for c in (g, o):
        c.pour_water(-3.0, 3.0, -3.0, 3.0, 0.05)
        c.pour_slurry(-1.0, 1.0, -1.0, 1.0, 0.02, "clay", 0.6)
    for k in range(4):
        for t in _tires(k, speed=2.0):
            g.push_tire(**t)
            o.push_tire(**t)
        g.step(1.0 / 30.0)
        o.step(1.0 / 30.0)
    gr, film, flow = g.publish()
    ofilm, oflow = o.publish()
    _close(film, ofilm, "film")
    _close(flow, oflow, "flow")
    p = g.planes()
    assert np.array_equal(gr[0::4], p["bed"]) and not gr[1::4].any()
    assert np.isfinite(flow).all()


def test_core_coarse_jacobi_fast_path_tracks_reference_sweep():
    """The fast path (fused kernel, snapshot confinement on the coarse sheets)
    against the reference's in-place sweep: 30 steps of the window scene with
    four tires keep volume within the reference's 1e-3 band of each other."""
    g = S.FluidCore((0.0, 0.0), path="fast")
    g.set_coarse_confine("jacobi")
    o = O.Core((0.0, 0.0))
    sc = scenes.window_scene(seed=2)
    for c in (g, o):
        c.map_water_bed(sc["bed"], sc["water"])
        c.lay_slurry(sc["slurry"], "clay", 0.6)
    for k in range(30):
        for t in _tires(k):
            g.push_tire(**t)
            o.push_tire(**t)
        g.step(1.0 / 30.0)
        o.step(1.0 / 30.0)
    lg, lo = g.ledger(), o.ledger()
    assert abs(lg["total_m3"] - lo["total_m3"]) < 1e-3 * lo["total_m3"]
    pg, po = g.planes(), o.planes()
    l1 = float(np.abs(pg["water"] - po["water"]).sum()) / float(np.abs(po["water"]).sum())
    assert l1 < 0.05, l1


def test_core_publish_surface_matches_oracle():
    """moisture / wet_edge / foam_crest (src/render_contract.rs:457-500,
    :526-569) from the same planes and foam on the GPU and in the oracle, bit
    for bit: a pool with a slurry patch, NaN and negative cells, NaN and
    out-of-range foam, random sheet moisture, with and without soil wetness
    and placed soil."""
    g, o = _pair()
    rng = np.random.default_rng(21)
    n = 200 * 200
    for c in (g, o):
        c.pour_water(-3.0, 3.0, -3.0, 3.0, 0.05)
        c.pour_slurry(-1.0, 1.0, -1.0, 1.0, 0.02, "clay", 0.6)
    for k in range(3):
        for t in _tires(k, speed=2.0):
            g.push_tire(**t)
        g.step(1.0 / 30.0)
    p = g.planes()
    p["moisture"] = rng.uniform(0.0, 0.6, n).astype(np.float32)
    p["water"][[5, 900, 20500]] = [np.nan, -0.01, 0.003]
    p["slurry"][[6, 901]] = [np.inf, 0.0039]
    g.set_planes(p)
    o.set_planes(p)
    # The tire steps left foam on the GPU; the oracle publishes the same.
    foam, turbid, _ = g.tracers()
    foam[[11, 12]] = [np.nan, 1.5]
    g.set_tracers(foam=foam)
    o.set_tracers(foam=foam, turbid=turbid)
    wet = rng.uniform(0.0, 0.6, n).astype(np.float32)
    wet[7] = np.nan
    placed = rng.uniform(-0.01, 0.05, n).astype(np.float32)
    for args in ((None, None), (wet, None), (None, placed), (wet, placed)):
        gm, ge, gf = g.publish_surface(*args)
        om, oe, of = o.publish_surface(*args)
        for a, b, name in zip((gm, ge, gf), (om, oe, of), ("moisture", "wet_edge", "foam_crest")):
            assert np.array_equal(a, b), (name, args[0] is not None, args[1] is not None)
        assert 0 < ge.sum() < n and np.isfinite(gf).all() and gf[0::2].any() and gf[0::2].max() <= 1.0
