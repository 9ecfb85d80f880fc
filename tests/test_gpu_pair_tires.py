"""Vehicles on the two-layer sheet handle (silt_pair_step_tires,
include/silt_pair.h): no tires is a plain step; the ruts are the oracle's
stamp_tread in order; the network's fluxes keep the FP64 ledger closed."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from silt_amd import drive as D
from silt_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0
RES, CELL = 256, 0.2
ORIGIN = (-0.5 * RES * CELL, -0.5 * RES * CELL)


def _scene(seed=3):
    bed = scenes.fbm_terrain(RES, CELL, ORIGIN, seed).ravel().astype(np.float32)
    water = np.full(RES * RES, 0.03, np.float32)
    slurry = np.zeros(RES * RES, np.float32)
    slurry[: RES * RES // 2] = 0.01
    return bed, water, slurry


def _pair(bed, water, slurry):
    p = S.SheetPair(RES, CELL, path="fast")
    z = np.zeros(RES * RES, np.float32)
    p.upload(bed, water, z, z, slurry, z, z, 0.5)
    p.set_ground(bed, ORIGIN)
    return p


def test_no_tires_is_a_plain_step():
    bed, water, slurry = _scene()
    a, b = _pair(bed, water, slurry), _pair(bed, water, slurry)
    try:
        for _ in range(5):
            a.step(DT)
            b.step_tires(DT, [])
        for x, y in zip(a.download(), b.download()):
            assert np.array_equal(x, y)
    finally:
        a.close()
        b.close()


def test_ruts_follow_the_oracle_and_the_ledger_closes_with_the_network():
    bed, water, slurry = _scene()
    p = _pair(bed, water, slurry)
    m = S.Mlp(11)
    p.set_surrogate(m.params())
    m.close()
    fleet = D.Fleet(seed=5, n=2, extent=RES * CELL, half=6.0)
    ob, op, init = bed.copy(), np.zeros_like(bed), bed.copy()
    try:
        for k in range(12):
            cs = fleet.contacts()
            for c in cs:
                O.stamp_tread(ob, op, init, RES, CELL, ORIGIN, c, DT)
            p.step_tires(DT, fleet.pushes(), 4, cs)
            fleet.advance(DT)
            led = p.ledger()
            for L in (led.water, led.slurry):
                assert abs(L.gap_m3) <= 1e-9 * L.initial_m3, (k, L.gap_m3, L.initial_m3)
            # The ledger's volume is kept by the flux apply (no recount of
            # the planes): it must equal the planes' own FP64 volume.
            if k % 4 == 3:
                h = p.download()
                for L, plane in ((led.water, h[0]), (led.slurry, h[3])):
                    vol = float(np.sum(np.maximum(plane, 0.0).astype(np.float64))) * float(np.float32(CELL)) ** 2
                    assert abs(L.volume_m3 - vol) <= 1e-11 * L.initial_m3, (k, L.volume_m3, vol)
        gb, gp, rut = p.ground()
        assert np.array_equal(gb, ob) and np.array_equal(gp, op)
        assert rut.max() > 0.005
        assert p.last_tire_ms() > 0.0
        h = p.download()
        assert h[0].min() >= 0.0 and h[3].min() >= 0.0
    finally:
        p.close()


def test_one_launch_per_stage_equals_one_vehicle_after_another():
    """Vehicles apart: each tire stage runs for every vehicle in one launch;
    one vehicle after another must give the same planes bit for bit."""
    bed, water, slurry = _scene(4)
    m = S.Mlp(12)
    params = m.params()
    m.close()
    pairs = [_pair(bed, water, slurry), _pair(bed, water, slurry)]
    fleets = [D.Fleet(seed=8, n=4, extent=RES * CELL, half=5.0) for _ in pairs]
    try:
        for p, on in zip(pairs, (True, False)):
            p.set_surrogate(params)
            p.set_tire_batching(on)
        for k in range(10):
            for p, f in zip(pairs, fleets):
                p.step_tires(DT, f.pushes(), 4, f.contacts())
                f.advance(DT)
            for x, y in zip(pairs[0].download(), pairs[1].download()):
                assert np.array_equal(x, y), k
        assert pairs[0].last_tire_ms() > 0.0
    finally:
        for p in pairs:
            p.close()
