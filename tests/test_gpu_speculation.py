"""Speculative launch of the last plan (silt_pair_step): a handle that
launches the previous step's plan before confirming it, and undoes it when
the statistics give another plan, must step exactly as one that always
plans first.  Puddles spreading over a dry bed change their work boxes
(undo and redo); a settled run keeps them (speculations hold); tires change
the substep count."""
import numpy as np
import pytest

import silt_amd as S
from silt_amd import drive as D
from silt_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0


def _puddles(res=256, cell=0.2, seed=4):
    origin = (-res * cell / 2.0, -res * cell / 2.0)
    bed = scenes.fbm_terrain(res, cell, origin, seed).ravel().astype(np.float32)
    water = np.zeros(res * res, np.float32)
    for cx, cz in ((-12.0, -8.0), (10.0, 9.0), (4.0, -14.0)):
        water += scenes.gaussian(res, cell, origin, np.float32(cx), np.float32(cz), 0.08, 2.0).ravel()
    water[water < 0.002] = 0.0
    slurry = np.zeros(res * res, np.float32)
    slurry.reshape(res, res)[100:140, 60:200] = 0.01
    return res, cell, bed, water, slurry


def _pair(res, cell, bed, water, slurry, spec):
    p = S.SheetPair(res, cell, path="fast")
    z = np.zeros(res * res, np.float32)
    p.upload(bed, water, z, z, slurry, z, z, 0.5)
    p.set_speculation(spec)
    return p


def test_speculation_steps_exactly_as_planning_first():
    res, cell, bed, water, slurry = _puddles()
    a, b = _pair(res, cell, bed, water, slurry, True), _pair(res, cell, bed, water, slurry, False)
    try:
        for k in range(80):
            na, nb = a.step(DT), b.step(DT)
            assert na == nb, k
            if k % 8 == 7:
                for x, y in zip(a.download(), b.download()):
                    assert np.array_equal(x, y), k
        la, lb = a.ledger(), b.ledger()
        assert la.water.gap_m3 == lb.water.gap_m3 and la.slurry.gap_m3 == lb.slurry.gap_m3
        hits, misses = a.speculation_counts()
        assert hits > 0 and misses > 0, (hits, misses)
        assert b.speculation_counts() == (0, 0)
    finally:
        a.close()
        b.close()


def test_speculation_with_tires_and_changing_substeps():
    res, cell, bed, water, slurry = _puddles(seed=5)
    water = np.maximum(water, np.float32(0.02))
    origin = (-res * cell / 2.0, -res * cell / 2.0)
    pairs = [_pair(res, cell, bed, water, slurry, on) for on in (True, False)]
    fleets = [D.Fleet(seed=3, n=2, extent=res * cell, half=6.0, speed=14.0) for _ in pairs]
    try:
        for p in pairs:
            p.set_ground(bed, origin)
        nsubs = set()
        for k in range(40):
            ns = [p.step_tires(DT, f.pushes(), 4, f.contacts()) for p, f in zip(pairs, fleets)]
            for f in fleets:
                f.advance(DT)
            assert ns[0] == ns[1], k
            nsubs.add(ns[0])
            for x, y in zip(pairs[0].download(), pairs[1].download()):
                assert np.array_equal(x, y), k
        assert pairs[0].ledger().water.gap_m3 == pairs[1].ledger().water.gap_m3
        # Every undone speculation pauses speculation for at least the next
        # step (the first step has no plan to speculate with).
        hits, misses = pairs[0].speculation_counts()
        assert hits + 2 * misses <= 39, (hits, misses)
    finally:
        for p in pairs:
            p.close()
