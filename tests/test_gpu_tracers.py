"""GPU parity of the foam and stirred-silt tracers (tracers.hip, step_tracers
and note_surface_foam, src/fluid_core.rs:6210-6750) against the FP32
restatement (oracle/core.c), step by step.

On a water-only window the sheet's water planes follow the reference exactly
(the core parity tests), so every tracer law sees the same inputs and foam and
silt must be equal value for value: four tires rolling and slipping through a
pond with a shore, the tracers carried and decaying after the tires leave,
the crest foam of note_surface_foam, and the tracers moving with the window.
The stirred and settled silt accounts, which the reference sums in f32 in list
order, are FP64 sums here, compared to 1e-5."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S

pytestmark = pytest.mark.gpu

N = 200 * 200
DT = 1.0 / 30.0


def _eq(a, b, what):
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, f"{what}: {bad.size} cells differ, first {bad[:5]} {a[bad[:5]]} vs {b[bad[:5]]}"


def _pond(c):
    """A 12 x 8 m pond 6 cm deep on a flat bed, with a shallow shelf."""
    bed = np.zeros(N, np.float32)
    water = np.zeros((200, 200), np.float32)
    water[70:130, 80:120] = 0.06
    water[70:130, 120:126] = 0.02  # a shelf where the shore foams and lifts silt
    c.map_water_bed(bed, water.ravel())


def _tires(k, speed=3.0, slip=0.3):
    z0 = -4.0 + 0.1 * k
    return [dict(x=x, z=z0 + dz, forward_x=0.0, forward_z=1.0, radius=0.38, width=0.28, speed=speed, slip=slip)
            for x in (-0.7, 0.7) for dz in (-1.2, 1.2)]


def _compare(g, o, what):
    gf, gt, (gs, gset) = g.tracers()
    of, ot, os_, oset = o.tracers()
    _eq(gf, of, f"foam {what}")
    _eq(gt, ot, f"silt {what}")
    assert abs(gs - os_) <= 1e-5 * max(abs(os_), 1e-12), (what, gs, os_)
    assert abs(gset - oset) <= 1e-5 * max(abs(oset), 1e-12), (what, gset, oset)
    return gf, gt


def _pair():
    g, o = S.FluidCore((0.0, 0.0), path="exact"), O.Core((0.0, 0.0))
    for c in (g, o):
        _pond(c)
    return g, o


def test_tires_through_a_pond_make_the_same_foam_and_silt():
    g, o = _pair()
    try:
        for k in range(14):
            for t in _tires(k):
                g.push_tire(**t)
                o.push_tire(**t)
            g.step(DT)
            o.step(DT)
            _eq(g.planes()["water"], o.planes()["water"], f"water step {k}")
            foam, silt = _compare(g, o, f"step {k}")
        assert foam.max() > 0.1 and silt.max() > 0.0
        # The tires leave: the foam is carried and decays, the silt settles.
        for k in range(12):
            g.step(DT)
            o.step(DT)
            _compare(g, o, f"coast {k}")
        # The crest foam of a seated sheet, from the last step's pushes.
        for t in _tires(20):
            g.push_tire(**t)
            o.push_tire(**t)
        g.step(DT)
        o.step(DT)
        p = g.planes()
        p["water"].reshape(200, 200)[100, 100] += 0.012
        g.set_planes(p)
        o.set_planes(p)
        g.note_surface_foam(DT)
        o.note_surface_foam(DT)
        _compare(g, o, "note_surface_foam")
    finally:
        g.close()


def test_tracers_move_with_the_window():
    g, o = _pair()
    try:
        for k in range(6):
            for t in _tires(k):
                g.push_tire(**t)
                o.push_tire(**t)
            g.step(DT)
            o.step(DT)
        for c in (g, o):
            c.recenter(1.3, -0.9)
        _compare(g, o, "after recenter")
        for k in range(4):
            g.step(DT)
            o.step(DT)
            _compare(g, o, f"after recenter step {k}")
    finally:
        g.close()
