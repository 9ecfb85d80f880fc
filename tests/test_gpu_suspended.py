"""GPU parity of the wheel-lifted suspended solids (suspended.hip:
entrain_wake and step_suspended, src/fluid_core.rs:8750-9003) against the
FP32 restatement (oracle/core.c), step by step.

On a water-only pond the water planes follow the reference exactly (the core
parity tests), so the lift behind each tire, the fall-out, the carry along v
and the cap see the same inputs.  The fall-out factor exp(-dt w_s / h) is
exp in double rounded to f32 on the GPU and the host's expf in the oracle,
which can differ in the last place: the plane is compared to 2e-6 relative
(a few ulps after the differences travel), the FP64 accounts (f32 running
sums in the reference) to 1e-5, and the contract's turbidity channel, which
carries the suspended and stirred solids over the water, to 2e-6."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S

pytestmark = pytest.mark.gpu

N = 200 * 200
DT = 1.0 / 30.0


def _pond(c):
    bed = np.zeros(N, np.float32)
    water = np.zeros((200, 200), np.float32)
    water[60:140, 80:120] = 0.06
    water[60:140, 120:126] = 0.02
    c.map_water_bed(bed, water.ravel())


def _tires(k, speed=7.5, slip=0.2):
    z0 = -6.0 + 0.25 * k
    return [dict(x=x, z=z0 + dz, forward_x=0.0, forward_z=1.0, radius=0.38, width=0.28, speed=speed, slip=slip)
            for x in (-0.7, 0.7) for dz in (-1.2, 1.2)]


def _close(a, b, what, rtol=2e-6):
    scale = max(float(np.abs(b).max()), 1e-12)
    bad = np.flatnonzero(np.abs(a.astype(np.float64) - b) > rtol * scale)
    assert bad.size == 0, f"{what}: {bad.size} cells differ, first {bad[:5]} {a[bad[:5]]} vs {b[bad[:5]]}"


def _compare(g, o, what):
    gs, (gent, gset) = g.suspended()
    os_, acc = o.suspended()
    _close(gs, os_, f"suspended {what}")
    assert abs(gent - acc["entrained_m3"]) <= 1e-5 * max(acc["entrained_m3"], 1e-12), (what, gent, acc)
    assert abs(gset - acc["settled_m3"]) <= 1e-5 * max(acc["settled_m3"], 1e-9), (what, gset, acc)
    _, gflow = g.publish()[1:]
    _, oflow = o.publish()
    _close(gflow[3::4], oflow[3::4], f"turbidity {what}")
    return gs, acc


def _pair():
    g, o = S.FluidCore((0.0, 0.0), path="exact"), O.Core((0.0, 0.0))
    for c in (g, o):
        _pond(c)
    return g, o


def test_tires_through_a_pond_lift_carry_and_settle_like_the_oracle():
    g, o = _pair()
    try:
        for k in range(20):
            for t in _tires(k):
                g.push_tire(**t)
                o.push_tire(**t)
            g.step(DT)
            o.step(DT)
            assert np.array_equal(g.planes()["water"], o.planes()["water"]), k
            susp, acc = _compare(g, o, f"step {k}")
        assert susp.max() > 1e-3 and acc["settled_m3"] > 0.0
        for k in range(15):
            g.step(DT)
            o.step(DT)
            _compare(g, o, f"coast {k}")
        # the account closes: column + settled + parked = entrained
        L = g.ledger_full()
        gap = L["suspended_m3"] + L["settled_suspended_m3"] + L["frozen_suspended_m3"] - L["entrained_m3"]
        assert L["entrained_m3"] > 1e-4 and abs(gap) <= 1e-9 * L["entrained_m3"]
        assert L["total_m3"] == pytest.approx(g.ledger()["total_m3"] + L["suspended_m3"]
                                              + L["settled_suspended_m3"] + L["frozen_suspended_m3"], rel=1e-12)
        cc = g.coarse_cell(100, 100)
        assert 0.0 <= cc["turbidity"] <= 1.0
    finally:
        g.close()


def test_a_set_plume_carried_by_v_matches_the_oracle():
    rng = np.random.default_rng(3)
    g, o = _pair()
    try:
        p = g.planes()
        p["vw"][:] = rng.uniform(-2.5, 2.5, N).astype(np.float32)
        g.set_planes(p)
        o.set_planes(p)
        s = np.where(rng.random(N) < 0.3, rng.uniform(0.0, 0.04, N), 0.0).astype(np.float32)
        g.set_suspended(s)
        o.set_suspended(s)
        for k in range(3):
            g.step(DT)
            o.step(DT)
            _compare(g, o, f"plume step {k}")
    finally:
        g.close()


def test_suspended_solids_move_with_the_window():
    g, o = _pair()
    try:
        for k in range(8):
            for t in _tires(k):
                g.push_tire(**t)
                o.push_tire(**t)
            g.step(DT)
            o.step(DT)
        for c in (g, o):
            c.recenter(0.0, 25.0)
        _compare(g, o, "after recenter")
        L = g.ledger_full()
        _, acc = o.suspended()
        assert L["frozen_suspended_m3"] > 0.0
        assert L["frozen_suspended_m3"] == pytest.approx(acc["frozen_m3"], rel=1e-5)
        for c in (g, o):
            c.recenter(0.0, 0.0)
        _compare(g, o, "back")
    finally:
        g.close()
