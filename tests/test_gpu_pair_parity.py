"""GPU parity of the two-layer sheet handle (silt_pair_*, include/silt_pair.h)
against the oracle, layer by layer.

The handle steps water and slurry as the reference's generic-res step_sheet
(src/fluid_core.rs:770-785) with FluidCore::step_pair's layer constants
(:8036-8125): water Gravity{0.08} at 0.9 Pa s, slurry Gravity{0.05} at its own
viscosity.  On the fast path both layers of a step share one launch of the
fused kernel; these tests hold each layer of that launch to the FP64 oracle
with that layer's constants, one step at a time from the GPU's own FP32 state:
|gpu - oracle64| <= 1e-5 * max|oracle64| per field.

* the 4096^2 workload that bench.py times (every cell wet: one box over the
  grid, the FULL instance of the kernel);
* 1000^2, which divides neither the 235-column strip nor the row chunk, with
  both layers wet everywhere and with different box lists per layer;
* a 2 cm pair, whose awake-tile boxes may touch or overlap (those layers take
  the stage kernels; fused_exact must then equal the FP32 oracle bit for bit);
* a layer that dries out while a speculative launch runs (undo path);
* 16 vehicles (64 contacts) on the 4096^2 sheet: ruts equal to the oracle's
  stamp, depths finite and non-negative, the FP64 ledger closed.
"""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from silt_amd import drive as D
from silt_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 1.0e-5
DT = 1.0 / 120.0
LAYERS = (("water", 0, 0.9, 0.08), ("slurry", 3, 0.5, 0.05))  # name, plane offset, nu, vorticity


@pytest.fixture(scope="module")
def large():
    return scenes.large_scene(res=4096, seed=7, puddles=64)


def _pair(sc, path="fast", nu_slurry=None):
    n = sc["res"] ** 2
    z = np.zeros(n, np.float32)
    p = S.SheetPair(sc["res"], sc["cell"], path=path)
    p.upload(sc["bed"], sc["water"], z, z, sc["slurry"], z, z, sc["nu_slurry"] if nu_slurry is None else nu_slurry)
    return p


def _boxes_fit(boxes):
    """The fused kernel's box rule: 1..8 boxes, pairwise neither touching nor
    overlapping (boxes are (i0, i1, j0, j1), inclusive)."""
    if not 1 <= len(boxes) <= 8:
        return False
    for a in range(len(boxes)):
        for b in range(a + 1, len(boxes)):
            p, q = boxes[a], boxes[b]
            if not (p[1] + 1 < q[0] or q[1] + 1 < p[0] or p[3] + 1 < q[2] or q[3] + 1 < p[2]):
                return False
    return True


def _oracle64(state, bed, res, cell, nu, vort):
    o = [a.astype(np.float64) for a in state]
    n = res * res
    nsub, boxes = O.step_sheet(*o, bed.astype(np.float64), np.full(n, nu, np.float64), np.ones(n, np.uint8), res,
                               cell, DT, ("gravity", vort), precision=64, want_boxes=True)
    return o, nsub, boxes


def _check_layer(gpu, ref, what, near_dry=None):
    """gpu (FP32) against the FP64 oracle.  near_dry masks the velocities of
    cells the dry cut (h < 1e-6 -> h = u = v = 0, the step's last operation)
    zeroed on one side only: a depth within rounding of 1e-6 m.  Their depths
    are still checked, and each step starts from the same FP32 state, so such
    a flip does not spread."""
    for a, b, name in zip(gpu, ref, "huv"):
        scale = float(np.abs(b).max())
        d = np.abs(a.astype(np.float64) - b)
        if near_dry is not None and name != "h":
            d = np.where(near_dry, 0.0, d)
        err = float(d.max())
        assert err <= TOL * max(scale, 1e-30), f"{what} {name}: max err {err:.3e} > {TOL} x {scale:.3e}"


def _step_and_check(p, sc, steps, mask_near_dry=False, expect_full=None):
    res, cell = sc["res"], sc["cell"]
    for k in range(steps):
        before = p.download()
        nsub = p.step(DT)
        after = p.download()
        assert nsub == 1, nsub
        for name, off, nu, vort in LAYERS:
            ref, onsub, boxes = _oracle64(before[off:off + 3], sc["bed"], res, cell, nu, vort)
            assert onsub == 1
            if expect_full is not None:
                assert (boxes == [(0, res - 1, 0, res - 1)]) == expect_full, boxes
            near = None
            if mask_near_dry:
                near = (after[off] == 0.0) != (ref[0] == 0.0)
                assert near.sum() <= max(16, 1e-3 * float((ref[0] > 1e-6).sum())), (name, int(near.sum()))
            _check_layer(after[off:off + 3], ref, f"{name} step {k}", near)
    return p


def test_benched_launch_at_4096_matches_fp64_per_layer(large):
    """The workload bench.py times: both layers of the 4096^2 sheet in one
    FULL launch of the fused kernel, each layer against the FP64 oracle."""
    p = _pair(large)
    try:
        _step_and_check(p, large, 2, expect_full=True)
        led = p.ledger()
        for L in (led.water, led.slurry):
            assert abs(L.gap_m3) <= 1e-9 * L.initial_m3, (L.gap_m3, L.initial_m3)
            assert abs(L.cut_m3) <= 1e-9 * L.initial_m3 and abs(L.residue_m3) <= 1e-6 * L.initial_m3
    finally:
        p.close()


def test_xcd_weighted_rows_step_as_uniform_chunks(large):
    """Tiles whose rows are scaled per XCD (silt_pair_set_xcd_weights) give
    the same planes bit for bit as uniform chunks: a cell's arithmetic does
    not depend on which tile computes it.  Ledger lines agree to FP64
    summation (their per-tile partial sums move with the split)."""
    w = [0.6, 1.4, 0.8, 1.2, 1.0, 0.7, 1.3, 1.0]
    a, b = _pair(large), _pair(large)
    try:
        b.set_xcd_weights(w)
        for _ in range(2):
            assert a.step(DT) == b.step(DT) == 1
        for x, y in zip(a.download(), b.download()):
            assert np.array_equal(x, y)
        la, lb = a.ledger(), b.ledger()
        for p_, q_ in ((la.water, lb.water), (la.slurry, lb.slurry)):
            assert abs(p_.volume_m3 - q_.volume_m3) <= 1e-12 * p_.initial_m3
            assert abs(p_.residue_m3 - q_.residue_m3) <= 1e-9 * p_.initial_m3
    finally:
        a.close()
        b.close()


def test_xcd_balance_restores_state_and_keeps_results():
    """silt_pair_balance runs calibration steps, restores the caller's planes
    and ledger, and leaves weights under which the steps equal uniform
    chunks bit for bit."""
    sc = scenes.large_scene(res=1000, seed=3, puddles=12)
    a, b = _pair(sc), _pair(sc)
    try:
        before = a.download()
        led0 = a.ledger()
        w, (t_uni, t_bal), us0, us1 = a.balance(DT, reps=2)
        for x, y in zip(before, a.download()):
            assert np.array_equal(x, y)
        led1 = a.ledger()
        assert led1.water.volume_m3 == led0.water.volume_m3 and led1.water.residue_m3 == led0.water.residue_m3
        assert t_uni > 0.0 and t_bal > 0.0
        assert all(u > 0.0 for u in us0) and all(u > 0.0 for u in us1)
        assert all(0.25 <= x <= 4.0 for x in w)
        assert a.speculation_counts() == (0, 0)
        for _ in range(2):
            assert a.step(DT) == b.step(DT) == 1
        for x, y in zip(a.download(), b.download()):
            assert np.array_equal(x, y)
    finally:
        a.close()
        b.close()


def test_two_layer_launch_at_1000_matches_fp64_per_layer():
    """A side that divides neither the strip width (235) nor the row chunk:
    partial strips and chunks on both layers of one launch."""
    sc = scenes.large_scene(res=1000, seed=3, puddles=12)
    p = _pair(sc)
    try:
        _step_and_check(p, sc, 3, expect_full=True)
    finally:
        p.close()


def test_layers_with_different_box_lists_match_fp64_per_layer():
    """Water puddles on a dry bed (several work boxes) and a slurry band (one
    box): the two descriptors of one launch carry different box lists."""
    res, cell = 1000, 0.2
    sc = scenes.large_scene(res=res, seed=9, puddles=5)
    origin = sc["origin"]
    water = np.zeros((res, res), np.float32)
    for cx, cz in ((-60.0, -50.0), (-55.0, 60.0), (40.0, -30.0), (70.0, 75.0)):
        water += scenes.gaussian(res, cell, origin, cx, cz, 0.10, 4.0)
    water[water < 1.0e-6] = 0.0
    slurry = np.zeros((res, res), np.float32)
    slurry[300:700, 100:450] = 0.02
    sc = dict(sc, water=water.ravel(), slurry=slurry.ravel())
    wb = O.work_boxes(sc["water"], res, 1e-6)
    assert 2 <= len(wb) <= 8 and _boxes_fit(wb), wb
    p = _pair(sc)
    try:
        _step_and_check(p, sc, 3, mask_near_dry=True, expect_full=False)
    finally:
        p.close()


def test_two_centimetre_pair_equals_the_fp32_oracle():
    """A 2 cm pair plans with awake tiles (src/fluid_core.rs:1537-1657), whose
    boxes may touch or number more than 8; such layers take the stage kernels.
    With fused_exact both paths keep the reference's arithmetic, so every step
    equals the FP32 oracle bit for bit."""
    res, cell = 160, 0.02
    rng = np.random.default_rng(21)
    X, Z = np.meshgrid(np.arange(res), np.arange(res))
    water = np.zeros((res, res), np.float32)
    slurry = np.zeros((res, res), np.float32)
    for plane, count, depth in ((water, 14, 0.03), (slurry, 9, 0.015)):
        for _ in range(count):
            ci, cj = rng.integers(8, res - 8, 2)
            r2 = (X - ci) ** 2 + (Z - cj) ** 2
            plane += np.where(r2 <= rng.integers(4, 30), np.float32(depth), np.float32(0.0))
    bed = (0.004 * np.sin(X / 9.0) * np.cos(Z / 7.0)).astype(np.float32).ravel()
    sc = {"res": res, "cell": cell, "bed": bed, "water": water.ravel(), "slurry": slurry.ravel(),
          "nu_slurry": 0.5}
    p = _pair(sc, path="fused_exact")
    n = res * res
    ref = [sc["water"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32),
           sc["slurry"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)]
    unfit = 0
    try:
        for k in range(6):
            p.step(1.0 / 60.0)
            for name, off, nu, vort in LAYERS:
                _, boxes = O.step_sheet(*ref[off:off + 3], bed, np.full(n, nu, np.float32), np.ones(n, np.uint8),
                                        res, cell, 1.0 / 60.0, ("gravity", vort), want_boxes=True)
                unfit += 0 if _boxes_fit(boxes) else 1
            got = p.download()
            for a, b, name in zip(got, ref, ("water", "uw", "vw", "slurry", "us", "vs")):
                bad = np.flatnonzero(a != b)
                assert bad.size == 0, f"step {k} {name}: {bad.size} cells differ, first {bad[:4]}"
        assert unfit > 0, "no step planned boxes the fused kernel does not cover"
        # The stage kernels book their dry cut and transport residue too.
        led = p.ledger()
        for L in (led.water, led.slurry):
            assert abs(L.gap_m3) <= 1e-9 * L.initial_m3, (L.gap_m3, L.initial_m3)
    finally:
        p.close()


def test_a_layer_that_dries_under_a_speculative_launch():
    """The slurry layer dries in its first step, so the next step's
    speculative launch of the last plan is undone: planes, statistics and
    ledger must be those of a handle that never speculated."""
    res, cell = 256, 0.2
    sc = scenes.large_scene(res=res, seed=4, puddles=3)
    n = res * res
    slurry = np.zeros((res, res), np.float32)
    us = np.zeros((res, res), np.float32)
    # Wet cells on even columns moving right at the cap: each gives about 13 %
    # through its right face and falls under the 1e-6 m dry cut, and so do
    # the cells it feeds.
    slurry[80:160, 80:160:2] = 1.1e-6
    us[80:160, 80:160:2] = 1.6
    z = np.zeros(n, np.float32)
    pairs = []
    for spec in (True, False):
        q = S.SheetPair(res, cell, path="fast")
        q.upload(sc["bed"], sc["water"], z, z, slurry.ravel(), us.ravel(), z, 0.5)
        q.set_speculation(spec)
        pairs.append(q)
    try:
        dt = 1.0 / 30.0
        for k in range(5):
            for q in pairs:
                q.step(dt)
            a, b = pairs[0].download(), pairs[1].download()
            for x, y, name in zip(a, b, ("water", "uw", "vw", "slurry", "us", "vs")):
                assert np.array_equal(x, y), (k, name)
            la, lb = pairs[0].ledger(), pairs[1].ledger()
            for L, M in ((la.water, lb.water), (la.slurry, lb.slurry)):
                assert L.volume_m3 == M.volume_m3 and L.gap_m3 == M.gap_m3, k
            area = float(np.float32(cell)) ** 2  # the handle's cell is an f32
            assert la.slurry.volume_m3 == pytest.approx(float(np.maximum(a[3], 0).astype(np.float64).sum()) * area,
                                                        rel=1e-12, abs=1e-18)
        assert float(pairs[0].download()[3].max()) == 0.0
        held, undone = pairs[0].speculation_counts()
        assert undone >= 1, (held, undone)
    finally:
        for q in pairs:
            q.close()


def test_sixteen_vehicles_on_the_4096_sheet(large):
    """The large workload's step (ruts on 64 contacts, the network per vehicle,
    the fused two-layer step, the network's fluxes) for a few steps: ruts equal
    to the oracle's stamp in contact order, depths finite and non-negative,
    the ledger closed per step with the network's fluxes in it."""
    res, cell = large["res"], large["cell"]
    origin = (-0.5 * res * cell, -0.5 * res * cell)
    p = _pair(large)
    m = S.Mlp(13)
    try:
        p.set_ground(large["bed"], origin)
        p.set_surrogate(m.params())
        fleet = D.Fleet(seed=7, n=16, extent=res * cell, half=25.0)
        ob, op, init = large["bed"].copy(), np.zeros_like(large["bed"]), large["bed"].copy()
        for k in range(4):
            cs = fleet.contacts()
            assert len(cs) == 64
            for c in cs:
                O.stamp_tread(ob, op, init, res, cell, origin, c, DT)
            p.step_tires(DT, fleet.pushes(), 4, cs)
            fleet.advance(DT)
            led = p.ledger()
            for L in (led.water, led.slurry):
                assert abs(L.gap_m3) <= 1e-9 * L.initial_m3, (k, L.gap_m3, L.initial_m3)
                assert 0.0 <= L.cut_m3 <= 1e-6 * L.initial_m3 and abs(L.residue_m3) <= 1e-6 * L.initial_m3
        planes = p.download()
        for a, name in zip(planes, ("water", "uw", "vw", "slurry", "us", "vs")):
            assert np.isfinite(a).all(), name
        assert planes[0].min() >= 0.0 and planes[3].min() >= 0.0
        area = float(np.float32(cell)) ** 2  # the handle's cell is an f32
        assert led.water.volume_m3 == pytest.approx(float(planes[0].astype(np.float64).sum()) * area, rel=1e-12)
        gb, gp, rut = p.ground()
        assert np.array_equal(gb, ob) and np.array_equal(gp, op)
        assert rut.max() > 0.005
    finally:
        m.close()
        p.close()
