"""Pins of the CPU oracle against the reference's own tests and closed forms.

The reference (Rust) cannot be built here, so the oracle is pinned by
(1) the reference's bit-exact equivalence tests re-run on the restatement,
(2) the reference's behavioural known-answer tests with their own thresholds,
(3) closed forms the mathematics fixes (a linear free surface, law end points).
Each test names the reference test it ports.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from xorshift import XorShift32

FINE_RES = 160


# --- (1) bit-exact equivalence tests ----------------------------------------

def _transport_case(rnd, res):
    """Inputs of transport_open_rows_match_the_serial_edges
    (src/fluid_core.rs:11274-11320), drawn in the reference's order."""
    n = res * res
    h = np.zeros(n, np.float32)
    vel = np.zeros(n, np.float32)
    gate = np.ones(n, np.uint8)
    bed = np.zeros(n, np.float32)
    f = np.float32
    for idx in range(n):
        r = rnd()
        if r < f(0.25):
            h[idx] = 0.0
        elif r < f(0.3):
            h[idx] = f(5.0e-7)
        elif r < f(0.33):
            h[idx] = f(5.0e-9)
        else:
            h[idx] = rnd() * f(0.3)
        q = rnd()
        if q < f(0.15):
            vel[idx] = 0.0
        elif q < f(0.2):
            vel[idx] = (rnd() - f(0.5)) * f(3.0e-5)
        else:
            vel[idx] = (rnd() - f(0.5)) * f(6.0)
        gate[idx] = 1 if rnd() > f(0.2) else 0
        bed[idx] = (rnd() - f(0.5)) * f(0.05)
    i0 = int(rnd() * f(3.0))
    j0 = int(rnd() * f(3.0))
    i1 = res - 1 - int(rnd() * f(3.0))
    j1 = res - 1 - int(rnd() * f(3.0))
    cell = f(0.02) + rnd() * f(0.2)
    dt = f(0.002) + rnd() * f(0.03)
    return h, vel, gate, bed, i0, i1, j0, j1, float(cell), float(dt)


@pytest.mark.parametrize("precision", [32, 64])
def test_transport_rows_match_the_serial_edges(precision):
    """transport_open_rows_match_the_serial_edges (src/fluid_core.rs:11274-11365):
    the branch-free row form equals the serial edge loop in depth and in the
    removed / added sums, on 96 seeded cases, with and without the bank test."""
    rnd = XorShift32(0x2545F491)
    dtype = np.float32 if precision == 32 else np.float64
    moved_cells = 0
    for case in range(96):
        res = [12, 37, 64, FINE_RES][case % 4]
        h, vel, gate, bed, i0, i1, j0, j1, cell, dt = _transport_case(rnd, res)
        h, vel, bed = h.astype(dtype), vel.astype(dtype), bed.astype(dtype)
        for along_x in (True, False):
            for with_bed in (False, True):
                b = bed if with_bed else None
                h_ref, h_new = h.copy(), h.copy()
                r_ref, a_ref = O.transport(h_ref, vel, gate, res, cell, dt, i0, i1, j0, j1, along_x,
                                           b, form="serial", precision=precision)
                r_new, a_new = O.transport(h_new, vel, gate, res, cell, dt, i0, i1, j0, j1, along_x,
                                           b, form="rows", precision=precision)
                assert np.array_equal(h_ref, h_new), f"depth case {case} x {along_x} bed {with_bed}"
                assert np.array_equal(r_ref, r_new), f"removed case {case}"
                assert np.array_equal(a_ref, a_new), f"added case {case}"
                moved_cells += int(np.count_nonzero(h_ref != h))
    assert moved_cells > 10000  # the cases really move water


def _bodies_case(rnd, res, case):
    """Inputs of run_bodies_match_the_flood_fill (src/fluid_core.rs:11232-11270)."""
    f = np.float32
    h = np.zeros(res * res, np.float32)
    blobs = 1 + int(rnd() * f(12.0))
    for _ in range(blobs):
        cx = rnd() * f(res)
        cz = rnd() * f(res)
        r = f(1.0) + rnd() * f(res) * f(0.2)
        for j in range(res):
            for i in range(res):
                dx = f(i) - cx
                dz = f(j) - cz
                if dx * dx + dz * dz < r * r and rnd() > f(0.15):
                    h[j * res + i] = f(0.01) + rnd() * f(0.1)
    if case % 7 == 0:
        for k in range(h.size):
            if rnd() < f(0.3):
                h[k] = f(0.02)
    return h


def test_run_bodies_match_the_flood_fill():
    """run_bodies_match_the_flood_fill (src/fluid_core.rs:11232-11270): the
    run-length bodies give the same boxes, in the same order, as a flood fill.
    Cases run at a quarter of the reference's 400 to keep the CPU suite short
    (the generator is pure Python); every grid size is still visited."""
    rnd = XorShift32(0x9E3779B9)
    split_cases = 0
    for case in range(100):
        res = [16, 40, 100, FINE_RES][case % 4]
        h = _bodies_case(rnd, res, case)
        want = O.work_boxes(h, res, 1.0e-4, True, form="flood")
        got = O.work_boxes(h, res, 1.0e-4, True, form="runs")
        assert want == got, f"case {case}"
        split_cases += len(got) > 1
    assert split_cases > 5  # the split path is exercised, not only the halo-3 box


def test_work_boxes_known_answers():
    """Two disks far apart split into two halo-6 boxes; a solid rectangle and a
    single body keep the halo-3 rectangle (src/fluid_core.rs:1345-1446)."""
    res = 64
    h = np.zeros(res * res, np.float32)
    h.reshape(res, res)[10:14, 10:14] = 0.1
    assert O.work_boxes(h, res, 1e-6) == [(7, 16, 7, 16)]
    h.reshape(res, res)[40:44, 45:50] = 0.1
    assert O.work_boxes(h, res, 1e-6) == [(4, 19, 4, 19), (39, 55, 34, 49)]
    # An L-shaped single body: not solid, but one body -> halo-3 bounding box.
    h[:] = 0
    h.reshape(res, res)[20:30, 20:22] = 0.1
    h.reshape(res, res)[28:30, 20:35] = 0.1
    assert O.work_boxes(h, res, 1e-6) == [(17, 37, 17, 32)]


# --- (2) behavioural known-answer tests ----------------------------------------

def test_slope_accelerates_downhill_closed_form():
    """slope_accelerates_downhill (src/fluid_core.rs:11367-11397).  The free
    surface is linear (eta = 0.05 - 0.04 j), so both symmetric 5-tap filters
    return it unchanged, the divergence correction and confinement vanish, and
    one step at (8, 8) gives v = g * (0.04 / cell) * mob(40) * dt *
    exp(-damp(40) dt) with mob(40) = 0.042 and damp(40) =
    1.7 * 0.42^0.3 * ((40/231)/0.42)^1.15 (the reference asserts v > 0.002)."""
    res, cell, dt = 16, 0.10, 0.05
    damp = 1.7 * 0.42 ** 0.3 * ((40.0 / 231.0) / 0.42) ** 1.15
    want = 9.81 * (0.04 / cell) * 0.042 * dt * math.exp(-damp * dt)
    for precision, tol in ((32, 2e-5), (64, 1e-6)):
        dtype = np.float32 if precision == 32 else np.float64
        n = res * res
        h = np.full(n, 0.05, dtype)
        u = np.zeros(n, dtype)
        v = np.zeros(n, dtype)
        bed = np.array([-0.04 * j for j in range(res) for _ in range(res)], np.float32).astype(dtype)
        nu = np.full(n, 40.0, dtype)
        gate = np.ones(n, np.uint8)
        O.step_sheet(h, u, v, bed, nu, gate, res, cell, dt, ("gravity", 0.0), precision=precision)
        got = v[8 * res + 8]
        assert got > 0.002
        assert abs(got - want) <= tol * want, (precision, got, want)
        assert abs(u[8 * res + 8]) < 1e-9  # no cross-slope push


def test_viscosity_orders_clay_silt_sand_and_thins():
    """viscosity_orders_clay_silt_sand_and_thins (src/fluid_core.rs:11399-11410)
    plus the end points of the laws (src/fluid_core.rs:150-192)."""
    clay = O.dynamic_viscosity("clay", 0.55)
    silt = O.dynamic_viscosity("silt", 0.55)
    sand = O.dynamic_viscosity("sand", 0.55)
    assert clay > silt * 1.8 and silt > sand * 1.5
    stiff, mid, thin = (O.dynamic_viscosity("clay", m) for m in (0.30, 0.55, 0.85))
    assert stiff > mid > thin
    # Dry end (moisture <= 0.22) is the stiff value, wet end (>= 0.92) the fluid one.
    for g, (s, fl) in {"clay": (420, 0.35), "silt": (220, 4.0), "sand": (28, 0.55),
                       "water": (0.9, 0.9), "other": (1.5, 0.08)}.items():
        assert O.dynamic_viscosity(g, 0.1, 64) == pytest.approx(s, rel=1e-6)
        assert O.dynamic_viscosity(g, 1.0, 64) == pytest.approx(fl, rel=1e-6)
    # Smoothstep midpoint: moisture 0.57 -> x = 0.5 -> thin = 0.5.
    assert O.dynamic_viscosity("clay", 0.57, 64) == pytest.approx(420 + (0.35 - 420) * 0.5, rel=1e-6)
    # Mobility: flat 0.042 at and above 12 Pa s, 0.042 * 2.6 at zero, raw value when negative.
    assert O.sheet_mobility(12.0) == pytest.approx(0.042, rel=1e-7)
    assert O.sheet_mobility(400.0) == pytest.approx(0.042, rel=1e-7)
    assert O.sheet_mobility(0.0) == pytest.approx(0.042 * 2.6, rel=1e-6)
    assert O.sheet_mobility(6.0, 64) == pytest.approx(0.042 * (1 + 0.25 * 1.6), rel=1e-6)
    assert O.sheet_mobility(-0.7) == pytest.approx(0.7, rel=1e-7)
    # Damping: both branches meet at rel = 0.42; power laws either side.
    knee = 0.42 * 231.0
    at = 1.7 * 0.42 ** 0.3
    assert O.sheet_damp(knee, 64) == pytest.approx(at, rel=1e-6)
    assert O.sheet_damp(knee * 0.999, 64) == pytest.approx(at * 0.999 ** 1.15, rel=1e-5)
    assert O.sheet_damp(231.0 * 2.0, 64) == pytest.approx(1.7 * 2.0 ** 0.3, rel=1e-6)
    assert O.sheet_damp(1e9, 64) == pytest.approx(1.7 * 12.0 ** 0.3, rel=1e-6)


def _sheet(res, dtype=np.float32):
    n = res * res
    return (np.zeros(n, dtype), np.zeros(n, dtype), np.zeros(n, dtype), np.zeros(n, dtype))


def test_sheet_conserves_and_runs_downhill():
    """sheet_conserves_and_runs_downhill (src/fluid_core.rs:11440-11502): 24^2 at
    0.2 m, clay at moisture 0.55, 60 steps of 0.05 s; volume within 1e-3 m^3 and
    the centroid moves more than 1.5 rows downhill."""
    res, cell = 24, 0.2
    h, u, v, bed = _sheet(res)
    nu = np.full(res * res, O.dynamic_viscosity("clay", 0.55), np.float32)
    gate = np.ones(res * res, np.uint8)
    H, B = h.reshape(res, res), bed.reshape(res, res)
    B[:] = (-0.16 * np.arange(res, dtype=np.float32))[:, None]
    H[4:9, 8:16] = 0.10
    before = O.column_sum(h, cell)

    def centroid(hh):
        m = float(hh.sum())
        return 0.0 if m <= 1e-6 else float((hh.reshape(res, res).sum(1) * np.arange(res)).sum()) / m

    c0 = centroid(h)
    for _ in range(60):
        O.step_sheet(h, u, v, bed, nu, gate, res, cell, 0.05, ("gravity", 0.02))
    after = O.column_sum(h, cell)
    assert abs(after - before) < 1e-3
    assert centroid(h) > c0 + 1.5


def _relief(h, bed, res, cell):
    c = res // 2
    x = (np.arange(res) + 0.5 - c) * cell
    X, Z = np.meshgrid(x, x)
    r = np.sqrt(X * X + Z * Z).ravel()
    eta = bed + h
    center = eta[r < 0.12]
    ring = eta[(r >= 0.48) & (r < 0.68)]
    return float(ring.mean() - center.mean())


def _rut_heal(grain, moisture):
    """rut_heal (src/fluid_core.rs:4673-4734) on the oracle."""
    res, cell, dt = 48, 0.06, 0.05
    h, u, v, bed = _sheet(res)
    nu = np.full(res * res, O.dynamic_viscosity(grain, moisture), np.float32)
    gate = np.ones(res * res, np.uint8)
    c = res // 2
    x = (np.arange(res, dtype=np.float32) + np.float32(0.5) - np.float32(c)) * np.float32(cell)
    X, Z = np.meshgrid(x, x)
    r = np.sqrt(X * X + Z * Z)
    bowl = np.exp(-(r * r) / np.float32(2.0 * 0.42 * 0.42)).astype(np.float32)
    bed[:] = (np.float32(-0.07) * bowl).ravel()
    h[:] = (np.float32(0.012) + np.float32(0.028) * (np.float32(1.0) - bowl)).ravel()
    relief0 = _relief(h, bed, res, cell)
    level_s, relief_8s, t = 20.0, relief0, 0.0
    while t < 20.0 - 1e-4:
        O.step_sheet(h, u, v, bed, nu, gate, res, cell, dt, ("gravity", 0.0))
        t += dt
        rr = _relief(h, bed, res, cell)
        if abs(t - 8.0) <= dt * 0.51:
            relief_8s = rr
        if level_s > 19.0 and rr < relief0 * 0.28:
            level_s = t
        if t >= 8.0 and level_s < 19.0:
            break
    return level_s, relief_8s > relief0 * 0.38


def test_rut_heal_times_match_the_mix():
    """rut_heal_times_match_the_mix (src/fluid_core.rs:11412-11438): clay and silt
    at moisture 0.55 level in 10-15 s and are still deep at 8 s; a thin clay mix
    (0.90) levels in 1-3 s."""
    clay_s, clay_deep = _rut_heal("clay", 0.55)
    silt_s, silt_deep = _rut_heal("silt", 0.55)
    thin_s, _ = _rut_heal("clay", 0.90)
    assert 10.0 <= clay_s <= 15.0, clay_s
    assert 10.0 <= silt_s <= 15.0, silt_s
    assert 1.0 <= thin_s <= 3.0, thin_s
    assert clay_deep and silt_deep


def test_fine_open():
    """fine_open (tests/e2e_niva.rs:956-1082): the 2 cm gravity sheet (one CFL
    pass, no projection, flat slow tiles asleep) keeps volume within 1e-3, stays
    finite and under 1.6 m/s, runs downhill (v > 0.02), and leaves a far
    sleeping tile bit-unchanged; the open wind plane moves its film downwind
    (mass-weighted u > 0.05) and keeps volume."""
    res, cell, dt = FINE_RES, 0.02, 1.0 / 30.0
    n = res * res
    h = np.full(n, 0.02, np.float32)
    u = np.full(n, 0.02, np.float32)
    v = np.zeros(n, np.float32)
    bed = np.zeros(n, np.float32)
    nu = np.ones(n, np.float32)
    gate = np.ones(n, np.uint8)
    H, B = h.reshape(res, res), bed.reshape(res, res)
    B[24:48, 24:48] = (-0.008 * np.arange(24, 48, dtype=np.float32))[:, None]
    H[24:48, 24:48] = 0.06
    sleep = (slice(152, 160), slice(152, 160))
    sh, su = H[sleep].copy(), u.reshape(res, res)[sleep].copy()
    before = float(np.maximum(h, 0).astype(np.float64).sum()) * cell * cell
    for _ in range(12):
        O.step_sheet(h, u, v, bed, nu, gate, res, cell, dt, ("gravity", 0.02))
    after = float(np.maximum(h, 0).astype(np.float64).sum()) * cell * cell
    assert before > 1e-4 and abs(after - before) / before < 1e-3
    assert np.isfinite(h).all() and np.isfinite(u).all() and np.isfinite(v).all()
    assert np.abs(u).max() <= 1.6 + 1e-3 and np.abs(v).max() <= 1.6 + 1e-3
    V = v.reshape(res, res)[24:80, 16:64]
    assert V[H[24:80, 16:64] > 1e-4].max() > 0.02
    assert np.array_equal(H[sleep], sh) and np.array_equal(u.reshape(res, res)[sleep], su)
    assert not v.reshape(res, res)[sleep].any()

    wres = 64
    wcell = 2.0 * (200 * 0.2 / 128)
    wn = wres * wres
    wh, wu, wv = np.zeros(wn, np.float32), np.zeros(wn, np.float32), np.zeros(wn, np.float32)
    wbed = np.zeros(wn, np.float32)
    wnu = np.full(wn, 0.4, np.float32)
    wgate = np.ones(wn, np.uint8)
    wh.reshape(wres, wres)[24:36, 12:24] = 0.04
    wb = float(wh.astype(np.float64).sum()) * wcell * wcell
    for _ in range(40):
        O.step_sheet(wh, wu, wv, wbed, wnu, wgate, wres, wcell, dt, ("wind", 1.2, -0.3, 0.35))
    wa = float(np.maximum(wh, 0).astype(np.float64).sum()) * wcell * wcell
    assert wb > 1e-4 and abs(wa - wb) / wb < 1e-3
    lim = max(math.hypot(1.2, -0.3), 1.6) + 1e-3
    assert np.abs(wu).max() <= lim and np.abs(wv).max() <= lim
    m = np.maximum(wh, 0).astype(np.float64)
    assert (m * wu).sum() / m.sum() > 0.05


# --- (3) invariants and FP32 vs FP64 -----------------------------------------

def test_transport_conserves_volume_exactly_in_fp64():
    """Each face moves dh out of one cell and into its neighbour, so in FP64
    the depth sum changes only through the max(0) clamp, which never fires
    because a face takes at most 0.45 h."""
    rnd = XorShift32(0x1234567)
    for case in range(8):
        res = [12, 37, 64, FINE_RES][case % 4]
        h, vel, gate, bed, i0, i1, j0, j1, cell, dt = _transport_case(rnd, res)
        h = h.astype(np.float64)
        vel = vel.astype(np.float64)
        before = h.sum()
        for along_x in (True, False):
            O.transport(h, vel, gate, res, cell, dt, i0, i1, j0, j1, along_x, None, precision=64)
        assert abs(h.sum() - before) <= 1e-12 * max(before, 1.0)
        assert (h >= 0).all()


def test_fp32_and_fp64_builds_of_the_oracle_agree_on_one_step():
    """One reference step of a 64^2 puddle on a slope: the FP32 and FP64 builds
    of the oracle (one source, two precisions) agree to 1e-5 of each field's
    scale, so the rounding of an FP32 step sits inside the tolerance the GPU's
    fast path is held to.  A consistency check of the two builds, not a pin
    of the arithmetic: the pins are the closed forms and known answers here
    and in test_oracle_pins.py."""
    res, cell, dt = 64, 0.2, 1.0 / 120.0
    rng = np.random.default_rng(5)
    n = res * res
    j, i = np.mgrid[0:res, 0:res]
    bed = (0.3 * np.sin(i / 7.0) * np.cos(j / 5.0) - 0.01 * j).astype(np.float32).ravel()
    h = (0.1 * np.exp(-((i - 32) ** 2 + (j - 30) ** 2) / (2 * 10.0 ** 2))).astype(np.float32).ravel()
    h[h < 1e-5] = 0
    u = (0.3 * rng.standard_normal(n) * (h > 0)).astype(np.float32)
    v = (0.3 * rng.standard_normal(n) * (h > 0)).astype(np.float32)
    nu = np.full(n, 0.9, np.float32)
    gate = np.ones(n, np.uint8)
    s32 = [x.copy() for x in (h, u, v)]
    s64 = [x.astype(np.float64) for x in (h, u, v)]
    O.step_sheet(*s32, bed, nu, gate, res, cell, dt, ("gravity", 0.08), precision=32)
    O.step_sheet(*s64, bed.astype(np.float64), nu.astype(np.float64), gate, res, cell, dt,
                 ("gravity", 0.08), precision=64)
    for a, b, name in zip(s32, s64, "huv"):
        scale = np.abs(b).max()
        err = np.abs(a.astype(np.float64) - b).max()
        assert err <= 1e-5 * scale, (name, err, scale)
