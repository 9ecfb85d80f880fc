"""GPU parity of the exact sheet path against the FP32 oracle.

The exact path follows the reference operation order with no fused
multiply-add, so on the same inputs it must equal the FP32 restatement value
for value (np.array_equal; -0.0 equals 0.0).  Every case steps the same
seeded state through both and compares h, u, v after every step.
"""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from hipmem import DeviceArray, synchronize
from silt_amd import scenes

pytestmark = pytest.mark.gpu


def _both(state, bed, nu, gate, res, cell, dt, drive, steps, opts=None, oracle_kw=None, check_each=True):
    opts = opts or {}
    oracle_kw = oracle_kw or {}
    g = [a.copy() for a in state]
    o = [a.copy() for a in state]
    for step in range(steps):
        ng = S.step_sheet(*g, bed, nu, gate, res, cell, dt, drive, **opts)
        no = O.step_sheet(*o, bed, nu, gate, res, cell, dt, drive, **oracle_kw)
        assert ng == no, f"substeps differ at step {step}: gpu {ng} oracle {no}"
        if check_each or step == steps - 1:
            for a, b, name in zip(g, o, "huv"):
                bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
                assert bad.size == 0, (f"{name} differs at step {step} in {bad.size} cells, first {bad[:5]} "
                                       f"gpu {a[bad[:5]]} oracle {b[bad[:5]]}")
    return g


def test_slope_step_matches_oracle():
    res, cell = 16, 0.10
    n = res * res
    h = np.full(n, 0.05, np.float32)
    bed = np.repeat(-0.04 * np.arange(res, dtype=np.float32), res)
    nu = np.full(n, 40.0, np.float32)
    gate = np.ones(n, np.uint8)
    _both([h, np.zeros(n, np.float32), np.zeros(n, np.float32)], bed, nu, gate, res, cell, 0.05,
          ("gravity", 0.0), 3)


def test_conserving_slump_matches_oracle_every_step():
    """sheet_conserves_and_runs_downhill's scene (src/fluid_core.rs:11440-11502),
    60 steps, confinement on (small box: the Gauss-Seidel sweep)."""
    res, cell = 24, 0.2
    n = res * res
    h = np.zeros((res, res), np.float32)
    h[4:9, 8:16] = 0.10
    bed = np.repeat(-0.16 * np.arange(res, dtype=np.float32), res)
    nu = np.full(n, O.dynamic_viscosity("clay", 0.55), np.float32)
    gate = np.ones(n, np.uint8)
    _both([h.ravel(), np.zeros(n, np.float32), np.zeros(n, np.float32)], bed, nu, gate, res, cell, 0.05,
          ("gravity", 0.02), 60)


@pytest.mark.parametrize("row_split", [True, False])
def test_window_scene_matches_oracle(row_split):
    """The 200^2 window scene (fBm bed, Gaussian puddle) at dt 1/120 for 30 steps:
    row_split True is a plain step_sheet call (snapshot Jacobi on the big box),
    False the layer inside FluidCore::step_pair (in-place Gauss-Seidel)."""
    sc = scenes.window_scene(seed=1)
    res, cell = sc["res"], sc["cell"]
    n = res * res
    nu = np.full(n, 0.9, np.float32)
    gate = np.ones(n, np.uint8)
    _both([sc["water"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)], sc["bed"], nu, gate, res, cell,
          1.0 / 120.0, ("gravity", 0.08), 30, opts={"row_split": row_split},
          oracle_kw={"inner_par": row_split})


def test_slurry_band_with_cell_viscosity_matches_oracle():
    """Per-cell viscosity: the damping factor uses the device exp/pow, so the
    comparison is to 2 ulp-scale tolerance instead of equality."""
    sc = scenes.window_scene(seed=2)
    res, cell = sc["res"], sc["cell"]
    n = res * res
    rng = np.random.default_rng(3)
    nu = (0.5 + 40.0 * rng.random(n)).astype(np.float32)
    gate = np.ones(n, np.uint8)
    g = [sc["slurry"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)]
    o = [a.copy() for a in g]
    for _ in range(10):
        S.step_sheet(*g, sc["bed"], nu, gate, res, cell, 1.0 / 120.0, ("gravity", 0.05))
        O.step_sheet(*o, sc["bed"], nu, gate, res, cell, 1.0 / 120.0, ("gravity", 0.05))
    for a, b in zip(g, o):
        scale = max(float(np.abs(b).max()), 1e-12)
        assert float(np.abs(a - b).max()) <= 1e-5 * scale


def test_split_boxes_and_gate_match_oracle():
    """Two separate puddles (two work boxes, stepped in order) and a gate plane
    with closed cells."""
    res, cell = 96, 0.2
    n = res * res
    H = np.zeros((res, res), np.float32)
    H[10:20, 10:22] = 0.08
    H[60:75, 50:70] = 0.05
    rng = np.random.default_rng(11)
    bed = (0.05 * rng.standard_normal((res, res))).astype(np.float32).ravel()
    gate = (rng.random(n) > 0.1).astype(np.uint8)
    nu = np.full(n, 0.9, np.float32)
    assert len(O.work_boxes(H.ravel(), res, 1e-6)) == 2
    _both([H.ravel(), np.zeros(n, np.float32), np.zeros(n, np.float32)], bed, nu, gate, res, cell, 1.0 / 60.0,
          ("gravity", 0.08), 20)


def test_fine_sheet_paces_match_oracle():
    """The 2 cm sheet: pace 0 (awake tiles, one pass), pace 2 (adaptive CFL on
    the wet rectangle) and pace 3 (one pressure pass + transport slices)."""
    res, cell = 160, 0.02
    n = res * res
    H = np.full((res, res), 0.02, np.float32)
    B = np.zeros((res, res), np.float32)
    B[24:48, 24:48] = (-0.008 * np.arange(24, 48, dtype=np.float32))[:, None]
    H[24:48, 24:48] = 0.06
    u = np.full(n, 0.02, np.float32)
    u.reshape(res, res)[60:80, 60:80] = 1.2  # fast enough for several CFL substeps
    nu = np.ones(n, np.float32)
    gate = np.ones(n, np.uint8)
    for pace in (0, 2, 3):
        _both([H.ravel().copy(), u.copy(), np.zeros(n, np.float32)], B.ravel(), nu, gate, res, cell, 1.0 / 30.0,
              ("gravity", 0.02), 4, opts={"pace": pace}, oracle_kw={"pace": pace})


def test_wind_plane_matches_oracle():
    wres = 64
    wcell = 2.0 * (200 * 0.2 / 128)
    wn = wres * wres
    wh = np.zeros((wres, wres), np.float32)
    wh[24:36, 12:24] = 0.04
    _both([wh.ravel(), np.zeros(wn, np.float32), np.zeros(wn, np.float32)], np.zeros(wn, np.float32),
          np.full(wn, 0.4, np.float32), np.ones(wn, np.uint8), wres, wcell, 1.0 / 30.0, ("wind", 1.2, -0.3, 0.35), 40)


def test_device_pointer_entry_matches_host_entry():
    sc = scenes.window_scene(seed=4)
    res, cell = sc["res"], sc["cell"]
    n = res * res
    nu = np.full(n, 0.9, np.float32)
    gate = np.ones(n, np.uint8)
    host = [sc["water"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)]
    dev = [DeviceArray(a) for a in host]
    bed_d = DeviceArray(sc["bed"])
    for _ in range(5):
        S.step_sheet(*host, sc["bed"], nu, gate, res, cell, 1.0 / 120.0, ("gravity", 0.08))
        S.step_sheet_device(*(d.ptr for d in dev), bed_d.ptr, 0, 0, res, cell, 1.0 / 120.0,
                            ("gravity", 0.08), nu_uniform=0.9)
    synchronize()
    for a, d in zip(host, dev):
        assert np.array_equal(a, d.numpy())
